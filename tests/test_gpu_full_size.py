"""GPU parity at the benchmark's sizes.

* Config 1 (256^2 dam break, MC limiter, walls, CFL 0.9, t = 0.5) run to the
  end on both FP64 engines: the identical step log (accepted/rejected, dt, CFL,
  time) and the identical final state as the oracle.
* 2048^2, 20 accepted steps of the same physics: FP64 fused engine bit for
  bit against the oracle.
* 8192^2 (config 2's size), where the oracle would take minutes: state round
  trips through the C ABI, and size-independent properties of an FP32 run of
  40 accepted steps (mass conservation in FP64 totals, mirror symmetry of the
  circular dam break, positivity, agreement of the device-controlled step
  sequence with the host-side controller's arithmetic).
* 16384^2, the largest size tested (byte offsets far past 2^32): the FP64
  fused engine bit for bit equal to the FP64 staged engine after 10 steps,
  FP32 close to both, mass conserved.
"""
from __future__ import annotations

import numpy as np
import pytest

from tests.helpers import sides_mask

pytestmark = pytest.mark.gpu


def _abi():
    from paraclaws_amd import abi

    abi.load()
    return abi


def _say(msg):
    print(msg, flush=True)


TKW = dict(t_final=0.5, dt0=1e-10, cfl_target=0.9, dt_max=1e-2, dt_min=1e-10, growth=1.2,
           shrink=0.5, n_snapshots=10)


def _oracle_cfg(O, n, t_final=0.5):
    cfg = O.Config(nx=n, ny=n, limiter=4, bc=(4, 4, 4, 4))
    ts = O.TimeSettings(t_final=t_final, dt0=1e-10, cfl_target=0.9, dt_max=1e-2, dt_min=1e-10,
                        growth=1.2, shrink=0.5, n_snap=10)
    return cfg, ts


@pytest.mark.parametrize("engine", ["fused", "staged"])
def test_config1_full_run_bitwise(oracle, engine):
    abi = _abi()
    O = oracle
    n = 256
    cfg, ts = _oracle_cfg(O, n)
    q = O.dam_break(cfg)
    q0 = q.copy()
    recs, tc = O.run(cfg, ts, q)
    _say(f"config 1 full run, fp64 {engine}: {len(recs)} attempts")
    with abi.Engine([abi.Domain(n, n, bc=cfg.bc)], precision=abi.FP64,
                    engine=abi.FUSED if engine == "fused" else abi.STAGED, limiter=4) as e:
        e.set_state(0, q0)
        e.fill_ghosts()
        e.time_init(**TKW)
        e.enable_log(len(recs) + 16)
        e.run(steps_per_graph=64)
        log = e.read_log(0)
        st = e.status(0)
        got = e.get_state(0)
    assert st.status == abi.STATUS_DONE
    assert log == recs
    assert st.time == tc.time
    m = sides_mask(n, n)
    assert np.array_equal(got[m], q[m])


def test_fp64_2048_twenty_steps_bitwise(oracle):
    abi = _abi()
    O = oracle
    n = 2048
    cfg, ts = _oracle_cfg(O, n)
    q = O.dam_break(cfg)
    q0 = q.copy()
    recs, tc = O.run(cfg, ts, q, max_accepted=20)
    _say("fp64 fused 2048^2, 20 accepted steps")
    with abi.Engine([abi.Domain(n, n, bc=cfg.bc)], precision=abi.FP64, limiter=4) as e:
        e.set_state(0, q0)
        e.fill_ghosts()
        e.time_init(**TKW)
        e.enable_log(64)
        e.run(steps_per_graph=32, max_accepted=20)
        log = e.read_log(0)
        got = e.get_state(0)
    assert log == recs
    m = sides_mask(n, n)
    assert np.array_equal(got[m], q[m])


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_state_round_trip_8192(prec):
    """pcw_set_state / pcw_get_state at 8192^2: FP64 exact, FP32 equal to
    the float32 rounding of the input (ghosts included)."""
    abi = _abi()
    n = 8192
    rng = np.random.default_rng(8192)
    q = (1.0 + rng.random((n + 4, n + 4, 3))).astype(np.float64)
    precision = abi.FP64 if prec == "fp64" else abi.FP32
    _say(f"round trip {prec} 8192^2")
    with abi.Engine([abi.Domain(n, n, bc=(4, 4, 4, 4))], precision=precision) as e:
        e.set_state(0, q)
        got = e.get_state(0)
    want = q if prec == "fp64" else q.astype(np.float32).astype(np.float64)
    assert np.array_equal(got, want)


def test_fp32_8192_properties():
    """40 accepted FP32 steps of the 8192^2 circular dam break with walls."""
    abi = _abi()
    from paraclaws_amd import ic

    n = 8192
    q0 = ic.dam_break(n, n)
    _say("fp32 8192^2 dam break, 40 accepted steps")
    with abi.Engine([abi.Domain(n, n, bc=(4, 4, 4, 4))], precision=abi.FP32, limiter=4) as e:
        e.set_state(0, q0)
        e.fill_ghosts()
        m0 = e.totals(0)[0]
        e.time_init(**TKW)
        e.enable_log(64)
        e.run(steps_per_graph=20, max_accepted=40)
        st = e.status(0)
        log = e.read_log(0)
        m1 = e.totals(0)[0]
        q = e.get_state(0)
    assert st.status == abi.STATUS_DONE and st.accepted == 40
    h = q[2:-2, 2:-2, 0]
    assert np.isfinite(q).all() and (h > 0).all()
    # mass: the walls are still at rest, interior faces telescope
    assert abs(m1 - m0) <= 1e-6 * m0, (m0, m1)
    # the front moved, and the state is the mirror image of itself in x and
    # y (normal momentum negated) to FP32 rounding
    assert np.abs(q[2:-2, 2:-2, 1]).max() > 0.0
    hu, hv = q[2:-2, 2:-2, 1], q[2:-2, 2:-2, 2]
    scale = np.abs(hu).max()
    for a, b in ((h, h[:, ::-1]), (h, h[::-1, :])):       # h even in x and y
        assert np.abs(a - b).max() <= 1e-5
    for a, b in ((hu, -hu[:, ::-1]), (hu, hu[::-1, :]),   # hu odd in x, even in y
                 (hv, hv[:, ::-1]), (hv, -hv[::-1, :])):  # hv even in x, odd in y
        assert np.abs(a - b).max() <= 1e-4 * scale
    # the device controller's steps follow the host controller's arithmetic:
    # growth by 1.2 from 1e-10 (the CFL stays far below the target), time
    # advanced by the next dt (src/time_control/MOD_time_control.f90:153-243)
    dt = 1e-10
    t = 0.0
    for acc, snap, dt_used, cfl, time, dt_next in log:
        assert acc == 1
        assert dt_used == dt
        dt = dt * 1.2
        t = t + dt
        assert dt_next == dt and time == t


def test_max_size_16384_fused_equals_staged():
    """16384^2, four times config 2's cells (byte offsets of the FP64 pool and
    of the staged engine's work arrays far past 2^32): 10 accepted steps of
    the dam break, the FP64 fused engine bit for bit equal to the FP64
    staged engine (the two share the per-cell arithmetic, not the indexing);
    the FP32 engine within 1e-5 relative L1 of it in h; mass conserved."""
    abi = _abi()
    from paraclaws_amd import ic

    n = 16384
    q0 = ic.dam_break(n, n)
    states, masses = {}, {}
    for name, prec, eng in (("fused", abi.FP64, abi.FUSED), ("staged", abi.FP64, abi.STAGED),
                            ("fp32", abi.FP32, abi.FUSED)):
        _say(f"16384^2 dam break, 10 accepted steps: {name}")
        with abi.Engine([abi.Domain(n, n, bc=(4, 4, 4, 4))], precision=prec, engine=eng,
                        limiter=4) as e:
            e.set_state(0, q0)
            e.fill_ghosts()
            m0 = e.totals(0)[0]
            e.time_init(**TKW)
            e.run(steps_per_graph=10, max_accepted=10)
            st = e.status(0)
            assert st.status == abi.STATUS_DONE and st.accepted == 10 and st.rejected == 0
            masses[name] = (m0, e.totals(0)[0])
            states[name] = e.get_state(0)
    del q0
    _say("16384^2: comparing")
    a, b = states["fused"].view(np.uint64), states["staged"].view(np.uint64)
    # every frame cell but the 2x2 ghost corners, which the reference never writes
    assert np.array_equal(a[2:-2], b[2:-2])
    assert np.array_equal(a[:2, 2:-2], b[:2, 2:-2]) and np.array_equal(a[-2:, 2:-2], b[-2:, 2:-2])
    del a, b
    h64 = states["fused"][2:-2, 2:-2, 0]
    h32 = states["fp32"][2:-2, 2:-2, 0]
    assert np.isfinite(h32).all()
    assert np.abs(h32 - h64).sum() <= 1e-5 * np.abs(h64).sum()
    for name, (a, b) in masses.items():
        assert abs(b - a) <= (1e-12 if name != "fp32" else 1e-6) * a, (name, a, b)
    assert np.abs(states["fused"][2:-2, 2:-2, 1]).max() > 0.0


def test_fp32_tile_height_capped_at_32bit_buffer_offsets():
    """FP32 tiles address their rows through 32-bit buffer offsets, so a tile
    spans less than 2^31 bytes of a plane.  16384 x 32768 (5.4e8 cells, a
    plane of 2.3e9 bytes) with tile_rows = 32768 asks for one band over the
    whole domain: the band is lowered to fit (two bands), and three accepted
    steps under the device controller end bit for bit where the default
    64-row tiles end."""
    abi = _abi()
    from paraclaws_amd import ic

    nx, ny = 16384, 32768
    q0 = ic.dam_break(nx, ny, xi=(-2.0, 2.0), eta=(-2.0, 2.0))
    out = {}
    for rows in (32768, 0):
        _say(f"{nx} x {ny} FP32 dam break, 3 accepted steps, tile_rows {rows}")
        with abi.Engine([abi.Domain(nx, ny, bc=(4, 4, 4, 4), xi=(-2.0, 2.0), eta=(-2.0, 2.0))],
                        precision=abi.FP32, limiter=4, tile_rows=rows) as e:
            e.set_state(0, q0)
            e.fill_ghosts()
            e.time_init(**TKW)
            e.run(steps_per_graph=3, max_accepted=3)
            st = e.status(0)
            assert st.status == abi.STATUS_DONE and st.accepted == 3 and st.rejected == 0
            out[rows] = (st.time, st.dt, e.get_state(0))
    del q0
    _say("tile height cap: comparing")
    (ta, da, a), (tb, db, b) = out[32768], out[0]
    assert ta == tb and da == db
    assert np.array_equal(a[2:-2].view(np.uint64), b[2:-2].view(np.uint64))
    assert np.abs(a[2:-2, 2:-2, 1]).max() > 0.0
