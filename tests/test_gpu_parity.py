"""GPU parity: the HIP engine through the C ABI against the oracle.

FP64 (fused and staged engines) must reproduce the oracle bit for bit: both
follow the reference's FP64 operation order with contraction disabled, and
the reference's CFL-0.9 scheme amplifies last-bit differences (a 1-ulp
perturbation reaches ~1e-10 relative by t = 0.5), so anything short of
bit-exactness would not stay within 1e-12.  FP32 is checked against the
FP64 oracle with the tolerance written at each assertion.
"""
from __future__ import annotations

import numpy as np
import pytest

from tests.helpers import rough_state, sides_mask, smooth_state, stable_dt

pytestmark = pytest.mark.gpu


def _abi():
    from paraclaws_amd import abi

    return abi


def _say(msg):
    print(msg, flush=True)


def _oracle_step(O, cfg, q, dt):
    """One accepted step of the oracle: solve, source, ghost fill."""
    ts = O.TimeSettings(dt0=dt, t_final=1e9, n_snap=1, cfl_target=0.9, shrink=1e-9)
    tc = ts.tc()
    rc = O.solve(cfg, tc, q)
    assert rc == 1
    O.source(cfg, q, None, dt)
    O.fill_ghosts(cfg, q)
    return tc.CFL


CASES = [
    # nx, ny, limiter(0 = off), higher_order, bc, state
    (37, 29, 4, True, (4, 4, 4, 4), "smooth"),
    (64, 64, 1, True, (2, 2, 2, 2), "rough"),
    (61, 33, 2, True, (3, 3, 3, 3), "rough"),
    (130, 71, 3, True, (4, 3, 2, 4), "rough"),
    (125, 9, 4, True, (2, 4, 3, 2), "smooth"),
    (8, 200, 4, False, (3, 2, 4, 3), "rough"),
    (70, 70, 0, True, (4, 4, 4, 4), "rough"),
    # the smallest grids the ABI accepts (nx, ny >= 2) and a one-strip, two-row one
    (2, 2, 4, True, (4, 4, 4, 4), "smooth"),
    (3, 7, 2, True, (3, 2, 4, 3), "rough"),
    (61, 2, 1, True, (2, 3, 3, 4), "rough"),
    # more grids whose metrics are all exactly 1 (the FP64 unit-metric kernel)
    (128, 32, 3, True, (3, 4, 2, 3), "rough"),
    (16, 64, 2, False, (4, 2, 3, 4), "smooth"),
]


@pytest.mark.parametrize("engine", ["fused", "staged"])
@pytest.mark.parametrize("case", CASES, ids=[f"{c[0]}x{c[1]}-lim{c[2]}-ho{int(c[3])}-bc{''.join(map(str, c[4]))}-{c[5]}" for c in CASES])
def test_fp64_step_bitwise(oracle, engine, case):
    abi = _abi()
    O = oracle
    nx, ny, lim, ho, bc, kind = case
    q0 = smooth_state(nx, ny, seed=nx + ny) if kind == "smooth" else rough_state(nx, ny, seed=nx * ny)
    cfg = O.Config(nx=nx, ny=ny, limiter=max(lim, 1), use_limiter=lim > 0, higher_order=ho, bc=bc)
    O.fill_ghosts(cfg, q0)
    dt = stable_dt(q0, nx, ny)
    ref = q0.copy()
    cfl_ref = _oracle_step(O, cfg, ref, dt)
    _say(f"fp64 {engine} step {case}")
    dom = abi.Domain(nx, ny, bc=bc)
    with abi.Engine([dom], precision=abi.FP64, engine=abi.FUSED if engine == "fused" else abi.STAGED,
                    higher_order=ho, use_limiter=lim > 0, limiter=max(lim, 1)) as e:
        e.set_state(0, q0)
        smax = e.step(dt)[0]
        e.accept()
        got = e.get_state(0)
    m = sides_mask(nx, ny)
    assert np.array_equal(got[m], ref[m]), np.abs(got[m] - ref[m]).max()
    dxi, deta = 2.0 / nx, 2.0 / ny
    cfl = max(smax[0] * dt / dxi, smax[1] * dt / deta)
    assert cfl == cfl_ref


def _operand_range_state(nx, ny, seed):
    """A smooth state whose momenta cover every magnitude class the FP64 cell
    physics tells apart, in row bands: plain values; exact zeros of both
    signs; values scaled to ~1e-100 (still where quotients by h and h*h share
    a reciprocal); and sprinkled tiny normals and denormals (1e-110 .. 5e-324,
    outside that range), so that some waves take each path."""
    rng = np.random.default_rng(seed)
    q = smooth_state(nx, ny, seed=seed)
    b = (ny + 4) // 4
    zeros = np.where(rng.random((b, nx + 4)) < 0.5, 0.0, -0.0)
    q[b:2 * b, :, 1] = zeros
    q[b:2 * b, :, 2] = np.where(rng.random((b, nx + 4)) < 0.5, -0.0, q[b:2 * b, :, 2])
    q[2 * b:3 * b, :, 1:] *= 1e-100
    tiny = np.array([1e-110, -1e-200, 1e-300, 5e-324, -3e-320, 2.0 ** -349, -(2.0 ** -351)])
    sel = rng.random((ny + 4 - 3 * b, nx + 4, 2)) < 0.3
    q[3 * b:, :, 1:] = np.where(sel, rng.choice(tiny, size=sel.shape), q[3 * b:, :, 1:])
    return q


@pytest.mark.parametrize("engine", ["fused", "staged"])
def test_fp64_step_bitwise_operand_ranges(oracle, engine):
    """One step from a state with momenta of every magnitude class (see
    _operand_range_state) equals the oracle bit for bit, signs of zero
    included."""
    abi = _abi()
    O = oracle
    nx, ny = 130, 52
    q0 = _operand_range_state(nx, ny, seed=21)
    cfg = O.Config(nx=nx, ny=ny, limiter=4, bc=(4, 3, 2, 4))
    O.fill_ghosts(cfg, q0)
    dt = stable_dt(q0, nx, ny)
    ref = q0.copy()
    cfl_ref = _oracle_step(O, cfg, ref, dt)
    _say(f"fp64 {engine} step, operand ranges")
    with abi.Engine([abi.Domain(nx, ny, bc=cfg.bc)], precision=abi.FP64,
                    engine=abi.FUSED if engine == "fused" else abi.STAGED, limiter=4) as e:
        e.set_state(0, q0)
        smax = e.step(dt)[0]
        e.accept()
        got = e.get_state(0)
    m = sides_mask(nx, ny)
    assert np.array_equal(got[m].view(np.uint64), ref[m].view(np.uint64))
    cfl = max(smax[0] * dt / (2.0 / nx), smax[1] * dt / (2.0 / ny))
    assert cfl == cfl_ref


FP32_CASES = CASES + [
    # grids with interior strips and long tiles, so that the FP32 kernel's
    # branch-free runs of rows are used; the last field is tile_rows
    (250, 301, 4, True, (4, 3, 2, 4), "rough", 64),
    (247, 190, 2, True, (3, 2, 4, 3), "smooth", 16),
    (181, 97, 1, True, (2, 4, 3, 2), "rough", 37),
    # nx a multiple of 60: the last strip fills its slot and its right ghost
    # columns sit in that slot's padding
    (120, 45, 4, True, (3, 3, 4, 2), "rough", 16),
    (60, 70, 2, True, (2, 4, 3, 3), "smooth"),
]


@pytest.mark.parametrize("case", FP32_CASES, ids=[f"{c[0]}x{c[1]}-lim{c[2]}-ho{int(c[3])}-bc{''.join(map(str, c[4]))}-{c[5]}" + (f"-t{c[6]}" if len(c) > 6 else "") for c in FP32_CASES])
def test_fp32_step_close_cases(oracle, case):
    """One FP32 fused step against the FP64 oracle on every BC combination,
    limiter and grid shape of the FP64 cases plus grids with interior strips:
    interior and ghost cells within 4e-6 of the field scale, sweep speeds
    within 1e-6 relative."""
    abi = _abi()
    O = oracle
    nx, ny, lim, ho, bc, kind = case[:6]
    tile_rows = case[6] if len(case) > 6 else 0
    q0 = smooth_state(nx, ny, seed=nx + ny) if kind == "smooth" else rough_state(nx, ny, seed=nx * ny)
    cfg = O.Config(nx=nx, ny=ny, limiter=max(lim, 1), use_limiter=lim > 0, higher_order=ho, bc=bc)
    O.fill_ghosts(cfg, q0)
    dt = stable_dt(q0, nx, ny)
    ref = q0.copy()
    cfl_ref = _oracle_step(O, cfg, ref, dt)
    _say(f"fp32 fused step {case}")
    with abi.Engine([abi.Domain(nx, ny, bc=bc)], precision=abi.FP32, higher_order=ho,
                    use_limiter=lim > 0, limiter=max(lim, 1), tile_rows=tile_rows) as e:
        e.set_state(0, q0)
        smax = e.step(dt)[0]
        e.accept()
        got = e.get_state(0)
    m = sides_mask(nx, ny)
    for k in range(3):
        d = np.abs(got[..., k][m] - ref[..., k][m])
        assert d.max() <= 4e-6 * max(1.0, np.abs(ref[..., k][m]).max()), (k, d.max())
    dxi, deta = 2.0 / nx, 2.0 / ny
    cfl = max(smax[0] * dt / dxi, smax[1] * dt / deta)
    assert abs(cfl - cfl_ref) <= 1e-6 * cfl_ref


@pytest.mark.parametrize("ixy", [1, 2])
def test_fp64_single_sweep_bitwise(oracle, ixy):
    abi = _abi()
    O = oracle
    nx, ny = 45, 38
    q0 = rough_state(nx, ny, seed=7)
    cfg = O.Config(nx=nx, ny=ny, limiter=4, bc=(3, 4, 2, 3))
    O.fill_ghosts(cfg, q0)
    dt = stable_dt(q0, nx, ny)
    ref = q0.copy()
    cfl_ref = O.sweep(cfg, ref, ixy, dt)
    _say(f"fp64 staged single sweep ixy={ixy}")
    with abi.Engine([abi.Domain(nx, ny, bc=cfg.bc)], precision=abi.FP64, engine=abi.STAGED) as e:
        e.set_state(0, q0)
        smax = e.sweep(ixy, dt)[0]
        got = e.get_state(0)
    assert np.array_equal(got, ref)
    assert smax * dt / (2.0 / (nx if ixy == 1 else ny)) == cfl_ref


def test_fp32_step_close(oracle):
    abi = _abi()
    O = oracle
    nx, ny = 150, 90
    q0 = smooth_state(nx, ny, seed=3)
    cfg = O.Config(nx=nx, ny=ny, limiter=4, bc=(4, 4, 4, 4))
    O.fill_ghosts(cfg, q0)
    dt = stable_dt(q0, nx, ny)
    ref = q0.copy()
    _oracle_step(O, cfg, ref, dt)
    _say("fp32 fused step")
    with abi.Engine([abi.Domain(nx, ny, bc=cfg.bc)], precision=abi.FP32, limiter=4) as e:
        e.set_state(0, q0)
        e.step(dt)
        e.accept()
        got = e.get_state(0)
    inner = (slice(2, -2), slice(2, -2))
    for k in range(3):
        d = np.abs(got[inner][..., k] - ref[inner][..., k])
        # FP32 state: per-cell error within a few ulp of the field scale
        assert d.max() <= 4e-6 * max(1.0, np.abs(ref[inner][..., k]).max()), (k, d.max())


def _config1_like(O, n, t_final):
    cfg = O.Config(nx=n, ny=n, limiter=4, bc=(4, 4, 4, 4))
    ts = O.TimeSettings(t_final=t_final, cfl_target=0.9, dt_max=1e-2, n_snap=10)
    return cfg, ts


@pytest.mark.parametrize("engine", ["fused", "staged"])
def test_fp64_device_controlled_trajectory(oracle, engine):
    """Config-1 physics (dam break, MC, walls, CFL 0.9) at 64^2 to t = 0.1:
    identical dt/CFL/time sequence and final state."""
    abi = _abi()
    O = oracle
    n = 64
    cfg, ts = _config1_like(O, n, 0.1)
    q = O.dam_break(cfg)
    q_init = q.copy()
    recs, tc = O.run(cfg, ts, q)
    _say(f"fp64 {engine} device-controlled run 64^2")
    with abi.Engine([abi.Domain(n, n, bc=cfg.bc)], precision=abi.FP64,
                    engine=abi.FUSED if engine == "fused" else abi.STAGED, limiter=4) as e:
        e.set_state(0, q_init)
        e.fill_ghosts()
        e.time_init(t_final=ts.t_final, dt0=ts.dt0, cfl_target=ts.cfl_target, dt_max=ts.dt_max,
                    dt_min=ts.dt_min, growth=ts.growth, shrink=ts.shrink, n_snapshots=ts.n_snap)
        e.enable_log(len(recs) + 16)
        e.run(steps_per_graph=16)
        log = e.read_log(0)
        st = e.status(0)
        got = e.get_state(0)
    assert st.status == abi.STATUS_DONE
    assert len(log) == len(recs)
    assert log == recs
    assert np.array_equal(got[sides_mask(n, n)], q[sides_mask(n, n)])


def test_fp64_host_controlled_trajectory_with_rejection(oracle):
    """Host-controlled loop (pcw_step/accept/reject) with a first step forced
    to be rejected (dt0 far above the CFL limit)."""
    abi = _abi()
    from paraclaws_amd.timectl import TimeControl, run_host_controlled

    O = oracle
    n = 48
    cfg = O.Config(nx=n, ny=n, limiter=2, bc=(2, 2, 2, 2))
    ts = O.TimeSettings(t_final=0.05, dt0=0.05, cfl_target=0.9, dt_max=1e-2, n_snap=5)
    q = O.dam_break(cfg)
    q_init = q.copy()
    recs, _ = O.run(cfg, ts, q)
    assert recs[0][0] == 0, "the first attempt must be rejected"
    _say("fp64 fused host-controlled run with rejection")
    tc = TimeControl(0.0, ts.t_final, ts.dt0, ts.n_snap, ts.cfl_target, ts.dt_max, ts.dt_min,
                     ts.growth, ts.shrink, True)
    with abi.Engine([abi.Domain(n, n, bc=cfg.bc)], precision=abi.FP64, limiter=2) as e:
        e.set_state(0, q_init)
        e.fill_ghosts()
        got_recs = run_host_controlled(e, tc, 2.0 / n, 2.0 / n)
        got = e.get_state(0)
    assert got_recs == recs
    assert np.array_equal(got[sides_mask(n, n)], q[sides_mask(n, n)])


@pytest.mark.parametrize("engine", ["fused", "staged"])
def test_fp64_device_controlled_rejection(oracle, engine):
    """Same forced rejection with the device controller: the eta sweep of the
    rejected attempt meets negative depths, which the reference never
    computes (it rejects after the xi sweep), so it must not stop the run."""
    abi = _abi()
    O = oracle
    n = 48
    cfg = O.Config(nx=n, ny=n, limiter=2, bc=(2, 2, 2, 2))
    ts = O.TimeSettings(t_final=0.05, dt0=0.05, cfl_target=0.9, dt_max=1e-2, n_snap=5)
    q = O.dam_break(cfg)
    q_init = q.copy()
    recs, _ = O.run(cfg, ts, q)
    _say(f"fp64 {engine} device-controlled run with rejection")
    with abi.Engine([abi.Domain(n, n, bc=cfg.bc)], precision=abi.FP64,
                    engine=abi.FUSED if engine == "fused" else abi.STAGED, limiter=2) as e:
        e.set_state(0, q_init)
        e.fill_ghosts()
        e.time_init(t_final=ts.t_final, dt0=ts.dt0, cfl_target=ts.cfl_target, dt_max=ts.dt_max,
                    dt_min=ts.dt_min, growth=ts.growth, shrink=ts.shrink, n_snapshots=ts.n_snap)
        e.enable_log(len(recs) + 8)
        e.run(steps_per_graph=8)
        log = e.read_log(0)
        got = e.get_state(0)
    assert log == recs
    assert np.array_equal(got[sides_mask(n, n)], q[sides_mask(n, n)])


# nx, ny, dam depth; 32 x 16 has mesh metrics that are exactly 1 (the FP64
# kernel skips them), the others compute them: one step runs both launches
BATCH_SIZES = [(40, 40, 1.8), (97, 61, 3.0), (64, 130, 2.5), (23, 18, 6.0), (32, 16, 2.0)]


def test_batch_equals_solo_and_oracle(oracle):
    """Instances of uneven sizes in one handle: each equals the same instance
    run alone (bit for bit, FP32 and FP64), and the FP64 batch equals the oracle."""
    abi = _abi()
    O = oracle
    doms, qs, cfgs = [], [], []
    for nx, ny, r in BATCH_SIZES:
        cfg = O.Config(nx=nx, ny=ny, limiter=4, bc=(4, 4, 4, 4))
        q = O.dam_break(cfg, h_in=r)
        O.fill_ghosts(cfg, q)
        doms.append(abi.Domain(nx, ny, bc=cfg.bc))
        qs.append(q)
        cfgs.append(cfg)
    tkw = dict(t_final=0.03, dt0=1e-4, cfl_target=0.9, dt_max=1e-2, dt_min=1e-10, growth=1.2,
               shrink=0.5, n_snapshots=3)
    for prec in (abi.FP64, abi.FP32):
        _say(f"batch precision={prec}")
        with abi.Engine(doms, precision=prec, limiter=4) as e:
            for k, q in enumerate(qs):
                e.set_state(k, q)
            e.time_init(**tkw)
            e.run(steps_per_graph=8)
            batch = [e.get_state(k) for k in range(len(doms))]
            stats = [e.status(k) for k in range(len(doms))]
        for k, d in enumerate(doms):
            assert stats[k].status == abi.STATUS_DONE
            with abi.Engine([d], precision=prec, limiter=4) as e1:
                e1.set_state(0, qs[k])
                e1.time_init(**tkw)
                e1.run(steps_per_graph=8)
                solo = e1.get_state(0)
            assert np.array_equal(batch[k], solo), (prec, k)
            if prec == abi.FP64:
                ref = qs[k].copy()
                ts = O.TimeSettings(t_final=tkw["t_final"], dt0=tkw["dt0"], n_snap=3)
                O.run(cfgs[k], ts, ref)
                m = sides_mask(d.nx, d.ny)
                assert np.array_equal(batch[k][m], ref[m]), k


def test_batch_tile_order_equals_solo(oracle):
    """Larger uneven instances with short tiles, so that the tile list spans
    every XCD range and holds interior and boundary tiles (the host reorders
    them per range): each instance of the batch equals the instance run alone,
    bit for bit in FP32 and FP64, and the FP64 batch equals the oracle."""
    abi = _abi()
    O = oracle
    sizes = [(300, 250, 2.2), (250, 301, 3.0), (181, 97, 1.6)]
    doms, qs, cfgs = [], [], []
    for nx, ny, r in sizes:
        cfg = O.Config(nx=nx, ny=ny, limiter=4, bc=(4, 3, 2, 4))
        q = O.dam_break(cfg, h_in=r)
        O.fill_ghosts(cfg, q)
        doms.append(abi.Domain(nx, ny, bc=cfg.bc))
        qs.append(q)
        cfgs.append(cfg)
    tkw = dict(t_final=0.004, dt0=1e-4, cfl_target=0.9, dt_max=1e-2, dt_min=1e-10, growth=1.2,
               shrink=0.5, n_snapshots=1)
    for prec in (abi.FP64, abi.FP32):
        _say(f"batch tile order precision={prec}")
        with abi.Engine(doms, precision=prec, limiter=4, tile_rows=16, waves_per_block=4) as e:
            for k, q in enumerate(qs):
                e.set_state(k, q)
            e.time_init(**tkw)
            e.run(steps_per_graph=8)
            batch = [e.get_state(k) for k in range(len(doms))]
            stats = [e.status(k) for k in range(len(doms))]
        for k, d in enumerate(doms):
            assert stats[k].status == abi.STATUS_DONE
            with abi.Engine([d], precision=prec, limiter=4, tile_rows=37, waves_per_block=2) as e1:
                e1.set_state(0, qs[k])
                e1.time_init(**tkw)
                e1.run(steps_per_graph=8)
                solo = e1.get_state(0)
            assert np.array_equal(batch[k], solo), (prec, k)
        if prec == abi.FP64:
            ref = qs[2].copy()
            ts = O.TimeSettings(t_final=tkw["t_final"], dt0=tkw["dt0"], n_snap=1)
            O.run(cfgs[2], ts, ref)
            m = sides_mask(doms[2].nx, doms[2].ny)
            assert np.array_equal(batch[2][m], ref[m])


def test_totals(oracle):
    abi = _abi()
    nx, ny = 300, 170
    q = smooth_state(nx, ny, seed=11)
    for prec in (abi.FP64, abi.FP32):
        with abi.Engine([abi.Domain(nx, ny)], precision=prec) as e:
            e.set_state(0, q)
            tot = e.totals(0)
        qq = q if prec == abi.FP64 else q.astype(np.float32).astype(np.float64)
        ref = qq[2:-2, 2:-2].reshape(-1, 3).sum(axis=0)
        assert np.allclose(tot, ref, rtol=1e-13, atol=1e-12), (prec, tot, ref)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_error_on_dry_cell(oracle, prec):
    """A dry interior cell (h = hu = hv = 0): the xi Jacobian holds 0/0, the
    reference stops ("isnan in jakobi", SMOD_jakobian.f90:55-58) in the xi
    sweep; both engines report PCW_ERR_STATE (FP32: h <= 0 in the sweep's
    range)."""
    abi = _abi()
    nx = ny = 20
    q = smooth_state(nx, ny, seed=1)
    q[10, 10, :] = 0.0
    precision = abi.FP64 if prec == "fp64" else abi.FP32
    with abi.Engine([abi.Domain(nx, ny)], precision=precision) as e:
        e.set_state(0, q)
        with pytest.raises(abi.PcwError) as ei:
            e.step(1e-4)
        assert ei.value.code == abi.PCW_ERR_STATE


BAD_CELLS = {"dry": (0.0, 0.0, 0.0), "dry_moving": (0.0, 0.3, -0.2),
             "negative": (-1e-3, 0.0, 0.0), "nan": (np.nan, 0.0, 0.0)}
BAD_PLACES = {"interior_tile": (130, 128), "bottom_row": (1, 128), "first_column": (128, 1),
              "last_column": (128, 256)}


@pytest.mark.parametrize("place", list(BAD_PLACES))
@pytest.mark.parametrize("kind", list(BAD_CELLS))
def test_fp32_bad_cell_found_from_the_speed_maximum(kind, place):
    """The FP32 step finds a cell with h <= 0 or NaN from its sweep's speed
    maximum (the cell's |u| + sqrt(g h) is +inf or NaN) instead of a test per
    cell; it reports PCW_ERR_STATE from the xi sweep wherever the cell is: in
    a tile that runs the unconditional body (256^2 in 32-row tiles: strips at
    columns 61-240, bands away from rows 1-2 and ny-1..ny), on the domain's
    first row, or on its first or last column (the general body).  A good
    state on the same tiles steps without error."""
    abi = _abi()
    n = 256
    q = smooth_state(n, n, seed=2)
    _say(f"fp32 bad cell {kind} at {place}")
    with abi.Engine([abi.Domain(n, n)], precision=abi.FP32, limiter=4, tile_rows=32) as e:
        e.set_state(0, q)
        e.step(1e-4)
        e.reject()
        j, i = BAD_PLACES[place]
        q[j + 1, i + 1] = BAD_CELLS[kind]
        e.set_state(0, q)
        with pytest.raises(abi.PcwError) as ei:
            e.step(1e-4)
        assert ei.value.code == abi.PCW_ERR_STATE


def test_fp64_zero_depth_with_momentum_reports_eta():
    """h = 0 with non-zero momenta: the xi Jacobian has infinite entries but
    no NaN, so the xi sweep itself does not stop; its speed is infinite (the
    step would be rejected) and the eta sweep then meets NaN.  pcw_step
    reports PCW_ERR_STATE_ETA and leaves the xi speed for the caller's
    validity test."""
    abi = _abi()
    nx = ny = 20
    q = smooth_state(nx, ny, seed=1)
    q[10, 10, 0] = 0.0
    with abi.Engine([abi.Domain(nx, ny)], precision=abi.FP64) as e:
        e.set_state(0, q)
        with pytest.raises(abi.PcwError) as ei:
            e.step(1e-4)
        assert ei.value.code == abi.PCW_ERR_STATE_ETA
        assert np.isinf(e.last_smax()[0][0])


def test_user_bc_rejected():
    abi = _abi()
    with pytest.raises(abi.PcwError) as ei:
        abi.Engine([abi.Domain(16, 16, bc=(1, 2, 2, 2))])
    assert ei.value.code == abi.PCW_ERR_BC


RESTORE_MODES = [("fp64", "fused"), ("fp64", "staged"), ("fp32", "fused")]


@pytest.mark.parametrize("prec,engine", RESTORE_MODES, ids=[f"{p}-{e}" for p, e in RESTORE_MODES])
def test_rejected_attempt_leaves_state_and_ghosts(prec, engine):
    """A rejected attempt leaves the current state as it was, ghost cells
    included, also when those ghosts are not what the boundary conditions
    would give (set without a ghost fill): under device control (one
    attempt, dt0 far above the CFL limit) and under host control (pcw_step,
    then pcw_reject).  The reference restores solution_old, ghosts included
    (src/main.f90:213-230)."""
    abi = _abi()
    n = 40
    q = rough_state(n, n, seed=5)  # random ghost cells
    p = abi.FP32 if prec == "fp32" else abi.FP64
    eng = abi.STAGED if engine == "staged" else abi.FUSED
    _say(f"rejected attempt keeps state and ghosts: {prec} {engine}")
    with abi.Engine([abi.Domain(n, n, bc=(4, 3, 2, 4))], precision=p, engine=eng, limiter=4) as e:
        e.set_state(0, q)
        ref = e.get_state(0)  # as stored (FP32: rounded once)
        e.time_init(t_final=1.0, dt0=0.5, cfl_target=0.9, dt_max=1.0, dt_min=1e-10)
        e.run(steps_per_graph=1, max_attempts=1)
        st = e.status(0)
        assert (st.accepted, st.rejected) == (0, 1)
        assert np.array_equal(e.get_state(0), ref)
        try:
            e.step(0.5)
        except abi.PcwError as err:  # the eta sweep may meet negative depths at this dt
            assert err.code == abi.PCW_ERR_STATE_ETA
        e.reject()
        assert np.array_equal(e.get_state(0), ref)


def test_finished_instance_state_untouched_by_later_runs():
    """An instance that has finished takes no attempt in later launches, and
    the FP64 ghost fill that follows each fused attempt leaves its state
    alone: a state set after it finished (random ghost cells, no ghost fill)
    is still there, bit for bit, after another run."""
    abi = _abi()
    n = 36
    _say("finished instance untouched: fp64 fused")
    with abi.Engine([abi.Domain(n, n, bc=(4, 4, 4, 4))], precision=abi.FP64, limiter=4) as e:
        e.set_state(0, smooth_state(n, n, seed=3))
        e.fill_ghosts()
        e.time_init(t_final=1.0, dt0=1e-4, cfl_target=0.9, dt_max=1e-2, dt_min=1e-10)
        e.run(steps_per_graph=2, max_accepted=2)
        st = e.status(0)
        assert st.status == abi.STATUS_DONE and st.accepted == 2
        q = rough_state(n, n, seed=9)
        e.set_state(0, q)
        e.run(steps_per_graph=2, max_accepted=2)
        st2 = e.status(0)
        assert (st2.accepted, st2.rejected) == (st.accepted, st.rejected)
        assert np.array_equal(e.get_state(0), q)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_autotune_mid_run_keeps_the_trajectory(prec):
    """pcw_autotune times candidate tiles on a handle in the middle of a
    device-controlled run: the run that follows is the one a fixed-tile
    handle makes, bit for bit (state, status, time and the step log).  Cell
    values do not depend on the tiling, and the autotuner restores the
    controller state it steps from."""
    abi = _abi()
    n = 200
    precision = abi.FP64 if prec == "fp64" else abi.FP32
    kw = dict(t_final=0.2, dt0=1e-6, cfl_target=0.9, dt_max=1e-2, dt_min=1e-10, n_snapshots=4)
    out = []
    for tune in (False, True):
        _say(f"autotune mid-run {prec} tune={tune}")
        with abi.Engine([abi.Domain(n, n, bc=(4, 2, 3, 4))], precision=precision, limiter=4,
                        tile_rows=64, waves_per_block=2) as e:
            e.set_state(0, smooth_state(n, n, seed=11))
            e.fill_ghosts()
            e.time_init(**kw)
            e.enable_log(400)
            e.run(steps_per_graph=8, max_accepted=12)
            if tune:
                rows, waves = e.autotune(steps=2)
                assert rows >= 1 and waves in (1, 2, 4)
            e.run(steps_per_graph=8, max_accepted=40)
            st = e.status(0)
            out.append((e.get_state(0), (st.status, st.accepted, st.rejected, st.time, st.dt),
                        e.read_log(0)))
    (q0, s0, l0), (q1, s1, l1) = out
    assert s0 == s1 and s0[1] == 40
    assert l0 == l1
    m = sides_mask(n, n)
    assert np.array_equal(q0[m].view(np.uint64), q1[m].view(np.uint64))


def test_host_steps_after_a_finished_run_leave_finished_instances():
    """Host stepping after a device-controlled run: the step kernels advance
    the running instances only, and pcw_accept switches those only, so an
    instance the run left DONE keeps its state (not the stale pending frame)."""
    abi = _abi()
    n = 40
    _say("host steps after a finished run")
    with abi.Engine([abi.Domain(n, n, bc=(4, 4, 4, 4)), abi.Domain(n, n, bc=(4, 4, 4, 4))],
                    precision=abi.FP32, limiter=4) as e:
        for k in range(2):
            e.set_state(k, smooth_state(n, n, seed=20 + k))
        e.fill_ghosts()
        e.time_init(t_final=1.0, dt0=1e-4, cfl_target=0.9, dt_max=1e-2, dt_min=1e-10)
        e.run(steps_per_graph=3, max_accepted=3)
        assert e.status(0).status == abi.STATUS_DONE
        done = [e.get_state(k) for k in range(2)]
        for _ in range(3):
            e.step(1e-4)
            e.accept()
        for k in range(2):
            assert np.array_equal(e.get_state(k), done[k])
