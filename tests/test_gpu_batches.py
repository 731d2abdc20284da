"""The batched workloads at their full sizes, and the mass budget of outflow
boundaries.

* The 512-instance dam-break ensemble (sizes 64^2..1024^2, each instance its
  own adaptive dt) to t = 0.1: every instance ends at t = 0.1, and members run
  alone give the batch's state bit for bit.  (The depth sum is not kept
  exactly even with walls: the eta sweep reads the ghost rows as they were
  before the step, src/main.f90:246, while the rows next to them are already
  xi-updated, so a wall face passes a small flux; about 7e-4 of the depth sum
  over this run.)
* The Manning + bathymetry batch of 2048 x 1024 instances: members run alone
  equal the batch bit for bit.
* Outflow (zero-gradient) sides: in conservation form the xi fluxes
  telescope, a zero-gradient face has zero waves, walls pass no mass and the
  bed-slope and friction terms leave h alone, so one step changes the depth
  sum by -dt/dxi * sum_j (hu(nx, j) - hu(1, j)) of the step's input state,
  up to the reference's eps leak (SMOD_approximate_Riemann_solvers.f90:60-101:
  s1 W1 + s2 W2 = df - eps q*, about 1e-10 per face), as long as the rows
  next to the y-walls stay at rest (else the stale ghost rows above pass a
  flux).  Checked step by step.
"""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

T_ENS = 0.1
TIME_KW = dict(t_final=T_ENS, dt0=1e-10, cfl_target=0.9, dt_max=1e-2, dt_min=1e-10, growth=1.2,
               shrink=0.5, n_snapshots=10)


def _say(msg):
    print(msg, flush=True)


def _ensemble_engine(abi, ic, members):
    doms = [abi.Domain(n, n, bc=(4, 4, 4, 4)) for n, _ in members]
    e = abi.Engine(doms, precision=abi.FP32, limiter=4)
    for k, (n, r) in enumerate(members):
        e.set_state(k, ic.dam_break(n, n, h_in=r))
    e.fill_ghosts()
    e.time_init(**TIME_KW)
    return e


def test_config4_ensemble_full_size():
    from paraclaws_amd import abi, ic

    members = ic.ensemble_config4()
    assert len(members) == 512
    _say("512-instance ensemble to t = 0.1")
    e = _ensemble_engine(abi, ic, members)
    m0 = np.array([e.totals(k)[0] for k in range(len(members))])
    e.run(steps_per_graph=64)
    steps = []
    for k in range(len(members)):
        st = e.status(k)
        assert st.status == abi.STATUS_DONE and st.error == 0, (k, st.status, st.error)
        assert st.time == pytest.approx(T_ENS, abs=1e-12), (k, st.time)
        steps.append(st.accepted)
    m1 = np.array([e.totals(k)[0] for k in range(len(members))])
    _say(f"  accepted steps {min(steps)}..{max(steps)}; depth sum changed by at most "
         f"{np.max(np.abs(m1 - m0) / m0):.2e} (stale ghost rows at the walls)")
    ns = np.array([n for n, _ in members])
    picks = [int(np.argmin(ns)), int(np.argmax(ns)), int(np.argsort(ns)[len(ns) // 2])]
    batch = {k: e.get_state(k) for k in picks}
    e.close()
    for k in picks:
        n, r = members[k]
        _say(f"  member {k} ({n}^2, ratio {r:.3f}) alone")
        with abi.Engine([abi.Domain(n, n, bc=(4, 4, 4, 4))], precision=abi.FP32, limiter=4) as s:
            s.set_state(0, ic.dam_break(n, n, h_in=r))
            s.fill_ghosts()
            s.time_init(**TIME_KW)
            s.run(steps_per_graph=64)
            st = s.status(0)
            assert st.status == abi.STATUS_DONE and st.accepted == steps[k]
            assert np.array_equal(s.get_state(0).view(np.uint64), batch[k].view(np.uint64))


def _friction_member(abi, ic, k, nx, ny):
    g = ic.SplitMix64(5)
    nman = [g.uniform(0.01, 0.05) for _ in range(k + 1)][k]
    dom = abi.Domain(nx, ny, bc=(2, 2, 4, 4), xi=(-2.0, 2.0), eta=(-1.0, 1.0),
                     source=abi.SRC_BATHY_FRICTION, manning_n=nman)
    b = ic.gaussian_bathymetry(nx, ny, seed=5000 + k, xi=(-2.0, 2.0), eta=(-1.0, 1.0))
    q = ic.still_water_with_step(b, nx, xi=(-2.0, 2.0), dry_tol=1e-6)
    return dom, b, q


def test_config5_members_equal_alone():
    """Four instances of the friction batch (2048 x 1024, Manning n from the
    batch's stream, bathymetry seeds 5000+k, superbee, outflow in x, walls in
    y) to t = 0.02 in one batch, and each alone: bit for bit."""
    from paraclaws_amd import abi, ic

    nx, ny, n_inst = 2048, 1024, 4
    kw = dict(TIME_KW, t_final=0.02)
    mem = [_friction_member(abi, ic, k, nx, ny) for k in range(n_inst)]
    _say("friction batch, 4 x 2048x1024 to t = 0.02")
    with abi.Engine([m[0] for m in mem], precision=abi.FP32, limiter=2, dry_tol=1e-6) as e:
        for k, (_, b, q) in enumerate(mem):
            e.set_state(k, q)
            e.set_bathymetry(k, b)
        e.fill_ghosts()
        e.time_init(**kw)
        e.run(steps_per_graph=32)
        sts = [e.status(k) for k in range(n_inst)]
        batch = [e.get_state(k) for k in range(n_inst)]
    for k, (dom, b, q) in enumerate(mem):
        assert sts[k].status == abi.STATUS_DONE and sts[k].error == 0
        with abi.Engine([dom], precision=abi.FP32, limiter=2, dry_tol=1e-6) as s:
            s.set_state(0, q)
            s.set_bathymetry(0, b)
            s.fill_ghosts()
            s.time_init(**kw)
            s.run(steps_per_graph=32)
            assert s.status(0).accepted == sts[k].accepted
            assert np.array_equal(s.get_state(0).view(np.uint64), batch[k].view(np.uint64)), k


BUDGET_MODES = [("fp64", "fused"), ("fp64", "staged"), ("fp32", "fused")]


@pytest.mark.parametrize("prec,engine", BUDGET_MODES, ids=[f"{p}-{e}" for p, e in BUDGET_MODES])
def test_outflow_mass_budget(prec, engine):
    """A config-5 style instance at 256 x 128 on [-2,2]x[-1,1] (Manning
    n = 0.03, superbee, outflow x-sides, walls in y): bathymetry (b <= 0.6, so
    that every cell stays wet), a dam-break step h += 1 on x < 0 and a flow
    u = 0.5 in the band |y| < 0.5, flat still water in the 32 rows next to
    each wall.  Each of 10 steps changes
    the depth sum by -dt/dxi * sum_j (hu(nx, j) - hu(1, j)) of its input state
    (FP64: to 1e-9 of the change plus the eps leak; FP32: to the rounding of
    the stored state), and the 12 rows next to each wall stay at rest (a
    disturbance spreads at most two rows per step)."""
    from paraclaws_amd import abi, ic

    nx, ny = 256, 128
    xi = (-2.0, 2.0)
    dxi = (xi[1] - xi[0]) / nx
    band = np.zeros(ny + 4, dtype=bool)
    band[2 + 32:2 + ny - 32] = True  # rows 33 .. ny-32: |y| < 0.5
    b = ic.gaussian_bathymetry(nx, ny, seed=5001, xi=xi, eta=(-1.0, 1.0))
    b *= 0.6 / b.max()
    b[~band] = 0.0
    q = ic.still_water_with_step(b, nx, xi=xi, step=0.0, dry_tol=1e-6)
    x = ic.centres(nx, *xi)  # frame columns -1 .. nx+2
    q[np.ix_(band, x < 0.0, [0])] += 1.0  # a dam-break step inside the band
    q[band, :, 1] = 0.5 * q[band, :, 0]
    p = abi.FP64 if prec == "fp64" else abi.FP32
    eng = abi.STAGED if engine == "staged" else abi.FUSED
    dom = abi.Domain(nx, ny, bc=(2, 2, 4, 4), xi=xi, eta=(-1.0, 1.0),
                     source=abi.SRC_BATHY_FRICTION, manning_n=0.03)
    _say(f"outflow budget {prec} {engine}")
    worst = 0.0
    with abi.Engine([dom], precision=p, engine=eng, limiter=2, dry_tol=1e-6) as e:
        e.set_state(0, q)
        e.set_bathymetry(0, b)
        e.fill_ghosts()
        quiet0 = e.get_state(0)[np.r_[2:14, ny - 10:ny + 2]]
        for k in range(10):
            s = e.get_state(0)[2:-2, 2:-2]
            assert s[..., 0].min() > 0.1  # wet: no cut-off changes h
            flux = float(np.sum(s[:, -1, 1]) - np.sum(s[:, 0, 1]))
            speed = np.max(np.abs(s[..., 1:]).max(axis=-1) / s[..., 0] + np.sqrt(9.81 * s[..., 0]))
            dt = 0.45 * dxi / speed
            m0 = e.totals(0)[0]
            e.step(dt)
            e.accept()
            m1 = e.totals(0)[0]
            expect = -dt / dxi * flux
            leak = 2.0 * (nx + 1) * ny * 1e-10 * 3.0 * dt / dxi  # eps q* per face, both sweeps
            tol = 1e-9 * abs(expect) + leak if prec == "fp64" else 2e-4
            assert abs((m1 - m0) - expect) <= tol, (k, m1 - m0, expect)
            worst = max(worst, abs((m1 - m0) - expect))
            assert abs(expect) > 0.5  # the budget is not trivially zero
        quiet1 = e.get_state(0)[np.r_[2:14, ny - 10:ny + 2]]
        # at rest up to the reference's eps leak of a state at rest (waves of
        # eps q / (2 s), about 1e-11; the FP32 engine drops those terms)
        assert np.abs(quiet1 - quiet0).max() <= 1e-9
    _say(f"  worst |dM - budget| = {worst:.3e}")


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_runs_stop_at_snapshots(prec):
    """pcw_stop_at_snapshots on a batch of three instances with their own dt:
    each device-controlled run stops every instance after the accepted step
    that reaches its next snapshot time (the instance's time is that snapshot
    time exactly, the last record of its log marks the snapshot, no earlier
    record does), and the next run continues it.  Runs stopped at every
    snapshot end in the states of one uninterrupted run, bit for bit, with the
    same accepted and rejected counts (the time loop of src/main.f90:193-283
    returning to the host for snapshot output at :256-268)."""
    from paraclaws_amd import abi, ic

    members = [(64, 2.0), (96, 4.0), (130, 1.5)]
    precision = abi.FP64 if prec == "fp64" else abi.FP32
    t_final, n_snap = 0.05, 5
    kw = dict(TIME_KW, t_final=t_final, n_snapshots=n_snap)
    interval = t_final / n_snap
    snaps = [i * interval for i in range(1, n_snap + 1)]
    snaps[-1] = t_final

    def engine():
        e = abi.Engine([abi.Domain(n, n, bc=(4, 4, 4, 4)) for n, _ in members], precision=precision,
                       limiter=4)
        for k, (n, r) in enumerate(members):
            e.set_state(k, ic.dam_break(n, n, h_in=r))
        e.fill_ghosts()
        e.time_init(**kw)
        return e

    _say(f"stop at snapshots {prec}: uninterrupted run")
    with engine() as ref:
        ref.run(steps_per_graph=16)
        want = [(ref.get_state(k), ref.status(k)) for k in range(len(members))]
    _say(f"stop at snapshots {prec}: runs stopped at every snapshot")
    with engine() as e:
        e.stop_at_snapshots(True)
        reached = {k: [] for k in range(len(members))}
        for _ in range(n_snap):
            e.enable_log(100000)
            e.run(steps_per_graph=16)
            for k in range(len(members)):
                log = e.read_log(k)
                assert log, (k, reached[k])
                assert all(r[1] == 0 for r in log[:-1])
                acc, snap, *_ = log[-1]
                assert acc == 1 and snap == len(reached[k]) + 1
                assert e.status(k).time == snaps[snap - 1]
                reached[k].append(snap)
        for k in range(len(members)):
            st = e.status(k)
            assert reached[k] == list(range(1, n_snap + 1))
            assert st.status == abi.STATUS_DONE and st.time == t_final
            q, st_ref = want[k]
            assert (st.accepted, st.rejected) == (st_ref.accepted, st_ref.rejected)
            assert np.array_equal(e.get_state(k).view(np.uint64), q.view(np.uint64))


TILE_CASES = [((37, 41), (4, 4, 4, 4)), ((130, 67), (2, 3, 4, 2)), ((61, 5), (3, 4, 2, 4)),
              ((64, 64), (4, 4, 4, 4))]


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_states_do_not_depend_on_tile_height(prec):
    """Cell values, ghost cells and the dt sequence do not depend on the
    tiling: 20 accepted steps of a dam break under the device controller
    with tile heights 2..64 and 1 or 4 waves per block (the default height
    shrinks small domains to bands of a few rows) end bit for bit equal.
    Odd heights put one-row-short bands next to the top or bottom boundary,
    where the FP32 step writes the ghost rows from the two rows next to it."""
    from paraclaws_amd import abi, ic

    p = abi.FP32 if prec == "fp32" else abi.FP64
    tkw = dict(TIME_KW, t_final=10.0)
    for (nx, ny), bc in TILE_CASES:
        _say(f"{prec} {nx}x{ny} bc {bc}: tile heights 0, 2, 3, 4, 5, 7, 16, 64")
        ref = None
        for rows, wv in ((0, 0), (2, 1), (3, 4), (4, 1), (5, 4), (7, 1), (16, 4), (64, 1)):
            with abi.Engine([abi.Domain(nx, ny, bc=bc)], precision=p, limiter=4, tile_rows=rows,
                            waves_per_block=wv) as e:
                e.set_state(0, ic.dam_break(nx, ny))
                e.fill_ghosts()
                e.time_init(**tkw)
                e.run(steps_per_graph=10, max_accepted=20)
                st = e.status(0)
                got = (st.accepted, st.rejected, st.time, e.get_state(0).view(np.uint64))
            if ref is None:
                ref = got
                assert got[0] == 20
                continue
            assert got[:3] == ref[:3], (rows, wv, got[:3], ref[:3])
            assert np.array_equal(got[3][2:-2], ref[3][2:-2]), (rows, wv)
            assert np.array_equal(got[3][:2, 2:-2], ref[3][:2, 2:-2]), (rows, wv)
            assert np.array_equal(got[3][-2:, 2:-2], ref[3][-2:, 2:-2]), (rows, wv)


def test_two_handles_on_two_threads_equal_one_thread():
    """Two handles driven from two host threads at once (each thread uploads
    its states, captures its step graphs and runs them while the other does
    the same) end in the states of the same instances run one after the
    other on one thread, bit for bit.  Each handle's transfers go through its
    own non-blocking stream, so one thread's graph capture and the other's
    copies do not meet on the legacy stream."""
    import threading

    from paraclaws_amd import abi, ic

    members = [m for m in ic.ensemble_config4()[:48] if m[0] <= 400]
    groups = [members[0::2], members[1::2]]
    _say(f"two handles of {len(groups[0])} and {len(groups[1])} dam breaks on two threads")

    def run(grp, out, slot):
        e = _ensemble_engine(abi, ic, grp)
        e.run(steps_per_graph=16)
        out[slot] = [(e.status(k).accepted, e.get_state(k)) for k in range(len(grp))]
        e.close()

    alone = [None, None]
    for g in range(2):
        run(groups[g], alone, g)
    both = [None, None]
    errs = []

    def guarded(g):
        try:
            run(groups[g], both, g)
        except Exception as exc:  # reported below, from the main thread
            errs.append(exc)

    ths = [threading.Thread(target=guarded, args=(g,)) for g in range(2)]
    for t in ths:
        t.start()
    for t in ths:
        t.join()
    assert not errs, errs
    for g in range(2):
        for (acc_a, qa), (acc_b, qb) in zip(alone[g], both[g]):
            assert acc_a == acc_b
            assert np.array_equal(qa.view(np.uint64), qb.view(np.uint64))


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_batch_in_pieces_equals_one_run(prec):
    """A batch runs in two lanes (two halves of its instances, their step
    graphs on two streams at once), its tile list rebuilt and split again as
    instances finish (in FP64 each lane's ghost-fill launch covers its own
    instances, so a lane graph is captured again when the split moves).
    Stopping it at accepted-step limits, retuning its tiles in between (the
    autotuner times both lanes at once) and going on with other graph lengths
    ends, bit for bit, where one uninterrupted run ends; two members run alone
    end there too."""
    from paraclaws_amd import abi, ic

    precision = abi.FP32 if prec == "fp32" else abi.FP64

    sizes = [(40, 40, 2.0), (70, 50, 3.5), (120, 90, 1.7), (33, 150, 6.0), (200, 64, 2.5),
             (90, 90, 4.0)]
    kw = dict(t_final=0.05, dt0=1e-5, cfl_target=0.9, dt_max=1e-2, dt_min=1e-10, growth=1.2,
              shrink=0.5, n_snapshots=2)

    def engine(sel):
        doms = [abi.Domain(sizes[k][0], sizes[k][1], bc=(4, 3, 2, 4)) for k in sel]
        e = abi.Engine(doms, precision=precision, limiter=4, tile_rows=16, waves_per_block=2)
        for j, k in enumerate(sel):
            nx, ny, r = sizes[k]
            e.set_state(j, ic.dam_break(nx, ny, h_in=r))
        e.fill_ghosts()
        e.time_init(**kw)
        return e

    def result(e, n):
        out = []
        for j in range(n):
            st = e.status(j)
            assert st.status == abi.STATUS_DONE, (j, st.status)
            out.append(((st.accepted, st.rejected, st.time), e.get_state(j)))
        return out

    everyone = list(range(len(sizes)))
    _say(f"{prec} batch of six, one run")
    with engine(everyone) as e:
        e.run(steps_per_graph=8)
        whole = result(e, len(sizes))
    _say(f"{prec} batch of six, in pieces with a retune")
    with engine(everyone) as e:
        e.run(steps_per_graph=8, max_accepted=10)
        e.autotune(steps=2)
        e.run(steps_per_graph=5, max_accepted=25)
        e.run(steps_per_graph=8)
        pieces = result(e, len(sizes))
    for k in everyone:
        assert whole[k][0] == pieces[k][0], k
        assert np.array_equal(whole[k][1].view(np.uint64), pieces[k][1].view(np.uint64)), k
    for k in (0, 4):
        _say(f"member {k} alone")
        with engine([k]) as e:
            e.run(steps_per_graph=8)
            (st, q), = result(e, 1)
        assert st == whole[k][0]
        assert np.array_equal(q.view(np.uint64), whole[k][1].view(np.uint64)), k
