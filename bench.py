#!/usr/bin/env python3
"""Benchmark: cell-updates/s of the 2D shallow-water wave-propagation step on
one MI355X.

Workload (one GPU, N = 1): circular dam break on [-1,1]^2, 8192 x 8192 cells,
reflective walls, HLL + MC limiter, CFL 0.9, adaptive dt from 1e-10, 200
accepted steps.  A step is one accepted time step of the whole domain (xi
sweep, eta sweep, source, ghost fill); a cell-update is one interior cell
advanced by one accepted step.

  value      FP32 fused engine, device-controlled run (graph-captured steps),
             whole-job cell-updates/s with the state resident in HBM.
  fp64_plain the plain FP64 path (one kernel per reference stage, HBM work
             arrays) on the same workload; speedup = value / fp64_plain.
  errors     per-field relative L1/Linf of FP32 against the FP64 path after
             the same accepted steps, FP32 replaying the FP64 dt sequence.
  roofline   the fused step kernel: 24 algorithmic bytes per cell-update
             (read h,hu,hv once, write once, FP32) x cells per launch / its
             average launch time from HIP events on the engine's stream.
  cpu_baseline  the FP64 CPU restatement (OpenMP) on a bounded sample of the
             same workload, on this machine's host cores.

Prints one JSON line on stdout; progress goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BYTES_PER_CELL_UPDATE_F32 = 24.0
BYTES_PER_CELL_UPDATE_F64 = 48.0
# the FP32 fused step kernel of this workload (MC limiter, second order, no
# metric arrays, no source terms)
KERNEL_F32 = "pcw::f32::k_step<4, true, false, 0>"
METRIC = "cell-updates/s/GPU (FP32 batched vs FP64 plain) at stated error; % HBM peak"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def sync():
    try:
        import torch

        if torch.cuda.is_available():
            torch.cuda.synchronize()
    except Exception:
        pass


T_ENS_BENCH = 0.1
TIME_KW = dict(t_final=0.5, dt0=1e-10, cfl_target=0.9, dt_max=1e-2, dt_min=1e-10, growth=1.2,
               shrink=0.5, n_snapshots=10)


def run_engine(abi, ic, n, precision, engine, steps, warmup, spg, autotune, log_steps=0,
               repeat=1, tiles=None, block=None):
    """Dam break n^2 from the initial state: `warmup` untimed accepted steps,
    then `repeat` timed runs of exactly `steps` accepted steps, each from the
    initial state.  The returned record is the run with the median wall time
    (walls of all runs in `walls_s`); the engine holds that run's state.
    block = (nx, ny): the lower-left nx x ny block of the n^2 dam break on
    [-1,1]^2 instead (walls on its four sides)."""
    if block:
        bx, by = block
        xi, eta = (-1.0, -1.0 + 2.0 * bx / n), (-1.0, -1.0 + 2.0 * by / n)
        dom = abi.Domain(bx, by, bc=(4, 4, 4, 4), xi=xi, eta=eta)
        q0 = ic.dam_break(bx, by, xi=xi, eta=eta)
    else:
        dom = abi.Domain(n, n, bc=(4, 4, 4, 4))
        q0 = ic.dam_break(n, n)
    e = abi.Engine([dom], precision=precision, engine=engine, limiter=4,
                   **(dict(tile_rows=tiles[0], waves_per_block=tiles[1]) if tiles else {}))

    def reset():
        e.set_state(0, q0)
        e.fill_ghosts()
        e.time_init(**TIME_KW)

    reset()
    tuned = tuple(tiles) if tiles else None
    if autotune and not tiles:
        tuned = e.autotune(steps=3)
        log(f"  autotune -> tile_rows={tuned[0]} waves_per_block={tuned[1]}")
    if warmup:
        e.run(steps_per_graph=spg, max_accepted=warmup)
    runs = []
    for k in range(repeat):
        reset()
        if log_steps:
            e.enable_log(log_steps)
        sync()
        t0 = time.perf_counter()
        e.run(steps_per_graph=spg, max_accepted=steps)
        sync()
        wall = time.perf_counter() - t0
        st = e.status(0)
        avg_ms, launches = e.kernel_stats()
        runs.append(dict(wall_s=wall, accepted=st.accepted, rejected=st.rejected, time=st.time,
                         avg_step_ms=avg_ms, graph_ms=e.last_elapsed_ms(), tuned=tuned))
    walls = [r["wall_s"] for r in runs]
    med = sorted(range(repeat), key=lambda k: walls[k])[repeat // 2]
    if med != repeat - 1:  # the engine's state and log are those of the last run
        reset()
        if log_steps:
            e.enable_log(log_steps)
        e.run(steps_per_graph=spg, max_accepted=steps)
    rec = dict(runs[med])
    rec["walls_s"] = walls
    return e, rec


def replay_fp32(abi, ic, n, dts):
    """FP32 engine stepped with a given dt sequence (host control)."""
    e = abi.Engine([abi.Domain(n, n, bc=(4, 4, 4, 4))], precision=abi.FP32, limiter=4)
    e.set_state(0, ic.dam_break(n, n))
    e.fill_ghosts()
    for dt in dts:
        e.step(dt)
        e.accept()
    return e


def run_ensemble(abi, ic, spg, tiles=None):
    """512 dam breaks of uneven sizes (64^2..1024^2, seed 4), reflective walls,
    own adaptive dt each, one step launch per graph node, FP32, to t = 0.1.
    tiles: (rows, waves per block) instead of the autotuner's."""
    members = ic.ensemble_config4()
    doms = [abi.Domain(n, n, bc=(4, 4, 4, 4)) for n, _ in members]
    e = abi.Engine(doms, precision=abi.FP32, limiter=4,
                   **(dict(tile_rows=tiles[0], waves_per_block=tiles[1]) if tiles else {}))
    for k, (n, r) in enumerate(members):
        e.set_state(k, ic.dam_break(n, n, h_in=r))
    e.fill_ghosts()
    tkw = dict(TIME_KW)
    tkw["t_final"] = T_ENS_BENCH
    e.time_init(**tkw)
    tuned = tuple(tiles) if tiles else e.autotune(steps=4)
    e.time_init(**tkw)
    sync()
    t0 = time.perf_counter()
    e.run(steps_per_graph=spg)
    sync()
    wall = time.perf_counter() - t0
    upd = 0
    acc_max = 0
    for k, (n, _) in enumerate(members):
        st = e.status(k)
        assert st.status == abi.STATUS_DONE, (k, st.status)
        upd += st.accepted * n * n
        acc_max = max(acc_max, st.accepted + st.rejected)
    e.close()
    value = upd / wall
    return dict(value=value, unit="cell-updates/s", instances=len(members),
                cells=int(sum(n * n for n, _ in members)), cell_updates=int(upd), wall_s=wall,
                max_attempts=acc_max, tuned=tuned,
                roofline_frac_wall_24B=value * BYTES_PER_CELL_UPDATE_F32 / 1e9 / HBM_PEAK_GBS)


def run_bathymetry(abi, ic, n, steps, spg, tiles=None, warmup=10):
    """Config 3: FP32 n x n over a seeded Gaussian-sum bathymetry (seed 3),
    still water h + b = 1 with dry cells (tol 1e-6), dam-break step h += 1 on
    x < 0, outflow (zero-gradient) BCs; FP64 mass total before and after.
    tiles: (rows, waves per block) instead of the autotuner's."""
    b = ic.gaussian_bathymetry(n, n, seed=3)
    q = ic.still_water_with_step(b, n, dry_tol=1e-6)
    dom = abi.Domain(n, n, bc=(2, 2, 2, 2), source=abi.SRC_BATHY)
    e = abi.Engine([dom], precision=abi.FP32, limiter=4, dry_tol=1e-6,
                   **(dict(tile_rows=tiles[0], waves_per_block=tiles[1]) if tiles else {}))
    e.set_state(0, q)
    e.set_bathymetry(0, b)
    e.fill_ghosts()
    del q, b
    m0 = e.totals(0)[0]
    e.time_init(**TIME_KW)
    tuned = tuple(tiles) if tiles else e.autotune(steps=3)
    if warmup:
        e.run(steps_per_graph=spg, max_accepted=warmup)
    e.set_state(0, ic.still_water_with_step(ic.gaussian_bathymetry(n, n, seed=3), n, dry_tol=1e-6))
    e.fill_ghosts()
    e.time_init(**TIME_KW)
    sync()
    t0 = time.perf_counter()
    e.run(steps_per_graph=spg, max_accepted=steps)
    sync()
    wall = time.perf_counter() - t0
    st = e.status(0)
    m1 = e.totals(0)[0]
    avg_ms, _ = e.kernel_stats()
    e.close()
    cells = n * n
    return dict(value=st.accepted * cells / wall, unit="cell-updates/s", steps=st.accepted,
                rejected=st.rejected, kernel_avg_ms=avg_ms,
                roofline_frac_28B=(28.0 * cells / (avg_ms * 1e-3) / 1e9 / HBM_PEAK_GBS) if avg_ms else None,
                mass_fp64_before=m0 * (2.0 / n) ** 2, mass_fp64_after=m1 * (2.0 / n) ** 2,
                tuned=tuned)


def fp32_errors_vs_fp64(abi, ic, dom, q, time_kw, engine_kw, bathy=None, max_accepted=0, spg=32):
    """One domain's FP32 error: the FP64 fused engine (the reference's
    arithmetic order) runs under its device controller, logging every attempt;
    the FP32 engine replays the accepted dt sequence from the same inputs.  Two
    floors the scheme itself sets, both FP64 replays of the same dt sequence:
    from the inputs rounded once to FP32 (state and bathymetry), and from a
    start whose depth carries a relative perturbation of 2^-24 (one FP32
    rounding) with a random sign per cell, which excites every mode even
    where the inputs are exact in FP32 (a dam break's two depths).  Per-field
    relative L1/Linf against the FP64 state.  At CFL 0.9 the reference's
    update amplifies perturbations (above CFL 0.5 it is not von Neumann
    stable), so one rounding can grow far above FP32 resolution; a perturbed
    replay that the engine rejects (the perturbation grew into a negative
    depth) is reported with the engine's message."""
    with abi.Engine([dom], precision=abi.FP64, **engine_kw) as e:
        e.set_state(0, q)
        if bathy is not None:
            e.set_bathymetry(0, bathy)
        e.fill_ghosts()
        e.time_init(**time_kw)
        e.enable_log(200000)
        e.run(steps_per_graph=spg, max_accepted=max_accepted)
        dts = [r[2] for r in e.read_log(0) if r[0] == 1]
        st = e.status(0)
        assert len(dts) == st.accepted
        ref = e.get_state(0)

    def replay(prec, qq, bb):
        with abi.Engine([dom], precision=prec, **engine_kw) as e:
            e.set_state(0, qq)
            if bb is not None:
                e.set_bathymetry(0, bb)
            e.fill_ghosts()
            for dt in dts:
                e.step(dt)
                e.accept()
            return e.get_state(0)

    r32 = lambda a: None if a is None else a.astype(np.float32).astype(np.float64)
    qp = q.copy()
    rng = np.random.default_rng(1)
    qp[..., 0] *= 1.0 + rng.choice([-1.0, 1.0], size=qp.shape[:2]) * 2.0 ** -24
    try:
        perturbed = ic.field_errors(replay(abi.FP64, qp, bathy), ref)
    except abi.PcwError as err:
        perturbed = f"FP64 replay stopped: {err}"
    return dict(errors=ic.field_errors(replay(abi.FP32, q, bathy), ref),
                fp64_from_fp32_rounded_inputs=ic.field_errors(replay(abi.FP64, r32(q), r32(bathy)), ref),
                fp64_from_2pow24_perturbed_depth=perturbed,
                accepted_steps=len(dts), time=st.time,
                reference="FP64 fused engine on the same inputs",
                fp32="replay of the reference's accepted dt sequence")


def bathymetry_errors(abi, ic, n, steps, spg):
    """Config 3's FP32 error and its floor (fp32_errors_vs_fp64) over `steps`
    accepted steps."""
    b = ic.gaussian_bathymetry(n, n, seed=3)
    q = ic.still_water_with_step(b, n, dry_tol=1e-6)
    dom = abi.Domain(n, n, bc=(2, 2, 2, 2), source=abi.SRC_BATHY)
    return fp32_errors_vs_fp64(abi, ic, dom, q, TIME_KW, dict(limiter=4, dry_tol=1e-6), bathy=b,
                               max_accepted=steps, spg=spg)


def ensemble_member_errors(abi, ic):
    """Config 4's FP32 error on the smallest, the median and the largest
    member run alone to t = 0.1 (a member alone equals the batch bit for bit),
    at the ensemble's CFL 0.9, and the largest member again at CFL 0.45, where
    the update is stable."""
    members = ic.ensemble_config4()
    order = sorted(range(len(members)), key=lambda k: members[k][0])
    out = []
    for k, cfl in ((order[0], 0.9), (order[len(order) // 2], 0.9), (order[-1], 0.9), (order[-1], 0.45)):
        n, r = members[k]
        rec = fp32_errors_vs_fp64(abi, ic, abi.Domain(n, n, bc=(4, 4, 4, 4)),
                                  ic.dam_break(n, n, h_in=r),
                                  dict(TIME_KW, t_final=T_ENS_BENCH, cfl_target=cfl), dict(limiter=4))
        rec.update(member=k, n=n, depth_ratio=r, cfl_target=cfl)
        out.append(rec)
    return out


def friction_member_errors(abi, ic, t_final=2.0):
    """Config 5's FP32 error on its first member (2048 x 1024, Manning n and
    bathymetry of seed 5000) run alone to t_final."""
    g = ic.SplitMix64(5)
    nman = g.uniform(0.01, 0.05)
    nx, ny = 2048, 1024
    dom = abi.Domain(nx, ny, bc=(2, 2, 4, 4), xi=(-2.0, 2.0), eta=(-1.0, 1.0),
                     source=abi.SRC_BATHY_FRICTION, manning_n=nman)
    b = ic.gaussian_bathymetry(nx, ny, seed=5000, xi=(-2.0, 2.0), eta=(-1.0, 1.0))
    q = ic.still_water_with_step(b, nx, xi=(-2.0, 2.0), dry_tol=1e-6)
    rec = fp32_errors_vs_fp64(abi, ic, dom, q, dict(TIME_KW, t_final=t_final),
                              dict(limiter=2, dry_tol=1e-6), bathy=b)
    rec.update(member=0, manning_n=nman)
    return rec


def run_friction_batch(abi, ic, spg, t_final=2.0, n_inst=64, tiles=None):
    """Config 5: 64 instances of 2048 x 1024 on [-2,2]x[-1,1], Manning n
    uniform in [0.01, 0.05] and Gaussian-sum bathymetry per instance (seed 5),
    superbee, outflow in x, walls in y, FP32 state, FP64 totals, to t = 2.0.
    tiles: (rows, waves per block) instead of the autotuner's."""
    g = ic.SplitMix64(5)
    nx, ny = 2048, 1024
    doms, bs, ns = [], [], []
    for k in range(n_inst):
        nman = g.uniform(0.01, 0.05)
        ns.append(nman)
        doms.append(abi.Domain(nx, ny, bc=(2, 2, 4, 4), xi=(-2.0, 2.0), eta=(-1.0, 1.0),
                               source=abi.SRC_BATHY_FRICTION, manning_n=nman))
    e = abi.Engine(doms, precision=abi.FP32, limiter=2, dry_tol=1e-6,
                   **(dict(tile_rows=tiles[0], waves_per_block=tiles[1]) if tiles else {}))
    m0 = 0.0
    for k in range(n_inst):
        b = ic.gaussian_bathymetry(nx, ny, seed=5000 + k, xi=(-2.0, 2.0), eta=(-1.0, 1.0))
        q = ic.still_water_with_step(b, nx, xi=(-2.0, 2.0), dry_tol=1e-6)
        e.set_state(k, q)
        e.set_bathymetry(k, b)
    e.fill_ghosts()
    for k in range(n_inst):
        m0 += e.totals(k)[0]
    tkw = dict(TIME_KW)
    tkw["t_final"] = t_final
    e.time_init(**tkw)
    tuned = tuple(tiles) if tiles else e.autotune(steps=4)
    e.time_init(**tkw)
    sync()
    t0 = time.perf_counter()
    e.run(steps_per_graph=spg)
    sync()
    wall = time.perf_counter() - t0
    upd = 0
    steps = []
    m1 = 0.0
    for k in range(n_inst):
        st = e.status(k)
        upd += st.accepted * nx * ny
        steps.append(st.accepted)
        m1 += e.totals(k)[0]
    e.close()
    cell = (4.0 / nx) * (2.0 / ny)
    return dict(value=upd / wall, unit="cell-updates/s", instances=n_inst, wall_s=wall,
                roofline_frac_wall_28B=upd / wall * 28.0 / 1e9 / HBM_PEAK_GBS,
                steps_min=min(steps), steps_max=max(steps), mass_fp64_before=m0 * cell,
                mass_fp64_after=m1 * cell, tuned=tuned)


def cpu_baseline(n, target_s=15.0):
    """FP64 CPU restatement (OpenMP) on the same dam break; as many steps from
    the initial state as fit in about target_s seconds (at least one)."""
    from oracle import oracle as O

    cores = int(os.environ.get("PCW_CPU_THREADS", "0")) or min(os.cpu_count() or 1, 16)
    cfg = O.Config(nx=n, ny=n, limiter=4, bc=(4, 4, 4, 4), nthreads=cores)
    ts = O.TimeSettings(**{"t_final": 0.5, "dt0": 1e-10, "cfl_target": 0.9, "dt_max": 1e-2,
                           "dt_min": 1e-10, "growth": 1.2, "shrink": 0.5, "n_snap": 10})
    q = O.dam_break(cfg)
    t0 = time.perf_counter()
    recs, tc = O.run(cfg, ts, q, max_accepted=3)
    one = (time.perf_counter() - t0) / 3
    steps = max(1, min(int(target_s / max(one, 1e-3)), 200))
    q = O.dam_break(cfg)
    t0 = time.perf_counter()
    recs, tc = O.run(cfg, ts, q, max_accepted=steps)
    el = time.perf_counter() - t0
    acc = sum(r[0] for r in recs)
    return dict(value=acc * n * n / el, unit="cell-updates/s", cores=cores, kind="port",
                sample=f"dam break {n}x{n}, first {acc} accepted steps from the initial state, "
                       f"FP64 CPU restatement with OpenMP ({el:.1f} s)",
                reference_structured=cpu_reference_structured(cores))


def cpu_reference_structured(cores, n=256, steps=10):
    """The reference's CPU cost structure: the restatement with LAPACK dgeev
    (workspace query, right eigenvectors) for every cell of every interface and
    BLAS dgemv for J*q, as SMOD_approximate_Riemann_solvers.f90:60-101 calls
    them (scipy's LAPACK, one BLAS thread as MKL_NUM_THREADS=1, OpenMP over the
    same threads), on the first `steps` accepted steps of config 1 (256^2);
    beside it the closed-form restatement on the same steps and the largest
    relative difference of their states."""
    from threadpoolctl import threadpool_limits

    from oracle import oracle as O

    cfg = O.Config(nx=n, ny=n, limiter=4, bc=(4, 4, 4, 4), nthreads=cores)
    ts = O.TimeSettings(t_final=0.5, dt0=1e-10, cfl_target=0.9, dt_max=1e-2, dt_min=1e-10,
                        growth=1.2, shrink=0.5, n_snap=10)
    q_ref = O.dam_break(cfg)
    with threadpool_limits(1):
        O.use_lapack(True)
        try:
            t0 = time.perf_counter()
            recs, _ = O.run(cfg, ts, q_ref, max_accepted=steps)
            el = time.perf_counter() - t0
        finally:
            O.use_lapack(False)
    q_cf = O.dam_break(cfg)
    t0 = time.perf_counter()
    O.run(cfg, ts, q_cf, max_accepted=steps)
    el_cf = time.perf_counter() - t0
    acc = sum(r[0] for r in recs)
    inner = (slice(2, -2), slice(2, -2))
    rel = float((np.abs(q_ref[inner] - q_cf[inner]).max(axis=(0, 1))
                 / np.abs(q_cf[inner]).max(axis=(0, 1))).max())
    return dict(value=acc * n * n / el, unit="cell-updates/s", cores=cores, kind="port",
                sample=f"dam break {n}x{n}, first {acc} accepted steps, per-interface LAPACK dgeev "
                       f"+ BLAS dgemv ({el:.2f} s)",
                closed_form_same_steps=acc * n * n / el_cf, max_rel_diff_vs_closed_form=rel)


def cpu_config1(abi, ic):
    """Config 1, the reference's own CPU case (256^2 dam break, MC, walls,
    CFL 0.9, to t = 0.5): the FP64 CPU restatement (OpenMP) timed over the
    whole run, and beside it the GPU engines over the same run, the FP64
    fused engine's final state compared bit for bit with the restatement's."""
    from oracle import oracle as O

    n = 256
    cores = int(os.environ.get("PCW_CPU_THREADS", "0")) or min(os.cpu_count() or 1, 16)
    cfg = O.Config(nx=n, ny=n, limiter=4, bc=(4, 4, 4, 4), nthreads=cores)
    ts = O.TimeSettings(t_final=0.5, dt0=1e-10, cfl_target=0.9, dt_max=1e-2, dt_min=1e-10,
                        growth=1.2, shrink=0.5, n_snap=10)
    q = O.dam_break(cfg)
    q0 = q.copy()
    t0 = time.perf_counter()
    recs, tc = O.run(cfg, ts, q)
    el = time.perf_counter() - t0
    acc = sum(r[0] for r in recs)
    out = {"cells": n * n, "accepted": acc, "attempts": len(recs), "t_final": tc.time,
           "cpu": {"value": acc * n * n / el, "wall_s": el, "cores": cores, "kind": "port"}}
    dts = [r[2] for r in recs if r[0] == 1]
    for name, prec in (("gpu_fp64_fused", abi.FP64), ("gpu_fp32_fused", abi.FP32)):
        with abi.Engine([abi.Domain(n, n, bc=(4, 4, 4, 4))], precision=prec, limiter=4) as e:
            e.set_state(0, q0)
            e.fill_ghosts()
            e.time_init(**TIME_KW)
            tuned = e.autotune(steps=3)
            walls = []
            for _ in range(3):
                e.set_state(0, q0)
                e.fill_ghosts()
                e.time_init(**TIME_KW)
                sync()
                t0 = time.perf_counter()
                e.run(steps_per_graph=64)
                sync()
                walls.append(time.perf_counter() - t0)
            st = e.status(0)
            got = e.get_state(0)
        rec = {"value": st.accepted * n * n / min(walls), "wall_s": min(walls),
               "accepted": st.accepted, "tuned": tuned}
        if prec == abi.FP64:
            rec["equals_cpu_bitwise"] = bool(np.array_equal(got[2:-2, 2:-2], q[2:-2, 2:-2]))
        else:
            # the CPU run's accepted dt sequence replayed (the same times; a
            # free FP32 run takes its own dt sequence)
            with abi.Engine([abi.Domain(n, n, bc=(4, 4, 4, 4))], precision=prec, limiter=4) as e:
                e.set_state(0, q0)
                e.fill_ghosts()
                for dt in dts:
                    e.step(dt)
                    e.accept()
                rec["errors_vs_fp64_same_dts"] = ic.field_errors(e.get_state(0), q)
        out[name] = rec
    out["fortran_host"] = fortran_host_config1(q)
    return out


FORTRAN_EXAMPLE = os.path.join(ROOT, "paraclaws_amd", "fortran", "examples", "dam_break")


def fortran_host_config1(q_cpu):
    """Config 1 through the Fortran host program (the drop-in: control.in of
    examples/dam_break, 10 snapshots written in the reference's formats), FP64
    fused engine, with the time loop on the host and on the device
    (custom.in line 3); the wall time of the time loop as the program reports
    it, and its final state against the CPU restatement's."""
    import re
    import subprocess
    import tempfile

    exe = os.path.join(FORTRAN_EXAMPLE, "build", "paraclaws_gpu")
    if not os.path.exists(exe):
        return {"skipped": "Fortran host not built"}
    out = {}
    for loop, name in ((0, "host_loop"), (1, "device_loop")):
        with tempfile.TemporaryDirectory() as d:
            with open(os.path.join(FORTRAN_EXAMPLE, "control.in")) as f:
                control = f.read()
            with open(os.path.join(d, "control.in"), "w") as f:
                f.write(control)
            with open(os.path.join(d, "custom.in"), "w") as f:
                f.write(f"0\n0\n{loop}\n")
            r = subprocess.run([exe], cwd=d, capture_output=True, text=True, timeout=600)
            m = re.search(r"accepted steps:\s*(\d+)\s+rejected:\s*(\d+)\s+wall:\s*([0-9.]+)", r.stdout)
            if r.returncode != 0 or not m:
                out[name] = {"failed": (r.stdout + r.stderr)[-500:]}
                continue
            got = np.fromfile(os.path.join(d, "output", "final_state.bin"), dtype=np.float64)
            got = got.reshape(q_cpu.shape)
            out[name] = {"wall_s": float(m.group(3)), "accepted": int(m.group(1)),
                         "rejected": int(m.group(2)), "snapshots_written": 10,
                         "equals_cpu_bitwise": bool(np.array_equal(got[2:-2, 2:-2], q_cpu[2:-2, 2:-2]))}
    return out


def run_mapped(abi, ic, n, steps, spg):
    """The reference's stretched example at n x n (x = 4 xi, y = 2 eta on
    [-1,1]^2: kappa = 8, gamma_xi = 2, gamma_eta = 4, given as arrays), dam
    break of radius 0.5 in physical space, walls, MC, FP32 metric-array kernel:
    36 algorithmic bytes per cell-update (state in and out, three metric
    planes in)."""
    dom = abi.Domain(n, n, bc=(4, 4, 4, 4))
    e = abi.Engine([dom], precision=abi.FP32, limiter=4)
    c = ic.centres(n, -1.0, 1.0)  # frame columns -1 .. n+2
    r = np.hypot(4.0 * c[None, :], 2.0 * c[:, None])
    q = np.zeros((n + 4, n + 4, 3))
    q[..., 0] = np.where(r < 0.5, 2.0, 1.0)
    del r
    shape = (n + 4, n + 4)
    e.set_metrics(0, np.full(shape, 8.0), np.full(shape, 2.0), np.full(shape, 4.0))
    e.set_state(0, q)
    e.fill_ghosts()
    e.time_init(**TIME_KW)
    tuned = e.autotune(steps=3)
    e.set_state(0, q)
    e.fill_ghosts()
    e.time_init(**TIME_KW)
    e.run(steps_per_graph=spg, max_accepted=10)
    e.set_state(0, q)
    e.fill_ghosts()
    e.time_init(**TIME_KW)
    del q
    sync()
    t0 = time.perf_counter()
    e.run(steps_per_graph=spg, max_accepted=steps)
    sync()
    wall = time.perf_counter() - t0
    st = e.status(0)
    avg_ms, _ = e.kernel_stats()
    e.close()
    cells = n * n
    return dict(value=st.accepted * cells / wall, unit="cell-updates/s", steps=st.accepted,
                rejected=st.rejected, kernel_avg_ms=avg_ms,
                roofline_frac_36B=(36.0 * cells / (avg_ms * 1e-3) / 1e9 / HBM_PEAK_GBS) if avg_ms else None,
                tuned=tuned)


def load_traffic(n):
    """HBM bytes per launch of the fused FP32 step from the committed rocprofv3
    FETCH_SIZE/WRITE_SIZE runs, if they exist for this grid size and kernel,
    with the tiles (rows, waves per block) they were measured at."""
    p = os.path.join(ROOT, "profiles", f"pmc_fused_step_f32_{n}.json")
    if os.path.exists(p):
        with open(p) as f:
            rec = json.load(f)
        if rec.get("kernel") == KERNEL_F32:
            return rec.get("traffic_bytes_per_launch"), [rec.get("tile_rows"), rec.get("waves_per_block")]
    return None, None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=8192)
    ap.add_argument("--steps-per-graph", type=int, default=50)
    ap.add_argument("--repeat", type=int, default=5,
                    help="timed runs of --steps steps each; the median run is reported")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--tiles", type=int, nargs=2, metavar=("ROWS", "WAVES"),
                    help="FP32 8192^2 leg: tile height and waves per block instead of the "
                         "autotuned ones (profiler runs, whose timing would mislead the autotuner)")
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--no-physics", action="store_true",
                    help="skip the bathymetry (8192^2) and friction-batch legs")
    ap.add_argument("--profile-only", action="store_true",
                    help="FP32 fused run only (for rocprofv3 kernel traces and counters)")
    ap.add_argument("--profile-bathymetry", action="store_true",
                    help="bathymetry + wet/dry (config 3) run only, for rocprofv3 traces and counters")
    ap.add_argument("--profile-fp64", action="store_true",
                    help="FP64 fused (bit-exact) run only, for rocprofv3 kernel traces")
    ap.add_argument("--profile-ensemble", action="store_true",
                    help="the 512-instance ensemble (config 4) only, for rocprofv3 kernel traces")
    ap.add_argument("--profile-friction", type=float, default=0.0, metavar="T",
                    help="the friction batch (config 5) to simulated time T only, for rocprofv3 "
                         "kernel traces")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="run a plain device copy of MIB MiB (FETCH/WRITE counter scale)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: the engine has no multi-GPU path "
              "(one domain per GPU; a sharded domain would exchange 2-row halo strips "
              "in place of the ghost fill).", file=sys.stderr)
        sys.exit(2)

    if args.copy_calibration:
        import torch

        nel = args.copy_calibration * (1 << 20) // 4
        a = torch.ones(nel, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        log(f"copied {args.copy_calibration} MiB 5 times")
        return

    from paraclaws_amd import abi, ic

    n = args.n
    cells = n * n
    if args.profile_bathymetry:
        log(f"bathymetry + wet/dry, FP32 {n}^2, {args.steps} accepted steps")
        bat = run_bathymetry(abi, ic, n, args.steps, args.steps_per_graph, tiles=args.tiles,
                             warmup=args.warmup)
        print(json.dumps({k: (float(v) if hasattr(v, "dtype") else v) for k, v in bat.items()}),
              flush=True)
        return
    if args.profile_ensemble:
        log("batched ensemble: 512 dam breaks 64^2..1024^2 (seed 4), FP32, to t = 0.1")
        print(json.dumps(run_ensemble(abi, ic, 64, tiles=args.tiles)), flush=True)
        return
    if args.profile_friction > 0:
        log(f"friction batch: 64 x 2048x1024, Manning + bathymetry, superbee, to t = "
            f"{args.profile_friction}")
        print(json.dumps(run_friction_batch(abi, ic, 64, t_final=args.profile_friction,
                                            tiles=args.tiles)), flush=True)
        return
    if args.profile_fp64:
        log(f"FP64 fused engine, dam break {n}^2, {args.steps} accepted steps")
        ef, rf = run_engine(abi, ic, n, abi.FP64, abi.FUSED, args.steps, args.warmup,
                            args.steps_per_graph, not args.no_autotune, repeat=args.repeat,
                            tiles=args.tiles)
        ef.close()
        print(json.dumps({"value": rf["accepted"] * cells / rf["wall_s"],
                          "step_avg_ms": rf["avg_step_ms"], "tuned": rf["tuned"],
                          "steps": rf["accepted"]}), flush=True)
        return
    log(f"FP32 fused engine, dam break {n}^2, {args.steps} accepted steps, "
        f"median of {args.repeat} runs")
    e32, r32 = run_engine(abi, ic, n, abi.FP32, abi.FUSED, args.steps, args.warmup,
                          args.steps_per_graph, not args.no_autotune, repeat=args.repeat,
                          tiles=args.tiles)
    value = r32["accepted"] * cells / r32["wall_s"]
    log(f"  wall {r32['wall_s'] * 1e3:.2f} ms  {value:.4e} cell-updates/s  "
        f"kernel avg {r32['avg_step_ms']:.4f} ms  rejected {r32['rejected']}")
    kernel_ms = r32["avg_step_ms"]
    achieved = BYTES_PER_CELL_UPDATE_F32 * cells / (kernel_ms * 1e-3) / 1e9 if kernel_ms > 0 else 0.0
    traffic, traffic_tiles = load_traffic(n)
    # the boundary's host transfers (start-up upload, snapshot download), timed
    # once outside the timed region: the rate of a run that includes them
    sync()
    t0 = time.perf_counter()
    qh = e32.get_state(0)
    t_get = time.perf_counter() - t0
    t0 = time.perf_counter()
    e32.set_state(0, qh)
    sync()
    t_set = time.perf_counter() - t0
    frame_gb = qh.nbytes / 1e9
    del qh
    host_transfers = {"frame_gb": frame_gb, "set_state_s": t_set, "get_state_s": t_get,
                      "value_with_one_upload_and_download":
                      r32["accepted"] * cells / (r32["wall_s"] + t_set + t_get)}
    log(f"  host transfers of the {frame_gb:.2f} GB frame: set_state {t_set * 1e3:.1f} ms, "
        f"get_state {t_get * 1e3:.1f} ms")
    e32.close()
    # what HBM sustains for the step's compulsory traffic: three planes of the
    # frame streamed in and three out, no stencil (pcw_stream_bandwidth)
    plane_bytes = 4 * 64 * ((n + 59) // 60 + 1) * (n + 4)
    att_best, att_med = abi.stream_bandwidth(plane_bytes, reps=10)
    log(f"  attainable (3-plane stream, {6 * plane_bytes / 1e9:.2f} GB per launch): "
        f"median {att_med:.0f} GB/s, best {att_best:.0f} GB/s; the step reaches "
        f"{achieved / att_med:.3f} of it")
    if args.profile_only:
        print(json.dumps({"value": value, "kernel_avg_ms": kernel_ms, "achieved_gbs": achieved,
                          "attainable_gbs": att_med,
                          "tuned": r32["tuned"], "steps": r32["accepted"]}), flush=True)
        return

    out_extra = {"timed_runs_wall_s": r32["walls_s"], "host_transfers": host_transfers}
    if n != 4096:
        log("FP32 fused engine, dam break 4096^2, same steps")
        e4, r4 = run_engine(abi, ic, 4096, abi.FP32, abi.FUSED, args.steps, args.warmup,
                            args.steps_per_graph, not args.no_autotune, repeat=args.repeat)
        e4.close()
        out_extra["fp32_4096"] = {"value": r4["accepted"] * 4096 * 4096 / r4["wall_s"],
                                  "unit": "cell-updates/s", "kernel_avg_ms": r4["avg_step_ms"],
                                  "roofline_frac_24B": BYTES_PER_CELL_UPDATE_F32 * 4096 * 4096
                                  / (r4["avg_step_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS,
                                  "tuned": r4["tuned"]}
        # one GPU's share of the n^2 domain in the block decompositions a
        # multi-GPU run would use (2: 1x2, 4: 2x2, 8: 4x2 blocks), the walls
        # standing in for the halo exchange
        blocks = {}
        for g, (bx, by) in ((2, (n, n // 2)), (4, (n // 2, n // 2)), (8, (n // 4, n // 2))):
            log(f"FP32 fused engine, one GPU's block of {n}^2 over {g} GPUs: {bx} x {by}")
            eb, rb = run_engine(abi, ic, n, abi.FP32, abi.FUSED, args.steps, args.warmup,
                                args.steps_per_graph, not args.no_autotune, repeat=args.repeat,
                                block=(bx, by))
            eb.close()
            blocks[str(g)] = {"block": [bx, by], "value": rb["accepted"] * bx * by / rb["wall_s"],
                              "unit": "cell-updates/s", "kernel_avg_ms": rb["avg_step_ms"],
                              "ms_per_step": rb["wall_s"] * 1e3 / max(1, rb["accepted"]),
                              "roofline_frac_24B": BYTES_PER_CELL_UPDATE_F32 * bx * by
                              / (rb["avg_step_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS,
                              "tuned": rb["tuned"]}
            log(f"  {blocks[str(g)]}")
        out_extra["fp32_rank_blocks"] = blocks
    if not args.no_fp64:
        log(f"FP64 plain (staged) engine, same workload")
        e64, r64 = run_engine(abi, ic, n, abi.FP64, abi.STAGED, args.steps, min(args.warmup, 5),
                              args.steps_per_graph, False)
        v64 = r64["accepted"] * cells / r64["wall_s"]
        log(f"  wall {r64['wall_s'] * 1e3:.2f} ms  {v64:.4e} cell-updates/s")
        staged = e64.get_state(0)
        e64.close()
        out_extra["fp64_plain"] = {"value": v64, "unit": "cell-updates/s",
                                   "ms_per_step": r64["wall_s"] * 1e3 / max(1, r64["accepted"]),
                                   "engine": "staged (one kernel per reference stage)"}
        out_extra["speedup_vs_fp64_plain"] = value / v64
        log(f"FP64 fused engine (bit-exact mode), same workload")
        ef, rf = run_engine(abi, ic, n, abi.FP64, abi.FUSED, args.steps, min(args.warmup, 5),
                            args.steps_per_graph, not args.no_autotune)
        vf = rf["accepted"] * cells / rf["wall_s"]
        fused = ef.get_state(0)
        # both FP64 engines follow the reference's operation order: the same
        # accepted steps give the same bits (ghost corners aside, which the
        # reference never writes)
        sides = np.ones(fused.shape[:2], dtype=bool)
        for a_ in (slice(0, 2), slice(-2, None)):
            for b_ in (slice(0, 2), slice(-2, None)):
                sides[a_, b_] = False
        same = bool(rf["accepted"] == r64["accepted"]
                    and np.array_equal(fused[sides].view(np.uint64), staged[sides].view(np.uint64)))
        del staged, fused
        log(f"  {vf:.4e} cell-updates/s, kernel {rf['avg_step_ms']:.4f} ms, equals staged bitwise: {same}")
        out_extra["fp64_fused"] = {"value": vf, "unit": "cell-updates/s",
                                   "kernel_avg_ms": rf["avg_step_ms"], "tuned": rf["tuned"],
                                   "equals_staged_bitwise": same,
                                   "roofline_frac_48B": BYTES_PER_CELL_UPDATE_F64 * cells
                                   / (rf["avg_step_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS,
                                   "fp32_speedup": value / vf}
        out_extra["free_running_time_fp32_vs_fp64"] = [r32["time"], rf["time"]]
        # error leg: at least 200 accepted steps whatever --steps is (the dt
        # ramp from 1e-10 takes about 70 of them), FP32 replaying the FP64
        # fused engine's accepted dt sequence
        err_steps = max(200, args.steps)
        log(f"FP64 fused reference run ({err_steps} accepted steps) and FP32 replay for the error")
        ef.close()
        ef, rr = run_engine(abi, ic, n, abi.FP64, abi.FUSED, err_steps, 0, args.steps_per_graph,
                            False, log_steps=4 * err_steps + 64, tiles=rf["tuned"])
        dts = [r[2] for r in ef.read_log(0) if r[0] == 1]
        ref = ef.get_state(0)
        ef.close()
        er = replay_fp32(abi, ic, n, dts)
        got = er.get_state(0)
        er.close()
        errs = ic.field_errors(got, ref)
        del got, ref
        log(f"  {len(dts)} steps to t = {rr['time']:.6e}: errors {errs}")
        out_extra["errors_fp32_vs_fp64"] = errs
        out_extra["errors_at"] = {"accepted_steps": len(dts), "time": rr["time"],
                                  "reference": "FP64 fused engine (bitwise equal to the oracle)",
                                  "fp32": "replay of the reference's accepted dt sequence"}

    if not args.no_ensemble:
        log("batched ensemble: 512 dam breaks 64^2..1024^2 (seed 4), FP32, to t = 0.1")
        ens = run_ensemble(abi, ic, 64)
        log(f"  {ens}")
        log("ensemble: FP32 error of the smallest, median and largest member against FP64")
        ens["errors_fp32_vs_fp64_members"] = ensemble_member_errors(abi, ic)
        log(f"  {ens['errors_fp32_vs_fp64_members']}")
        out_extra["batched_ensemble"] = ens

    if not args.no_physics:
        log(f"bathymetry + wet/dry, FP32 {n}^2, outflow, {args.steps} steps")
        bat = run_bathymetry(abi, ic, n, args.steps, args.steps_per_graph)
        log(f"  {bat}")
        log(f"bathymetry: FP64 fused run ({args.steps} accepted steps) and FP32 replay for the error")
        bat["errors_fp32_vs_fp64"] = bathymetry_errors(abi, ic, n, args.steps, args.steps_per_graph)
        log(f"  {bat['errors_fp32_vs_fp64']}")
        out_extra["bathymetry"] = bat
        log("friction batch: 64 x 2048x1024, Manning + bathymetry, superbee, to t = 2.0")
        fr = run_friction_batch(abi, ic, 64)
        log("friction batch: FP32 error of the first member against FP64, to t = 2.0")
        fr["errors_fp32_vs_fp64_member0"] = friction_member_errors(abi, ic)
        log(f"  {fr}")
        out_extra["friction_batch"] = fr
        log(f"stretched grid (x = 4 xi, y = 2 eta), metric arrays, FP32 {n}^2, {args.steps} steps")
        mp = run_mapped(abi, ic, n, args.steps, args.steps_per_graph)
        log(f"  {mp}")
        out_extra["mapped_metric_arrays"] = mp

    cpu = None
    if not args.no_cpu:
        log("CPU baseline (FP64 restatement, OpenMP)")
        cpu = cpu_baseline(n)
        log(f"  {cpu}")
        log("config 1 (256^2 to t = 0.5): CPU restatement and GPU engines over the whole run")
        c1 = cpu_config1(abi, ic)
        log(f"  {c1}")
        out_extra["config1_full_run"] = c1

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": r32["accepted"],
        "warmup": args.warmup,
        "ms_per_step": r32["wall_s"] * 1e3 / max(1, r32["accepted"]),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"circular dam break {n}x{n} on [-1,1]^2, reflective walls, "
                               f"HLL + MC limiter, CFL 0.9, adaptive dt from 1e-10, "
                               f"{args.steps} accepted steps, FP32 fused engine",
                   "cells": cells, "precision": "fp32 state/flux, fp64 dt/time",
                   "tile_rows_waves": r32["tuned"]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_tile_rows_waves": traffic_tiles,
                     "kernel": KERNEL_F32, "kernel_avg_ms": kernel_ms,
                     "bytes_per_cell_update": BYTES_PER_CELL_UPDATE_F32,
                     "attainable": att_med, "attainable_best": att_best,
                     "attainable_kernel": "pcw_stream_bandwidth: 3 planes in, 3 out, 16-byte "
                                          "non-temporal accesses, median of 10 launches",
                     "frac_of_attainable": achieved / att_med},
        "cpu_baseline": cpu,
    }
    line.update(out_extra)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
