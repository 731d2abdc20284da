"""Mapped grids (the reference's cartesian_stretching example and a curved
map): capacity and length-ratio arrays through the oracle and the engine."""
from __future__ import annotations

import numpy as np
import pytest

from tests.helpers import rough_state, sides_mask, stable_dt

STRETCH = lambda a, b: (4 * a, 2 * b)  # noqa: E731  x = 4 xi, y = 2 eta


def CURVED(a, b):
    return a + 0.08 * np.sin(np.pi * b), b + 0.05 * a * a


def _cfg(oracle, nx, ny, mapping, **kw):
    from paraclaws_amd import ic

    k, gx, ge, x, y = ic.mapped_metrics(nx, ny, mapping)
    cfg = oracle.Config(nx=nx, ny=ny, kappa=np.ascontiguousarray(k),
                        gamma_xi=np.ascontiguousarray(gx), gamma_eta=np.ascontiguousarray(ge), **kw)
    return cfg, (k, gx, ge, x, y)


def test_identity_metric_arrays_equal_builtin(oracle):
    """Passing the identity map's metrics as arrays changes nothing (bitwise)."""
    nx, ny = 23, 17
    ident = lambda a, b: (a, b)  # noqa: E731
    cfg_a, _ = _cfg(oracle, nx, ny, ident, limiter=3, bc=(3, 4, 2, 2))
    cfg_b = oracle.Config(nx=nx, ny=ny, limiter=3, bc=(3, 4, 2, 2))
    m = oracle.metrics(cfg_b)
    assert np.array_equal(m["kappa"], cfg_a.kappa)
    assert np.array_equal(m["gamma_xi"], cfg_a.gamma_xi)
    assert np.array_equal(m["gamma_eta"], cfg_a.gamma_eta)
    q = rough_state(nx, ny, seed=4)
    oracle.fill_ghosts(cfg_b, q)
    a, b = q.copy(), q.copy()
    dt = stable_dt(q, nx, ny)
    oracle.sweep(cfg_a, a, 1, dt)
    oracle.sweep(cfg_b, b, 1, dt)
    assert np.array_equal(a, b)


def test_stretched_metrics_closed_form():
    from paraclaws_amd import ic

    k, gx, ge, x, y = ic.mapped_metrics(20, 20, STRETCH)
    assert np.allclose(k, 8.0, rtol=1e-14)
    assert np.allclose(gx, 2.0, rtol=1e-14) and np.allclose(ge, 4.0, rtol=1e-14)


@pytest.mark.parametrize("mapping", [STRETCH, CURVED], ids=["stretch", "curved"])
def test_mapped_oracle_matches_reference_structured(oracle, mapping):
    from tests import refstruct
    from tests.helpers import GRAV

    nx, ny = 9, 8
    cfg, (k, gx, ge, _, _) = _cfg(oracle, nx, ny, mapping, limiter=4, bc=(2, 3, 4, 2))
    q = rough_state(nx, ny, seed=6)
    oracle.fill_ghosts(cfg, q)
    dt = 0.5 * stable_dt(q, nx, ny)
    a, b = q.copy(), q.copy()
    tc = oracle.TimeSettings(dt0=dt, shrink=1e-9).tc()
    assert oracle.solve(cfg, tc, a) == 1
    rc, _ = refstruct.solve(b, dt, nx, ny, GRAV, shrink=1e-9, metrics=(k, gx, ge))
    assert rc == 1
    assert np.allclose(a, b, rtol=1e-13, atol=1e-13)


@pytest.mark.gpu
@pytest.mark.parametrize("engine", ["fused", "staged"])
@pytest.mark.parametrize("mapping", [STRETCH, CURVED], ids=["stretch", "curved"])
def test_gpu_fp64_mapped_bitwise(oracle, engine, mapping):
    from paraclaws_amd import abi

    nx, ny = 75, 66
    cfg, (k, gx, ge, _, _) = _cfg(oracle, nx, ny, mapping, limiter=4, bc=(4, 3, 2, 4))
    q = rough_state(nx, ny, seed=8)
    oracle.fill_ghosts(cfg, q)
    dt = 0.5 * stable_dt(q, nx, ny)
    ref = q.copy()
    tc = oracle.TimeSettings(dt0=dt, shrink=1e-9).tc()
    assert oracle.solve(cfg, tc, ref) == 1
    oracle.fill_ghosts(cfg, ref)
    print(f"gpu fp64 mapped {engine}", flush=True)
    with abi.Engine([abi.Domain(nx, ny, bc=cfg.bc)], precision=abi.FP64,
                    engine=abi.FUSED if engine == "fused" else abi.STAGED, limiter=4) as e:
        e.set_state(0, q)
        e.set_metrics(0, k, gx, ge)
        smax = e.step(dt)[0]
        e.accept()
        got = e.get_state(0)
    m = sides_mask(nx, ny)
    assert np.array_equal(got[m], ref[m])
    assert max(smax[0] * dt / (2 / nx), smax[1] * dt / (2 / ny)) == tc.CFL


@pytest.mark.gpu
def test_gpu_mixed_batch_mapped_and_plain(oracle):
    """One handle with a mapped instance, an identity-map instance whose
    metrics the FP64 kernel computes and one whose metrics are exactly 1
    (three kernel variants per FP64 step, two per FP32 step) equals each
    instance run alone."""
    from paraclaws_amd import abi

    n = 48
    cfg_m, (k, gx, ge, _, _) = _cfg(oracle, n, n, STRETCH, limiter=4, bc=(4, 4, 4, 4))
    qa = oracle.dam_break(oracle.Config(nx=n, ny=n))
    qb = oracle.dam_break(oracle.Config(nx=n, ny=n), h_in=3.0)
    qc = oracle.dam_break(oracle.Config(nx=64, ny=64), h_in=2.5)
    sizes = [n, n, 64]
    tkw = dict(t_final=0.05, dt0=1e-4, n_snapshots=2)
    for prec in (abi.FP64, abi.FP32):
        print(f"gpu mixed batch precision={prec}", flush=True)
        with abi.Engine([abi.Domain(s_, s_, bc=(4, 4, 4, 4)) for s_ in sizes],
                        precision=prec, limiter=4) as e:
            e.set_state(0, qa)
            e.set_state(1, qb)
            e.set_state(2, qc)
            e.set_metrics(0, k, gx, ge)
            e.fill_ghosts()
            e.time_init(**tkw)
            e.run(steps_per_graph=8)
            every = [e.get_state(i) for i in range(3)]
        for idx, (q0, met) in enumerate(((qa, True), (qb, False), (qc, False))):
            with abi.Engine([abi.Domain(sizes[idx], sizes[idx], bc=(4, 4, 4, 4))],
                            precision=prec, limiter=4) as e1:
                e1.set_state(0, q0)
                if met:
                    e1.set_metrics(0, k, gx, ge)
                e1.fill_ghosts()
                e1.time_init(**tkw)
                e1.run(steps_per_graph=8)
                solo = e1.get_state(0)
            assert np.array_equal(every[idx], solo), (prec, idx)


@pytest.mark.gpu
@pytest.mark.parametrize("nx,ny", [(64, 32), (37, 29)], ids=["unit-metrics", "computed-metrics"])
def test_gpu_fp64_metric_kinds_agree(oracle, nx, ny):
    """The FP64 fused kernel on the identity map -- with the metrics skipped
    (all exactly 1, 64 x 32) or computed in the kernel (37 x 29) -- equals the
    same grid run with the identity map's metrics passed as arrays, bit for
    bit, over a device-controlled run."""
    from paraclaws_amd import abi

    cfg = oracle.Config(nx=nx, ny=ny, limiter=4, bc=(4, 3, 2, 4))
    m = oracle.metrics(cfg)
    unit = all(bool(np.all(m[k] == 1.0)) for k in ("kappa", "gamma_xi", "gamma_eta"))
    assert unit == ((nx, ny) == (64, 32))
    q0 = rough_state(nx, ny, seed=nx + ny)
    oracle.fill_ghosts(cfg, q0)
    tkw = dict(t_final=0.02, dt0=1e-4, n_snapshots=2)
    out = []
    for arrays in (False, True):
        print(f"fp64 metric kinds {nx}x{ny} arrays={arrays}", flush=True)
        with abi.Engine([abi.Domain(nx, ny, bc=cfg.bc)], precision=abi.FP64, limiter=4) as e:
            e.set_state(0, q0)
            if arrays:
                e.set_metrics(0, m["kappa"], m["gamma_xi"], m["gamma_eta"])
            e.time_init(**tkw)
            e.run(steps_per_graph=8)
            st = e.status(0)
            out.append((e.get_state(0), st.accepted, st.rejected))
    assert out[0][1:] == out[1][1:]
    msk = sides_mask(nx, ny)
    assert np.array_equal(out[0][0][msk].view(np.uint64), out[1][0][msk].view(np.uint64))


@pytest.mark.gpu
@pytest.mark.parametrize("mapping", [STRETCH, CURVED], ids=["stretch", "curved"])
def test_gpu_fp32_mapped_close(oracle, mapping):
    """FP32 fused step with metric arrays (FP32 frames in 64-element slot rows,
    metric planes in the same rows) against the FP64 oracle: interior and
    ghost cells within 4e-6 of the field scale, CFL within 1e-6 relative.
    nx = 175 gives two strips that fill their slots and one shifted to end at
    column nx."""
    from paraclaws_amd import abi

    nx, ny = 175, 66
    cfg, (k, gx, ge, _, _) = _cfg(oracle, nx, ny, mapping, limiter=4, bc=(4, 3, 2, 4))
    q = rough_state(nx, ny, seed=9)
    oracle.fill_ghosts(cfg, q)
    dt = 0.5 * stable_dt(q, nx, ny)
    ref = q.copy()
    tc = oracle.TimeSettings(dt0=dt, shrink=1e-9).tc()
    assert oracle.solve(cfg, tc, ref) == 1
    oracle.fill_ghosts(cfg, ref)
    print("gpu fp32 mapped", flush=True)
    with abi.Engine([abi.Domain(nx, ny, bc=cfg.bc)], precision=abi.FP32, limiter=4) as e:
        e.set_state(0, q)
        e.set_metrics(0, k, gx, ge)
        smax = e.step(dt)[0]
        e.accept()
        got = e.get_state(0)
    m = sides_mask(nx, ny)
    for c in range(3):
        d = np.abs(got[..., c][m] - ref[..., c][m])
        assert d.max() <= 4e-6 * max(1.0, np.abs(ref[..., c][m]).max()), (c, d.max())
    cfl = max(smax[0] * dt / (2 / nx), smax[1] * dt / (2 / ny))
    assert abs(cfl - tc.CFL) <= 1e-6 * tc.CFL


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["fp64-fused", "fp64-staged", "fp32"])
def test_gpu_mapped_with_bathymetry(oracle, prec):
    """A mapped grid (the stretched map) with the bathymetry source and the
    wet/dry cut-off: one step against the oracle, bit for bit in FP64 (both
    engines), within 4e-6 of the field scale in FP32 (the kernel variant with
    metric arrays and source terms)."""
    from paraclaws_amd import abi, ic

    nx, ny = 130, 70
    cfg, (k, gx, ge, _, _) = _cfg(oracle, nx, ny, STRETCH, limiter=4, bc=(2, 3, 4, 2),
                                  source=1, dry_tol=1e-6)
    b = ic.gaussian_bathymetry(nx, ny, seed=6)
    q = ic.still_water_with_step(b, nx, dry_tol=1e-6)
    q[..., 0] = np.maximum(q[..., 0], 0.0)
    oracle.fill_ghosts(cfg, q)
    h = q[2:-2, 2:-2, 0]
    dt = 0.3 * (2.0 / max(nx, ny)) / float(np.sqrt(9.81 * h.max()) * 4.0)
    ref = q.copy()
    tc = oracle.TimeSettings(dt0=dt, shrink=1e-9).tc()
    assert oracle.solve(cfg, tc, ref) == 1
    oracle.source(cfg, ref, np.ascontiguousarray(b), dt)
    oracle.fill_ghosts(cfg, ref)
    precision = abi.FP32 if prec == "fp32" else abi.FP64
    engine = abi.STAGED if prec == "fp64-staged" else abi.FUSED
    print(f"gpu mapped + bathymetry {prec}", flush=True)
    with abi.Engine([abi.Domain(nx, ny, bc=cfg.bc, source=1)], precision=precision,
                    engine=engine, limiter=4, dry_tol=1e-6) as e:
        e.set_state(0, q)
        e.set_metrics(0, k, gx, ge)
        e.set_bathymetry(0, b)
        e.step(dt)
        e.accept()
        got = e.get_state(0)
    m = sides_mask(nx, ny)
    if precision == abi.FP64:
        assert np.array_equal(got[m].view(np.uint64), ref[m].view(np.uint64))
    else:
        for c in range(3):
            d = np.abs(got[..., c][m] - ref[..., c][m])
            assert d.max() <= 4e-6 * max(1.0, np.abs(ref[..., c][m]).max()), (c, d.max())
