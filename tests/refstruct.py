"""Reference-structured restatement of one solve_dimensional_split, used only
to pin the oracle.

It follows the reference's structure rather than the oracle's: per
interface a 3x3 Jacobian matrix (examples/2D_dam_break/cartesian/
SMOD_jakobian.f90:44-53), eigenvalues from LAPACK dgeev with right
eigenvectors requested (SMOD_eigenvalue_problem.f90:50-66, the complex
check of :72-83), fluxes from BLAS dgemv (SMOD_matrix_operations.f90:5-23),
work arrays zeroed after every sweep (MOD_work_arrays.f90:78-109), the serial
limiter with its wnorm2r carry (MOD_flux_correction.f90:54-123), and the
split_flux mask form (MOD_flux_calculation.f90:105-120).  Plain Python loops:
tiny grids only.
"""
from __future__ import annotations

import numpy as np
from scipy.linalg import blas, lapack

EPS = 1e-10


class ComplexEigen(RuntimeError):
    pass


def get_jakobi(ixy, q, g):
    h, hu, hv = q
    J = np.zeros((3, 3))
    if ixy == 1:
        J[0] = [0.0, 1.0, 0.0]
        J[1] = [-(hu ** 2 / h ** 2) + g * h, 2.0 * hu / h, 0.0]
        J[2] = [-(hu * hv / h ** 2), hv / h, hu / h]
    else:
        J[0] = [0.0, 0.0, 1.0]
        J[1] = [-(hu * hv / h ** 2), hv / h, hu / h]
        J[2] = [-(hv ** 2 / h ** 2) + g * h, 0.0, 2.0 * hv / h]
    if np.isnan(J).any():
        raise FloatingPointError("isnan in jakobi")
    return J


def eigen_real(J):
    wr, wi, vl, vr, info = lapack.dgeev(np.asfortranarray(J), compute_vl=0, compute_vr=1)
    if info != 0:
        raise RuntimeError(f"dgeev info={info}")
    if (wi > EPS).any():
        raise ComplexEigen("ABORT")
    return wr


def matvec(J, v):
    return blas.dgemv(1.0, np.asfortranarray(J), np.asarray(v, dtype=np.float64))


class Work:
    def __init__(self, nxf, nyf):
        shp = (nyf, nxf)
        self.lam = np.zeros(shp + (2,))
        self.waves = np.zeros(shp + (2, 3))  # (j, i, mw, me)
        self.Am = np.zeros(shp + (3,))
        self.Ap = np.zeros(shp + (3,))
        self.wl = np.zeros(shp + (2,))
        self.wr = np.zeros(shp + (2,))
        self.f = np.zeros(shp + (3,))

    def zero(self):
        for a in (self.lam, self.waves, self.Am, self.Ap, self.wl, self.wr, self.f):
            a[...] = 0.0


def _metrics(nx, ny, xlo, xhi, ylo, yhi):
    """kappa and length ratios as SMOD_mesh_manipulation.f90:36-173 computes them."""
    dxi, deta = (xhi - xlo) / nx, (yhi - ylo) / ny
    kap = np.zeros((ny + 4, nx + 4))
    gx = np.zeros((ny + 4, nx + 4))
    ge = np.zeros((ny + 4, nx + 4))
    for j in range(-1, ny + 3):
        for i in range(-1, nx + 3):
            xc = [xlo + (i - 1) * dxi]
            xc.append(xc[0] + dxi)
            xc += [xc[1], xc[0], xc[0]]
            yc = [ylo + (j - 1) * deta]
            yc.append(yc[0])
            yc.append(yc[1] + deta)
            yc += [yc[2], yc[0]]
            gx[j + 1, i + 1] = np.hypot(xc[3] - xc[0], yc[3] - yc[0]) / deta
            ge[j + 1, i + 1] = np.hypot(xc[1] - xc[0], yc[1] - yc[0]) / dxi
            a = 0.0
            for m in range(4):
                a = a + 0.5 * (yc[m] + yc[m + 1]) * (xc[m] - xc[m + 1])
            kap[j + 1, i + 1] = a / (dxi * deta)
    return dxi, deta, kap, gx, ge


def _phi(code, r):
    if code == 1:
        return max(0.0, min(1.0, r))
    if code == 2:
        return max(0.0, min(1.0, 2.0 * r), min(2.0, r))
    if code == 3:
        return (r + abs(r)) / (1.0 + abs(r))
    if code == 4:
        return max(0.0, min((1.0 + r) / 2.0, 2.0, 2.0 * r))
    raise ValueError("invalid limiter choice => ABORTING")


def _sweep(q, ixy, dt, nx, ny, met, work, g, ho, use_lim, lim, cfl_local):
    dxi, deta, kap, gx, ge = met
    Q = lambda i, j: q[j + 1, i + 1]  # noqa: E731  (view of vars(:,i,j))
    W = work
    ix = lambda i, j: (j + 1, i + 1)  # noqa: E731
    d = dxi if ixy == 1 else deta
    if ixy == 1:
        rng = [(i, j) for j in range(1, ny + 1) for i in range(0, nx + 2)]
    else:
        rng = [(i, j) for i in range(1, nx + 1) for j in range(0, ny + 2)]
    for (i, j) in rng:
        L = (i - 1, j) if ixy == 1 else (i, j - 1)
        qL, qR = Q(*L).copy(), Q(i, j).copy()
        JL, JR = get_jakobi(ixy, qL, g), get_jakobi(ixy, qR, g)
        lL, lR = eigen_real(JL), eigen_real(JR)
        s2 = np.max(np.abs(np.concatenate([lL, lR])))
        s1 = -s2
        W.lam[ix(i, j)] = [s1, s2]
        if abs(s1) > EPS and abs(s2) > EPS:
            fL, fR = matvec(JL, qL), matvec(JR, qR)
            qs = (fR - fL - s2 * qR + s1 * qL) / (EPS + s1 - s2)
            W.waves[ix(i, j)][0] = qs - qL
            W.waves[ix(i, j)][1] = qR - qs
        else:
            W.lam[ix(i, j)] = 0.0
            W.waves[ix(i, j)] = 0.0
        for mw in range(2):
            cfl_local = max(cfl_local, W.lam[ix(i, j)][mw] * dt / d)
    # split_flux
    for j in range(0, ny + 2):
        for i in range(0, nx + 2):
            for mw in range(2):
                s = W.lam[ix(i, j)][mw]
                for me in range(3):
                    w = W.waves[ix(i, j)][mw][me]
                    mask = 1.0 if s < 0.0 else 0.0
                    W.Am[ix(i, j)][me] = W.Am[ix(i, j)][me] + mask * s * w
                    W.Ap[ix(i, j)][me] = W.Ap[ix(i, j)][me] + (1.0 - mask) * s * w
    dtod = dt / d
    for j in range(1, ny + 1):
        for i in range(1, nx + 1):
            if ixy == 1:
                gl, gr, R = gx[ix(i, j)], gx[ix(i + 1, j)], ix(i + 1, j)
            else:
                gl, gr, R = ge[ix(i, j)], ge[ix(i, j + 1)], ix(i, j + 1)
            for me in range(3):
                q[ix(i, j)][me] = q[ix(i, j)][me] - dtod / kap[ix(i, j)] * (
                    gl * W.Ap[ix(i, j)][me] + gr * W.Am[R][me])
    if ho:
        if use_lim:
            if ixy == 1:
                for j in range(1, ny + 1):
                    for mw in range(2):
                        for me in range(3):
                            W.wr[ix(0, j)][mw] += W.waves[ix(0, j)][mw][me] * W.waves[ix(1, j)][mw][me]
            else:
                for i in range(1, nx + 1):
                    for mw in range(2):
                        for me in range(3):
                            W.wr[ix(i, 0)][mw] += W.waves[ix(i, 0)][mw][me] * W.waves[ix(i, 1)][mw][me]
            for j in range(1, ny + 1):
                for i in range(1, nx + 1):
                    for mw in range(2):
                        W.wl[ix(i, j)][mw] = W.wr[ix(i - 1, j)][mw] if ixy == 1 else W.wr[ix(i, j - 1)][mw]
                        W.wr[ix(i, j)][mw] = 0.0
                        nr = ix(i + 1, j) if ixy == 1 else ix(i, j + 1)
                        wn2 = 0.0
                        for me in range(3):
                            wn2 = wn2 + W.waves[ix(i, j)][mw][me] ** 2
                            W.wr[ix(i, j)][mw] += W.waves[ix(i, j)][mw][me] * W.waves[nr][mw][me]
                        up = W.wl[ix(i, j)][mw] if W.lam[ix(i, j)][mw] > 0.0 else W.wr[ix(i, j)][mw]
                        if not abs(wn2) <= EPS:
                            phi = _phi(lim, up / wn2)
                            W.waves[ix(i, j)][mw] = phi * W.waves[ix(i, j)][mw]
        if ixy == 1:
            rngf = [(i, j, ix(i - 1, j)) for j in range(1, ny + 1) for i in range(0, nx + 2)]
        else:
            rngf = [(i, j, ix(i, j - 1)) for j in range(0, ny + 2) for i in range(1, nx + 1)]
        for (i, j, P) in rngf:
            for mw in range(2):
                al = abs(W.lam[ix(i, j)][mw])
                for me in range(3):
                    W.f[ix(i, j)][me] = W.f[ix(i, j)][me] + 0.5 * al * (
                        1.0 - al * dtod / (0.5 * (kap[P] + kap[ix(i, j)]))) * W.waves[ix(i, j)][mw][me]
        for j in range(1, ny + 1):
            for i in range(1, nx + 1):
                R = ix(i + 1, j) if ixy == 1 else ix(i, j + 1)
                for me in range(3):
                    q[ix(i, j)][me] = q[ix(i, j)][me] - dtod / kap[ix(i, j)] * (W.f[R][me] - W.f[ix(i, j)][me])
    W.zero()
    return cfl_local


def sweep(q, ixy, dt, nx, ny, g, ho=True, use_lim=True, lim=4, bounds=(-1.0, 1.0, -1.0, 1.0)):
    met = _metrics(nx, ny, *bounds)
    work = Work(nx + 4, ny + 4)
    return _sweep(q, ixy, dt, nx, ny, met, work, g, ho, use_lim, lim, 0.0)


def solve(q, dt, nx, ny, g, cfl_target=0.9, shrink=0.5, **kw):
    """Both sweeps with the validity test between them; returns (ierr, CFL).
    metrics=(kappa, gamma_xi, gamma_eta) frames replace the identity map's."""
    bounds = kw.pop("bounds", (-1.0, 1.0, -1.0, 1.0))
    met = _metrics(nx, ny, *bounds)
    if kw.get("metrics") is not None:
        k, gx, ge = kw["metrics"]
        met = (met[0], met[1], k, gx, ge)
    work = Work(nx + 4, ny + 4)
    ho, use_lim, lim = kw.get("ho", True), kw.get("use_lim", True), kw.get("lim", 4)
    cfl = _sweep(q, 1, dt, nx, ny, met, work, g, ho, use_lim, lim, 0.0)
    if not cfl_target / cfl >= shrink:
        return -1, cfl
    cfl = _sweep(q, 2, dt, nx, ny, met, work, g, ho, use_lim, lim, cfl)
    return (1 if cfl_target / cfl >= shrink else -1), cfl
