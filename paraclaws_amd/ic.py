"""Initial conditions and synthetic inputs in the reference's frame layout.

Frames are numpy arrays of shape (ny+4, nx+4, 3) holding
solution%vars(3, -1:nx+2, -1:ny+2) (equation index fastest).
"""
from __future__ import annotations

import numpy as np


def centres(n: int, lo: float, hi: float) -> np.ndarray:
    """Cell centres of frame indices -1..n+2, rounded as
    src/domain/SMOD_mesh_manipulation.f90:50-51 computes them."""
    d = (hi - lo) / n
    i = np.arange(-1, n + 3, dtype=np.float64)
    return (lo + (i - 1) * d) + d / 2


def dam_break(nx: int, ny: int, xi=(-1.0, 1.0), eta=(-1.0, 1.0), h_in=2.0, h_out=1.0,
              radius=0.5) -> np.ndarray:
    """Circular dam break over the whole frame
    (examples/2D_dam_break/cartesian/SMOD_set_initial_conditions.f90:34-46)."""
    x = centres(nx, *xi)
    y = centres(ny, *eta)
    q = np.zeros((ny + 4, nx + 4, 3))
    r = np.sqrt(x[None, :] * x[None, :] + y[:, None] * y[:, None])
    q[..., 0] = np.where(r < radius, h_in, h_out)
    return q


class SplitMix64:
    """splitmix64 (Steele, Lea, Flood 2014): a portable seeded generator so
    synthetic inputs are reproducible from the seed alone."""

    MASK = (1 << 64) - 1

    def __init__(self, seed: int):
        self.s = seed & self.MASK

    def next_u64(self) -> int:
        self.s = (self.s + 0x9E3779B97F4A7C15) & self.MASK
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & self.MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & self.MASK
        return z ^ (z >> 31)

    def uniform(self, lo: float = 0.0, hi: float = 1.0) -> float:
        return lo + (hi - lo) * ((self.next_u64() >> 11) * (1.0 / (1 << 53)))


def gaussian_bathymetry(nx, ny, seed, xi=(-1.0, 1.0), eta=(-1.0, 1.0), n_bumps=8,
                        amp=(0.0, 0.8), width=(0.05, 0.3)) -> np.ndarray:
    """b = sum_k A_k exp(-|x - x_k|^2 / (2 s_k^2)) over the frame (ny+4, nx+4)."""
    g = SplitMix64(seed)
    bumps = []
    for _ in range(n_bumps):
        a = g.uniform(*amp)
        s = g.uniform(*width)
        cx = g.uniform(*xi)
        cy = g.uniform(*eta)
        bumps.append((a, s, cx, cy))
    x = centres(nx, *xi)[None, :]
    y = centres(ny, *eta)[:, None]
    b = np.zeros((ny + 4, nx + 4))
    for a, s, cx, cy in bumps:
        b += a * np.exp(-((x - cx) ** 2 + (y - cy) ** 2) / (2 * s * s))
    return b


def still_water_with_step(b: np.ndarray, nx: int, xi=(-1.0, 1.0), level=1.0, step=1.0,
                          dry_tol=1e-6) -> np.ndarray:
    """h = max(0, level - b), dry below dry_tol, then h += step where x < 0;
    hu = hv = 0."""
    ny4, nx4 = b.shape
    q = np.zeros((ny4, nx4, 3))
    h = np.maximum(0.0, level - b)
    h[h < dry_tol] = 0.0
    x = centres(nx, *xi)[None, :]
    h = h + np.where(x < 0.0, step, 0.0)
    q[..., 0] = h
    return q


def field_errors(got: np.ndarray, ref: np.ndarray) -> dict:
    """Per-field relative L1 and Linf over the interior; absolute values when
    the reference field is identically zero."""
    out = {}
    g = got[2:-2, 2:-2]
    r = ref[2:-2, 2:-2]
    for k, name in enumerate(("h", "hu", "hv")):
        d = np.abs(g[..., k] - r[..., k])
        s1 = np.abs(r[..., k]).sum()
        sm = np.abs(r[..., k]).max()
        out[name] = {
            "l1_rel": float(d.sum() / s1) if s1 > 0 else float(d.sum()),
            "linf_rel": float(d.max() / sm) if sm > 0 else float(d.max()),
        }
    return out


def ensemble_config4(n_inst=512, seed=4, n_min=64, n_max=1024, ratio=(1.5, 10.0)):
    """Batched dam-break ensemble: square sizes log-uniform in [n_min, n_max]
    and inner depth uniform in ratio, from a splitmix64 stream of `seed`.
    Returns a list of (n, depth_ratio)."""
    import math

    g = SplitMix64(seed)
    out = []
    for _ in range(n_inst):
        n = int(round(math.exp(g.uniform(math.log(n_min), math.log(n_max)))))
        r = g.uniform(*ratio)
        out.append((n, r))
    return out


def mapped_metrics(nx, ny, mapping, xi=(-1.0, 1.0), eta=(-1.0, 1.0)):
    """capacity and face length ratios of the mapped grid x, y = mapping(xi, eta),
    as src/domain/SMOD_mesh_manipulation.f90:36-48 and :147-173 compute them:
    corners in counter-clockwise order, length ratios of the faces 1->4 (xi
    face) and 1->2 (eta face), shoelace area over the capacity.  Returns
    (kappa, gamma_xi, gamma_eta, x, y), frames of shape (ny+4, nx+4); x, y are
    the mapped cell centres."""
    dxi = (xi[1] - xi[0]) / nx
    deta = (eta[1] - eta[0]) / ny
    i = np.arange(-1, nx + 3, dtype=np.float64)[None, :]
    j = np.arange(-1, ny + 3, dtype=np.float64)[:, None]
    xc1 = xi[0] + (i - 1) * dxi + 0 * j
    xc2 = xc1 + dxi
    ec1 = eta[0] + (j - 1) * deta + 0 * i
    ec3 = ec1 + deta
    cx = [xc1, xc2, xc2, xc1, xc1]
    cy = [ec1, ec1, ec3, ec3, ec1]
    px, py = zip(*[mapping(a, b) for a, b in zip(cx, cy)])
    fdx, fdy = px[3] - px[0], py[3] - py[0]
    gam_xi = np.sqrt(fdx * fdx + fdy * fdy) / deta
    fdx, fdy = px[1] - px[0], py[1] - py[0]
    gam_eta = np.sqrt(fdx * fdx + fdy * fdy) / dxi
    area = np.zeros_like(xc1)
    for m in range(4):
        area = area + 0.5 * (py[m] + py[m + 1]) * (px[m] - px[m + 1])
    kappa = area / (dxi * deta)
    xcen = (xi[0] + (i - 1) * dxi) + dxi / 2 + 0 * j
    ecen = (eta[0] + (j - 1) * deta) + deta / 2 + 0 * i
    x, y = mapping(xcen, ecen)
    return kappa, gam_xi, gam_eta, x, y
