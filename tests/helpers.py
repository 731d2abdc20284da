"""Shared test inputs: seeded initial states in the reference's frame layout."""
from __future__ import annotations

import numpy as np

GRAV = float(np.float32(9.81))


def centres(n, lo=-1.0, hi=1.0):
    d = (hi - lo) / n
    i = np.arange(-1, n + 3)
    return lo + (i - 1) * d + d / 2


def smooth_state(nx, ny, seed=0, amp=0.3, vel=0.3):
    """A smooth, strictly positive depth with nonzero momenta (frame incl. ghosts)."""
    rng = np.random.default_rng(seed)
    x = centres(nx)[None, :]
    y = centres(ny)[:, None]
    k = rng.uniform(1.0, 3.0, size=4)
    ph = rng.uniform(0, 2 * np.pi, size=4)
    h = 1.5 + amp * np.sin(k[0] * np.pi * x + ph[0]) * np.cos(k[1] * np.pi * y + ph[1])
    hu = vel * h * np.sin(k[2] * np.pi * y + ph[2]) * np.ones_like(x)
    hv = vel * h * np.cos(k[3] * np.pi * x + ph[3]) * np.ones_like(y)
    q = np.zeros((ny + 4, nx + 4, 3))
    q[..., 0] = h
    q[..., 1] = hu
    q[..., 2] = hv
    return q


def rough_state(nx, ny, seed=0):
    """Piecewise state with jumps in every field (exercises every limiter
    branch).  Depths stay in [1, 1.25] so that linear extrapolation into the
    ghost cells (BC 3: 3 q1 - 2 q2) keeps them positive."""
    rng = np.random.default_rng(seed)
    q = np.zeros((ny + 4, nx + 4, 3))
    q[..., 0] = rng.choice([1.0, 1.1, 1.2], size=(ny + 4, nx + 4)) + 0.05 * rng.random((ny + 4, nx + 4))
    q[..., 1] = rng.normal(0.0, 0.4, size=(ny + 4, nx + 4))
    q[..., 2] = rng.normal(0.0, 0.4, size=(ny + 4, nx + 4))
    return q


def stable_dt(q, nx, ny, cfl=0.45, lo=-1.0, hi=1.0):
    h = q[2:-2, 2:-2, 0]
    u = np.abs(q[2:-2, 2:-2, 1] / h)
    v = np.abs(q[2:-2, 2:-2, 2] / h)
    c = np.sqrt(GRAV * h)
    d = (hi - lo) / max(nx, ny)
    return cfl * d / float(max((u + c).max(), (v + c).max()))


def sides_mask(nx, ny):
    """Frame cells the reference ever writes: everything except the 4x2x2 ghost corners."""
    m = np.ones((ny + 4, nx + 4), dtype=bool)
    for a in (slice(0, 2), slice(-2, None)):
        for b in (slice(0, 2), slice(-2, None)):
            m[a, b] = False
    return m
