"""The engine's tile list, planned on the host (pcw_plan_tiles, no GPU).

A step launch hands each wave one tile: a strip of 60 output columns (the
last strip shifted to end at column nx) and a band of rows.  Every interior
cell of every instance must be produced, each row of a strip exactly once,
by bands of at least two rows (the FP32 step finishes the eta ghost rows from
the two rows next to the boundary inside one tile).  Each instance's tiles are
contiguous in the list, because the completion of an attempt counts the tiles
from the instance's first one.  A fused batch's list is split into two lanes
at the instance boundary nearest its middle; single instances and the staged
engine's lists are not split.  FP32 tiles span fewer than 2^31 bytes of a plane, the range of
the buffer offsets the step kernel addresses them with.
"""
from __future__ import annotations

import itertools

import numpy as np
import pytest

STRIP = 60


def _abi():
    from paraclaws_amd import abi

    abi.load()
    return abi


def _check_instance(tiles, nx, ny):
    """Strips cover columns 1..nx, and each strip's bands partition 1..ny."""
    x0s = sorted(set(tiles[:, 1].tolist()))
    covered = np.zeros(nx + 1, dtype=bool)
    for x0 in x0s:
        assert 1 <= x0 and (x0 + STRIP - 1 <= nx or x0 == 1), (x0, nx)
        covered[x0:min(x0 + STRIP - 1, nx) + 1] = True
        bands = sorted((int(a), int(b)) for a, b in tiles[tiles[:, 1] == x0][:, 2:4])
        assert bands[0][0] == 1 and bands[-1][1] == ny, bands
        for (a0, a1), (b0, _) in zip(bands, bands[1:]):
            assert b0 == a1 + 1, (a0, a1, b0)
        assert min(b1 - b0 + 1 for b0, b1 in bands) >= min(2, ny), bands
    assert covered[1:].all()
    # strips start at multiples of 60 columns, except the last one shifted to end at nx
    assert len(x0s) == (nx + STRIP - 1) // STRIP


SIZES = [(2, 2), (3, 7), (61, 2), (60, 5), (61, 61), (121, 37), (300, 250), (1000, 17),
         (59, 400)]


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
@pytest.mark.parametrize("tile_rows", [0, 1, 2, 3, 5, 16, 37, 70, 256])
@pytest.mark.parametrize("wpb", [1, 4])
def test_tiles_cover_every_cell_once_per_strip(prec, tile_rows, wpb):
    abi = _abi()
    precision = abi.FP32 if prec == "fp32" else abi.FP64
    p = abi.make_params(precision, limiter=4, tile_rows=tile_rows, waves_per_block=wpb)
    doms = [abi.Domain(nx, ny) for nx, ny in SIZES]
    tiles, split = abi.plan_tiles(p, doms, lanes=True)
    # each instance's tiles are contiguous, in instance order
    inst = tiles[:, 0]
    starts = [int(np.argmax(inst == k)) for k in range(len(doms))]
    for k, (nx, ny) in enumerate(SIZES):
        mine = np.where(inst == k)[0]
        assert mine.size and (np.diff(mine) == 1).all(), k
        assert mine[0] == starts[k]
        _check_instance(tiles[mine], nx, ny)
    assert 0 < split < len(tiles) and tiles[split - 1, 0] != tiles[split, 0]


@pytest.mark.parametrize("n_inst", [1, 2, 3, 7, 40])
def test_lanes_split_at_the_instance_boundary_nearest_the_middle(n_inst):
    abi = _abi()
    rng = np.random.default_rng(n_inst)
    sizes = [(int(rng.integers(2, 400)), int(rng.integers(2, 400))) for _ in range(n_inst)]
    doms = [abi.Domain(nx, ny) for nx, ny in sizes]
    p = abi.make_params(abi.FP32, limiter=4, tile_rows=32, waves_per_block=4)
    tiles, split = abi.plan_tiles(p, doms, lanes=True)
    T = len(tiles)
    if n_inst == 1:
        assert split == 0
        return
    bounds = [i for i in range(1, T) if tiles[i, 0] != tiles[i - 1, 0]]
    assert split in bounds
    assert abs(2 * split - T) == min(abs(2 * b - T) for b in bounds)
    # one lane when the caller does not ask for lanes, and for the staged engine
    assert abi.plan_tiles(p, doms, lanes=False)[1] == 0
    st = abi.make_params(abi.FP64, engine=abi.STAGED, limiter=4, tile_rows=32, waves_per_block=4)
    assert abi.plan_tiles(st, doms, lanes=True)[1] == 0
    # FP64 fused batches split alike (their own list: no shortened edge tiles)
    p64 = abi.make_params(abi.FP64, limiter=4, tile_rows=32, waves_per_block=4)
    t64, s64 = abi.plan_tiles(p64, doms, lanes=True)
    b64 = [i for i in range(1, len(t64)) if t64[i, 0] != t64[i - 1, 0]]
    assert s64 in b64 and abs(2 * s64 - len(t64)) == min(abs(2 * b - len(t64)) for b in b64)
    # the split moves no tile: the same tiles in both plans
    same = abi.plan_tiles(p, doms, lanes=False)[0]
    key = lambda a: sorted(map(tuple, a.tolist()))  # noqa: E731
    assert key(same) == key(tiles)


def test_fp32_tiles_stay_within_32_bit_plane_offsets():
    """A one-band request on a 16384 x 32768 FP32 domain is lowered so that a
    tile's rows (with the four extra rows it marches) span fewer than 2^31
    bytes of a plane; the FP64 engine addresses with 64-bit offsets and keeps
    the request."""
    abi = _abi()
    nx, ny = 16384, 32768
    pitch = 64 * ((nx + STRIP - 1) // STRIP + 1)  # slot rows, in elements
    p = abi.make_params(abi.FP32, limiter=4, tile_rows=ny, waves_per_block=4)
    tiles, _ = abi.plan_tiles(p, [abi.Domain(nx, ny)])
    rows = tiles[:, 3] - tiles[:, 2] + 1
    assert ((rows + 5) * pitch * 4 < 2 ** 31).all()
    assert rows.max() > 1000  # lowered, not shortened to small bands
    _check_instance(tiles, nx, ny)
    p64 = abi.make_params(abi.FP64, limiter=4, tile_rows=ny, waves_per_block=4)
    t64, _ = abi.plan_tiles(p64, [abi.Domain(nx, ny)])
    assert (t64[:, 2] == 1).all() and (t64[:, 3] == ny).all()


def test_fp32_edge_tiles_are_shorter():
    """FP32 tiles on a domain boundary run the kernel's general body, about
    1.2x an interior tile's time per row, and get about four fifths of the
    rows: every band of the two edge strips, the first and last band of the
    others."""
    abi = _abi()
    nx, ny, ty = 600, 700, 70
    p = abi.make_params(abi.FP32, limiter=4, tile_rows=ty, waves_per_block=4)
    tiles, _ = abi.plan_tiles(p, [abi.Domain(nx, ny)])
    rows = tiles[:, 3] - tiles[:, 2] + 1
    edge_strip = (tiles[:, 1] == 1) | (tiles[:, 1] + STRIP - 1 >= nx)
    edge_band = (tiles[:, 2] == 1) | (tiles[:, 3] == ny)
    te = (4 * ty + 2) // 5
    assert rows[edge_strip].max() <= te
    assert rows[edge_band].max() <= te
    inner = ~(edge_strip | edge_band)
    assert rows[inner].max() <= ty and rows[inner].min() > te


@pytest.mark.parametrize("bad", ["nx", "ny", "count", "precision", "cap", "huge"])
def test_plan_rejects_bad_arguments(bad):
    import ctypes as C

    abi = _abi()
    L = abi.load()
    p = abi.make_params(abi.FP32, limiter=4)
    d = abi.Domain(100, 100).c()
    n, cap, out = 1, 0, None
    if bad == "nx":
        d.nx = 1
    elif bad == "ny":
        d.ny = 0
    elif bad == "count":
        n = 0
    elif bad == "precision":
        p.precision = 7
    elif bad == "cap":
        cap = 10  # no output array
    elif bad == "huge":  # more tiles than list positions allow
        p.tile_rows = 3
        n = 2000
        d = (abi.PcwDomain * n)(*[abi.Domain(200000, 200000).c() for _ in range(n)])
        rc = L.pcw_plan_tiles(C.byref(p), d, n, 0, out, cap, None)
        assert rc == abi.PCW_ERR_ARG
        return
    rc = L.pcw_plan_tiles(C.byref(p), C.byref(d), n, 0, out, cap, None)
    assert rc == abi.PCW_ERR_ARG


def test_plan_is_deterministic_and_wave_count_only_reorders():
    """The same request gives the same list; another wave count per block
    changes the order (the XCD ranges of the list) but not the tiles."""
    abi = _abi()
    doms = [abi.Domain(nx, ny) for nx, ny in itertools.islice(SIZES, 3, 8)]
    key = lambda a: sorted(map(tuple, a.tolist()))  # noqa: E731
    plans = []
    for wpb in (1, 2, 4):
        p = abi.make_params(abi.FP32, limiter=4, tile_rows=24, waves_per_block=wpb)
        a, _ = abi.plan_tiles(p, doms, lanes=True)
        b, _ = abi.plan_tiles(p, doms, lanes=True)
        assert np.array_equal(a, b)
        plans.append(a)
    assert key(plans[0]) == key(plans[1]) == key(plans[2])
