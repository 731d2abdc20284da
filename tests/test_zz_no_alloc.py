"""pcw_step_noalloc (the entry point that allocates nothing) against pcw_step.

Device frames for pcw_step_noalloc come from the HIP runtime the engine
library is linked with (/opt/rocm/lib/libamdhip64.so.7, through ctypes), not
from the library.  Each case runs two calls in a row, the second on a smaller
domain, through the same device scratch array; every result must equal
pcw_step's on the same input bit for bit (interior and physical ghost cells;
the ghost corners are never written), and so must the sweep maxima.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from tests.helpers import sides_mask, smooth_state, stable_dt

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _hip():
    from paraclaws_amd import abi

    abi.load()  # binds the runtime first; the CDLL below returns the same library
    hip = C.CDLL("/opt/rocm/lib/libamdhip64.so.7")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipDeviceSynchronize.argtypes = []
    return hip


class DeviceFrames:
    """Two device frames (in, out) of one domain's layout, freed on exit."""

    def __init__(self, hip, nbytes):
        self.hip = hip
        self.ptrs = []
        for _ in range(2):
            p = C.c_void_p()
            assert hip.hipMalloc(C.byref(p), nbytes) == 0
            assert p.value % 256 == 0
            self.ptrs.append(p.value)
        self.nbytes = nbytes

    def put(self, k, arr):
        a = np.ascontiguousarray(arr)
        assert a.nbytes == self.nbytes
        assert self.hip.hipMemcpy(self.ptrs[k], a.ctypes.data, a.nbytes, H2D) == 0

    def get(self, k, dtype):
        a = np.empty(self.nbytes // np.dtype(dtype).itemsize, dtype=dtype)
        assert self.hip.hipMemcpy(a.ctypes.data, self.ptrs[k], a.nbytes, D2H) == 0
        return a

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.hip.hipDeviceSynchronize()
        for p in self.ptrs:
            self.hip.hipFree(p)


# (nx, ny, bc, source, seed): the second case of each precision is smaller
CASES = [(150, 97, (4, 3, 2, 4), 0, 1), (61, 40, (2, 4, 4, 3), 2, 2)]


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_step_noalloc_equals_step(prec):
    from paraclaws_amd import abi

    hip = _hip()
    precision = abi.FP64 if prec == "fp64" else abi.FP32
    for nx, ny, bc, src, seed in CASES:
        print(f"noalloc {prec} {nx}x{ny} bc={bc} source={src}", flush=True)
        q = smooth_state(nx, ny, seed=seed)
        dt = stable_dt(q, nx, ny)
        dom = abi.Domain(nx, ny, bc=bc, source=src, manning_n=0.03 if src else 0.0)
        with abi.Engine([dom], precision=precision, limiter=4) as e:
            e.set_state(0, q)
            smax_ref = e.step(dt)[0].copy()
            e.accept()
            ref = e.get_state(0)
        planes = abi.frame_to_planes(q, precision)
        with DeviceFrames(hip, planes.nbytes) as f:
            f.put(0, planes)
            f.put(1, np.full(planes.shape, np.nan, dtype=planes.dtype))  # every written cell shows
            rc, smax, msg = abi.step_noalloc(abi.make_params(precision, limiter=4), dom,
                                             f.ptrs[0], f.ptrs[1], dt)
            assert rc == abi.PCW_OK, msg
            got = abi.planes_to_frame(f.get(1, planes.dtype), precision, nx, ny)
            inp = abi.planes_to_frame(f.get(0, planes.dtype), precision, nx, ny)
        m = sides_mask(nx, ny)
        assert np.array_equal(smax, smax_ref)
        assert np.array_equal(got[m].view(np.uint64), ref[m].view(np.uint64))
        assert np.array_equal(inp, abi.planes_to_frame(planes, precision, nx, ny))  # input untouched
