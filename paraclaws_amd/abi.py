"""ctypes binding of the engine's C ABI (``include/pcw.h``).

This is a thin host-side mirror used by the tests, the benchmark and
scripted runs; the Fortran host binds the same entry points with
``iso_c_binding``.  There is no CPU fallback: loading fails loudly when the
HIP library has not been built.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpcw_hip.so")

PCW_OK, PCW_REJECTED = 0, 1
PCW_ERR_ARG, PCW_ERR_HIP, PCW_ERR_STATE, PCW_ERR_BC, PCW_ERR_CONTROL = -1, -2, -3, -4, -5
PCW_ERR_STATE_ETA = -6
FP64, FP32 = 0, 1
FUSED, STAGED = 0, 1
BC_ZERO_GRADIENT, BC_CONSTANT_GRADIENT, BC_REFLECTIVE = 2, 3, 4
SRC_NONE, SRC_BATHY, SRC_FRICTION, SRC_BATHY_FRICTION = 0, 1, 2, 3
PCW_MAX_SIDE, PCW_MAX_DOMAINS = 262136, 65535   # size limits of pcw_create
STATUS_RUNNING, STATUS_DONE, STATUS_ERROR = 0, 1, 2

# g = 9.81 as a default-kind (REAL*4) literal widened to double, as the
# reference's dam-break setup stores it.
GRAV_REF = float(np.float32(9.81))


class PcwDomain(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("bc", C.c_int32 * 4),
                ("source", C.c_int32), ("pad_", C.c_int32),
                ("xi_min", C.c_double), ("xi_max", C.c_double),
                ("eta_min", C.c_double), ("eta_max", C.c_double),
                ("manning_n", C.c_double)]


class PcwParams(C.Structure):
    _fields_ = [("precision", C.c_int32), ("engine", C.c_int32),
                ("higher_order", C.c_int32), ("use_limiter", C.c_int32),
                ("limiter", C.c_int32), ("tile_rows", C.c_int32),
                ("waves_per_block", C.c_int32), ("pad_", C.c_int32),
                ("grav", C.c_double), ("eps", C.c_double), ("dry_tol", C.c_double)]


class PcwTime(C.Structure):
    _fields_ = [("t0", C.c_double), ("t_final", C.c_double), ("dt0", C.c_double),
                ("cfl_target", C.c_double), ("dt_max", C.c_double), ("dt_min", C.c_double),
                ("growth", C.c_double), ("shrink", C.c_double),
                ("n_snapshots", C.c_int32), ("adaptive", C.c_int32)]


class PcwStepRecord(C.Structure):
    _fields_ = [("accepted", C.c_int32), ("snapshot", C.c_int32), ("dt_used", C.c_double),
                ("cfl", C.c_double), ("time", C.c_double), ("dt_next", C.c_double)]


class PcwStatus(C.Structure):
    _fields_ = [("status", C.c_int32), ("error", C.c_int32), ("accepted", C.c_int64),
                ("rejected", C.c_int64), ("time", C.c_double), ("dt", C.c_double),
                ("cfl_last", C.c_double)]


# every entry point declared in include/pcw.h, with its ctypes signature
_P = C.POINTER
_dp = _P(C.c_double)
_h = C.c_void_p
SIGNATURES = {
    "pcw_create": ([_P(PcwParams), _P(PcwDomain), C.c_int32, _P(C.c_void_p)], C.c_int),
    "pcw_destroy": ([_h], C.c_int),
    "pcw_last_error": ([_h], C.c_char_p),
    "pcw_device_bytes": ([_h, _P(C.c_int64)], C.c_int),
    "pcw_set_state": ([_h, C.c_int32, _dp], C.c_int),
    "pcw_get_state": ([_h, C.c_int32, _dp], C.c_int),
    "pcw_set_bathymetry": ([_h, C.c_int32, _dp], C.c_int),
    "pcw_set_metrics": ([_h, C.c_int32, _dp, _dp, _dp], C.c_int),
    "pcw_step": ([_h, _dp, _dp], C.c_int),
    "pcw_accept": ([_h], C.c_int),
    "pcw_reject": ([_h], C.c_int),
    "pcw_sweep": ([_h, C.c_int32, _dp, _dp], C.c_int),
    "pcw_source": ([_h, _dp], C.c_int),
    "pcw_fill_ghosts": ([_h], C.c_int),
    "pcw_time_init": ([_h, _P(PcwTime)], C.c_int),
    "pcw_enable_log": ([_h, C.c_int32], C.c_int),
    "pcw_read_log": ([_h, C.c_int32, _P(PcwStepRecord), C.c_int32, _P(C.c_int32)], C.c_int),
    "pcw_run": ([_h, C.c_int32, C.c_int64, C.c_int64], C.c_int),
    "pcw_stop_at_snapshots": ([_h, C.c_int32], C.c_int),
    "pcw_get_status": ([_h, C.c_int32, _P(PcwStatus)], C.c_int),
    "pcw_totals": ([_h, C.c_int32, _dp], C.c_int),
    "pcw_last_elapsed_ms": ([_h, _dp], C.c_int),
    "pcw_kernel_stats": ([_h, _dp, _P(C.c_int64)], C.c_int),
    "pcw_autotune": ([_h, C.c_int32, _P(C.c_int32), _P(C.c_int32)], C.c_int),
    "pcw_stream_bandwidth": ([C.c_int64, C.c_int32, _dp, _dp], C.c_int),
    "pcw_plan_tiles": ([_P(PcwParams), _P(PcwDomain), C.c_int32, C.c_int32, _P(C.c_int32), C.c_int64,
                        _P(C.c_int32)], C.c_int64),
    "pcw_step_noalloc": ([_P(PcwParams), _P(PcwDomain), C.c_void_p, C.c_void_p, C.c_double, _dp,
                          C.c_void_p], C.c_int),
}

_lib = None


class PcwError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"pcw error {code}: {msg}")
        self.code = code


def load(path: str = LIB_PATH):
    """Load libpcw_hip.so; raises when it is missing (no fallback exists)."""
    global _lib
    if _lib is None:
        if not os.path.exists(path):
            raise RuntimeError(
                f"{path} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`"
                " (hipcc, gfx950); there is no CPU fallback")
        L = C.CDLL(path)
        for name, (args, res) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def _ptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def frame_layout(precision, nx, ny):
    """Device layout of one state frame (pcw_step_noalloc's state_in/out):
    (pitch, element positions of the frame columns -1..nx+2 within a row,
    numpy dtype).  Each of the planes h, hu, hv is ny+4 rows of pitch."""
    if precision == FP64:
        pitch = (nx + 4 + 63) // 64 * 64
        return pitch, np.arange(nx + 4), np.float64
    nslot = (nx + 59) // 60
    i = np.arange(-1, nx + 3)
    s = np.minimum(np.maximum(i - 1, 0) // 60, nslot - 1)
    pos = np.where(i <= 0, 63 + i, 64 * (s + 1) + (i - 1 - 60 * s))
    return 64 * (nslot + 1), pos, np.float32


def frame_to_planes(q, precision):
    """Frame (ny+4, nx+4, 3) -> contiguous planes (3, ny+4, pitch) of the
    device layout; unused elements are 0."""
    ny, nx = q.shape[0] - 4, q.shape[1] - 4
    pitch, pos, dt = frame_layout(precision, nx, ny)
    out = np.zeros((3, ny + 4, pitch), dtype=dt)
    for k in range(3):
        out[k][:, pos] = q[..., k]
    return out


def planes_to_frame(planes, precision, nx, ny):
    """Inverse of frame_to_planes, as doubles."""
    pitch, pos, _ = frame_layout(precision, nx, ny)
    p = np.asarray(planes).reshape(3, ny + 4, pitch)
    q = np.zeros((ny + 4, nx + 4, 3))
    for k in range(3):
        q[..., k] = p[k][:, pos]
    return q


def step_noalloc(params: "PcwParams", dom: "Domain", state_in: int, state_out: int, dt: float,
                 stream: int | None = None):
    """pcw_step_noalloc on device addresses state_in/state_out (ints);
    returns (rc, smax[2], message)."""
    L = load()
    smax = np.zeros(2)
    d = dom.c()
    rc = L.pcw_step_noalloc(C.byref(params), C.byref(d), C.c_void_p(state_in),
                            C.c_void_p(state_out), dt, _ptr(smax), C.c_void_p(stream))
    msg = L.pcw_last_error(None).decode() if rc else ""
    return rc, smax, msg


def stream_bandwidth(plane_bytes: int, reps: int = 10):
    """pcw_stream_bandwidth: (best, median) GB/s of a 3-in/3-out plane stream."""
    best, med = C.c_double(), C.c_double()
    rc = load().pcw_stream_bandwidth(plane_bytes, reps, C.byref(best), C.byref(med))
    if rc != PCW_OK:
        raise PcwError(rc, "pcw_stream_bandwidth failed")
    return best.value, med.value


def plan_tiles(params, domains, lanes=False):
    """pcw_plan_tiles: the tile list a handle of `domains` steps with, computed
    on the host (no device).  Returns (tiles, split): an (n, 4) int32 array of
    (instance, x0, y0, y1) and the list position of the second lane (0: one)."""
    L = load()
    n = len(domains)
    arr = (PcwDomain * n)(*[d.c() for d in domains])
    split = C.c_int32(0)
    total = L.pcw_plan_tiles(C.byref(params), arr, n, int(bool(lanes)), None, 0, C.byref(split))
    if total < 0:
        raise PcwError(int(total), "pcw_plan_tiles: bad argument")
    out = np.zeros((max(total, 1), 4), dtype=np.int32)
    got = L.pcw_plan_tiles(C.byref(params), arr, n, int(bool(lanes)),
                           out.ctypes.data_as(_P(C.c_int32)), total, C.byref(split))
    assert got == total
    return out[:total], split.value


def make_params(precision=FP64, engine=FUSED, higher_order=True, use_limiter=True, limiter=4,
                grav=GRAV_REF, eps=1e-10, dry_tol=0.0, tile_rows=0, waves_per_block=0):
    p = PcwParams()
    p.precision, p.engine = precision, engine
    p.higher_order, p.use_limiter, p.limiter = int(higher_order), int(use_limiter), limiter
    p.tile_rows, p.waves_per_block = tile_rows, waves_per_block
    p.grav, p.eps, p.dry_tol = grav, eps, dry_tol
    return p


class Domain:
    """One instance's settings (control.in's mesh and BC lines)."""

    def __init__(self, nx, ny, bc=(2, 2, 2, 2), xi=(-1.0, 1.0), eta=(-1.0, 1.0),
                 source=SRC_NONE, manning_n=0.0):
        self.nx, self.ny, self.bc, self.xi, self.eta = nx, ny, tuple(bc), xi, eta
        self.source, self.manning_n = source, manning_n

    def c(self) -> PcwDomain:
        d = PcwDomain()
        d.nx, d.ny = self.nx, self.ny
        for k in range(4):
            d.bc[k] = int(self.bc[k])
        d.source = self.source
        d.xi_min, d.xi_max = self.xi
        d.eta_min, d.eta_max = self.eta
        d.manning_n = self.manning_n
        return d

    def frame(self) -> np.ndarray:
        return np.zeros((self.ny + 4, self.nx + 4, 3))


class Engine:
    """Owns one pcw_handle (one or more domains on the current GPU)."""

    def __init__(self, domains, precision=FP64, engine=FUSED, higher_order=True,
                 use_limiter=True, limiter=4, grav=GRAV_REF, eps=1e-10, dry_tol=0.0,
                 tile_rows=0, waves_per_block=0):
        self.L = load()
        self.domains = list(domains)
        p = make_params(precision, engine, higher_order, use_limiter, limiter, grav, eps, dry_tol,
                        tile_rows, waves_per_block)
        arr = (PcwDomain * len(self.domains))(*[d.c() for d in self.domains])
        h = C.c_void_p()
        rc = self.L.pcw_create(C.byref(p), arr, len(self.domains), C.byref(h))
        self.h = h
        if rc != PCW_OK:
            msg = self.L.pcw_last_error(h).decode() if h.value else "pcw_create failed"
            if h.value:
                self.L.pcw_destroy(h)
            self.h = None
            raise PcwError(rc, msg)
        self.n = len(self.domains)

    def _chk(self, rc):
        if rc not in (PCW_OK,):
            raise PcwError(rc, self.L.pcw_last_error(self.h).decode())
        return rc

    def close(self):
        if getattr(self, "h", None):
            self.L.pcw_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # state
    def set_state(self, k, q):
        q = np.ascontiguousarray(q, dtype=np.float64)
        self._chk(self.L.pcw_set_state(self.h, k, _ptr(q)))

    def get_state(self, k=0):
        q = self.domains[k].frame()
        self._chk(self.L.pcw_get_state(self.h, k, _ptr(q)))
        return q

    def set_bathymetry(self, k, b):
        b = np.ascontiguousarray(b, dtype=np.float64)
        self._chk(self.L.pcw_set_bathymetry(self.h, k, _ptr(b)))

    def set_metrics(self, k, kappa, gamma_xi, gamma_eta):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (kappa, gamma_xi, gamma_eta)]
        self._chk(self.L.pcw_set_metrics(self.h, k, *[_ptr(a) for a in arrs]))

    # host-controlled stepping
    def step(self, dt):
        dts = np.ascontiguousarray(np.broadcast_to(np.asarray(dt, dtype=np.float64), (self.n,)))
        smax = np.zeros(2 * self.n)
        rc = self.L.pcw_step(self.h, _ptr(dts), _ptr(smax))
        self._last_smax = smax.reshape(self.n, 2)
        self._chk(rc)
        return self._last_smax

    def last_smax(self):
        """Sweep maxima of the last pcw_step (also set when it reports an
        eta-sweep error)."""
        return self._last_smax

    def accept(self):
        self._chk(self.L.pcw_accept(self.h))

    def reject(self):
        self._chk(self.L.pcw_reject(self.h))

    def sweep(self, ixy, dt):
        dts = np.ascontiguousarray(np.broadcast_to(np.asarray(dt, dtype=np.float64), (self.n,)))
        smax = np.zeros(self.n)
        self._chk(self.L.pcw_sweep(self.h, ixy, _ptr(dts), _ptr(smax)))
        return smax

    def source(self, dt):
        dts = np.ascontiguousarray(np.broadcast_to(np.asarray(dt, dtype=np.float64), (self.n,)))
        self._chk(self.L.pcw_source(self.h, _ptr(dts)))

    def fill_ghosts(self):
        self._chk(self.L.pcw_fill_ghosts(self.h))

    # device-controlled run
    def time_init(self, t0=0.0, t_final=0.5, dt0=1e-10, cfl_target=0.9, dt_max=1e-2,
                  dt_min=1e-10, growth=1.2, shrink=0.5, n_snapshots=10, adaptive=True):
        t = PcwTime()
        t.t0, t.t_final, t.dt0, t.cfl_target = t0, t_final, dt0, cfl_target
        t.dt_max, t.dt_min, t.growth, t.shrink = dt_max, dt_min, growth, shrink
        t.n_snapshots, t.adaptive = n_snapshots, int(adaptive)
        self._chk(self.L.pcw_time_init(self.h, C.byref(t)))

    def enable_log(self, cap):
        self._chk(self.L.pcw_enable_log(self.h, cap))

    def read_log(self, k=0, cap=100000):
        recs = (PcwStepRecord * cap)()
        n = C.c_int32()
        self._chk(self.L.pcw_read_log(self.h, k, recs, cap, C.byref(n)))
        return [(r.accepted, r.snapshot, r.dt_used, r.cfl, r.time, r.dt_next) for r in recs[:n.value]]

    def run(self, steps_per_graph=32, max_attempts=0, max_accepted=0):
        self._chk(self.L.pcw_run(self.h, steps_per_graph, max_attempts, max_accepted))

    def stop_at_snapshots(self, on=True):
        self._chk(self.L.pcw_stop_at_snapshots(self.h, 1 if on else 0))

    def status(self, k=0):
        s = PcwStatus()
        self._chk(self.L.pcw_get_status(self.h, k, C.byref(s)))
        return s

    def totals(self, k=0):
        out = np.zeros(3)
        self._chk(self.L.pcw_totals(self.h, k, _ptr(out)))
        return out

    def last_elapsed_ms(self):
        v = C.c_double()
        self._chk(self.L.pcw_last_elapsed_ms(self.h, C.byref(v)))
        return v.value

    def kernel_stats(self):
        v = C.c_double()
        n = C.c_int64()
        self._chk(self.L.pcw_kernel_stats(self.h, C.byref(v), C.byref(n)))
        return v.value, n.value

    def autotune(self, steps=5):
        r, w = C.c_int32(), C.c_int32()
        self._chk(self.L.pcw_autotune(self.h, steps, C.byref(r), C.byref(w)))
        return r.value, w.value

    def device_bytes(self):
        v = C.c_int64()
        self._chk(self.L.pcw_device_bytes(self.h, C.byref(v)))
        return v.value
