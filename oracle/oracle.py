"""ctypes wrapper around the oracle library (``oracle/liboracle.so``).

TEST INFRASTRUCTURE ONLY: the oracle is the checker.  Only the test suite,
``__graft_entry__.smoke()`` and the CPU-baseline leg of ``bench.py`` import
this module; the product path never does.

Arrays are Fortran frames ``vars(3, -1:nx+2, -1:ny+2)`` held in numpy as
``shape (ny+4, nx+4, 3)`` C-order, which is the same memory as the Fortran
column-major frame (equation index fastest).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

# g = 9.81 written as a default-kind (REAL*4) literal and stored in a double
# (examples/2D_dam_break/cartesian/SMOD_setup.f90:21).
GRAV_REF = float(np.float32(9.81))
UD_PRECISION = 1e-10  # src/constant/MOD_constants.f90:18


class OrcCfg(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32),
        ("higher_order", C.c_int32), ("use_limiter", C.c_int32),
        ("limiter", C.c_int32), ("bc", C.c_int32 * 4),
        ("source", C.c_int32), ("nthreads", C.c_int32),
        ("xi_min", C.c_double), ("xi_max", C.c_double),
        ("eta_min", C.c_double), ("eta_max", C.c_double),
        ("grav", C.c_double), ("eps", C.c_double),
        ("manning_n", C.c_double), ("dry_tol", C.c_double),
        ("kappa", C.c_void_p), ("gamma_xi", C.c_void_p), ("gamma_eta", C.c_void_p),
    ]


class OrcTc(C.Structure):
    _fields_ = [
        ("time", C.c_double), ("time_initial", C.c_double), ("time_final", C.c_double),
        ("dt", C.c_double), ("dt_previous", C.c_double), ("dt_max", C.c_double),
        ("dt_min", C.c_double), ("time_precision", C.c_double), ("CFL", C.c_double),
        ("CFL_target", C.c_double), ("CFL_previous_timestep", C.c_double),
        ("shrink_limit", C.c_double), ("growth_limit", C.c_double),
        ("emergency_shrink_ratio", C.c_double),
        ("adaptive", C.c_int32), ("snapshot_exception", C.c_int32),
        ("current_snapshot_index", C.c_int32), ("number_of_snapshots", C.c_int32),
        ("status", C.c_int32), ("pad_", C.c_int32),
        ("snapshot_intervals", C.c_double * 4096),
    ]


class OrcStepRec(C.Structure):
    _fields_ = [
        ("accepted", C.c_int32), ("snapshot", C.c_int32),
        ("dt_used", C.c_double), ("cfl", C.c_double),
        ("time", C.c_double), ("dt_next", C.c_double),
    ]


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc, -ffp-contract=off)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-B" if force else "-s", "-C", _HERE, "liboracle.so"],
                       check=True, stdout=subprocess.DEVNULL)
    return _LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        P = C.POINTER
        dp = P(C.c_double)
        L.orc_metrics.argtypes = [P(OrcCfg), dp, dp, dp, dp, dp]
        L.orc_init_dam_break.argtypes = [P(OrcCfg), dp, C.c_double, C.c_double, C.c_double]
        L.orc_fill_ghosts.argtypes = [P(OrcCfg), dp]
        L.orc_sweep.argtypes = [P(OrcCfg), dp, C.c_int, C.c_double, dp]
        L.orc_solve_dimensional_split.argtypes = [P(OrcCfg), P(OrcTc), dp]
        L.orc_source.argtypes = [P(OrcCfg), dp, dp, C.c_double]
        L.orc_tc_initialize.argtypes = [P(OrcTc), C.c_double, C.c_double, C.c_double,
                                        C.c_int32, C.c_double, C.c_double, C.c_double,
                                        C.c_double, C.c_double, C.c_int32]
        L.orc_tc_update_time.argtypes = [P(OrcTc), C.c_int32]
        L.orc_tc_update_time.restype = None
        L.orc_tc_update_max_cfl.argtypes = [P(OrcTc), C.c_double]
        L.orc_tc_update_max_cfl.restype = None
        L.orc_tc_check_validity.argtypes = [P(OrcTc)]
        L.orc_tc_reset_cfl.argtypes = [P(OrcTc)]
        L.orc_tc_reset_cfl.restype = None
        L.orc_run.argtypes = [P(OrcCfg), P(OrcTc), dp, dp, C.c_int32, C.c_int32,
                              P(OrcStepRec), C.c_int32, C.c_int32]
        L.orc_cell_phys.argtypes = [P(OrcCfg), C.c_int, dp, dp, dp]
        L.orc_hll.argtypes = [P(OrcCfg), C.c_int, dp, dp, dp]
        L.orc_phi.argtypes = [C.c_int, C.c_double]
        L.orc_phi.restype = C.c_double
        L.orc_totals.argtypes = [P(OrcCfg), dp, dp]
        L.orc_totals.restype = None
        L.orc_use_lapack.argtypes = [C.c_void_p, C.c_void_p]
        L.orc_use_lapack.restype = None
        _lib = L
    return _lib


def use_lapack(on: bool = True) -> None:
    """Switch the restatement to the reference's per-interface cost structure
    (SMOD_approximate_Riemann_solvers.f90:60-101: LAPACK dgeev for the
    eigenvalues of each cell's Jacobian, with its workspace query and right
    eigenvectors, and BLAS dgemv for J*q), with scipy's LAPACK and BLAS; off:
    the closed forms.  Process-wide.  Callers limit the BLAS threads to one,
    as the reference's MKL_NUM_THREADS=1 (set_env.sh:24)."""
    L = lib()
    if not on:
        L.orc_use_lapack(None, None)
        return
    from scipy.linalg import cython_blas, cython_lapack

    name = C.pythonapi.PyCapsule_GetName
    name.restype = C.c_char_p
    name.argtypes = [C.py_object]
    ptr = C.pythonapi.PyCapsule_GetPointer
    ptr.restype = C.c_void_p
    ptr.argtypes = [C.py_object, C.c_char_p]

    def fn(mod, sym):
        cap = mod.__pyx_capi__[sym]
        return ptr(cap, name(cap))

    L.orc_use_lapack(fn(cython_lapack, "dgeev"), fn(cython_blas, "dgemv"))


def _dp(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


@dataclass
class Config:
    """Solver settings; field names follow control.in (MOD_IO.f90:61-94)."""
    nx: int = 32
    ny: int = 32
    xi_min: float = -1.0
    xi_max: float = 1.0
    eta_min: float = -1.0
    eta_max: float = 1.0
    higher_order: bool = True
    use_limiter: bool = True
    limiter: int = 4
    bc: tuple = (2, 2, 2, 2)
    grav: float = GRAV_REF
    eps: float = UD_PRECISION
    source: int = 0
    manning_n: float = 0.0
    dry_tol: float = 0.0
    # OpenMP threads: the machine's CPU share for one GPU (16) at most
    nthreads: int = min(16, os.cpu_count() or 1)
    # optional metric frames (ny+4, nx+4) of a mapped grid; None = identity map
    kappa: object = None
    gamma_xi: object = None
    gamma_eta: object = None

    def c(self) -> OrcCfg:
        o = OrcCfg()
        o.nx, o.ny = self.nx, self.ny
        o.higher_order, o.use_limiter, o.limiter = int(self.higher_order), int(self.use_limiter), self.limiter
        for k in range(4):
            o.bc[k] = int(self.bc[k])
        o.source, o.nthreads = self.source, self.nthreads
        o.xi_min, o.xi_max, o.eta_min, o.eta_max = self.xi_min, self.xi_max, self.eta_min, self.eta_max
        o.grav, o.eps, o.manning_n, o.dry_tol = self.grav, self.eps, self.manning_n, self.dry_tol
        for name in ("kappa", "gamma_xi", "gamma_eta"):
            a = getattr(self, name)
            if a is not None:
                assert a.dtype == np.float64 and a.flags.c_contiguous
                assert a.shape == (self.ny + 4, self.nx + 4)
                setattr(o, name, a.ctypes.data)
        return o

    def frame(self) -> np.ndarray:
        return np.zeros((self.ny + 4, self.nx + 4, 3), dtype=np.float64)


@dataclass
class TimeSettings:
    """time_control settings (MOD_time_control.f90:70-151)."""
    t0: float = 0.0
    t_final: float = 0.5
    dt0: float = 1e-10
    n_snap: int = 10
    cfl_target: float = 0.9
    dt_max: float = 1e-2
    dt_min: float = 1e-10
    growth: float = 1.2
    shrink: float = 0.5
    adaptive: bool = True

    def tc(self) -> OrcTc:
        t = OrcTc()
        rc = lib().orc_tc_initialize(C.byref(t), self.t0, self.t_final, self.dt0, self.n_snap,
                                     self.cfl_target, self.dt_max, self.dt_min, self.growth,
                                     self.shrink, int(self.adaptive))
        if rc:
            raise ValueError(f"orc_tc_initialize failed rc={rc}")
        return t


def metrics(cfg: Config):
    shape = (cfg.ny + 4, cfg.nx + 4)
    out = [np.zeros(shape) for _ in range(5)]
    c = cfg.c()
    lib().orc_metrics(C.byref(c), *[_dp(a) for a in out])
    return dict(zip(["xi", "eta", "kappa", "gamma_xi", "gamma_eta"], out))


def dam_break(cfg: Config, h_in=2.0, h_out=1.0, radius=0.5) -> np.ndarray:
    q = cfg.frame()
    c = cfg.c()
    lib().orc_init_dam_break(C.byref(c), _dp(q), h_in, h_out, radius)
    return q


def fill_ghosts(cfg: Config, q: np.ndarray) -> None:
    c = cfg.c()
    rc = lib().orc_fill_ghosts(C.byref(c), _dp(q))
    if rc:
        raise RuntimeError(f"orc_fill_ghosts rc={rc}")


def sweep(cfg: Config, q: np.ndarray, ixy: int, dt: float, cfl: float = 0.0) -> float:
    c = cfg.c()
    v = C.c_double(cfl)
    rc = lib().orc_sweep(C.byref(c), _dp(q), ixy, dt, C.byref(v))
    if rc:
        raise RuntimeError(f"orc_sweep rc={rc}")
    return v.value


def solve(cfg: Config, tc: OrcTc, q: np.ndarray) -> int:
    c = cfg.c()
    return lib().orc_solve_dimensional_split(C.byref(c), C.byref(tc), _dp(q))


def source(cfg: Config, q: np.ndarray, bathy, dt: float) -> None:
    c = cfg.c()
    rc = lib().orc_source(C.byref(c), _dp(q), _dp(bathy), dt)
    if rc:
        raise RuntimeError(f"orc_source rc={rc}")


def run(cfg: Config, ts: TimeSettings, q: np.ndarray, bathy=None, max_accepted=0,
        max_attempts=0, tc: OrcTc | None = None):
    """Main loop; returns (records, tc).  q is advanced in place.  A run given
    the controller `tc` of an earlier one continues it: the ghost fill that
    precedes a run's loop is skipped, so that changes made to q between the
    two (hooks) stay as the reference's would."""
    c = cfg.c()
    t = ts.tc() if tc is None else tc
    cap = max(max_attempts, max_accepted * 4, 1) if (max_attempts or max_accepted) else 200000
    recs = (OrcStepRec * cap)()
    n = lib().orc_run(C.byref(c), C.byref(t), _dp(q), _dp(bathy), max_accepted, max_attempts,
                      recs, cap, 1 if tc is None else 0)
    if n < 0:
        raise RuntimeError(f"orc_run rc={n}")
    rows = [(r.accepted, r.snapshot, r.dt_used, r.cfl, r.time, r.dt_next) for r in recs[:min(n, cap)]]
    return rows, t


def totals(cfg: Config, q: np.ndarray):
    out = np.zeros(3)
    c = cfg.c()
    lib().orc_totals(C.byref(c), _dp(q), _dp(out))
    return out


def cell_phys(cfg: Config, ixy: int, q):
    """(rc, speed, flux) of one cell for sweep ixy."""
    q = np.ascontiguousarray(q, dtype=np.float64)
    s = C.c_double()
    f = np.zeros(3)
    c = cfg.c()
    rc = lib().orc_cell_phys(C.byref(c), ixy, _dp(q), C.byref(s), _dp(f))
    return rc, s.value, f


def hll(cfg: Config, ixy: int, qL, qR):
    """(rc, lambda[2], W1[3], W2[3]) of one interface."""
    qL = np.ascontiguousarray(qL, dtype=np.float64)
    qR = np.ascontiguousarray(qR, dtype=np.float64)
    out = np.zeros(8)
    c = cfg.c()
    rc = lib().orc_hll(C.byref(c), ixy, _dp(qL), _dp(qR), _dp(out))
    return rc, out[0:2], out[2:5], out[5:8]


def phi(code: int, r: float) -> float:
    return lib().orc_phi(code, r)
