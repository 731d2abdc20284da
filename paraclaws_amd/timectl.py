"""Host-side time controller and main loop for the host-controlled path.

``TimeControl`` mirrors the reference's ``time_control`` type
(src/time_control/MOD_time_control.f90:16-279) in FP64, field for field;
``run_host_controlled`` is the main loop of src/main.f90:193-283 driving an
``abi.Engine`` through pcw_step / pcw_accept / pcw_reject.  The same logic
runs on the device for pcw_run (paraclaws_amd/csrc/pcw_engine.hpp).
"""
from __future__ import annotations

from dataclasses import dataclass, field


class EndProgram(RuntimeError):
    """end_program (src/utility/MOD_error_handling.f90:20-35)."""


@dataclass
class TimeControl:
    time_initial: float
    time_final: float
    dt: float
    number_of_snapshots: int
    CFL_target: float
    dt_max: float
    dt_min: float
    timestep_growth_limit: float
    timestep_shrink_limit: float
    adaptive_timestepping: bool = True
    time: float = 0.0
    dt_previous: float = 0.0
    CFL: float = 0.0
    CFL_previous_timestep: float = 0.0
    time_precision: float = 0.01
    emergency_shrink_ratio: float = 0.25
    snapshot_exception: bool = False
    current_snapshot_index: int = 1
    snapshot_intervals: list = field(default_factory=list)

    def __post_init__(self):
        # initialize, MOD_time_control.f90:97-150
        self.time = self.time_initial
        self.dt_previous = self.dt
        if self.dt_min > self.dt_max:
            raise EndProgram("ABORT: Invalid timestep bounds!")
        if self.number_of_snapshots < 1:
            self.number_of_snapshots = 1
        interval = (self.time_final - self.time_initial) / self.number_of_snapshots
        self.snapshot_intervals = [self.time_initial + i * interval
                                   for i in range(1, self.number_of_snapshots + 1)]
        self.snapshot_intervals[-1] = self.time_final

    def update_time(self, step_valid: bool) -> None:
        """updateTime, MOD_time_control.f90:153-243."""
        self.dt_previous = self.dt
        dt_step = self.dt
        if step_valid:
            if self.adaptive_timestepping:
                cf = self.CFL_target / self.CFL if self.CFL > 0.0 else 1.0
                if cf >= self.timestep_growth_limit:
                    cf = self.timestep_growth_limit
                elif abs(cf - 1.0) <= self.time_precision:
                    cf = 1.0
                dt_step = self.dt_previous * cf
            snap = self.snapshot_intervals[self.current_snapshot_index - 1]
            if self.time + dt_step > snap:
                dt_step = snap - self.time
                self.snapshot_exception = True
            else:
                self.snapshot_exception = False
        else:
            if not self.adaptive_timestepping:
                raise EndProgram("constant time steps do not allow for adaptive correction to "
                                 "severe CFL violation!")
            dt_step = self.emergency_shrink_ratio * self.dt
        if dt_step > self.dt_max:
            dt_step = self.dt_max
        if dt_step < self.dt_min:
            if not self.adaptive_timestepping and not self.snapshot_exception:
                raise EndProgram("minimum time step size reached, but control still detects CFL "
                                 "violation - constant dt not compatible!")
            dt_step = self.dt_min
        if step_valid:
            self.time = self.time + dt_step
        self.dt = dt_step

    def update_max_CFL(self, c: float) -> None:
        self.CFL = self.CFL if self.CFL > c else c

    def check_timestep_validity(self) -> bool:
        with_inf = self.CFL_target / self.CFL if self.CFL != 0.0 else float("inf")
        return with_inf >= self.timestep_shrink_limit

    def reset_CFL(self) -> None:
        self.CFL_previous_timestep = self.CFL
        self.CFL = 0.0


def solve_dimensional_split(engine, tc: TimeControl, dxi: float, deta: float) -> int:
    """One attempt (MOD_solve.f90:23-175): both sweeps on the GPU, then the
    CFL bookkeeping and validity tests in the reference's order.  Returns
    1 (accepted) or -1 (rejected); the engine's pending state is kept or
    dropped by the caller."""
    from paraclaws_amd import abi

    eta_failed = False
    try:
        smax = engine.step(tc.dt)[0]
    except abi.PcwError as e:
        if e.code != abi.PCW_ERR_STATE_ETA:
            raise
        eta_failed = True  # fatal only if the xi sweep alone accepts the step
        smax = engine.last_smax()[0]
    tc.update_max_CFL(smax[0] * tc.dt / dxi)
    if not tc.check_timestep_validity():
        return -1
    if eta_failed:
        raise abi.PcwError(abi.PCW_ERR_STATE, "NaN in the Jacobian or complex eigenvalues")
    tc.update_max_CFL(smax[1] * tc.dt / deta)
    return 1 if tc.check_timestep_validity() else -1


def run_host_controlled(engine, tc: TimeControl, dxi: float, deta: float, max_accepted=0,
                        max_attempts=0):
    """Main loop of src/main.f90:193-283 for a single-instance engine.
    Returns the per-attempt records (accepted, snapshot, dt_used, cfl, time, dt_next)."""
    recs = []
    accepted = attempts = 0
    while tc.time < tc.time_final:
        if max_accepted and accepted >= max_accepted:
            break
        if max_attempts and attempts >= max_attempts:
            break
        dt_used = tc.dt
        ierr = solve_dimensional_split(engine, tc, dxi, deta)
        cfl = tc.CFL
        if ierr < 0:
            tc.update_time(False)
            engine.reject()
            tc.reset_CFL()
            recs.append((0, 0, dt_used, cfl, tc.time, tc.dt))
        else:
            engine.accept()  # source and ghost fill ran inside the fused step
            tc.update_time(True)
            snap = 0
            if (tc.current_snapshot_index <= tc.number_of_snapshots
                    and tc.time >= tc.snapshot_intervals[tc.current_snapshot_index - 1]):
                snap = tc.current_snapshot_index
                tc.current_snapshot_index += 1
            tc.reset_CFL()
            accepted += 1
            recs.append((1, snap, dt_used, cfl, tc.time, tc.dt))
        attempts += 1
    return recs
