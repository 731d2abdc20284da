// pcw_engine.hpp -- instance descriptors, device-resident per-instance state
// and the FP64 time controller (a replica of time_control's updateTime,
// src/time_control/MOD_time_control.f90:153-279) that runs both on the host
// and, for device-controlled runs, in the last workgroup of each step.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "../../include/pcw.h"

namespace pcw {

constexpr int kStripOut = 60;   // output columns per wave (64 lanes - 2x2 halo)
constexpr int kMinBandRows = 3; // smallest requested band height (bands of >= 2 rows)

// Layout of a frame row, columns i = -1 .. nx+2 (ghost, interior, ghost).
// Plain rows (FP64 engines): column i at element i+1.  Slot rows (FP32
// engine): the row is a sequence of 64-element slots; slot 0 holds the ghost
// columns -1, 0 in its last two elements, slot s+1 holds the interior columns
// 60s+1 .. 60s+60 (kStripOut per wave) and four padding elements, the last
// slot also the ghost columns nx+1, nx+2 right after column nx.  A wave that
// finishes the 60 columns of a strip then writes one aligned, whole
// 256-byte slot per row and plane (partial cache-line writes cost gfx950's
// memory system about half again the time of whole lines).
__host__ __device__ inline int col_pos(int slot_rows, int nslot, int i) {
  if (!slot_rows) return i + 1;
  if (i <= 0) return 63 + i;
  int s = (i - 1) / kStripOut;
  if (s > nslot - 1) s = nslot - 1;
  return 64 * (s + 1) + (i - 1 - kStripOut * s);
}

// Read-only description of one instance (domain).
struct KInst {
  int nx, ny, pitch, ntiles;
  int bc[4];
  int source, tile0;   // tile0: index of the instance's first tile in the tile list
  int slot_rows;       // 1: frame rows in 64-element slots (FP32 engine), 0: plain rows
  int nslot;           // slot rows: interior slots, ceil(nx / kStripOut)
  int unit_met;        // identity map whose rounded metrics are all exactly 1 (FP64 engine)
  long long plane;     // elements per field plane (pitch * (ny+4))
  long long off[2];    // element offset of the h-plane frame origin (i=-1,j=-1) of buffer 0/1
  long long boff;      // bathymetry frame offset, -1 when absent
  long long moff;      // metric planes (kappa, gamma_xi, gamma_eta) offset, -1: identity map
  double dxi, deta, xi_min, eta_min;
  double manning;
  // FP32 extrapolation ratios (x_ghost - x_near) / (x_far - x_near) of the
  // linear-extrapolation boundary, per side (left, right, bottom, top) and
  // ghost layer (1, 2), from the reference's cell centres
  float gr[8];
};

// element offset of frame column i within a row
__host__ __device__ inline int col_pos(const KInst& I, int i) { return col_pos(I.slot_rows, I.nslot, i); }

// Tile height of an instance for a requested height ty.  The FP32 step
// addresses a tile's frame rows y0-1 .. y1+3 through buffer descriptors with
// 32-bit byte offsets, so those rows of one plane must span less than 2^31
// bytes; the FP64 kernels index with 64-bit offsets.
inline int tile_rows_for(const KInst& I, int ty) {
  if (!I.slot_rows) return ty;
  const long long cap = 0x7fffffffLL / (4LL * I.pitch) - 5;
  return (int)std::max(1LL, std::min((long long)ty, cap));
}

// time_control fields (MOD_time_control.f90:16-49), FP64.
struct TimeCtl {
  double time, time_initial, time_final, dt, dt_previous, dt_max, dt_min;
  double time_precision, CFL, CFL_target, CFL_previous, shrink, growth, emergency;
  int adaptive, snapshot_exception, snap_idx, n_snap;
  int ended;  // end_program reached
  int pad;
  const double* snaps;  // snapshot_intervals (device or host pointer)
};

// FP32 engine: one of kSlots partial accumulators of an instance's attempt
// (tile t of the instance uses slot t % kSlots), one per 64-byte line so
// that the tiles' atomics spread over many addresses.
constexpr int kSlots = 64;
struct TileSlot {
  unsigned int mx, my;  // maxima of the sweep speeds (float bits)
  int bad;              // error bits
  unsigned int count;   // tiles of this slot done in the running attempt
  unsigned int pad[12];
};

// Device-resident per-instance state.
struct InstState {
  TimeCtl tc;
  unsigned long long smax_x, smax_y;  // running max of the sweep speeds (double bits)
  double last_smax_x, last_smax_y;    // maxima of the last completed attempt
  double cfl_last;
  long long accepted, rejected;
  long long max_accepted;             // stop after this many accepted steps (0 = no limit)
  unsigned int tiles_done;            // completion counter of the running attempt
  unsigned int smax32_x, smax32_y;    // FP32 engine: running maxima of the sweep speeds (float bits)
  int cur;                            // index of the current buffer
  int status;                         // PCW_STATUS_*
  int error;                          // PCW_ERR_* when status is ERROR
  int bad;                            // set by tiles: bit 1 NaN/complex eigenvalue in the xi sweep, bit 2 in the eta sweep
  int last_bad;                       // bits of the last completed attempt
  int controlled;                     // 1: the last tile runs the controller
  int last_ok;                        // staged engine: 1 when the last attempt was kept
  int ghosts_due;                     // FP64 fused engine: 1 when the attempt just completed was accepted
  int stop_snap;                      // 1: a device-controlled run stops the instance at each snapshot
  int log_n, log_cap;
  pcw_step_record* log;
};

// --------------------------------------------------------- controller
// initialize, MOD_time_control.f90:70-151 (snapshot table written by the caller)
__host__ __device__ inline void tc_init(TimeCtl& t, const pcw_time& s) {
  t.time_initial = s.t0;
  t.time_final = s.t_final;
  t.time = s.t0;
  t.dt = s.dt0;
  t.dt_previous = s.dt0;
  t.dt_max = s.dt_max;
  t.dt_min = s.dt_min;
  t.adaptive = s.adaptive;
  t.n_snap = s.n_snapshots < 1 ? 1 : s.n_snapshots;
  t.snap_idx = 1;
  t.CFL = 0.0;
  t.CFL_previous = 0.0;
  t.CFL_target = s.cfl_target;
  t.time_precision = 0.01;
  t.growth = s.growth;
  t.shrink = s.shrink;
  t.emergency = 0.25;
  t.snapshot_exception = 0;
  t.ended = 0;
}

// snapshot_intervals, MOD_time_control.f90:133-139
__host__ __device__ inline void tc_snapshots(const pcw_time& s, double* out) {
  int n = s.n_snapshots < 1 ? 1 : s.n_snapshots;
  double interval = (s.t_final - s.t0) / n;
  for (int i = 1; i <= n; ++i) out[i - 1] = s.t0 + i * interval;
  out[n - 1] = s.t_final;
}

// updateTime, MOD_time_control.f90:153-243
__host__ __device__ inline void tc_update_time(TimeCtl& t, int step_valid) {
  double change_factor, dt_step;
  t.dt_previous = t.dt;
  dt_step = t.dt;
  if (step_valid) {
    if (t.adaptive) {
      if (t.CFL > 0.0) change_factor = t.CFL_target / t.CFL;
      else change_factor = 1.0;
      if (change_factor >= t.growth) change_factor = t.growth;
      else if (fabs(change_factor - 1.0) <= t.time_precision) change_factor = 1.0;
      dt_step = t.dt_previous * change_factor;
    }
    double snap = t.snaps[t.snap_idx - 1];
    if (t.time + dt_step > snap) {
      dt_step = snap - t.time;
      t.snapshot_exception = 1;
    } else {
      t.snapshot_exception = 0;
    }
  } else {
    if (!t.adaptive) { t.ended = 1; return; }
    dt_step = t.emergency * t.dt;
  }
  if (dt_step > t.dt_max) dt_step = t.dt_max;
  if (dt_step < t.dt_min) {
    if (!t.adaptive && !t.snapshot_exception) { t.ended = 1; return; }
    dt_step = t.dt_min;
  }
  if (step_valid) t.time = t.time + dt_step;
  t.dt = dt_step;
}

__host__ __device__ inline void tc_update_max_cfl(TimeCtl& t, double c) {
  t.CFL = t.CFL > c ? t.CFL : c;
}
__host__ __device__ inline int tc_valid(const TimeCtl& t) {
  return t.CFL_target / t.CFL >= t.shrink;
}
__host__ __device__ inline void tc_reset_cfl(TimeCtl& t) {
  t.CFL_previous = t.CFL;
  t.CFL = 0.0;
}

// One attempted step of the main loop (src/main.f90:201-277) given the two
// sweep maxima: CFL of each sweep as in MOD_solve.f90:81/134, the validity
// tests after each sweep (:109-120, :161-174), accept/reject, updateTime,
// snapshot index, reset_CFL.  Returns 1 when accepted.
__host__ __device__ inline int tc_attempt(TimeCtl& t, double smax_x, double smax_y,
                                          double dxi, double deta, pcw_step_record* rec) {
  const double dt = t.dt;
  double cx = smax_x * dt / dxi;
  double cy = smax_y * dt / deta;
  tc_update_max_cfl(t, cx);
  int valid = tc_valid(t);
  if (valid) {
    tc_update_max_cfl(t, cy);
    valid = tc_valid(t);
  }
  if (rec) {
    rec->accepted = valid;
    rec->snapshot = 0;
    rec->dt_used = dt;
    rec->cfl = t.CFL;
  }
  if (!valid) {
    tc_update_time(t, 0);
    tc_reset_cfl(t);
  } else {
    tc_update_time(t, 1);
    if (!t.ended && t.snap_idx <= t.n_snap && t.time >= t.snaps[t.snap_idx - 1]) {
      if (rec) rec->snapshot = t.snap_idx;
      t.snap_idx += 1;
    }
    tc_reset_cfl(t);
  }
  if (rec) {
    rec->time = t.time;
    rec->dt_next = t.dt;
  }
  return valid;
}

// Error semantics of one attempt with the device controller.  The reference
// stops on a NaN Jacobian or complex eigenvalues of a sweep it runs
// (SMOD_jakobian.f90:55-70, SMOD_eigenvalue_problem.f90:72-83), but it skips
// the eta sweep when the xi sweep alone already rejects the step
// (MOD_solve.f90:109-120).  Returns 1 when the instance stopped with an
// error; else *ok is the accept decision (tc_attempt).
__host__ __device__ inline int finish_checks(InstState* S, double sx, double sy, double dxi,
                                             double deta, int badbits, pcw_step_record* rec,
                                             int* ok) {
  if (badbits & 1) {
    S->status = PCW_STATUS_ERROR;
    S->error = PCW_ERR_STATE;
    return 1;
  }
  if (badbits & 2) {
    // is the step rejected by the xi sweep alone?
    TimeCtl probe = S->tc;
    tc_update_max_cfl(probe, sx * probe.dt / dxi);
    if (tc_valid(probe)) {
      S->status = PCW_STATUS_ERROR;
      S->error = PCW_ERR_STATE;
      return 1;
    }
  }
  *ok = tc_attempt(S->tc, sx, sy, dxi, deta, rec);
  return 0;
}

// Bookkeeping after a device-controlled attempt (accept counters already
// updated): the step log, then the instance's status.  It stops (DONE) when
// the time loop of src/main.f90:193 ends (time >= time_final), at the
// accepted-step limit of the run, or, when stop_snap is set, after the step
// that reached a snapshot time (main.f90:256-268: the host writes the
// snapshot, and the next run continues).
__host__ __device__ inline void after_attempt(InstState* S, const pcw_step_record& rec) {
  if (S->log && S->log_n < S->log_cap) S->log[S->log_n++] = rec;
  if (S->tc.ended) {
    S->status = PCW_STATUS_ERROR;
    S->error = PCW_ERR_CONTROL;
  } else if (!(S->tc.time < S->tc.time_final) ||
             (S->max_accepted > 0 && S->accepted >= S->max_accepted) ||
             (S->stop_snap && rec.accepted && rec.snapshot > 0)) {
    S->status = PCW_STATUS_DONE;
  }
}

}  // namespace pcw
