// pcw_tile.hpp -- end of a fused-step wave: publish the wave's speed maxima
// and error bits, and let the last tile of an instance finish the attempt.
#pragma once

#include "pcw_device.hpp"
#include "pcw_engine.hpp"

namespace pcw {

// mx, my: the wave's maxima of the xi / eta interface speeds (already reduced
// over the wave); bad_x / bad_y: a lane of the wave met a NaN Jacobian or a
// complex eigenvalue in that sweep.  The instance's last tile (atomic
// counter) reads the instance maxima, resets the accumulators and, under
// device control, runs the time controller: the CFL bookkeeping and validity
// tests of src/solver/MOD_solve.f90:60-175, updateTime of
// src/time_control/MOD_time_control.f90:153-243, the buffer flip on accept.
//
// Ordering without fences: a device-scope fence on this GPU writes back the
// issuing XCD's L2, which per wave costs more than the wave's work.  The
// partial results are device-scope atomics, and each wave waits for them to
// complete (their return values are consumed) before it increments the tile
// counter, so by the time the last tile's increment returns, every other
// tile's maxima and error bits are in place at the atomics' coherence point,
// where the last tile reads them with atomics.  The grid's stores of the new
// state need no ordering here: the next launch is what reads them.
__device__ __forceinline__ void tile_finish(InstState* S, const KInst& I, int cur, double mx,
                                            double my, bool bad_x, bool bad_y, int lane) {
  unsigned int prev = 0;
  if (lane == 0) {
    const unsigned long long o1 =
        atomicMax(&S->smax_x, (unsigned long long)__double_as_longlong(mx));
    const unsigned long long o2 =
        atomicMax(&S->smax_y, (unsigned long long)__double_as_longlong(my));
    int o3 = 0;
    if (bad_x | bad_y) o3 = atomicOr(&S->bad, (bad_x ? 1 : 0) | (bad_y ? 2 : 0));
    asm volatile("" ::"v"(o1), "v"(o2), "v"(o3));  // wait until the three have completed
    prev = atomicAdd(&S->tiles_done, 1u);
  }
  prev = __shfl(prev, 0, 64);
  if (prev + 1 != (unsigned)I.ntiles) return;
  // last tile of this instance: finish the attempt
  if (lane != 0) return;
  // read and reset the accumulators for the next attempt
  const double sx = __longlong_as_double((long long)atomicExch(&S->smax_x, 0ull));
  const double sy = __longlong_as_double((long long)atomicExch(&S->smax_y, 0ull));
  const int badbits = atomicExch(&S->bad, 0);
  atomicExch(&S->tiles_done, 0u);
  S->last_smax_x = sx;
  S->last_smax_y = sy;
  S->last_bad = badbits;
  if (!S->controlled) return;  // host control: the host reads last_smax_* and last_bad
  pcw_step_record rec;
  int ok;
  if (finish_checks(S, sx, sy, I.dxi, I.deta, badbits, &rec, &ok)) return;
  S->cfl_last = rec.cfl;
  if (ok) {
    S->cur = cur ^ 1;
    S->accepted += 1;
  } else {
    S->rejected += 1;
  }
  if (S->log && S->log_n < S->log_cap) S->log[S->log_n++] = rec;
  if (S->tc.ended) {
    S->status = PCW_STATUS_ERROR;
    S->error = PCW_ERR_CONTROL;
  } else if (!(S->tc.time < S->tc.time_final) ||
             (S->max_accepted > 0 && S->accepted >= S->max_accepted)) {
    S->status = PCW_STATUS_DONE;
  }
}

}  // namespace pcw
