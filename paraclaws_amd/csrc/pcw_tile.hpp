// pcw_tile.hpp -- end of a fused-step wave: publish the wave's speed maxima
// and error bits, and let the last tile of an instance finish the attempt
// (one level for the FP64 kernel, two levels of accumulators for the FP32
// kernel).
#pragma once

#include "pcw_device.hpp"
#include "pcw_engine.hpp"

namespace pcw {

// Completion of one attempt of an instance from its speed maxima sx, sy and
// error bits (bit 1: NaN Jacobian or complex eigenvalue in the xi sweep, bit
// 2: in the eta sweep): record them for the host and, under device control,
// run the time controller -- the CFL bookkeeping and validity tests of
// src/solver/MOD_solve.f90:60-175, updateTime of
// src/time_control/MOD_time_control.f90:153-243 and the buffer flip on accept.
// cur is the buffer index the attempt read.
__device__ __forceinline__ void finish_attempt(InstState* S, const KInst& I, int cur, double sx,
                                               double sy, int badbits) {
  S->last_smax_x = sx;
  S->last_smax_y = sy;
  S->last_bad = badbits;
  if (!S->controlled) return;  // host control: the host reads last_smax_* and last_bad
  S->ghosts_due = 0;
  pcw_step_record rec;
  int ok;
  if (finish_checks(S, sx, sy, I.dxi, I.deta, badbits, &rec, &ok)) return;
  S->cfl_last = rec.cfl;
  S->ghosts_due = ok;
  if (ok) {
    S->cur = cur ^ 1;
    S->accepted += 1;
  } else {
    S->rejected += 1;
  }
  after_attempt(S, rec);
}

// mx, my: the wave's maxima of the xi / eta interface speeds (already reduced
// over the wave); bad_x / bad_y: a lane of the wave met a NaN Jacobian or a
// complex eigenvalue in that sweep.  The instance's last tile (atomic
// counter) reads the instance maxima, resets the accumulators and completes
// the attempt (finish_attempt).
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
  finish_attempt(S, I, cur, sx, sy, badbits);
}

// FP32 engine: read and reset the instance accumulators, complete the attempt
__device__ __forceinline__ void complete_f32(InstState* S, const KInst& I, int cur) {
  const float sx = __builtin_bit_cast(float, atomicExch(&S->smax32_x, 0u));
  const float sy = __builtin_bit_cast(float, atomicExch(&S->smax32_y, 0u));
  int bad = atomicExch(&S->bad, 0);
  atomicExch(&S->tiles_done, 0u);
  if (sx != sx) bad |= 1;  // a NaN speed (NaN momentum) is reported like a NaN Jacobian
  if (sy != sy) bad |= 2;
  finish_attempt(S, I, cur, (double)sx, (double)sy, bad);
}

// Instances of at most kDirectTiles tiles complete on one level: the two
// extra dependent atomic round trips of the second level cost about 1 us per
// step, more than that many tiles' atomics on one address (at 256 the
// 512-instance ensemble, most of whose instances are that small, ran 0.4 %
// slower).
constexpr int kDirectTiles = 128;

// Two-level completion (FP32 engine, many tiles per instance): tile t of the
// instance (t = 0 .. ntiles-1) publishes its maxima (float bits; as unsigned
// integers a NaN wins) and error bits to slot t % kSlots, waits for those
// atomics and counts itself in the slot.  The slot's last tile moves the
// slot's results to the instance accumulators the same way and counts the
// slot; the last slot's last tile completes the attempt.  Every address sees
// at most ntiles / kSlots + 1 tiles' atomics, instead of all of them.
__device__ __forceinline__ void tile_finish_slots(InstState* S, const KInst& I, TileSlot* slots,
                                                  int t, int cur, float mx, float my, int badbits,
                                                  int lane) {
  if (I.ntiles <= kDirectTiles) {
    // few tiles: one level, every tile on the instance's accumulators
    unsigned int prev = 0;
    if (lane == 0) {
      const unsigned o1 = atomicMax(&S->smax32_x, __builtin_bit_cast(unsigned, mx));
      const unsigned o2 = atomicMax(&S->smax32_y, __builtin_bit_cast(unsigned, my));
      int o3 = 0;
      if (badbits) o3 = atomicOr(&S->bad, badbits);
      asm volatile("" ::"v"(o1), "v"(o2), "v"(o3));  // wait until the three have completed
      prev = atomicAdd(&S->tiles_done, 1u);
    }
    prev = __shfl(prev, 0, 64);
    if (prev + 1 != (unsigned)I.ntiles) return;
    if (lane != 0) return;
    complete_f32(S, I, cur);
    return;
  }
  const int nslots = I.ntiles < kSlots ? I.ntiles : kSlots;
  const int s = t % kSlots;
  const unsigned per = (unsigned)(I.ntiles / kSlots + (s < I.ntiles % kSlots ? 1 : 0));
  TileSlot* L = slots + s;
  unsigned int prev = 0;
  if (lane == 0) {
    const unsigned o1 = atomicMax(&L->mx, __builtin_bit_cast(unsigned, mx));
    const unsigned o2 = atomicMax(&L->my, __builtin_bit_cast(unsigned, my));
    int o3 = 0;
    if (badbits) o3 = atomicOr(&L->bad, badbits);
    asm volatile("" ::"v"(o1), "v"(o2), "v"(o3));  // wait until the three have completed
    prev = atomicAdd(&L->count, 1u);
  }
  prev = __shfl(prev, 0, 64);
  if (prev + 1 != per) return;
  if (lane != 0) return;
  // last tile of the slot: move the slot's results to the instance
  const unsigned smx = atomicExch(&L->mx, 0u);
  const unsigned smy = atomicExch(&L->my, 0u);
  const int sbad = atomicExch(&L->bad, 0);
  atomicExch(&L->count, 0u);
  const unsigned p1 = atomicMax(&S->smax32_x, smx);
  const unsigned p2 = atomicMax(&S->smax32_y, smy);
  int p3 = 0;
  if (sbad) p3 = atomicOr(&S->bad, sbad);
  asm volatile("" ::"v"(p1), "v"(p2), "v"(p3));
  if (atomicAdd(&S->tiles_done, 1u) + 1 != (unsigned)nslots) return;
  complete_f32(S, I, cur);  // last slot: complete the attempt
}

}  // namespace pcw
