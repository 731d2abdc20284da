// pcw_fused.hip -- FP64 step kernel: xi sweep, eta sweep and source term of
// one time step, for any number of instances at once.  The physical ghost
// fill of the new state follows in a small kernel of its own (k_ghosts), so
// that this one stays within 168 VGPRs (three waves per SIMD; four, within
// 128 VGPRs, for instances whose mesh metrics are exactly 1).
//
// Mapping (gfx950, wave64): a wave owns a strip of 60 output columns plus
// two halo lanes on each side and marches down the rows of its tile.  Each
// row is loaded once (coalesced SoA h/hu/hv), swept in xi with lane shifts
// through the LDS crossbar carrying the neighbour states, and fed into a
// register window that runs the eta sweep two rows behind.  The eta sweep reads the
// xi-updated interior and the ghost rows as they were before the step
// (the reference refreshes ghosts once per step, src/main.f90:246).
// Results go to the instance's other buffer (ping-pong), so a rejected
// attempt is rolled back by not flipping the buffer index.
//
// Reference stages fused here (src/solver/MOD_solve.f90:23-175 and callees):
// wave_decomposition/HLL, split_flux, integrate_flux_difference, limiter,
// second_order_flux_correction, integrator_flux_correction_increment, the
// CFL max (as max wave speed) and integrate_source_term.
#include "pcw_device.hpp"
#include "pcw_engine.hpp"
#include "pcw_kernels.hpp"
#include "pcw_tile.hpp"

namespace pcw {

// Neighbour lanes in xi through the LDS crossbar (ds_bpermute, two dwords
// per double) rather than DPP, whose wave-shift forms issue at half the rate
// of arithmetic on the VALU.  addr = 4 * source lane.  Lane 0 receives lane
// 63's value and lane 63 lane 0's: both are halo lanes, whose results are
// never stored and never enter a speed maximum or an error flag.
__device__ __forceinline__ double bperm64(int addr, double v) {
  const uint64_t u = __builtin_bit_cast(uint64_t, v);
  const int lo = __builtin_amdgcn_ds_bpermute(addr, (int)(uint32_t)u);
  const int hi = __builtin_amdgcn_ds_bpermute(addr, (int)(uint32_t)(u >> 32));
  return __builtin_bit_cast(double, ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}

template <int LIM, int MET>
__device__ __forceinline__ void fused_step(const FusedArgs<double>& a) {
  using T = double;
  const int lane = threadIdx.x & 63;
  const int wpb = blockDim.x >> 6;
  // XCD-aware order: block b runs on XCD b % 8; give each XCD a contiguous
  // range of the live tiles so neighbouring strips share that XCD's L2.
  // Blocks beyond the live list (the grid was sized for a longer one) exit.
  // The launch takes the nlive tiles from list position base on.
  const int nlive = __builtin_amdgcn_readfirstlane(a.live[0]);
  const int base = __builtin_amdgcn_readfirstlane(a.live[1]);
  const int nb = ((nlive + wpb - 1) / wpb + 7) & ~7, per = nb >> 3;
  if ((int)blockIdx.x >= nb) return;
  const int bb = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
  const int lt = __builtin_amdgcn_readfirstlane(bb * wpb + (threadIdx.x >> 6));
  if (lt >= nlive) return;
  const int tid = base + lt;  // position in the tile list
  const int4 t = a.tiles[tid];
  const int in = __builtin_amdgcn_readfirstlane(t.x);
  const KInst& I = a.inst[in];
  InstState* S = a.st + in;
  if (MET != metric_mode(I)) return;  // each kind of metric handling has a launch of its own
  if (S->status != PCW_STATUS_RUNNING) {
    // no attempt this launch: the ghost fill that follows must not run
    if (tid == I.tile0 && lane == 0) S->ghosts_due = 0;
    return;
  }

  const int nx = I.nx, ny = I.ny, pitch = I.pitch;
  const long long plane = I.plane;
  const int cur = S->cur;
  const double dt = S->tc.dt;
  const T* __restrict__ qin = a.pool + I.off[cur];
  T* __restrict__ qout = a.pool + I.off[cur ^ 1];
  const T* __restrict__ mq = MET == kMetArrays ? a.mpool + I.moff : nullptr;  // kappa, gamma_xi, gamma_eta planes
  const int x0 = t.y, y0 = t.z, y1 = t.w;
  const int i = x0 - 2 + lane;                 // this lane's column (1-based interior)
  const int adl = ((lane + 63) & 63) << 2;     // left neighbour lane (column i-1)
  const int adr = ((lane + 1) & 63) << 2;      // right neighbour lane (column i+1)
  auto from_left = [adl](double v) { return bperm64(adl, v); };   // value of column i-1
  auto from_right = [adr](double v) { return bperm64(adr, v); };  // value of column i+1
  const int fc = i + 1 < nx + 3 ? i + 1 : nx + 3;  // frame column, clamped for loads
  const bool out_lane = lane >= 2 && lane <= 61 && i >= 1 && i <= nx;
  const bool xface = lane >= 1 && i >= 0 && i <= nx + 1;  // interface i-1/2 in the xi CFL range
  const bool xlim = i >= 1 && i <= nx;         // interface limited (Q7 of the reference)
  const bool check_cell = i >= -1 && i <= nx + 1;

  const double dtod_x = dt / I.dxi, dtod_y = dt / I.deta;
  const T g = a.g, dry = a.dry;
  Phys P;
  P.g = a.g;
  P.eps = a.eps;
  P.dry = a.dry;
  Mesh M;
  M.xi_min = I.xi_min;
  M.eta_min = I.eta_min;
  M.dxi = I.dxi;
  M.deta = I.deta;
  const int src = I.source;
  const bool do_src = (src != PCW_SRC_NONE) || a.wetdry;
  const T n2 = (T)(I.manning * I.manning);
  const T* __restrict__ bb_ = (I.boff >= 0) ? a.bpool + I.boff : nullptr;
  const bool use_b = bb_ && (src == PCW_SRC_BATHY || src == PCW_SRC_BATHY_FRICTION);
  using MT = double;
  const MT gam_eta_col = MET == kMetUnit ? MT(1) : gamma_eta_exact(M, i);

  T smx = T(0), smy = T(0);
  int bad = 0;   // bit 1: NaN/complex in the xi sweep, bit 2: in the eta sweep

  // eta-sweep register window
  T yq[3] = {T(0), T(0), T(0)};      // row r-1 (xi-updated)
  CellQ<T> yc;                        // eta physics of row r-1
  yc.s = T(0); yc.f0 = yc.f1 = yc.f2 = T(0); yc.bad = 0;
  Waves<T> Wa, Wb;                    // interfaces r-1, r-2 (unlimited)
  Wa.s2 = Wb.s2 = T(0);
  for (int k = 0; k < 3; ++k) { Wa.w1[k] = Wa.w2[k] = Wb.w1[k] = Wb.w2[k] = T(0); }
  T Fb[3] = {T(0), T(0), T(0)};       // correction flux at interface r-2
  T qpp[3] = {T(0), T(0), T(0)};      // row r-2 (xi-updated)
  MT kap_p = MT(1), kap_pp = MT(1);   // kappa(i, r-1), kappa(i, r-2)
  MT gy_p = MT(1), gy_pp = MT(1);     // gamma_eta(i, r-1), gamma_eta(i, r-2) (metric arrays)
  T bm = T(0), b0 = T(0);             // bathymetry rows r-3, r-2 of this column

  // prefetch first row
  long long roff = (long long)(y0 - 2 + 1) * pitch + fc;
  T nh = qin[roff], nhu = qin[plane + roff], nhv = qin[2 * plane + roff];
  T nk = T(1), ngx = T(1), ngy = T(1);
  if (MET == kMetArrays) {
    nk = mq[roff];
    ngx = mq[plane + roff];
    ngy = mq[2 * plane + roff];
  }

  for (int r = y0 - 2; r <= y1 + 2; ++r) {
    T q[3] = {nh, nhu, nhv};
    const T mk_r = nk, mgx_r = ngx, mgy_r = ngy;
    if (r < y1 + 2) {  // prefetch the next row
      long long o = (long long)(r + 2) * pitch + fc;
      nh = qin[o];
      nhu = qin[plane + o];
      nhv = qin[2 * plane + o];
      if (MET == kMetArrays) {
        nk = mq[o];
        ngx = mq[plane + o];
        ngy = mq[2 * plane + o];
      }
    }
    T bnew = T(0);
    if (use_b && r >= y0) bnew = bb_[(long long)r * pitch + fc];  // b(i, r-1)

    const MT kap_r = MET == kMetArrays ? mk_r : (MET == kMetUnit ? MT(1) : kappa_exact(M, i, r));
    const MT gy_r = MET == kMetArrays ? (MT)mgy_r : gam_eta_col;
    T qx[3] = {q[0], q[1], q[2]};

    // ------------------------------------------------------ xi sweep
    if (r >= 1 && r <= ny) {
      CellQ<T> c;
      c = cell_exact<1>(q[0], q[1], q[2], P);
      bad |= (c.bad && check_cell) ? 1 : 0;
      T qL[3] = {from_left(q[0]), from_left(q[1]), from_left(q[2])};
      CellQ<T> cL;
      cL.s = from_left(c.s);
      cL.f0 = qL[1];
      cL.f1 = from_left(c.f1);
      cL.f2 = from_left(c.f2);
      Waves<T> w;
      w = hll_exact(qL, cL, q, c, P.eps);
      smx = (xface && w.s2 > smx) ? w.s2 : smx;
      T w1R[3], F[3] = {T(0), T(0), T(0)};
#pragma unroll
      for (int k = 0; k < 3; ++k) w1R[k] = from_right(w.w1[k]);
      // metric neighbours (all lane shifts in wave-uniform control flow)
      const MT kap_l = MET == kMetUnit ? MT(1) : from_left(kap_r);  // kappa(i-1, r)
      const MT gx_l = MET == kMetArrays ? mgx_r : (MET == kMetUnit ? MT(1) : gamma_xi_exact(M, r));
      const MT gx_r = MET == kMetArrays ? from_right(gx_l) : gx_l;    // gamma_xi(i+1, r)
      if (a.ho) {
        T w2L[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) w2L[k] = from_left(w.w2[k]);
        T p1 = T(1), p2 = T(1);
        if (LIM > 0 && xlim) {
          p1 = limit_exact<LIM>(w.w1, w1R, P.eps);
          p2 = limit_exact<LIM>(w.w2, w2L, P.eps);
        }
        const double al = w.s2;
        const double coef = 0.5 * al * (1.0 - al * dtod_x / (0.5 * (kap_l + kap_r)));
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          double t1 = coef * (p1 * w.w1[k]);
          double t2 = coef * (p2 * w.w2[k]);
          F[k] = t1 + t2;
        }
      }
      const T s2R = from_right(w.s2);
      T FR[3];
#pragma unroll
      for (int k = 0; k < 3; ++k) FR[k] = from_right(F[k]);
      if (i >= 1 && i <= nx) {
        const double fac = dtod_x / kap_r;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          double ap = w.s2 * w.w2[k];
          double am = (-s2R) * w1R[k];
          double q1 = q[k] - fac * (gx_l * ap + gx_r * am);
          qx[k] = a.ho ? q1 - dtod_x / kap_r * (FR[k] - F[k]) : q1;
        }
      }
    }

    // ------------------------------------------------------ eta sweep
    CellQ<T> cy;
    cy = cell_exact<2>(qx[0], qx[1], qx[2], P);
    // the eta sweep reads rows [-1, ny+1] of columns [1, nx]
    bad |= (cy.bad && out_lane && r <= ny + 1) ? 2 : 0;
    Waves<T> Wc;
    Wc.s2 = T(0);
#pragma unroll
    for (int k = 0; k < 3; ++k) { Wc.w1[k] = T(0); Wc.w2[k] = T(0); }
    if (r >= y0 - 1) {
      Wc = hll_exact(yq, yc, qx, cy, P.eps);
      smy = (out_lane && r >= 0 && r <= ny + 1 && Wc.s2 > smy) ? Wc.s2 : smy;
    }
    T Fa[3] = {T(0), T(0), T(0)};
    if (r >= y0 + 1 && a.ho) {  // correction flux at interface r-1
      T p1 = T(1), p2 = T(1);
      if (LIM > 0 && r - 1 >= 1 && r - 1 <= ny) {
        p1 = limit_exact<LIM>(Wa.w1, Wc.w1, P.eps);
        p2 = limit_exact<LIM>(Wa.w2, Wb.w2, P.eps);
      }
      const double al = Wa.s2;
      const double coef = 0.5 * al * (1.0 - al * dtod_y / (0.5 * (kap_pp + kap_p)));
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        double t1 = coef * (p1 * Wa.w1[k]);
        double t2 = coef * (p2 * Wa.w2[k]);
        Fa[k] = t1 + t2;
      }
    }
    if (r >= y0 + 2) {  // finish row j = r-2
      const int j = r - 2;
      T qn[3];
      const double fac = dtod_y / kap_pp;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        double ap = Wb.s2 * Wb.w2[k];
        double am = (-Wa.s2) * Wa.w1[k];
        double q1 = qpp[k] - fac * (gy_pp * ap + gy_p * am);
        qn[k] = a.ho ? q1 - dtod_y / kap_pp * (Fa[k] - Fb[k]) : q1;
      }
      if (do_src) {
        T bxm = from_left(b0), bxp = from_right(b0);
        apply_source<T>(qn[0], qn[1], qn[2], src, dry, g, (T)I.dxi, (T)I.deta, (T)dt, n2, bxm,
                        bxp, bm, bnew);
      } else {
        reference_source<T>(qn[0], qn[1], qn[2]);
      }
      const long long o = (long long)(j + 1) * pitch + (i + 1);
      if (out_lane) {
        // default cache policy: non-temporal stores made this arithmetic-bound
        // step 1 % slower (they speed up the memory-bound FP32 step)
        qout[o] = qn[0];
        qout[plane + o] = qn[1];
        qout[2 * plane + o] = qn[2];
      }
    }
    // shift the window
    Wb = Wa;
    Wa = Wc;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      Fb[k] = Fa[k];
      qpp[k] = yq[k];
      yq[k] = qx[k];
    }
    yc = cy;
    kap_pp = kap_p;
    kap_p = kap_r;
    gy_pp = gy_p;
    gy_p = gy_r;
    bm = b0;
    b0 = bnew;
  }

  // ------------------------------------------------ reductions + controller
  const T mx = wave_max(smx), my = wave_max(smy);
  const int bad_x = __any(bad & 1), bad_y = __any(bad & 2);
  tile_finish(S, I, cur, (double)mx, (double)my, bad_x != 0, bad_y != 0, lane);
}

template <int LIM, int MET>
__global__ void __launch_bounds__(256) k_fused_step(FusedArgs<double> a) {
  fused_step<LIM, MET>(a);
}

// Without the metric arithmetic the step fits 128 VGPRs: four waves per SIMD.
template <int LIM>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4)))
k_fused_step_unit(FusedArgs<double> a) {
  fused_step<LIM, kMetUnit>(a);
}

// ------------------------------------------------------------- launcher
template <int MET>
static hipError_t launch_lim(int lim, const FusedArgs<double>& a, int blocks, int threads,
                             hipStream_t s) {
  if constexpr (MET == kMetUnit) {
    switch (lim) {
      case 0: k_fused_step_unit<0><<<blocks, threads, 0, s>>>(a); break;
      case 1: k_fused_step_unit<1><<<blocks, threads, 0, s>>>(a); break;
      case 2: k_fused_step_unit<2><<<blocks, threads, 0, s>>>(a); break;
      case 3: k_fused_step_unit<3><<<blocks, threads, 0, s>>>(a); break;
      default: k_fused_step_unit<4><<<blocks, threads, 0, s>>>(a); break;
    }
  } else {
    switch (lim) {
      case 0: k_fused_step<0, MET><<<blocks, threads, 0, s>>>(a); break;
      case 1: k_fused_step<1, MET><<<blocks, threads, 0, s>>>(a); break;
      case 2: k_fused_step<2, MET><<<blocks, threads, 0, s>>>(a); break;
      case 3: k_fused_step<3, MET><<<blocks, threads, 0, s>>>(a); break;
      default: k_fused_step<4, MET><<<blocks, threads, 0, s>>>(a); break;
    }
  }
  return hipGetLastError();
}

hipError_t launch_fused_f64(int lim, int met_mode, const FusedArgs<double>& a, int blocks,
                            int threads, hipStream_t s) {
  switch (met_mode) {
    case kMetArrays: return launch_lim<kMetArrays>(lim, a, blocks, threads, s);
    case kMetUnit: return launch_lim<kMetUnit>(lim, a, blocks, threads, s);
    default: return launch_lim<kMetComputed>(lim, a, blocks, threads, s);
  }
}

}  // namespace pcw
