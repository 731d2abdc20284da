// pcw_fused_f32.hip -- the FP32 fused step: one kernel per time step (xi
// sweep, eta sweep, source term, physical ghost fill) for any number of
// instances, built for gfx950's memory system and VALU.
//
// Mapping (wave64): as in the FP64 fused kernel, a wave owns a strip of 60
// output columns plus two halo lanes on each side and marches down the rows
// of its tile; neighbour states in xi come from lane shifts through the LDS
// crossbar (ds_bpermute), and the eta sweep runs in a register window two
// rows behind the xi sweep.  Frame rows are in 64-element slots (col_pos in
// pcw_engine.hpp): a wave's 60 output columns fill one slot, so each finished
// row is one aligned 256-byte store per plane.
//
// What differs from the FP64 kernel:
//   * Fluctuations are combined per interface: G = A+dQ - F goes to the
//     cell on the interface's high side and H = A-dQ + F to the cell on its
//     low side, so a cell update is q - dt/d * (G_i + H_{i+1}) and xi needs
//     one lane shift of H instead of shifts of s2, w1 and F.
//   * The eta window's per-row values live in three rotating slots (row r,
//     r-1, r-2); the row loop is unrolled by three so that the rotation is a
//     renaming and costs no register moves.  Input rows are prefetched six
//     rows ahead (three with metric arrays or source terms) into a ring of
//     registers that rotates with the slots.
//   * Rows whose every condition holds (interior rows of interior strips,
//     away from the tile's first rows and the domain's top and bottom rows)
//     run a body without any of the range tests; the other rows, and the
//     strips on the domain's xi boundaries, run the general body.
//   * Divisions are reciprocals, square roots are v_sqrt_f32 (1 ulp).
//     Without metric arrays dt/d is folded into the correction
//     coefficients, which saves one instruction per interface and sweep.
//   * A cell with h <= 0 (or NaN) in a sweep's range is reported like the
//     reference's NaN-in-Jacobian / complex-eigenvalue stop
//     (SMOD_jakobian.f90:55-58, SMOD_eigenvalue_problem.f90:72-83), unless
//     the wet/dry cut-off is enabled.  It is found from the sweep's speed
//     maximum (such a cell's speed is +inf or NaN), with no test per cell.
//   * The kernel is VALU-issue bound as much as HBM bound (VALUBusy ~99 % at
//     ~0.87 of the bandwidth a plain 3-plane stream sustains), so
//     instruction count matters: 142 VALU instructions per row in the
//     unconditional body.
//
// Reference stages fused here (src/solver/MOD_solve.f90:23-175 and callees):
// wave_decomposition/HLL (SMOD_approximate_Riemann_solvers.f90:60-101),
// split_flux, integrate_flux_difference, the limiter
// (MOD_flux_correction.f90:54-123), second_order_flux_correction,
// integrator_flux_correction_increment, the CFL maximum (as the maximum wave
// speed), integrate_source_term and apply_physical_boundaries.
#include "pcw_device.hpp"
#include "pcw_engine.hpp"
#include "pcw_kernels.hpp"
#include "pcw_tile.hpp"

namespace pcw {
namespace f32 {

__device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float vsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

// max of two speeds (non-negative floats, +inf or NaN) as unsigned integers:
// one instruction, and +inf or a NaN of either sign wins instead of being
// dropped as by fmaxf.  A cell with h <= 0 or NaN has a speed |u| + c that
// is +inf or NaN (h = 0: u = hu * (1/0) is inf or NaN; h < 0: c = sqrt(g h)
// is NaN), so a maximum whose bits reach those of +inf reports the cell
// (kNotFinite) without a test per cell.
constexpr unsigned kNotFinite = 0x7f800000u;
__device__ __forceinline__ float smax2(float a, float b) {
  return __builtin_bit_cast(float, max(__builtin_bit_cast(unsigned, a), __builtin_bit_cast(unsigned, b)));
}
__device__ __forceinline__ float wave_smax(float v) {
  for (int o = 32; o >= 1; o >>= 1) v = smax2(v, __shfl_xor(v, o, 64));
  return v;
}

// Neighbour lanes in xi through the LDS crossbar (ds_bpermute): the shifts
// do not occupy the VALU, whose DPP forms issue at half the rate of plain
// FP32 arithmetic on gfx950.  Lane 0 receives lane 63 and vice versa; both
// are halo lanes, whose results are not stored.
__device__ __forceinline__ float bperm(int addr, float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_ds_bpermute(addr, __builtin_bit_cast(int, v)));
}

// Buffer descriptors over one field plane of a tile's frame rows y0-1 ..
// y1+3: row offsets go to the scalar offset, the lane's column to the 32-bit
// vector offset.  A lane whose vector offset is beyond the descriptor's size
// has its store dropped by the hardware range check, which masks stores
// without a branch.
using rsrc_t = __amdgpu_buffer_rsrc_t;
constexpr unsigned kOff = 0x80000000u;  // out-of-range vector offset
__device__ __forceinline__ rsrc_t make_rsrc(const float* base, int bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(base), 0, bytes, 0x00020000);
}
__device__ __forceinline__ float bload(rsrc_t r, unsigned voff, int soff) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, soff, 0));
}
// The new state is stored non-temporal (cache policy nt): the kernel never
// reads what it stores, and the next step reads it back only after the XCDs'
// L2s (4 MB each) have turned over many times (a step moves 1.6 GB at
// 8192^2).  With nt the 8192^2 step runs 4.5 % faster (391 -> 373 us).  Loads
// keep the default policy: neighbouring strips and tiles re-read their halos
// from L2 (nt loads made the step 518 us).
constexpr int kStoreNT = 2;  // gfx950 buffer-store cache policy bit nt
__device__ __forceinline__ void bstore(float v, rsrc_t r, unsigned voff, int soff) {
  __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, (int)voff, soff,
                                        kStoreNT);
}

// One row's cell values and the interface below it, carried by the eta window.
struct Slot {
  float q[3];              // xi-updated state of the row
  float s, f1, f2;         // eta speed |v| + c and flux components 1, 2 (component 0 is q[2])
  float kap, fac;          // capacity and dt/deta/capacity of the row (metric arrays)
  float gy;                // gamma_eta of the interface below the row (metric arrays)
  float s2, w1[3], w2[3];  // unlimited waves of the interface below the row
};

// A prefetched input row.
struct Pre {
  float h, hu, hv;
  float k, gx, gy;  // metric arrays
};

struct Ctx {
  rsrc_t qi[3], qo[3];  // input and output planes h, hu, hv
  rsrc_t mi[3];         // metric planes kappa, gamma_xi, gamma_eta (metric arrays)
  rsrc_t bi;            // bathymetry
  bool has_b;
  int p4;               // row pitch in bytes
  int rbase;            // frame row at the descriptors' origin (y0 - 1)
  unsigned lo;          // load offset: this lane's position in the slot row (clamped) * 4
  unsigned so;          // store offset: position * 4 (whole slot, or out of range off the output lanes)
  int lane, i, nx, ny, y0, y1;
  int bl, br;           // ds_bpermute byte addresses of the left / right neighbour lane
  int src, wetdry, x0, l_nx;
  int nslot;            // interior slots of a frame row (slot rows)
  int bc[4];
  float gr[8];       // extrapolation ratios of the linear-extrapolation boundary
  float dtdx, dtdy, eps, g, dry;
  float sx, sy, fk;  // source-term constants of the step (see source)
  bool out_lane, xface, xlim;
};
__device__ __forceinline__ float lft(const Ctx& C, float v) { return bperm(C.bl, v); }
__device__ __forceinline__ float rgt(const Ctx& C, float v) { return bperm(C.br, v); }

__device__ __forceinline__ float readlane(float v, int l) {
  return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), l));
}

// ghost value of one field from the first (near) and second (far) interior
// cells next to a boundary: zero gradient (2), linear extrapolation (3) or
// reflective wall (4, the normal momentum negated); layer kk = 1 or 2
__device__ __forceinline__ float ghost(int bc, float near, float far, int kk, bool normal,
                                       float ratio) {
  if (bc == 2) return near;
  if (bc == 3) return fmaf(far - near, ratio, near);
  const float v = kk == 1 ? near : far;
  return normal ? -v : v;
}

// per-lane maxima of the interface speeds; smy over all lanes, masked to the
// output lanes at the end of the tile
struct Red {
  float smx, smy;
};

// eta-direction (DIR 2) or xi-direction (DIR 1) speed and flux of one cell
template <int DIR, bool WET>
__device__ __forceinline__ void cell(float h, float hu, float hv, float g, float dry, float& s,
                                     float& f1, float& f2) {
  // normal velocity un; the transverse flux is un * (transverse momentum)
  const float rh = rcp(h);
  float un = (DIR == 1 ? hu : hv) * rh;
  const float gh = g * h;
  float gc = gh;
  if (WET) {
    un = !(h > dry) ? 0.0f : un;
    gc = gh > 0.0f ? gh : 0.0f;
  }
  const float c = vsqrt(gc);
  s = fabsf(un) + c;
  if (DIR == 1) {
    f1 = fmaf(hu, un, gh * h);
    f2 = un * hv;
  } else {
    f1 = un * hu;
    f2 = fmaf(hv, un, gh * h);
  }
}

// HLL with the reference's symmetric speeds s1 = -s2 and regularised
// denominator D = eps + s1 - s2 (SMOD_approximate_Riemann_solvers.f90:
// 189-224): with dq = qR - qL and df = fR - fL, the middle state's waves are
//   w1 = qs - qL = (df - s2 dq - eps qL) / D,  w2 = qR - qs = -(df + s2 dq - eps qR) / D.
// The eps q terms (below FP32 resolution next to the others, and exactly
// what the reference leaks into waves of a quiescent state) are dropped, so
// a state at rest has exactly zero waves, and then w1 + w2 = dq * 2 s2 /
// (2 s2 - eps), which is dq to FP32 resolution: w2 = dq - w1.
// DRY: the state may hold dry cells (speed 0), so waves are zero when
// s2 <= eps, through a zero reciprocal: no branch; the speed is returned as
// it is, and at s2 <= eps (1e-10) its share of the CFL maximum is below any
// live interface's.  Without dry cells s2 <= eps needs h < 1e-21 on both
// sides, and the waves are the formula's, as in the reference.
template <bool DRY>
__device__ __forceinline__ float hll(const float qL[3], const float fL[3], float sL,
                                     const float qR[3], const float fR[3], float sR, float eps,
                                     float w1[3], float w2[3]) {
  const float s2 = smax2(sL, sR);
  const bool live = !DRY || s2 > eps;
  const float rD = live ? rcp(fmaf(-2.0f, s2, eps)) : 0.0f;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float dq = qR[k] - qL[k], df = fR[k] - fL[k];
    w1[k] = fmaf(-s2, dq, df) * rD;
    w2[k] = DRY ? fmaf(dq, live ? 1.0f : 0.0f, -w1[k]) : dq - w1[k];
  }
  return s2;
}

// Half the limiter function, phi(theta) / 2, of one wave
// (MOD_flux_correction.f90:76-120; 1 minmod, 2 superbee, 3 van Leer, 4 MC):
// theta = (upwind wave . wave) / |wave|^2, and a wave with |wave|^2 <= eps is
// left unlimited (phi = 1).  In halves, the bounds of minmod, superbee and MC
// are the [0, 1] clamp that gfx950 applies as an output modifier of the
// instruction producing the value; the caller doubles the flux coefficient
// instead, which is exact.
__device__ __forceinline__ float clamp01(float x) { return __builtin_amdgcn_fmed3f(x, 0.0f, 1.0f); }
template <int LIM>
__device__ __forceinline__ float half_phi(const float w[3], const float nb[3], float eps) {
  const float n = fmaf(w[2], w[2], fmaf(w[1], w[1], w[0] * w[0]));
  const float d = fmaf(nb[2], w[2], fmaf(nb[1], w[1], nb[0] * w[0]));
  const float t = d * rcp(n);
  float p;
  if (LIM == 1) {
    p = 0.5f * clamp01(t);                                   // max(0, min(1, t)) / 2
  } else if (LIM == 2) {
    p = fmaxf(fminf(0.5f, t), clamp01(0.5f * t));            // max(0, min(1, 2t), min(2, t)) / 2
  } else if (LIM == 3) {
    const float a = fabsf(t);
    p = (0.5f * (t + a)) * rcp(1.0f + a);                    // (t + |t|) / (1 + |t|) / 2
  } else {
    p = clamp01(fminf(fmaf(0.25f, t, 0.25f), t));            // max(0, min((1+t)/2, 2, 2t)) / 2
  }
  return n <= eps ? 0.5f : p;
}

// correction flux of one interface: coef (p1 w1 + p2 w2) with
// coef = |s| (1 - |s| dt/d / kbar) and the halved limiter values p1, p2
// (the reference's 0.5 |s| (1 - |s| dt/d / kbar) phi)
__device__ __forceinline__ void corr_flux(float coef, float p1, float p2, const float w1[3],
                                          const float w2[3], float F[3]) {
  const float c1 = coef * p1, c2 = coef * p2;
#pragma unroll
  for (int k = 0; k < 3; ++k) F[k] = fmaf(c1, w1[k], c2 * w2[k]);
}

// source term of one finished cell (wet/dry cut-off, bathymetry slope,
// semi-implicit Manning friction), FP32 with reciprocals.  The instance's
// constants come folded: sx = -dt g / dxi, sy = -dt g / deta (slope, times
// the central difference bx = b(i+1) - b(i-1), by = b(j+1) - b(j-1)), and
// fk = dt g n^2 (friction: hu / (1 + fk |u| h^(-4/3))).
__device__ __forceinline__ void source(float& h, float& hu, float& hv, int kind, float dry,
                                       float sx, float sy, float fk, float bx, float by) {
  if (!(h > dry)) {
    h = h > 0.0f ? h : 0.0f;
    hu = 0.0f;
    hv = 0.0f;
    return;
  }
  float nhu = hu, nhv = hv;
  if (kind == PCW_SRC_BATHY || kind == PCW_SRC_BATHY_FRICTION) {
    nhu = fmaf(sx * bx, h, hu);
    nhv = fmaf(sy * by, h, hv);
  }
  if (kind == PCW_SRC_FRICTION || kind == PCW_SRC_BATHY_FRICTION) {
    const float rh = rcp(h);
    const float u = hu * rh, v = hv * rh;
    const float sp = vsqrt(fmaf(u, u, v * v));
    const float hm43 = __builtin_amdgcn_exp2f(-1.33333333f * __builtin_amdgcn_logf(h));  // h^(-4/3)
    const float rden = rcp(fmaf(fk, sp * hm43, 1.0f));
    nhu = nhu * rden;
    nhv = nhv * rden;
  }
  hu = nhu;
  hv = nhv;
}

template <int LIM, bool HO, bool MET, int PHYS, bool GEN>
__device__ __forceinline__ void finish(const Ctx& C, int r, const Slot& N, const Slot& A,
                                       const Slot& B, float acc[3], float bm, float b0,
                                       float bnew, float finm[3]);

// Input rows are prefetched kPF rows ahead, through a ring of kPF buffers
// (6 for the plain variants; 3 where metric arrays or source terms take the
// registers).
template <bool MET, int PHYS>
struct Ring {
  static constexpr int kPF = (MET || PHYS) ? 3 : 6;
};

// xi sweep of row r: the xi-updated state qx of the row's cells from the
// prefetched input P (P itself for rows outside 1..ny in the general body).
template <int LIM, bool HO, bool MET, int PHYS, bool GEN>
__device__ __forceinline__ void xi_row(const Ctx& C, int r, const Pre& P, Red& R, float qx[3]) {
  const float q[3] = {P.h, P.hu, P.hv};
  const float mk = P.k, mgx = P.gx;
  constexpr bool wet = PHYS == 2;
  qx[0] = q[0];
  qx[1] = q[1];
  qx[2] = q[2];
  if (!GEN || (r >= 1 && r <= C.ny)) {
    float s, f1, f2;
    if (wet) cell<1, true>(q[0], q[1], q[2], C.g, C.dry, s, f1, f2);
    else cell<1, false>(q[0], q[1], q[2], C.g, C.dry, s, f1, f2);
    const float qL[3] = {lft(C, q[0]), lft(C, q[1]), lft(C, q[2])};
    const float fL[3] = {qL[1], lft(C, f1), lft(C, f2)};
    const float fR[3] = {q[1], f1, f2};
    const float sL = lft(C, s);
    float w1[3], w2[3];
    const float s2 = hll<(PHYS != 0)>(qL, fL, sL, q, fR, s, C.eps, w1, w2);
    R.smx = smax2(R.smx, (!GEN || C.xface) ? s2 : 0.0f);
    float p1 = 0.5f, p2 = 0.5f;  // halved limiter values (unlimited: 1 / 2)
    if (HO && LIM > 0) {
      const float w1R[3] = {rgt(C, w1[0]), rgt(C, w1[1]), rgt(C, w1[2])};
      const float w2L[3] = {lft(C, w2[0]), lft(C, w2[1]), lft(C, w2[2])};
      p1 = half_phi<LIM>(w1, w1R, C.eps);
      p2 = half_phi<LIM>(w2, w2L, C.eps);
      if (GEN) {
        p1 = C.xlim ? p1 : 0.5f;
        p2 = C.xlim ? p2 : 0.5f;
      }
    }
    if (HO && !MET) {
      // dt/dxi folded into the coefficients: nu = s dt/dxi, the correction
      // coefficients c = nu (1 - nu) p, and per field
      //   dt/dxi H = (c1 - nu) w1 + c2 w2     (to the interface's low cell)
      //   q - dt/dxi G = q + c1 w1 - (nu - c2) w2
      // (the waves die before the shifted values arrive)
      const float nu = C.dtdx * s2, cf = fmaf(-nu, nu, nu);
      const float c1 = cf * p1, c2 = cf * p2, e = c1 - nu, am = nu - c2;
      float loc[3], hr[3];
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        loc[k] = fmaf(-am, w2[k], fmaf(c1, w1[k], q[k]));
        hr[k] = rgt(C, fmaf(e, w1[k], c2 * w2[k]));
      }
#pragma unroll
      for (int k = 0; k < 3; ++k) qx[k] = loc[k] - hr[k];
      return;
    }
    float F[3] = {0.0f, 0.0f, 0.0f};
    const float kap_l = MET ? shr1(mk) : 1.0f;
    if (HO) {
      // s (1 - s dt/dxi / kbar)
      const float coef = s2 * fmaf(-s2 * C.dtdx, rcp(0.5f * (kap_l + mk)), 1.0f);
      corr_flux(coef, p1, p2, w1, w2, F);
    }
    const float sg = MET ? mgx * s2 : s2;
    float G[3], HR[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      G[k] = fmaf(sg, w2[k], -F[k]);
      HR[k] = rgt(C, fmaf(-sg, w1[k], F[k]));
    }
    const float fac = MET ? C.dtdx * rcp(mk) : C.dtdx;
#pragma unroll
    for (int k = 0; k < 3; ++k) qx[k] = fmaf(-fac, G[k] + HR[k], q[k]);
  }
}

// load row r+kPF (clamped to the tile's last row) into P
template <bool MET, int PHYS>
__device__ __forceinline__ void prefetch(const Ctx& C, int r, Pre& P) {
  const int so = (min(r + Ring<MET, PHYS>::kPF, C.y1 + 2) + 1 - C.rbase) * C.p4;
  P.h = bload(C.qi[0], C.lo, so);
  P.hu = bload(C.qi[1], C.lo, so);
  P.hv = bload(C.qi[2], C.lo, so);
  if (MET) {
    P.k = bload(C.mi[0], C.lo, so);
    P.gx = bload(C.mi[1], C.lo, so);
    P.gy = bload(C.mi[2], C.lo, so);
  }
}

// eta sweep of row r from its xi-updated state qx: the row's eta cell
// values into slot N, the face below it (rows r-1 | r), and from row y0+1
// on the finishing of face r-1 (finish).  mk, mgy: the row's metric values.
template <int LIM, bool HO, bool MET, int PHYS, bool GEN>
__device__ __forceinline__ void eta_row(const Ctx& C, int r, const float qx[3], float mk,
                                        float mgy, Slot& N, Slot& A, Slot& B, float acc[3],
                                        Red& R, float& bm, float& b0, float bnew, float finm[3]) {
  constexpr bool wet = PHYS == 2;
#pragma unroll
  for (int k = 0; k < 3; ++k) N.q[k] = qx[k];
  if (wet) cell<2, true>(qx[0], qx[1], qx[2], C.g, C.dry, N.s, N.f1, N.f2);
  else cell<2, false>(qx[0], qx[1], qx[2], C.g, C.dry, N.s, N.f1, N.f2);
  if (MET) {
    N.kap = mk;
    N.fac = C.dtdy * rcp(mk);
    N.gy = mgy;
  }
  // interface r (rows r-1 | r)
  if (!GEN || r >= C.y0 - 1) {
    const float fA[3] = {A.q[2], A.f1, A.f2}, fN[3] = {qx[2], N.f1, N.f2};
    N.s2 = hll<(PHYS != 0)>(A.q, fA, A.s, qx, fN, N.s, C.eps, N.w1, N.w2);
    R.smy = smax2(R.smy, (!GEN || (r >= 0 && r <= C.ny + 1)) ? N.s2 : 0.0f);
  } else {
    N.s2 = 0.0f;
#pragma unroll
    for (int k = 0; k < 3; ++k) N.w1[k] = N.w2[k] = 0.0f;
  }
  // interface r-1: limiter, correction flux, G and H
  if (!GEN || r >= C.y0 + 1) finish<LIM, HO, MET, PHYS, GEN>(C, r, N, A, B, acc, bm, b0, bnew, finm);
  if (PHYS) {
    bm = b0;
    b0 = bnew;
  }
}

// One row r of a wave's march, sweeps in sequence.  N: slot of row r
// (written), A: row r-1, B: row r-2.  P: this row's prefetched input,
// reloaded with row r+kPF.  GEN: the general body with every range test;
// otherwise the row must be an interior row of an interior strip (see the
// kernel's chunk test).
template <int LIM, bool HO, bool MET, int PHYS, bool GEN>
__device__ __forceinline__ void row(const Ctx& C, int r, Slot& N, Slot& A, Slot& B, Pre& P,
                                    float acc[3], Red& R, float& bm, float& b0, float finm[3]) {
  // rows are not interleaved by the scheduler: that would move the
  // prefetch loads to the end of a chunk and waste their head start
  __builtin_amdgcn_sched_barrier(0);
  // nor is the row's arithmetic started earlier, in the previous row (the
  // input is in registers long before): the empty asm, ordered after the
  // previous row's stores, is where the row's input values come from
  asm volatile("" : "+v"(P.h), "+v"(P.hu), "+v"(P.hv));
  const float mk = P.k, mgy = P.gy;
  float bnew = 0.0f;
  if (PHYS && C.has_b && r >= C.y0) bnew = bload(C.bi, C.lo, (r - C.rbase) * C.p4);  // b(i, r-1)
  float qx[3];
  xi_row<LIM, HO, MET, PHYS, GEN>(C, r, P, R, qx);
  // the prefetch goes after the row's last use of P, so that the load can
  // write the same registers; without metric arrays it is pinned between
  // the sweeps so that the row's loads and stores are issued apart
  if (!MET) {
    __builtin_amdgcn_sched_barrier(0);
    prefetch<MET, PHYS>(C, r, P);
    __builtin_amdgcn_sched_barrier(0);
  }
  eta_row<LIM, HO, MET, PHYS, GEN>(C, r, qx, mk, mgy, N, A, B, acc, R, bm, b0, bnew, finm);
  if (MET) prefetch<MET, PHYS>(C, r, P);  // the eta sweep used this row's metric values
}

// The part of a row that needs interface r-1 (from row y0+1 on): its
// limiter and correction flux, row r-2's last contribution, source term,
// stores and ghosts, and row r-1's first contribution.
template <int LIM, bool HO, bool MET, int PHYS, bool GEN>
__device__ __forceinline__ void finish(const Ctx& C, int r, const Slot& N, const Slot& A,
                                       const Slot& B, float acc[3], float bm, float b0,
                                       float bnew, float finm[3]) {
  // dH: H of interface r-1 (to row r-2), times dt/deta where that is folded
  // into the coefficients (PS); aq: row r-1's xi state minus dt/deta G of the
  // same interface
  constexpr bool PS = HO && !MET;
  float dH[3], aq[3];
  float p1 = 0.5f, p2 = 0.5f;  // halved limiter values (unlimited: 1 / 2)
  if (HO && LIM > 0 && (!GEN || (r - 1 >= 1 && r - 1 <= C.ny))) {
    p1 = half_phi<LIM>(A.w1, N.w1, C.eps);
    p2 = half_phi<LIM>(A.w2, B.w2, C.eps);
  }
  if (PS) {
    // dt/deta folded into the coefficients, as in the xi sweep
    const float nu = C.dtdy * A.s2, cf = fmaf(-nu, nu, nu);
    const float c1 = cf * p1, c2 = cf * p2, e = c1 - nu, am = nu - c2;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      dH[k] = fmaf(e, A.w1[k], c2 * A.w2[k]);
      aq[k] = fmaf(-am, A.w2[k], fmaf(c1, A.w1[k], A.q[k]));
    }
  } else {
    float Fa[3] = {0.0f, 0.0f, 0.0f};
    if (HO) {
      const float coef = A.s2 * fmaf(-A.s2 * C.dtdy, rcp(0.5f * (B.kap + A.kap)), 1.0f);
      corr_flux(coef, p1, p2, A.w1, A.w2, Fa);
    }
    const float sg = MET ? A.gy * A.s2 : A.s2;
    const float facA = MET ? A.fac : C.dtdy;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      dH[k] = fmaf(-sg, A.w1[k], Fa[k]);
      aq[k] = fmaf(-facA, fmaf(sg, A.w2[k], -Fa[k]), A.q[k]);
    }
  }
  if (!GEN || r >= C.y0 + 2) {  // finish row j = r-2
    const int j = r - 2;
    const float facB = MET ? B.fac : C.dtdy;
    float qn[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) qn[k] = PS ? acc[k] - dH[k] : fmaf(-facB, dH[k], acc[k]);
    if (PHYS == 2 || (PHYS == 1 && C.src != PCW_SRC_NONE)) {
      // b(i+1) - b(i-1) through the LDS crossbar (lanes 0 and 63, which
      // receive each other's values, are halo lanes)
      const float bx = rgt(C, b0) - lft(C, b0), by = bnew - bm;
      source(qn[0], qn[1], qn[2], C.src, C.dry, C.sx, C.sy, C.fk, bx, by);
    }
    const int sj = (j + 1 - C.rbase) * C.p4;
    // an interior tile starts in the unconditional body: its first rows
    // finish halo rows below the tile, which are not stored (out-of-range
    // offset, no branch)
    const unsigned so = (GEN || r >= C.y0 + 2) ? C.so : kOff;
#pragma unroll
    for (int k = 0; k < 3; ++k) bstore(qn[k], C.qo[k], so, sj);
    if (GEN) {
      // xi-boundary ghosts of row j from this row's finished cells (near:
      // column 1 or nx, far: column 2 or nx-1, read as wave-uniform values)
      if (C.x0 == 1) {
        float nr[3], fr[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) { nr[k] = readlane(qn[k], 2); fr[k] = readlane(qn[k], 3); }
        if (C.lane <= 1) {
          const int kk = 2 - C.lane;  // lane 1: ghost column 0, lane 0: column -1
#pragma unroll
          for (int k = 0; k < 3; ++k)
            bstore(ghost(C.bc[0], nr[k], fr[k], kk, k == 1, kk == 1 ? C.gr[0] : C.gr[1]), C.qo[k],
                   4u * (unsigned)col_pos(1, C.nslot, C.lane - 1), sj);
        }
      }
      if (C.l_nx >= 3 && C.l_nx <= 61) {
        float nr[3], fr[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) { nr[k] = readlane(qn[k], C.l_nx); fr[k] = readlane(qn[k], C.l_nx - 1); }
        const int kk = C.lane - C.l_nx;
        if (kk == 1 || kk == 2) {
#pragma unroll
          for (int k = 0; k < 3; ++k)
            bstore(ghost(C.bc[1], nr[k], fr[k], kk, k == 1, kk == 1 ? C.gr[2] : C.gr[3]), C.qo[k],
                   4u * (unsigned)col_pos(1, C.nslot, C.nx + kk), sj);
        }
      }
      // eta-boundary ghosts; finm holds the finished row j-1 (rows next to
      // the domain's bottom and top rows are finished in the general body)
      if (j == 2 && C.y0 == 1) {
#pragma unroll
        for (int kk = 1; kk <= 2; ++kk) {
          const int sg = (2 - kk - C.rbase) * C.p4;  // ghost row 1-kk
#pragma unroll
          for (int k = 0; k < 3; ++k)
            bstore(ghost(C.bc[2], finm[k], qn[k], kk, k == 2, C.gr[3 + kk]), C.qo[k], C.so, sg);
        }
      }
      if (j == C.ny) {
#pragma unroll
        for (int kk = 1; kk <= 2; ++kk) {
          const int sg = (C.ny + kk + 1 - C.rbase) * C.p4;  // ghost row ny+kk
#pragma unroll
          for (int k = 0; k < 3; ++k)
            bstore(ghost(C.bc[3], qn[k], finm[k], kk, k == 2, C.gr[5 + kk]), C.qo[k], C.so, sg);
        }
      }
#pragma unroll
      for (int k = 0; k < 3; ++k) finm[k] = qn[k];
    }
  }
  // row r-1: xi-updated state minus the contribution of its lower interface
#pragma unroll
  for (int k = 0; k < 3; ++k) acc[k] = aq[k];
}

// Four waves per SIMD except with metric arrays: the source-term variant
// with the MC limiter needs 129 VGPRs unbounded (three waves); bounded to 128
// it spills 8 bytes per lane and runs config 3's step 2.7 % faster.
template <int LIM, bool HO, bool MET, int PHYS>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(MET ? 1 : 4)))
k_step(FusedArgs<float> a) {
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
  if (S->status != PCW_STATUS_RUNNING) return;
  if (MET != (I.moff >= 0)) return;  // instances with and without metric arrays run in separate launches
  const bool phys_inst = I.source != PCW_SRC_NONE || a.wetdry;
  if ((PHYS != 0) != phys_inst) return;  // likewise instances with and without source terms

  Ctx C;
  const int cur = S->cur;
  const double dt = S->tc.dt;
  C.nx = I.nx;
  C.ny = I.ny;
  C.x0 = t.y;
  C.y0 = t.z;
  C.y1 = t.w;
  C.lane = lane;
  C.bl = ((lane + 63) & 63) * 4;
  C.br = ((lane + 1) & 63) * 4;
  C.i = C.x0 - 2 + lane;                              // this lane's column (1-based interior)
  C.l_nx = C.nx - C.x0 + 2;                           // lane of column nx
  C.out_lane = lane >= 2 && lane <= 61 && C.i >= 1 && C.i <= C.nx;
  C.p4 = I.pitch * 4;
  C.rbase = C.y0 - 1;
  // frame rows are in 64-element slots (col_pos); the host lays out every
  // FP32 instance that way
  C.nslot = I.nslot;
  C.lo = 4u * (unsigned)col_pos(1, C.nslot, C.i < C.nx + 2 ? C.i : C.nx + 2);
  if ((C.x0 - 1) % kStripOut == 0 && C.x0 + kStripOut - 1 <= C.nx) {
    // the strip fills slot s+1: all 64 lanes store, the four halo lanes
    // into the slot's padding, so that every row store is one aligned
    // 256-byte slot
    const int s = (C.x0 - 1) / kStripOut;
    C.so = 4u * (unsigned)(64 * (s + 1) + ((lane - 2) & 63));
  } else {
    // the strip shifted to end at column nx: output lanes only
    C.so = C.out_lane ? 4u * (unsigned)col_pos(1, C.nslot, C.i) : kOff;
  }
  {
    const long long o = (long long)C.rbase * I.pitch;  // descriptor origin within a plane
    const int bytes = (C.y1 - C.y0 + 5) * C.p4;
    const float* qin = a.pool + I.off[cur] + o;
    const float* qout = a.pool + I.off[cur ^ 1] + o;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      C.qi[k] = make_rsrc(qin + k * I.plane, bytes);
      C.qo[k] = make_rsrc(qout + k * I.plane, bytes);
      C.mi[k] = MET ? make_rsrc(a.mpool + I.moff + o + k * I.plane, bytes) : C.qi[k];
    }
    const bool use_b = I.boff >= 0 && (I.source == PCW_SRC_BATHY || I.source == PCW_SRC_BATHY_FRICTION);
    C.has_b = PHYS && use_b;
    C.bi = C.has_b ? make_rsrc(a.bpool + I.boff + o, bytes) : C.qi[0];
  }
  C.xface = lane >= 1 && C.i >= 0 && C.i <= C.nx + 1;  // interface i-1/2 in the xi CFL range
  C.xlim = C.i >= 1 && C.i <= C.nx;                   // interface limited
  C.src = I.source;
  C.wetdry = a.wetdry;
  for (int k = 0; k < 4; ++k) C.bc[k] = I.bc[k];
  for (int k = 0; k < 8; ++k) C.gr[k] = I.gr[k];
  C.dtdx = (float)(dt / I.dxi);
  C.dtdy = (float)(dt / I.deta);
  C.eps = (float)a.eps;
  C.g = (float)a.g;
  C.dry = (float)a.dry;
  C.sx = (float)(-dt * a.g / I.dxi);
  C.sy = (float)(-dt * a.g / I.deta);
  C.fk = (float)(dt * a.g * I.manning * I.manning);

  Red R;
  R.smx = R.smy = 0.0f;
  Slot X, Y, Z;
  X.s2 = Y.s2 = Z.s2 = 0.0f;
  X.s = Y.s = Z.s = 0.0f;
  X.f1 = Y.f1 = Z.f1 = X.f2 = Y.f2 = Z.f2 = 0.0f;
  X.kap = Y.kap = Z.kap = X.fac = Y.fac = Z.fac = X.gy = Y.gy = Z.gy = 1.0f;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    X.q[k] = Y.q[k] = Z.q[k] = 0.0f;
    X.w1[k] = Y.w1[k] = Z.w1[k] = X.w2[k] = Y.w2[k] = Z.w2[k] = 0.0f;
  }
  float acc[3] = {0.0f, 0.0f, 0.0f};
  float finm[3] = {0.0f, 0.0f, 0.0f};
  float bm = 0.0f, b0 = 0.0f;

  // prefetch rows y0-2 .. y0-2+kPF-1 (clamped to y1+2: a tile of a two-row
  // domain marches six rows)
  constexpr int kPF = Ring<MET, PHYS>::kPF;
  Pre P[kPF];
#pragma unroll
  for (int d = 0; d < kPF; ++d) {
    const int so = (min(C.y0 - 2 + d, C.y1 + 2) + 1 - C.rbase) * C.p4;
    P[d].h = bload(C.qi[0], C.lo, so);
    P[d].hu = bload(C.qi[1], C.lo, so);
    P[d].hv = bload(C.qi[2], C.lo, so);
    P[d].k = P[d].gx = P[d].gy = 1.0f;
    if (MET) {
      P[d].k = bload(C.mi[0], C.lo, so);
      P[d].gx = bload(C.mi[1], C.lo, so);
      P[d].gy = bload(C.mi[2], C.lo, so);
    }
  }

  // Rows y0-2 .. y1+2 in chunks of six (the slots rotate every three rows,
  // the input buffers every six).  A chunk runs the unconditional body when
  // all its rows satisfy every range condition: for a tile away from the
  // domain's boundaries (interior strip, rows y0-2 .. y1+2 inside 1 .. ny)
  // that is every full chunk, whose start-up rows compute waves and fluxes
  // below the tile that are discarded (the store is skipped); for other
  // tiles, the chunks between the tile's start-up rows and the domain's
  // bottom ghost rows (rs) and its top rows (re), away from the xi
  // boundaries.  Other chunks run the general body.  A run of unconditional
  // chunks is entered with no memory operation in flight, so that the wait
  // for each prefetched row counts only the run's own loads and stores.
  const bool edge = C.x0 == 1 || (C.l_nx >= 3 && C.l_nx <= 61);
  const int rl = C.y1 + 2;
  const bool interior = !edge && C.y0 >= 3 && rl <= C.ny;
  const int rs = interior ? C.y0 - 2 : C.y0 + 2 + (C.y0 == 1 ? 2 : 0);
  const int re = interior ? rl : min(rl, C.ny);
  int r = C.y0 - 2;
  while (r <= rl) {
    if (edge || r < rs || r + 5 > re) {
      row<LIM, HO, MET, PHYS, true>(C, r, X, Y, Z, P[0 % kPF], acc, R, bm, b0, finm);
      if (r + 1 <= rl) row<LIM, HO, MET, PHYS, true>(C, r + 1, Z, X, Y, P[1 % kPF], acc, R, bm, b0, finm);
      if (r + 2 <= rl) row<LIM, HO, MET, PHYS, true>(C, r + 2, Y, Z, X, P[2 % kPF], acc, R, bm, b0, finm);
      if (r + 3 <= rl) row<LIM, HO, MET, PHYS, true>(C, r + 3, X, Y, Z, P[3 % kPF], acc, R, bm, b0, finm);
      if (r + 4 <= rl) row<LIM, HO, MET, PHYS, true>(C, r + 4, Z, X, Y, P[4 % kPF], acc, R, bm, b0, finm);
      if (r + 5 <= rl) row<LIM, HO, MET, PHYS, true>(C, r + 5, Y, Z, X, P[5 % kPF], acc, R, bm, b0, finm);
      r += 6;
      continue;
    }
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
    do {
      row<LIM, HO, MET, PHYS, false>(C, r, X, Y, Z, P[0 % kPF], acc, R, bm, b0, finm);
      row<LIM, HO, MET, PHYS, false>(C, r + 1, Z, X, Y, P[1 % kPF], acc, R, bm, b0, finm);
      row<LIM, HO, MET, PHYS, false>(C, r + 2, Y, Z, X, P[2 % kPF], acc, R, bm, b0, finm);
      row<LIM, HO, MET, PHYS, false>(C, r + 3, X, Y, Z, P[3 % kPF], acc, R, bm, b0, finm);
      row<LIM, HO, MET, PHYS, false>(C, r + 4, Z, X, Y, P[4 % kPF], acc, R, bm, b0, finm);
      row<LIM, HO, MET, PHYS, false>(C, r + 5, Y, Z, X, P[5 % kPF], acc, R, bm, b0, finm);
      r += 6;
    } while (r + 5 <= re);
  }

  // wave maxima and error bits; the instance's last tile completes the attempt
  const float mx = wave_smax(R.smx), my = wave_smax(C.out_lane ? R.smy : 0.0f);
  const bool chk = PHYS != 2;  // with the wet/dry cut-off h <= 0 is a dry cell
  const int bad = chk ? (__builtin_bit_cast(unsigned, mx) >= kNotFinite ? 1 : 0) |
                            (__builtin_bit_cast(unsigned, my) >= kNotFinite ? 2 : 0)
                      : 0;
  tile_finish_slots(S, I, a.slots + (size_t)in * kSlots, tid - I.tile0, cur, mx, my, bad, lane);
}

template <bool HO, bool MET, int PHYS>
static hipError_t launch_lim(int lim, const FusedArgs<float>& a, int blocks, int threads,
                             hipStream_t s) {
  switch (lim) {
    case 0: k_step<0, HO, MET, PHYS><<<blocks, threads, 0, s>>>(a); break;
    case 1: k_step<1, HO, MET, PHYS><<<blocks, threads, 0, s>>>(a); break;
    case 2: k_step<2, HO, MET, PHYS><<<blocks, threads, 0, s>>>(a); break;
    case 3: k_step<3, HO, MET, PHYS><<<blocks, threads, 0, s>>>(a); break;
    default: k_step<4, HO, MET, PHYS><<<blocks, threads, 0, s>>>(a); break;
  }
  return hipGetLastError();
}

template <bool MET, int PHYS>
static hipError_t launch_ho(int lim, const FusedArgs<float>& a, int blocks, int threads,
                            hipStream_t s) {
  if (!a.ho) {
    k_step<0, false, MET, PHYS><<<blocks, threads, 0, s>>>(a);
    return hipGetLastError();
  }
  return launch_lim<true, MET, PHYS>(lim, a, blocks, threads, s);
}

template <bool MET, int PHYS>
static int occupancy_ho(int lim, bool ho, int threads) {
  int n = 0;
  hipError_t e;
  if (!ho) {
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_step<0, false, MET, PHYS>, threads, 0);
  } else {
    switch (lim) {
      case 0: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_step<0, true, MET, PHYS>, threads, 0); break;
      case 1: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_step<1, true, MET, PHYS>, threads, 0); break;
      case 2: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_step<2, true, MET, PHYS>, threads, 0); break;
      case 3: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_step<3, true, MET, PHYS>, threads, 0); break;
      default: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_step<4, true, MET, PHYS>, threads, 0); break;
    }
  }
  return e == hipSuccess ? n : 0;
}

}  // namespace f32

// phys: 0 no source terms, 1 source terms without the wet/dry cut-off, 2 with it
int fused_f32_blocks_per_cu(int lim, bool ho, bool met, int phys, int threads) {
  if (met) return phys == 2 ? f32::occupancy_ho<true, 2>(lim, ho, threads)
                  : phys == 1 ? f32::occupancy_ho<true, 1>(lim, ho, threads)
                              : f32::occupancy_ho<true, 0>(lim, ho, threads);
  return phys == 2 ? f32::occupancy_ho<false, 2>(lim, ho, threads)
         : phys == 1 ? f32::occupancy_ho<false, 1>(lim, ho, threads)
                     : f32::occupancy_ho<false, 0>(lim, ho, threads);
}

hipError_t launch_fused_f32(int lim, bool met, int phys, const FusedArgs<float>& a, int blocks,
                            int threads, hipStream_t s) {
  if (met) {
    return phys == 2 ? f32::launch_ho<true, 2>(lim, a, blocks, threads, s)
           : phys == 1 ? f32::launch_ho<true, 1>(lim, a, blocks, threads, s)
                       : f32::launch_ho<true, 0>(lim, a, blocks, threads, s);
  }
  return phys == 2 ? f32::launch_ho<false, 2>(lim, a, blocks, threads, s)
         : phys == 1 ? f32::launch_ho<false, 1>(lim, a, blocks, threads, s)
                     : f32::launch_ho<false, 0>(lim, a, blocks, threads, s);
}

}  // namespace pcw
