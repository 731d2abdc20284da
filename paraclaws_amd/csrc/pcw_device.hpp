// pcw_device.hpp -- per-cell and per-interface arithmetic of the 2D
// shallow-water wave-propagation update, for gfx950.
//
// Exact (double) arithmetic: the reference's FP64 operations in the
// reference's order (Jacobian entries of
// examples/2D_dam_break/cartesian/SMOD_jakobian.f90:44-53, J*q in BLAS column
// order, true IEEE divisions, reference-rounded mesh metrics), so the FP64
// mode reproduces the CPU restatement bit for bit.  Every translation unit is
// compiled with -ffp-contract=off.  The FP32 kernel's arithmetic lives in
// pcw_fused_f32.hip.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/pcw.h"

namespace pcw {

// --------------------------------------------------------------- lane shifts
// DPP wave_shr:1 / wave_shl:1 (GFX9 family): lane l receives lane l-1 / l+1.
// The FP32 kernel takes its metric and bathymetry neighbours this way.
// Lanes shifted in from outside the wave receive 0 (bound_ctrl), so every
// lane is written and no old value has to be set up.
__device__ __forceinline__ int shr1_i(int v) {
  return __builtin_amdgcn_update_dpp(0, v, 0x138, 0xf, 0xf, true);
}
__device__ __forceinline__ int shl1_i(int v) {
  return __builtin_amdgcn_update_dpp(0, v, 0x130, 0xf, 0xf, true);
}
__device__ __forceinline__ float shr1(float v) {
  return __builtin_bit_cast(float, shr1_i(__builtin_bit_cast(int, v)));
}
__device__ __forceinline__ float shl1(float v) {
  return __builtin_bit_cast(float, shl1_i(__builtin_bit_cast(int, v)));
}

// wave-wide max of a non-negative value (butterfly over 64 lanes)
template <typename T>
__device__ __forceinline__ T wave_max(T v) {
  for (int o = 32; o >= 1; o >>= 1) {
    T w = __shfl_xor(v, o, 64);
    v = v > w ? v : w;
  }
  return v;
}

// ---------------------------------------------------------------- constants
struct Phys {
  double g;       // gravity (REAL*4 9.81 widened in the reference)
  double eps;     // ud_precision
  double dry;     // wet/dry threshold (0 = reference behaviour)
};

// Per-cell quantities of one sweep direction: speed s = max|eig(J)|, flux f.
template <typename T>
struct CellQ {
  T s, f0, f1, f2;
  int bad;  // 1: NaN in J or complex eigenvalues (reference error stop)
};

// ------------------------------------------------ quotients by one divisor
// gfx950 divides in FP64 with v_div_scale (scaling when an operand lies near
// the ends of the exponent range), v_rcp_f64 and two Newton steps on the
// reciprocal of the divisor, the quotient with one fma correction
// (v_div_fmas), then v_div_fixup (zero, infinite and NaN operands).  When no
// scaling applies, the reciprocal depends on the divisor alone, so quotients
// with one divisor can share it: each then costs a multiply, two fmas and the
// fixup, and has the bits of the full division.  No scaling applies when the
// divisor's binary exponent (frexp) lies in [-299, 300] and the dividend is
// zero or has its exponent in [-700, 440]: the exponent difference stays
// below 768, the divisor, its reciprocal and the quotient are normal, and
// the dividend is far from the denormal range.  For zero, infinite and NaN
// operands v_div_fixup returns its answer whatever the estimate.
struct Divisor {
  double d, r;
};
__device__ __forceinline__ Divisor divisor(double d) {
  double r = __builtin_amdgcn_rcp(d);
  double e = fma(-d, r, 1.0);
  r = fma(r, e, r);
  e = fma(-d, r, 1.0);
  r = fma(r, e, r);
  return {d, r};
}
__device__ __forceinline__ double quot(double n, const Divisor& D) {
  const double q = n * D.r;
  const double rem = fma(-D.d, q, n);
  return __builtin_amdgcn_div_fixup(fma(rem, D.r, q), D.d, n);
}
// frexp exponent of x (0 for zero) within [lo, hi]
__device__ __forceinline__ bool exp_in(double x, int lo, int hi) {
  return (unsigned)(__builtin_amdgcn_frexp_exp(x) - lo) <= (unsigned)(hi - lo);
}

// --------------------------------------------------------- Exact (double)
// Cell physics for the xi sweep (dir 1) or eta sweep (dir 2):
// SMOD_jakobian.f90:44-53 entries, eigenvalues {u-c,u,u+c}, J*q column order.
// When every lane of the wave has h in [2^-150, 2^150) and hu, hv zero or in
// [2^-350, 2^220), the divisors h and h*h and the dividends hu, hv, hu*hu,
// hu*hv, hv*hv are in the range where quotients share a reciprocal; then
// also 2*hu/h = 2*(hu/h) exactly.
template <int DIR>
__device__ __forceinline__ CellQ<double> cell_exact(double h, double hu, double hv,
                                                    const Phys& P) {
  CellQ<double> r;
  r.bad = 0;
  const double g = P.g;
  if (P.dry > 0.0 && !(h > P.dry)) {  // this build's wet/dry cut-off
    double hp = h > 0.0 ? h : 0.0;
    r.s = sqrt(g * hp);
    if (DIR == 1) { r.f0 = hu; r.f1 = g * h * h; r.f2 = 0.0; }
    else          { r.f0 = hv; r.f1 = 0.0;       r.f2 = g * h * h; }
    return r;
  }
  double gh = g * h, cc = 0.0;
  if (gh < 0.0) {
    if (sqrt(-gh) > P.eps) r.bad = 1;
  } else {
    cc = sqrt(gh);
  }
  const bool shared = __all(exp_in(h, -149, 150) && exp_in(hu, -349, 220) && exp_in(hv, -349, 220));
  if (DIR == 1) {
    double J21, J22, J31, J32, J33;
    if (shared) {
      const Divisor Dh = divisor(h), Dhh = divisor(h * h);
      J21 = -quot(hu * hu, Dhh) + g * h;
      J33 = quot(hu, Dh);
      J22 = 2.0 * J33;
      J31 = -quot(hu * hv, Dhh);
      J32 = quot(hv, Dh);
    } else {
      J21 = -(hu * hu / (h * h)) + g * h;
      J22 = 2.0 * hu / h;
      J31 = -(hu * hv / (h * h));
      J32 = hv / h;
      J33 = hu / h;
    }
    if (isnan(J21) || isnan(J22) || isnan(J31) || isnan(J32) || isnan(J33)) r.bad = 1;
    r.s = fabs(J33) + cc;
    r.f0 = hu;
    r.f1 = J21 * h + J22 * hu;
    r.f2 = (J31 * h + J32 * hu) + J33 * hv;
  } else {
    double J21, J22, J23, J31, J33;
    if (shared) {
      const Divisor Dh = divisor(h), Dhh = divisor(h * h);
      J21 = -quot(hu * hv, Dhh);
      J22 = quot(hv, Dh);
      J23 = quot(hu, Dh);
      J31 = -quot(hv * hv, Dhh) + g * h;
      J33 = 2.0 * J22;
    } else {
      J21 = -(hu * hv / (h * h));
      J22 = hv / h;
      J23 = hu / h;
      J31 = -(hv * hv / (h * h)) + g * h;
      J33 = 2.0 * hv / h;
    }
    if (isnan(J21) || isnan(J22) || isnan(J23) || isnan(J31) || isnan(J33)) r.bad = 1;
    r.s = fabs(J22) + cc;
    r.f0 = hv;
    r.f1 = (J21 * h + J22 * hu) + J23 * hv;
    r.f2 = J31 * h + J33 * hv;
  }
  return r;
}

// ----------------------------------------------------------------- waves
// HLL of SMOD_approximate_Riemann_solvers.f90:60-101 for L | R.
template <typename T>
struct Waves {
  T s2;             // lambda(2) = s2, lambda(1) = -s2 (0,0 when skipped)
  T w1[3], w2[3];   // waves 1 and 2
};

__device__ __forceinline__ Waves<double> hll_exact(const double qL[3], const CellQ<double>& cL,
                                                   const double qR[3], const CellQ<double>& cR,
                                                   double eps) {
  Waves<double> w;
  double s2 = cL.s > cR.s ? cL.s : cR.s;
  double s1 = -s2;
  if (fabs(s1) > eps && fabs(s2) > eps) {
    double den = eps + s1 - s2;
    double fL[3] = {cL.f0, cL.f1, cL.f2}, fR[3] = {cR.f0, cR.f1, cR.f2};
    double num[3], qs[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) num[k] = fR[k] - fL[k] - s2 * qR[k] + s1 * qL[k];
    // the three quotients share a reciprocal when the wave's operands allow
    if (__all(exp_in(den, -299, 300) && exp_in(num[0], -700, 440) && exp_in(num[1], -700, 440) &&
              exp_in(num[2], -700, 440))) {
      const Divisor D = divisor(den);
#pragma unroll
      for (int k = 0; k < 3; ++k) qs[k] = quot(num[k], D);
    } else {
#pragma unroll
      for (int k = 0; k < 3; ++k) qs[k] = num[k] / den;
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      w.w1[k] = qs[k] - qL[k];
      w.w2[k] = qR[k] - qs[k];
    }
    w.s2 = s2;
  } else {
    w.s2 = 0.0;
#pragma unroll
    for (int k = 0; k < 3; ++k) { w.w1[k] = 0.0; w.w2[k] = 0.0; }
  }
  return w;
}

// ------------------------------------------------------------- limiters
// MOD_flux_correction.f90:101-112 (1 minmod, 2 superbee, 3 van Leer, 4 MC).
template <int LIM>
__device__ __forceinline__ double phi_exact(double r) {
  if (LIM == 1) return fmax(0.0, fmin(1.0, r));
  if (LIM == 2) return fmax(fmax(0.0, fmin(1.0, 2.0 * r)), fmin(2.0, r));
  if (LIM == 3) return (r + fabs(r)) / (1.0 + fabs(r));
  return fmax(0.0, fmin(fmin((1.0 + r) / 2.0, 2.0), 2.0 * r));
}

// Limiter factor of one wave (MOD_flux_correction.f90:76-120): theta uses
// the unlimited dot product with the upwind neighbour; a wave whose squared
// norm is <= ud_precision is left unlimited.
template <int LIM>
__device__ __forceinline__ double limit_exact(const double w[3], const double nb[3],
                                              double eps) {
  double wnorm2 = w[0] * w[0] + w[1] * w[1];
  wnorm2 = wnorm2 + w[2] * w[2];
  double up = nb[0] * w[0] + nb[1] * w[1];
  up = up + nb[2] * w[2];
  if (fabs(wnorm2) <= eps) return 1.0;
  return phi_exact<LIM>(up / wnorm2);
}

// ---------------------------------------------------------- mesh metrics
// Identity-mapped Cartesian mesh, rounded as
// src/domain/SMOD_mesh_manipulation.f90:36-51 and :147-173 compute it.
struct Mesh {
  double xi_min, eta_min, dxi, deta;
};

__host__ __device__ __forceinline__ double corner_x1(const Mesh& m, int i) {
  return m.xi_min + (double)(i - 1) * m.dxi;
}
__host__ __device__ __forceinline__ double corner_y1(const Mesh& m, int j) {
  return m.eta_min + (double)(j - 1) * m.deta;
}
// capacity kappa(i,j)
__host__ __device__ __forceinline__ double kappa_exact(const Mesh& m, int i, int j) {
  double x1 = corner_x1(m, i), x2 = x1 + m.dxi, x3 = x2, x4 = x1, x5 = x1;
  double y1 = corner_y1(m, j), y2 = y1, y3 = y2 + m.deta, y4 = y3, y5 = y1;
  double a = 0.0;
  a = a + 0.5 * (y1 + y2) * (x1 - x2);
  a = a + 0.5 * (y2 + y3) * (x2 - x3);
  a = a + 0.5 * (y3 + y4) * (x3 - x4);
  a = a + 0.5 * (y4 + y5) * (x4 - x5);
  return a / (m.dxi * m.deta);
}
// length_ratio_xi(1,i,j): depends on j only (face 1->4 is vertical)
__host__ __device__ __forceinline__ double gamma_xi_exact(const Mesh& m, int j) {
  double y1 = corner_y1(m, j), y3 = y1 + m.deta;
  double fdx = 0.0, fdy = y3 - y1;
  return sqrt(fdx * fdx + fdy * fdy) / m.deta;
}
// length_ratio_eta(1,i,j): depends on i only (face 1->2 is horizontal)
__host__ __device__ __forceinline__ double gamma_eta_exact(const Mesh& m, int i) {
  double x1 = corner_x1(m, i), x2 = x1 + m.dxi;
  double fdx = x2 - x1, fdy = 0.0;
  return sqrt(fdx * fdx + fdy * fdy) / m.dxi;
}
// cell centre, SMOD_mesh_manipulation.f90:50-51
__host__ __device__ __forceinline__ double centre(double lo, double d, int i) {
  return lo + (double)(i - 1) * d + d / 2;
}

// ---------------------------------------------------- boundary conditions
// Ghost value from the two interior cells next to a boundary
// (SMOD_boundary_conditions.f90:92-105): near = first interior cell, far =
// second, k = 1 for the first ghost layer, 2 for the second.  For BC 3 the
// derivative is taken in the reference's centre coordinates (x_near,
// x_far, x_ghost).  normal: this field is the normal momentum (negated by
// the reflective wall, which mirrors near for k = 1 and far for k = 2).
template <typename T>
__device__ __forceinline__ T ghost_value(int bc, T near, T far, int k, bool normal, double xn,
                                         double xf, double xg) {
  if (bc == 2) return near;
  if (bc == 3) {
    double d = ((double)far - (double)near) / (xf - xn);
    return (T)((double)near + d * (xg - xn));
  }
  T v = (k == 1) ? near : far;
  return normal ? -v : v;
}

// The reference's integrate_source_term with its source_function phi = 0
// (src/solver/MOD_solve.f90:319-327, SMOD_source_function.f90:15-28): q + 0*dt
// on every interior cell and field.  The sum is q itself except that an
// interior -0.0 becomes +0.0 (round to nearest); it is kept, not folded away,
// so that signs of zero follow the reference too.
template <typename T>
__device__ __forceinline__ void reference_source(T& h, T& hu, T& hv) {
  h = h + T(0);
  hu = hu + T(0);
  hv = hv + T(0);
}

// Source term of this build's physics on one cell (no reference counterpart;
// the reference's source_function returns 0): wet/dry cut-off, bathymetry
// slope (balancing the reference flux's pressure g h^2), semi-implicit Manning.
// Every result ends with the reference's + 0*dt (reference_source), so that a
// zero bathymetry slope or a zero Manning coefficient gives the bits of the
// reference's phi = 0 step, signs of zero included.
template <typename T>
__device__ __forceinline__ void apply_source(T& h, T& hu, T& hv, int kind, T dry, T g, T dxi,
                                             T deta, T dt, T n2, T bxm, T bxp, T bym, T byp) {
  if (!(h > dry)) {
    h = h > T(0) ? h : T(0);
    hu = T(0);
    hv = T(0);
    return;
  }
  T nhu = hu, nhv = hv;
  if (kind == PCW_SRC_BATHY || kind == PCW_SRC_BATHY_FRICTION) {
    T bx = bxp - bxm, by = byp - bym;
    nhu = hu + dt * (-g * h * bx / dxi);
    nhv = hv + dt * (-g * h * by / deta);
  }
  if (kind == PCW_SRC_FRICTION || kind == PCW_SRC_BATHY_FRICTION) {
    T u = hu / h, v = hv / h;
    T sp = sqrt(u * u + v * v);
    T h43 = cbrt(h) * h;
    T den = T(1) + dt * g * n2 * sp / h43;
    nhu = nhu / den;
    nhv = nhv / den;
  }
  hu = nhu;
  hv = nhv;
  reference_source(h, hu, hv);
}

}  // namespace pcw
