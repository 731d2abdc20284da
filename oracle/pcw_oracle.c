/*
 * pcw_oracle.c -- CPU restatement (FP64, OpenMP) of paraCLAWs' per-step
 * dimension-split wave-propagation update, its time controller and its main
 * loop, for the 2D shallow-water dam-break physics.
 *
 * TEST INFRASTRUCTURE ONLY: this is the checker.  It is loaded by the test
 * suite, the smoke check and the benchmark's CPU-baseline leg, never by the
 * product path.
 *
 * Build with -ffp-contract=off: every expression below is written in the
 * evaluation order the Fortran source implies (left to right within equal
 * precedence), so that the product's FP64 mode can be compared bit for bit.
 *
 * The reference computes eigenvalues with LAPACK dgeev and J*q with BLAS
 * dgemv (src/utility/SMOD_eigenvalue_problem.f90:56-66,
 * SMOD_matrix_operations.f90:22).  Here the eigenvalues are the closed form
 * {u-c, u, u+c}, c = sqrt(g h), so max|lambda| = |u| + c, and J*q is
 * evaluated in reference-BLAS column order (y = ((A1*x1) + A2*x2) + A3*x3).
 * The closed form is checked against LAPACK dgeev in the tests.
 */
#include "pcw_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define NG 2 /* num_ghost */

/* ---------------------------------------------------------------- layout */
typedef struct {
  int nx, ny, nxf, nyf;
} frame_t;

static frame_t frame_of(const orc_cfg* c) {
  frame_t f;
  f.nx = c->nx;
  f.ny = c->ny;
  f.nxf = c->nx + 2 * NG;
  f.nyf = c->ny + 2 * NG;
  return f;
}
/* cell (i,j), i in [-1, nx+2], j in [-1, ny+2] -> linear frame index */
static inline long F(const frame_t* f, int i, int j) {
  return (long)(i + 1) + (long)f->nxf * (long)(j + 1);
}
#define Q(v, f, me, i, j) (v)[(me) + 3 * F((f), (i), (j))]

/* ------------------------------------------------------------ mesh metrics
 * src/domain/SMOD_mesh_manipulation.f90:28-51 (coordinates) and :147-173
 * (length ratios, shoelace area, capacity) for the identity map
 * (src/user_subroutines/SMOD_coordinate_mapping.f90:35-43). */
static double d_xi_of(const orc_cfg* c) { return (c->xi_max - c->xi_min) / c->nx; }
static double d_eta_of(const orc_cfg* c) { return (c->eta_max - c->eta_min) / c->ny; }

/* cell centre, SMOD_mesh_manipulation.f90:50-51 */
static inline double centre(double lo, double d, int i) {
  return lo + (double)(i - 1) * d + d / 2;
}

typedef struct {
  double kappa, gamma_xi, gamma_eta;
} metric_t;

static inline metric_t metric_at(double xi_min, double d_xi, double eta_min,
                                 double d_eta, int i, int j) {
  metric_t m;
  /* corners, SMOD_mesh_manipulation.f90:36-48 */
  double x1 = xi_min + (double)(i - 1) * d_xi;
  double x2 = x1 + d_xi;
  double x3 = x2, x4 = x1, x5 = x1;
  double y1 = eta_min + (double)(j - 1) * d_eta;
  double y2 = y1;
  double y3 = y2 + d_eta;
  double y4 = y3, y5 = y1;
  /* xi-face 1->4 (:147-153); norm2 of (0, dy) is |dy| */
  double fdx = x4 - x1, fdy = y4 - y1;
  m.gamma_xi = sqrt(fdx * fdx + fdy * fdy) / d_eta;
  /* eta-face 1->2 (:156-162) */
  fdx = x2 - x1;
  fdy = y2 - y1;
  m.gamma_eta = sqrt(fdx * fdx + fdy * fdy) / d_xi;
  /* shoelace area (:165-170), then capacity (:173) */
  double a = 0.0;
  a = a + 0.5 * (y1 + y2) * (x1 - x2);
  a = a + 0.5 * (y2 + y3) * (x2 - x3);
  a = a + 0.5 * (y3 + y4) * (x3 - x4);
  a = a + 0.5 * (y4 + y5) * (x4 - x5);
  m.kappa = a / (d_xi * d_eta);
  return m;
}

/* metrics of cell (i,j): the caller's arrays, or the identity map's */
static inline double kap_of(const orc_cfg* c, const frame_t* f, double dx, double dy, int i, int j) {
  if (c->kappa) return c->kappa[F(f, i, j)];
  return metric_at(c->xi_min, dx, c->eta_min, dy, i, j).kappa;
}
static inline double gxi_of(const orc_cfg* c, const frame_t* f, double dx, double dy, int i, int j) {
  if (c->gamma_xi) return c->gamma_xi[F(f, i, j)];
  return metric_at(c->xi_min, dx, c->eta_min, dy, i, j).gamma_xi;
}
static inline double geta_of(const orc_cfg* c, const frame_t* f, double dx, double dy, int i, int j) {
  if (c->gamma_eta) return c->gamma_eta[F(f, i, j)];
  return metric_at(c->xi_min, dx, c->eta_min, dy, i, j).gamma_eta;
}

int orc_metrics(const orc_cfg* c, double* xi, double* eta, double* kappa,
                double* gamma_xi, double* gamma_eta) {
  frame_t f = frame_of(c);
  double dx = d_xi_of(c), dy = d_eta_of(c);
  for (int j = -1; j <= c->ny + 2; ++j)
    for (int i = -1; i <= c->nx + 2; ++i) {
      long k = F(&f, i, j);
      metric_t m = metric_at(c->xi_min, dx, c->eta_min, dy, i, j);
      if (xi) xi[k] = centre(c->xi_min, dx, i);
      if (eta) eta[k] = centre(c->eta_min, dy, j);
      if (kappa) kappa[k] = m.kappa;
      if (gamma_xi) gamma_xi[k] = m.gamma_xi;
      if (gamma_eta) gamma_eta[k] = m.gamma_eta;
    }
  return 0;
}

/* ------------------------------------------------------ initial condition
 * examples/2D_dam_break/cartesian/SMOD_set_initial_conditions.f90:34-46 */
int orc_init_dam_break(const orc_cfg* c, double* vars, double h_in,
                       double h_out, double radius) {
  frame_t f = frame_of(c);
  double dx = d_xi_of(c), dy = d_eta_of(c);
  for (int j = -1; j <= c->ny + 2; ++j)
    for (int i = -1; i <= c->nx + 2; ++i) {
      double x = centre(c->xi_min, dx, i);
      double y = centre(c->eta_min, dy, j);
      Q(vars, &f, 0, i, j) = (sqrt(x * x + y * y) < radius) ? h_in : h_out;
      Q(vars, &f, 1, i, j) = 0.0;
      Q(vars, &f, 2, i, j) = 0.0;
    }
  return 0;
}

/* ---------------------------------------------------------- ghost fill
 * src/user_subroutines/SMOD_boundary_conditions.f90:58-213 (BC 2 and 3),
 * deriv_2p src/utility/SMOD_derivatives.f90:5-11.  Corners are never filled.
 * BC 4 (reflective wall: mirror the first two interior cells, negate the
 * normal momentum) is this build's addition. */
static inline double deriv_2p(double x1, double x2, double y1, double y2) {
  return (y2 - y1) / (x2 - x1);
}

int orc_fill_ghosts(const orc_cfg* c, double* vars) {
  frame_t f = frame_of(c);
  const int nx = c->nx, ny = c->ny;
  double dx = d_xi_of(c), dy = d_eta_of(c);
  for (int side = 0; side < 4; ++side)
    if (c->bc[side] < ORC_BC_ZERO_GRADIENT || c->bc[side] > ORC_BC_REFLECTIVE)
      return -10 - side; /* abort_program */
  /* left */
  for (int j = 1; j <= ny; ++j)
    for (int i = -1; i <= 0; ++i)
      for (int me = 0; me < 3; ++me) {
        double v;
        switch (c->bc[0]) {
          case ORC_BC_ZERO_GRADIENT: v = Q(vars, &f, me, 1, j); break;
          case ORC_BC_CONSTANT_GRADIENT: {
            double x1 = centre(c->xi_min, dx, 1), x2 = centre(c->xi_min, dx, 2);
            double xg = centre(c->xi_min, dx, i);
            v = Q(vars, &f, me, 1, j) +
                deriv_2p(x1, x2, Q(vars, &f, me, 1, j), Q(vars, &f, me, 2, j)) * (xg - x1);
          } break;
          default: v = Q(vars, &f, me, 1 - i, j); if (me == 1) v = -v; break;
        }
        Q(vars, &f, me, i, j) = v;
      }
  /* right */
  for (int j = 1; j <= ny; ++j)
    for (int i = nx + 1; i <= nx + 2; ++i)
      for (int me = 0; me < 3; ++me) {
        double v;
        switch (c->bc[1]) {
          case ORC_BC_ZERO_GRADIENT: v = Q(vars, &f, me, nx, j); break;
          case ORC_BC_CONSTANT_GRADIENT: {
            double x1 = centre(c->xi_min, dx, nx - 1), x2 = centre(c->xi_min, dx, nx);
            double xg = centre(c->xi_min, dx, i);
            v = Q(vars, &f, me, nx, j) +
                deriv_2p(x1, x2, Q(vars, &f, me, nx - 1, j), Q(vars, &f, me, nx, j)) * (xg - x2);
          } break;
          default: v = Q(vars, &f, me, 2 * nx + 1 - i, j); if (me == 1) v = -v; break;
        }
        Q(vars, &f, me, i, j) = v;
      }
  /* bottom */
  for (int j = -1; j <= 0; ++j)
    for (int i = 1; i <= nx; ++i)
      for (int me = 0; me < 3; ++me) {
        double v;
        switch (c->bc[2]) {
          case ORC_BC_ZERO_GRADIENT: v = Q(vars, &f, me, i, 1); break;
          case ORC_BC_CONSTANT_GRADIENT: {
            double y1 = centre(c->eta_min, dy, 1), y2 = centre(c->eta_min, dy, 2);
            double yg = centre(c->eta_min, dy, j);
            v = Q(vars, &f, me, i, 1) +
                deriv_2p(y1, y2, Q(vars, &f, me, i, 1), Q(vars, &f, me, i, 2)) * (yg - y1);
          } break;
          default: v = Q(vars, &f, me, i, 1 - j); if (me == 2) v = -v; break;
        }
        Q(vars, &f, me, i, j) = v;
      }
  /* top */
  for (int j = ny + 1; j <= ny + 2; ++j)
    for (int i = 1; i <= nx; ++i)
      for (int me = 0; me < 3; ++me) {
        double v;
        switch (c->bc[3]) {
          case ORC_BC_ZERO_GRADIENT: v = Q(vars, &f, me, i, ny); break;
          case ORC_BC_CONSTANT_GRADIENT: {
            double y1 = centre(c->eta_min, dy, ny - 1), y2 = centre(c->eta_min, dy, ny);
            double yg = centre(c->eta_min, dy, j);
            v = Q(vars, &f, me, i, ny) +
                deriv_2p(y1, y2, Q(vars, &f, me, i, ny - 1), Q(vars, &f, me, i, ny)) * (yg - y2);
          } break;
          default: v = Q(vars, &f, me, i, 2 * ny + 1 - j); if (me == 2) v = -v; break;
        }
        Q(vars, &f, me, i, j) = v;
      }
  return 0;
}

/* ------------------------------------------- physics: J, eigenvalues, J*q
 * examples/2D_dam_break/cartesian/SMOD_jakobian.f90:44-53 (J entries),
 * closed-form max|eig(J)| = |u| + sqrt(g h), and J*q in BLAS column order.
 * Returns nonzero when the reference would stop (NaN in J, complex
 * eigenvalues).  With a wet/dry threshold (this build's physics), a cell
 * with h <= dry_tol has u = v = 0. */
static inline int phys_x(const orc_cfg* c, const double q[3], double* s,
                         double fl[3]) {
  const double g = c->grav, h = q[0], hu = q[1], hv = q[2];
  if (c->dry_tol > 0.0 && !(h > c->dry_tol)) {
    double hp = h > 0.0 ? h : 0.0;
    *s = sqrt(g * hp);
    fl[0] = hu;
    fl[1] = g * h * h;
    fl[2] = 0.0;
    return 0;
  }
  double J21 = -(hu * hu / (h * h)) + g * h;
  double J22 = 2.0 * hu / h;
  double J31 = -(hu * hv / (h * h));
  double J32 = hv / h;
  double J33 = hu / h;
  if (isnan(J21) || isnan(J22) || isnan(J31) || isnan(J32) || isnan(J33)) return 1;
  /* dgeev returns u +- i*sqrt(-g h) for g h < 0; the reference stops only
   * when that imaginary part exceeds ud_precision
   * (src/utility/SMOD_eigenvalue_problem.f90:72-83) */
  double gh = g * h, cc = 0.0;
  if (gh < 0.0) { if (sqrt(-gh) > c->eps) return 2; } else cc = sqrt(gh);
  *s = fabs(J33) + cc;
  fl[0] = hu;
  fl[1] = J21 * h + J22 * hu;
  fl[2] = (J31 * h + J32 * hu) + J33 * hv;
  return 0;
}

static inline int phys_y(const orc_cfg* c, const double q[3], double* s,
                         double fl[3]) {
  const double g = c->grav, h = q[0], hu = q[1], hv = q[2];
  if (c->dry_tol > 0.0 && !(h > c->dry_tol)) {
    double hp = h > 0.0 ? h : 0.0;
    *s = sqrt(g * hp);
    fl[0] = hv;
    fl[1] = 0.0;
    fl[2] = g * h * h;
    return 0;
  }
  double J21 = -(hu * hv / (h * h));
  double J22 = hv / h;
  double J23 = hu / h;
  double J31 = -(hv * hv / (h * h)) + g * h;
  double J33 = 2.0 * hv / h;
  if (isnan(J21) || isnan(J22) || isnan(J23) || isnan(J31) || isnan(J33)) return 1;
  double gh = g * h, cc = 0.0;
  if (gh < 0.0) { if (sqrt(-gh) > c->eps) return 2; } else cc = sqrt(gh);
  *s = fabs(J22) + cc;
  fl[0] = hv;
  fl[1] = (J21 * h + J22 * hu) + J23 * hv;
  fl[2] = J31 * h + J33 * hv;
  return 0;
}

/* ----------------------------- the reference's per-interface structure
 * Optional (orc_use_lapack): each cell of an interface gets the reference's
 * Jacobian (examples/2D_dam_break/cartesian/SMOD_jakobian.f90:44-53, column
 * major), its eigenvalues from LAPACK dgeev exactly as calculateEigenSolution
 * calls it (src/utility/SMOD_eigenvalue_problem.f90:5-89: a workspace query,
 * an allocated work array, right eigenvectors requested, complex check) and
 * J*q from BLAS dgemv (SMOD_matrix_operations.f90:5-23).  This is the cost
 * structure of the reference's CPU path; its values agree with the closed
 * forms above to rounding (the eigenvalues are LAPACK's, J*q is BLAS's). */
typedef void (*orc_dgeev_fn)(char*, char*, int*, double*, int*, double*, double*, double*, int*,
                             double*, int*, double*, int*, int*);
typedef void (*orc_dgemv_fn)(char*, int*, int*, double*, double*, int*, double*, int*, double*,
                             double*, int*);
static orc_dgeev_fn g_dgeev = 0;
static orc_dgemv_fn g_dgemv = 0;

void orc_use_lapack(void* dgeev, void* dgemv) {
  g_dgeev = (dgeev && dgemv) ? (orc_dgeev_fn)dgeev : 0;
  g_dgemv = (dgeev && dgemv) ? (orc_dgemv_fn)dgemv : 0;
}

static int phys_lapack(const orc_cfg* c, int ixy, const double q[3], double* s, double fl[3]) {
  const double g = c->grav, h = q[0], hu = q[1], hv = q[2];
  double J[9] = {0.0};  /* J(k, l) at J[(l - 1) * 3 + (k - 1)] */
#define JM(k, l) J[((l) - 1) * 3 + ((k) - 1)]
  if (ixy == 1) {
    JM(1, 2) = 1.0;
    JM(2, 1) = -(hu * hu / (h * h)) + g * h;
    JM(2, 2) = 2.0 * hu / h;
    JM(3, 1) = -(hu * hv / (h * h));
    JM(3, 2) = hv / h;
    JM(3, 3) = hu / h;
  } else {
    JM(1, 3) = 1.0;
    JM(2, 1) = -(hu * hv / (h * h));
    JM(2, 2) = hv / h;
    JM(2, 3) = hu / h;
    JM(3, 1) = -(hv * hv / (h * h)) + g * h;
    JM(3, 3) = 2.0 * hv / h;
  }
#undef JM
  for (int k = 0; k < 9; ++k)
    if (isnan(J[k])) return 1;
  char jobvl = 'N', jobvr = 'V';
  int n = 3, lda = 3, ldvl = 1, ldvr = 3, lwork = -1, info = 0;
  double A[9], wr[3] = {0.0}, wi[3] = {0.0}, vl[1] = {0.0}, R[9] = {0.0}, worktmp[1] = {0.0};
  memcpy(A, J, sizeof A);
  double* work = malloc(sizeof(double));
  g_dgeev(&jobvl, &jobvr, &n, A, &lda, wr, wi, vl, &ldvl, R, &ldvr, worktmp, &lwork, &info);
  lwork = (int)worktmp[0];
  free(work);
  work = calloc((size_t)(lwork > 1 ? lwork : 1), sizeof(double));
  g_dgeev(&jobvl, &jobvr, &n, A, &lda, wr, wi, vl, &ldvl, R, &ldvr, work, &lwork, &info);
  free(work);
  for (int k = 0; k < 3; ++k)
    if (wi[k] > c->eps) return 2;
  double sm = 0.0;
  for (int k = 0; k < 3; ++k) sm = fabs(wr[k]) > sm ? fabs(wr[k]) : sm;
  *s = sm;
  char tr = 'N';
  int inc = 1;
  double one = 1.0, zero = 0.0, x[3] = {q[0], q[1], q[2]};
  g_dgemv(&tr, &n, &n, &one, J, &lda, x, &inc, &zero, fl, &inc);
  return 0;
}

/* --------------------------------------------------------- work arrays
 * src/solver/MOD_work_arrays.f90:14-70; zeroed after every sweep (:78-109). */
typedef struct {
  double *lambda, *waves, *Am, *Ap, *wl, *wr, *f;
  long n;
} work_t;

static int work_alloc(work_t* w, const frame_t* f) {
  w->n = (long)f->nxf * f->nyf;
  w->lambda = calloc((size_t)(2 * w->n), sizeof(double));
  w->waves = calloc((size_t)(6 * w->n), sizeof(double));
  w->Am = calloc((size_t)(3 * w->n), sizeof(double));
  w->Ap = calloc((size_t)(3 * w->n), sizeof(double));
  w->wl = calloc((size_t)(2 * w->n), sizeof(double));
  w->wr = calloc((size_t)(2 * w->n), sizeof(double));
  w->f = calloc((size_t)(3 * w->n), sizeof(double));
  return (w->lambda && w->waves && w->Am && w->Ap && w->wl && w->wr && w->f) ? 0 : -1;
}
static void work_free(work_t* w) {
  free(w->lambda); free(w->waves); free(w->Am); free(w->Ap);
  free(w->wl); free(w->wr); free(w->f);
}
static void work_zero(work_t* w) {
#pragma omp parallel for schedule(static)
  for (long k = 0; k < w->n; ++k) {
    w->lambda[2 * k] = w->lambda[2 * k + 1] = 0.0;
    w->wl[2 * k] = w->wl[2 * k + 1] = 0.0;
    w->wr[2 * k] = w->wr[2 * k + 1] = 0.0;
    for (int e = 0; e < 3; ++e) {
      w->Am[3 * k + e] = 0.0;
      w->Ap[3 * k + e] = 0.0;
      w->f[3 * k + e] = 0.0;
    }
    for (int e = 0; e < 6; ++e) w->waves[6 * k + e] = 0.0;
  }
}
#define LAM(w, k, mw) (w)->lambda[2 * (k) + (mw)]
#define WAV(w, k, me, mw) (w)->waves[6 * (k) + 3 * (mw) + (me)]

/* ------------------------------------------------------------- HLL
 * src/solver/SMOD_approximate_Riemann_solvers.f90:15-102 for the interface
 * between cell L (i-1 or j-1) and R (i or j). */
static inline int hll(const orc_cfg* c, int ixy, const double* qL,
                      const double* qR, work_t* w, long k) {
  double sL, sR, fL[3], fR[3];
  int e1, e2;
  if (g_dgeev && !(c->dry_tol > 0.0)) {  /* the reference's structure (no wet/dry in it) */
    e1 = phys_lapack(c, ixy, qL, &sL, fL);
    e2 = phys_lapack(c, ixy, qR, &sR, fR);
  } else {
    e1 = (ixy == 1) ? phys_x(c, qL, &sL, fL) : phys_y(c, qL, &sL, fL);
    e2 = (ixy == 1) ? phys_x(c, qR, &sR, fR) : phys_y(c, qR, &sR, fR);
  }
  if (e1 || e2) return 1;
  double s2 = sL > sR ? sL : sR;
  double s1 = -s2;
  LAM(w, k, 1) = s2;
  LAM(w, k, 0) = s1;
  const double eps = c->eps;
  if (fabs(s1) > eps && fabs(s2) > eps) {
    double den = eps + s1 - s2;
    for (int me = 0; me < 3; ++me) {
      double qs = (fR[me] - fL[me] - s2 * qR[me] + s1 * qL[me]) / den;
      WAV(w, k, me, 0) = qs - qL[me];
      WAV(w, k, me, 1) = qR[me] - qs;
    }
  } else {
    LAM(w, k, 0) = 0.0;
    LAM(w, k, 1) = 0.0;
    for (int me = 0; me < 3; ++me) {
      WAV(w, k, me, 0) = 0.0;
      WAV(w, k, me, 1) = 0.0;
    }
  }
  return 0;
}

/* limiter functions, src/solver/MOD_flux_correction.f90:101-112 */
static inline double phi_of(int code, double r) {
  switch (code) {
    case 1: return fmax(0.0, fmin(1.0, r));
    case 2: return fmax(fmax(0.0, fmin(1.0, 2.0 * r)), fmin(2.0, r));
    case 3: return (r + fabs(r)) / (1.0 + fabs(r));
    default: return fmax(0.0, fmin(fmin((1.0 + r) / 2.0, 2.0), 2.0 * r));
  }
}

/* ---------------------------------------------------------------- sweep
 * solve_dimensional_split's per-sweep body, src/solver/MOD_solve.f90:71-105
 * (xi) and :124-157 (eta), with its callees:
 *   wave_decomposition  MOD_flux_calculation.f90:61-79
 *   split_flux          MOD_flux_calculation.f90:82-121
 *   integrate_flux_difference       MOD_solve.f90:177-238
 *   limiter             MOD_flux_correction.f90:17-125
 *   second_order_flux_correction    MOD_flux_correction.f90:127-187
 *   integrator_flux_correction_increment  MOD_solve.f90:240-293 */
static int sweep_impl(const orc_cfg* c, work_t* w, double* vars, int ixy,
                      double dt, double* cfl_local) {
  frame_t fr = frame_of(c);
  const frame_t* f = &fr;
  const int nx = c->nx, ny = c->ny;
  const double dxi = d_xi_of(c), deta = d_eta_of(c);
  const double delta = (ixy == 1) ? dxi : deta;
  int err = 0;
  double cfl = *cfl_local;
  /* invalid limiter choice => error stop (MOD_flux_correction.f90:110-111) */
  if (c->higher_order && c->use_limiter && (c->limiter < 1 || c->limiter > 4))
    return -3;

  /* wave decomposition + CFL (MOD_solve.f90:72-86 / :126-139) */
  if (ixy == 1) {
#pragma omp parallel for reduction(max : cfl) reduction(| : err) schedule(static)
    for (int j = 1; j <= ny; ++j)
      for (int i = 0; i <= nx + 1; ++i) {
        long k = F(f, i, j);
        err |= hll(c, 1, &Q(vars, f, 0, i - 1, j), &Q(vars, f, 0, i, j), w, k);
        for (int mw = 0; mw < 2; ++mw) {
          double v = LAM(w, k, mw) * dt / dxi;
          cfl = cfl > v ? cfl : v;
        }
      }
  } else {
#pragma omp parallel for reduction(max : cfl) reduction(| : err) schedule(static)
    for (int i = 1; i <= nx; ++i)
      for (int j = 0; j <= ny + 1; ++j) {
        long k = F(f, i, j);
        err |= hll(c, 2, &Q(vars, f, 0, i, j - 1), &Q(vars, f, 0, i, j), w, k);
        for (int mw = 0; mw < 2; ++mw) {
          double v = LAM(w, k, mw) * dt / deta;
          cfl = cfl > v ? cfl : v;
        }
      }
  }
  if (err) return -2;
  *cfl_local = cfl;

  /* split_flux over j in [0, ny+1], i in [0, nx+1] for both sweeps */
#pragma omp parallel for schedule(static)
  for (int j = 0; j <= ny + 1; ++j)
    for (int i = 0; i <= nx + 1; ++i) {
      long k = F(f, i, j);
      for (int mw = 0; mw < 2; ++mw) {
        double s = LAM(w, k, mw);
        for (int me = 0; me < 3; ++me) {
          double wv = WAV(w, k, me, mw);
          double mask = (s < 0.0) ? 1.0 : 0.0;
          w->Am[3 * k + me] = w->Am[3 * k + me] + mask * s * wv;
          w->Ap[3 * k + me] = w->Ap[3 * k + me] + (1.0 - mask) * s * wv;
        }
      }
    }

  /* integrate_flux_difference */
  const double dtod = dt / delta;
  const double d_xi = dxi, d_eta = deta;
#pragma omp parallel for schedule(static)
  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i) {
      const double kap = kap_of(c, f, d_xi, d_eta, i, j);
      double gl, gr;
      long kr;
      if (ixy == 1) {
        gl = gxi_of(c, f, d_xi, d_eta, i, j);
        gr = gxi_of(c, f, d_xi, d_eta, i + 1, j);
        kr = F(f, i + 1, j);
      } else {
        gl = geta_of(c, f, d_xi, d_eta, i, j);
        gr = geta_of(c, f, d_xi, d_eta, i, j + 1);
        kr = F(f, i, j + 1);
      }
      long k = F(f, i, j);
      for (int me = 0; me < 3; ++me)
        Q(vars, f, me, i, j) = Q(vars, f, me, i, j) -
            dtod / kap * (gl * w->Ap[3 * k + me] + gr * w->Am[3 * kr + me]);
    }

  if (c->higher_order) {
    /* limiter (serial in the reference; each row/column here is an
     * independent chain, so rows run in parallel with identical results) */
    if (c->use_limiter) {
      if (ixy == 1) {
#pragma omp parallel for schedule(static)
        for (int j = 1; j <= ny; ++j) {
          long k0 = F(f, 0, j), k1 = F(f, 1, j);
          for (int mw = 0; mw < 2; ++mw)
            for (int me = 0; me < 3; ++me)
              w->wr[2 * k0 + mw] = w->wr[2 * k0 + mw] + WAV(w, k0, me, mw) * WAV(w, k1, me, mw);
        }
      } else {
#pragma omp parallel for schedule(static)
        for (int i = 1; i <= nx; ++i) {
          long k0 = F(f, i, 0), k1 = F(f, i, 1);
          for (int mw = 0; mw < 2; ++mw)
            for (int me = 0; me < 3; ++me)
              w->wr[2 * k0 + mw] = w->wr[2 * k0 + mw] + WAV(w, k0, me, mw) * WAV(w, k1, me, mw);
        }
      }
      const int outer_n = (ixy == 1) ? ny : nx;
      const int inner_n = (ixy == 1) ? nx : ny;
#pragma omp parallel for schedule(static)
      for (int o = 1; o <= outer_n; ++o)
        for (int in = 1; in <= inner_n; ++in) {
          int i = (ixy == 1) ? in : o;
          int j = (ixy == 1) ? o : in;
          long k = F(f, i, j);
          long kprev = (ixy == 1) ? F(f, i - 1, j) : F(f, i, j - 1);
          long knext = (ixy == 1) ? F(f, i + 1, j) : F(f, i, j + 1);
          for (int mw = 0; mw < 2; ++mw) {
            double wnorm2 = 0.0;
            w->wl[2 * k + mw] = w->wr[2 * kprev + mw];
            w->wr[2 * k + mw] = 0.0;
            for (int me = 0; me < 3; ++me) {
              wnorm2 = wnorm2 + WAV(w, k, me, mw) * WAV(w, k, me, mw);
              w->wr[2 * k + mw] = w->wr[2 * k + mw] + WAV(w, k, me, mw) * WAV(w, knext, me, mw);
            }
            double up = (LAM(w, k, mw) > 0.0) ? w->wl[2 * k + mw] : w->wr[2 * k + mw];
            if (!(fabs(wnorm2) <= c->eps)) {
              double r = up / wnorm2;
              double phi = phi_of(c->limiter, r);
              for (int me = 0; me < 3; ++me) WAV(w, k, me, mw) = phi * WAV(w, k, me, mw);
            }
          }
        }
    }
    /* second_order_flux_correction */
    if (ixy == 1) {
#pragma omp parallel for schedule(static)
      for (int j = 1; j <= ny; ++j)
        for (int i = 0; i <= nx + 1; ++i) {
          long k = F(f, i, j);
          double ka = kap_of(c, f, d_xi, d_eta, i - 1, j);
          double kb = kap_of(c, f, d_xi, d_eta, i, j);
          for (int mw = 0; mw < 2; ++mw) {
            double al = fabs(LAM(w, k, mw));
            for (int me = 0; me < 3; ++me)
              w->f[3 * k + me] = w->f[3 * k + me] +
                  0.5 * al * (1.0 - al * dtod / (0.5 * (ka + kb))) * WAV(w, k, me, mw);
          }
        }
    } else {
#pragma omp parallel for schedule(static)
      for (int j = 0; j <= ny + 1; ++j)
        for (int i = 1; i <= nx; ++i) {
          long k = F(f, i, j);
          double ka = kap_of(c, f, d_xi, d_eta, i, j - 1);
          double kb = kap_of(c, f, d_xi, d_eta, i, j);
          for (int mw = 0; mw < 2; ++mw) {
            double al = fabs(LAM(w, k, mw));
            for (int me = 0; me < 3; ++me)
              w->f[3 * k + me] = w->f[3 * k + me] +
                  0.5 * al * (1.0 - al * dtod / (0.5 * (ka + kb))) * WAV(w, k, me, mw);
          }
        }
    }
    /* integrator_flux_correction_increment */
#pragma omp parallel for schedule(static)
    for (int j = 1; j <= ny; ++j)
      for (int i = 1; i <= nx; ++i) {
        double kap = kap_of(c, f, d_xi, d_eta, i, j);
        long k = F(f, i, j);
        long kr = (ixy == 1) ? F(f, i + 1, j) : F(f, i, j + 1);
        for (int me = 0; me < 3; ++me)
          Q(vars, f, me, i, j) = Q(vars, f, me, i, j) -
              dtod / kap * (w->f[3 * kr + me] - w->f[3 * k + me]);
      }
  }
  return 0;
}

int orc_sweep(const orc_cfg* c, double* vars, int ixy, double dt,
              double* cfl_local) {
#ifdef _OPENMP
  if (c->nthreads > 0) omp_set_num_threads(c->nthreads);
#endif
  frame_t f = frame_of(c);
  work_t w;
  if (work_alloc(&w, &f)) { work_free(&w); return -1; }
  int rc = sweep_impl(c, &w, vars, ixy, dt, cfl_local);
  work_free(&w);
  return rc;
}

/* ---------------------------------------------------------- time control
 * src/time_control/MOD_time_control.f90 */
int orc_tc_initialize(orc_tc* tc, double t0, double t1, double dt0,
                      int32_t n_snap, double cfl_target, double dt_max,
                      double dt_min, double growth, double shrink,
                      int32_t adaptive) {
  memset(tc, 0, sizeof(*tc));
  tc->time_initial = t0;
  tc->time_final = t1;
  tc->time = t0;
  tc->dt = dt0;
  tc->dt_previous = dt0;
  tc->dt_max = dt_max;
  tc->dt_min = dt_min;
  tc->adaptive = adaptive;
  tc->emergency_shrink_ratio = 0.25;
  if (tc->dt_min > tc->dt_max) { tc->status = 1; return -1; } /* :114-119 */
  tc->number_of_snapshots = n_snap < 1 ? 1 : n_snap;       /* :122-130 */
  if (tc->number_of_snapshots > 4096) return -2;
  double interval = (tc->time_final - tc->time_initial) / tc->number_of_snapshots;
  for (int i = 1; i <= tc->number_of_snapshots; ++i)
    tc->snapshot_intervals[i - 1] = tc->time_initial + i * interval;
  tc->snapshot_intervals[tc->number_of_snapshots - 1] = tc->time_final;
  tc->current_snapshot_index = 1;
  tc->CFL = 0.0;
  tc->CFL_previous_timestep = 0.0;
  tc->CFL_target = cfl_target;
  tc->time_precision = 0.01;
  tc->growth_limit = growth;
  tc->shrink_limit = shrink;
  return 0;
}

/* updateTime, MOD_time_control.f90:153-243 */
void orc_tc_update_time(orc_tc* tc, int32_t step_valid) {
  double change_factor, dt_step;
  tc->dt_previous = tc->dt;
  dt_step = tc->dt;
  if (step_valid) {
    if (tc->adaptive) {
      if (tc->CFL > 0.0) change_factor = tc->CFL_target / tc->CFL;
      else change_factor = 1.0;
      if (change_factor >= tc->growth_limit) change_factor = tc->growth_limit;
      else if (fabs(change_factor - 1.0) <= tc->time_precision) change_factor = 1.0;
      dt_step = tc->dt_previous * change_factor;
    }
    double snap = tc->snapshot_intervals[tc->current_snapshot_index - 1];
    if (tc->time + dt_step > snap) {
      dt_step = snap - tc->time;
      tc->snapshot_exception = 1;
    } else {
      tc->snapshot_exception = 0;
    }
  } else {
    if (!tc->adaptive) { tc->status = 1; return; } /* end_program */
    dt_step = tc->emergency_shrink_ratio * tc->dt;
  }
  if (dt_step > tc->dt_max) dt_step = tc->dt_max;
  if (dt_step < tc->dt_min) {
    if (!tc->adaptive && !tc->snapshot_exception) { tc->status = 1; return; }
    dt_step = tc->dt_min;
  }
  if (step_valid) tc->time = tc->time + dt_step;
  tc->dt = dt_step;
}

void orc_tc_update_max_cfl(orc_tc* tc, double cfl) {
  tc->CFL = tc->CFL > cfl ? tc->CFL : cfl; /* Fortran max */
}

int32_t orc_tc_check_validity(const orc_tc* tc) {
  return tc->CFL_target / tc->CFL >= tc->shrink_limit;
}

void orc_tc_reset_cfl(orc_tc* tc) {
  tc->CFL_previous_timestep = tc->CFL;
  tc->CFL = 0.0;
}

/* ------------------------------------------- solve_dimensional_split
 * src/solver/MOD_solve.f90:23-175 on one rank (find_min_max_scalar is the
 * identity there). */
static int solve_impl(const orc_cfg* c, work_t* w, orc_tc* tc, double* vars) {
  double cfl_local = 0.0;
  int rc = sweep_impl(c, w, vars, 1, tc->dt, &cfl_local);
  if (rc) return rc - 100;
  orc_tc_update_max_cfl(tc, cfl_local);
  work_zero(w);
  int valid = orc_tc_check_validity(tc) ? 1 : 0;
  if (!valid) return -1;
  rc = sweep_impl(c, w, vars, 2, tc->dt, &cfl_local);
  if (rc) return rc - 100;
  orc_tc_update_max_cfl(tc, cfl_local);
  work_zero(w);
  valid = valid * (orc_tc_check_validity(tc) ? 1 : 0);
  return valid ? 1 : -1;
}

int orc_solve_dimensional_split(const orc_cfg* c, orc_tc* tc, double* vars) {
#ifdef _OPENMP
  if (c->nthreads > 0) omp_set_num_threads(c->nthreads);
#endif
  frame_t f = frame_of(c);
  work_t w;
  if (work_alloc(&w, &f)) { work_free(&w); return -1000; }
  int rc = solve_impl(c, &w, tc, vars);
  work_free(&w);
  return rc;
}

/* --------------------------------------------------------- source term
 * integrate_source_term, src/solver/MOD_solve.f90:295-336: forward Euler
 * q += phi*dt over the interior.  The reference's phi is 0
 * (src/user_subroutines/SMOD_source_function.f90:15-28); q + 0*dt leaves
 * every value as it is except an interior -0.0, which becomes +0.0.  The
 * other kinds are this build's physics:
 *   bathymetry  phi_hu = -g h (b(i+1)-b(i-1))/dxi, phi_hv likewise in eta
 *               (pressure term g h^2 of the reference flux, so the balancing
 *               slope term carries 2 g h db/dx),
 *   friction    semi-implicit Manning: hu /= 1 + dt g n^2 |u| / h^(4/3),
 *   wet/dry     cells with h <= dry_tol get hu = hv = 0 and h = max(h, 0). */
int orc_source(const orc_cfg* c, double* vars, const double* bathy, double dt) {
  frame_t fr = frame_of(c);
  const frame_t* f = &fr;
  if (c->source == ORC_SRC_NONE && !(c->dry_tol > 0.0)) {
    /* the reference's q = q + phi*dt with phi = 0 (MOD_solve.f90:319-327) */
    const double zero_dt = 0.0 * dt;
#pragma omp parallel for schedule(static)
    for (int j = 1; j <= c->ny; ++j)
      for (int i = 1; i <= c->nx; ++i)
        for (int me = 0; me < 3; ++me) Q(vars, f, me, i, j) = Q(vars, f, me, i, j) + zero_dt;
    return 0;
  }
  const double g = c->grav, dxi = d_xi_of(c), deta = d_eta_of(c);
  const int use_b = (c->source == ORC_SRC_BATHY || c->source == ORC_SRC_BATHY_FRICTION);
  const int use_fr = (c->source == ORC_SRC_FRICTION || c->source == ORC_SRC_BATHY_FRICTION);
  if (use_b && !bathy) return -1;
  const double n2 = c->manning_n * c->manning_n;
#pragma omp parallel for schedule(static)
  for (int j = 1; j <= c->ny; ++j)
    for (int i = 1; i <= c->nx; ++i) {
      double h = Q(vars, f, 0, i, j), hu = Q(vars, f, 1, i, j), hv = Q(vars, f, 2, i, j);
      if (!(h > c->dry_tol)) {
        Q(vars, f, 0, i, j) = h > 0.0 ? h : 0.0;
        Q(vars, f, 1, i, j) = 0.0;
        Q(vars, f, 2, i, j) = 0.0;
        continue;
      }
      double nhu = hu, nhv = hv;
      if (use_b) {
        double bx = bathy[F(f, i + 1, j)] - bathy[F(f, i - 1, j)];
        double by = bathy[F(f, i, j + 1)] - bathy[F(f, i, j - 1)];
        nhu = hu + dt * (-g * h * bx / dxi);
        nhv = hv + dt * (-g * h * by / deta);
      }
      if (use_fr) {
        double u = hu / h, v = hv / h;
        double sp = sqrt(u * u + v * v);
        double h43 = cbrt(h) * h;
        double den = 1.0 + dt * g * n2 * sp / h43;
        nhu = nhu / den;
        nhv = nhv / den;
      }
      /* every kind ends with the reference's + 0*dt, so that a zero slope or
       * a zero Manning coefficient gives the phi = 0 step's bits */
      Q(vars, f, 0, i, j) = h + 0.0;
      Q(vars, f, 1, i, j) = nhu + 0.0;
      Q(vars, f, 2, i, j) = nhv + 0.0;
    }
  return 0;
}

/* ------------------------------------------------------------ main loop
 * src/main.f90:193-283 without output: copy, solve, reject/rollback,
 * source, ghost fill, updateTime, snapshot index, reset_CFL. */
int orc_run(const orc_cfg* c, orc_tc* tc, double* vars, const double* bathy,
            int32_t max_accepted, int32_t max_attempts, orc_step_rec* rec,
            int32_t rec_cap, int32_t fill_first) {
#ifdef _OPENMP
  if (c->nthreads > 0) omp_set_num_threads(c->nthreads);
#endif
  frame_t f = frame_of(c);
  long nv = 3L * f.nxf * f.nyf;
  double* old = malloc((size_t)nv * sizeof(double));
  work_t w;
  if (!old || work_alloc(&w, &f)) { free(old); return -1000; }
  /* the ghost fill before the loop (src/main.f90:161, :189); a run that
   * continues an earlier one (fill_first = 0) starts from the state as it is */
  int rc = fill_first ? orc_fill_ghosts(c, vars) : 0;
  int attempts = 0, accepted = 0;
  while (rc == 0 && tc->time < tc->time_final) {
    if (max_accepted > 0 && accepted >= max_accepted) break;
    if (max_attempts > 0 && attempts >= max_attempts) break;
    memcpy(old, vars, (size_t)nv * sizeof(double));
    double dt_used = tc->dt;
    int ierr = solve_impl(c, &w, tc, vars);
    if (ierr < -1) { rc = ierr; break; }
    orc_step_rec r;
    memset(&r, 0, sizeof r);
    r.dt_used = dt_used;
    r.cfl = tc->CFL;
    if (ierr < 0) {
      orc_tc_update_time(tc, 0);
      memcpy(vars, old, (size_t)nv * sizeof(double));
      orc_tc_reset_cfl(tc);
      r.accepted = 0;
    } else {
      rc = orc_source(c, vars, bathy, tc->dt);
      if (rc) break;
      rc = orc_fill_ghosts(c, vars);
      if (rc) break;
      orc_tc_update_time(tc, 1);
      r.accepted = 1;
      if (tc->status == 0 &&
          tc->current_snapshot_index <= tc->number_of_snapshots &&
          tc->time >= tc->snapshot_intervals[tc->current_snapshot_index - 1]) {
        r.snapshot = tc->current_snapshot_index;
        tc->current_snapshot_index += 1;
      }
      orc_tc_reset_cfl(tc);
      ++accepted;
    }
    r.time = tc->time;
    r.dt_next = tc->dt;
    if (rec && attempts < rec_cap) rec[attempts] = r;
    ++attempts;
    if (tc->status) break;
  }
  work_free(&w);
  free(old);
  return rc ? rc : attempts;
}

void orc_totals(const orc_cfg* c, const double* vars, double out[3]) {
  frame_t f = frame_of(c);
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  for (int j = 1; j <= c->ny; ++j)
    for (int i = 1; i <= c->nx; ++i) {
      s0 += Q(vars, &f, 0, i, j);
      s1 += Q(vars, &f, 1, i, j);
      s2 += Q(vars, &f, 2, i, j);
    }
  out[0] = s0;
  out[1] = s1;
  out[2] = s2;
}

/* ------------------------------------------------------------ test hooks */
int orc_cell_phys(const orc_cfg* c, int ixy, const double q[3], double* s, double f[3]) {
  return (ixy == 1) ? phys_x(c, q, s, f) : phys_y(c, q, s, f);
}

int orc_hll(const orc_cfg* c, int ixy, const double qL[3], const double qR[3], double out[8]) {
  double lam[2], wav[6];
  work_t w;
  w.lambda = lam;
  w.waves = wav;
  int rc = hll(c, ixy, qL, qR, &w, 0);
  out[0] = lam[0];
  out[1] = lam[1];
  for (int k = 0; k < 6; ++k) out[2 + k] = wav[k];
  return rc;
}

double orc_phi(int code, double r) { return phi_of(code, r); }
