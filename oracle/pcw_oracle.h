/*
 * pcw_oracle.h -- CPU restatement (FP64) of paraCLAWs' 2D wave-propagation
 * finite-volume update.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only the test
 * suite, the smoke check and the benchmark's CPU-baseline leg may load it,
 * and never as the product path.
 *
 * Layout: every field is a Fortran-style frame vars(3, -1:nx+2, -1:ny+2)
 * (equation index fastest, then i, then j; column-major), exactly the
 * reference's solution%vars with ilow=jlow=1 and num_ghost=2
 * (src/domain/SMOD_domain_allocation.f90:84-100).  Work arrays follow
 * src/solver/MOD_work_arrays.f90:14-70.
 */
#ifndef PCW_ORACLE_H
#define PCW_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Boundary codes (src/user_subroutines/SMOD_boundary_conditions.f90:13).
 * 4 and 5 are additions of this build (the reference has no wall BC). */
#define ORC_BC_USER 1
#define ORC_BC_ZERO_GRADIENT 2
#define ORC_BC_CONSTANT_GRADIENT 3
#define ORC_BC_REFLECTIVE 4

/* Source kinds.  0 is the reference's phi = 0 (SMOD_source_function.f90:15-28);
 * the others are physics of this build with no reference counterpart. */
#define ORC_SRC_NONE 0
#define ORC_SRC_BATHY 1      /* -g h grad(b) with wet/dry cut-off            */
#define ORC_SRC_FRICTION 2   /* Manning friction                              */
#define ORC_SRC_BATHY_FRICTION 3

typedef struct orc_cfg {
  int32_t nx, ny;                 /* m_xi, m_eta                              */
  int32_t higher_order;           /* add_higher_order_flux_correction         */
  int32_t use_limiter;            /* use_limiters                             */
  int32_t limiter;                /* 1 minmod, 2 superbee, 3 van Leer, 4 MC    */
  int32_t bc[4];                  /* xi_low, xi_high, eta_low, eta_high        */
  int32_t source;                 /* ORC_SRC_*                                 */
  int32_t nthreads;               /* OpenMP threads (0 = runtime default)      */
  double xi_min, xi_max, eta_min, eta_max;
  double grav;                    /* MOD_solver_parameters grav               */
  double eps;                     /* MOD_constants ud_precision (1e-10)       */
  double manning_n;               /* friction coefficient (source 2/3)        */
  double dry_tol;                 /* wet/dry threshold (source 1/3)           */
  /* Optional mesh metrics of a mapped grid, frame arrays (nx+4)*(ny+4):
   * capacity, length_ratio_xi(1,:,:), length_ratio_eta(1,:,:)
   * (src/domain/SMOD_mesh_manipulation.f90:153,162,173).  NULL: the
   * identity map's metrics are computed in place. */
  const double* kappa;
  const double* gamma_xi;
  const double* gamma_eta;
} orc_cfg;

/* time_control (src/time_control/MOD_time_control.f90:16-66). */
typedef struct orc_tc {
  double time, time_initial, time_final;
  double dt, dt_previous, dt_max, dt_min;
  double time_precision, CFL, CFL_target, CFL_previous_timestep;
  double shrink_limit, growth_limit, emergency_shrink_ratio;
  int32_t adaptive, snapshot_exception;
  int32_t current_snapshot_index; /* 1-based as in the reference */
  int32_t number_of_snapshots;
  int32_t status;                 /* 0 ok; 1 end_program was reached          */
  int32_t pad_;
  double snapshot_intervals[4096];
} orc_tc;

typedef struct orc_step_rec {
  int32_t accepted;               /* 1 accepted, 0 rejected                    */
  int32_t snapshot;               /* snapshot index written at this step, 0 if none */
  double dt_used;                 /* the dt the sweeps used                    */
  double cfl;                     /* t_ctrl%CFL at the end of the attempt      */
  double time;                    /* t_ctrl%time after updateTime              */
  double dt_next;                 /* t_ctrl%dt after updateTime                */
} orc_step_rec;

/* Mesh metrics of the identity-mapped Cartesian mesh, rounded as the
 * reference computes them; every output is a frame array (nx+4)*(ny+4). */
int orc_metrics(const orc_cfg* c, double* xi, double* eta, double* kappa,
                double* gamma_xi, double* gamma_eta);

/* Circular dam break: h = h_in inside the radius, h_out outside, over the
 * whole frame including ghosts; hu = hv = 0. */
int orc_init_dam_break(const orc_cfg* c, double* vars, double h_in,
                       double h_out, double radius);

/* update_rank_boundaries on a single rank: physical BCs only. */
int orc_fill_ghosts(const orc_cfg* c, double* vars);

/* One sweep (ixy = 1 xi, 2 eta), in place, with the reference's stage order.
 * cfl_local is max-accumulated (not reset).  Returns 0, or <0 on a NaN
 * Jacobian (the reference's "isnan in jakobi" error stop). */
int orc_sweep(const orc_cfg* c, double* vars, int ixy, double dt,
              double* cfl_local);

/* solve_dimensional_split: returns 1 accepted / -1 rejected, <-1 on error. */
int orc_solve_dimensional_split(const orc_cfg* c, orc_tc* tc, double* vars);

/* integrate_source_term with the built-in source kinds; bathy may be NULL
 * for ORC_SRC_NONE / ORC_SRC_FRICTION. */
int orc_source(const orc_cfg* c, double* vars, const double* bathy, double dt);

/* time_control procedures. */
int orc_tc_initialize(orc_tc* tc, double t0, double t1, double dt0,
                      int32_t n_snap, double cfl_target, double dt_max,
                      double dt_min, double growth, double shrink,
                      int32_t adaptive);
void orc_tc_update_time(orc_tc* tc, int32_t step_valid);
void orc_tc_update_max_cfl(orc_tc* tc, double cfl);
int32_t orc_tc_check_validity(const orc_tc* tc);
void orc_tc_reset_cfl(orc_tc* tc);

/* The main loop of src/main.f90:193-283 (without IO): runs until
 * time >= time_final, or max_accepted accepted steps, or max_attempts
 * attempts.  Writes one record per attempt (up to rec_cap). Returns the
 * number of attempts, or <0 on error. */
int orc_run(const orc_cfg* c, orc_tc* tc, double* vars, const double* bathy,
            int32_t max_accepted, int32_t max_attempts, orc_step_rec* rec,
            int32_t rec_cap, int32_t fill_first);

/* Test hooks: per-cell speed and flux of one sweep direction (returns the
 * reference's error condition), one HLL interface (out: lambda1, lambda2,
 * W1[3], W2[3]) and the limiter function phi(code, r). */
int orc_cell_phys(const orc_cfg* c, int ixy, const double q[3], double* s, double f[3]);
int orc_hll(const orc_cfg* c, int ixy, const double qL[3], const double qR[3], double out[8]);
double orc_phi(int code, double r);

/* FP64 totals of the interior: sum h, sum hu, sum hv (cell-area weighted
 * by kappa*dxi*deta). */
void orc_totals(const orc_cfg* c, const double* vars, double out[3]);

/* The reference's per-interface cost structure (LAPACK dgeev for the
 * eigenvalues, BLAS dgemv for J*q, per cell of every interface) instead of the
 * closed forms: the Fortran-convention dgeev and dgemv of a LAPACK/BLAS the
 * caller supplies; NULL switches back to the closed forms.  Process-wide. */
void orc_use_lapack(void* dgeev, void* dgemv);

#ifdef __cplusplus
}
#endif
#endif
