/*
 * pcw.h -- C ABI of the MI355X (gfx950) wave-propagation engine.
 *
 * One handle holds one or more independent 2D shallow-water domains
 * ("instances"), each a frame of nx*ny interior cells with two ghost layers,
 * resident in HBM as structure-of-arrays planes h, hu, hv.  All entry points
 * are plain C: pointers, sizes and status codes, no C++ or torch types.
 *
 * The entry points replace the reference's per-step path behind its Fortran
 * module interfaces (paths relative to the paraCLAWs repository):
 *   pcw_step + pcw_accept/pcw_reject
 *       -> solve_dimensional_split (src/solver/MOD_solve.f90:23-175) with the
 *          copy/rollback of src/main.f90:201,221; the fused step also applies
 *          integrate_source_term (MOD_solve.f90:295-336) and the physical BC
 *          fill of update_rank_boundaries
 *          (src/user_subroutines/SMOD_boundary_conditions.f90:17-213)
 *   pcw_sweep            -> one sweep of solve_dimensional_split (:71-105, :124-157)
 *   pcw_fill_ghosts      -> update_rank_boundaries (:17-56) on one rank
 *   pcw_source           -> integrate_source_term (MOD_solve.f90:295-336)
 *   pcw_set_state/get_state -> solution%vars(3,ilow-2:ihigh+2,jlow-2:jhigh+2)
 *                          (src/domain/SMOD_domain_allocation.f90:84-100)
 *   pcw_run              -> the time loop of src/main.f90:193-283 with the
 *                          controller of src/time_control/MOD_time_control.f90
 *                          replicated on the device
 *   pcw_stop_at_snapshots -> the loop's return to the host for snapshot
 *                          output (src/main.f90:256-268)
 *
 * Return codes: 0 ok, 1 step rejected (solve_dimensional_split's ierr = -1),
 * negative = fatal (the host maps it to abort_program,
 * src/utility/MOD_error_handling.f90:37-50); pcw_last_error() describes it.
 */
#ifndef PCW_H
#define PCW_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PCW_OK 0
#define PCW_REJECTED 1
#define PCW_ERR_ARG (-1)
#define PCW_ERR_HIP (-2)
#define PCW_ERR_STATE (-3)   /* NaN Jacobian / complex eigenvalues (reference error stop) */
#define PCW_ERR_BC (-4)      /* BC code 1 (user) or unknown: abort_program */
#define PCW_ERR_CONTROL (-5) /* end_program reached by the time controller */
#define PCW_ERR_STATE_ETA (-6) /* pcw_step/pcw_sweep: NaN/complex eigenvalues in the eta sweep
                                  only; fatal unless the xi sweep alone rejects the step (the
                                  reference then never runs the eta sweep) */

#define PCW_FP64 0           /* reference arithmetic order, bit-exact with the FP64 CPU restatement */
#define PCW_FP32 1           /* FP32 state and flux arithmetic, FP64 time/dt/totals */

#define PCW_ENGINE_FUSED 0   /* one LDS-free register-streaming kernel per step */
#define PCW_ENGINE_STAGED 1  /* FP64 only: one kernel per reference stage, HBM work arrays */

#define PCW_BC_ZERO_GRADIENT 2
#define PCW_BC_CONSTANT_GRADIENT 3
#define PCW_BC_REFLECTIVE 4  /* wall: mirror two interior cells, negate normal momentum (no reference counterpart) */

#define PCW_SRC_NONE 0
#define PCW_SRC_BATHY 1
#define PCW_SRC_FRICTION 2
#define PCW_SRC_BATHY_FRICTION 3

#define PCW_STATUS_RUNNING 0
#define PCW_STATUS_DONE 1
#define PCW_STATUS_ERROR 2

/* Size limits, checked by pcw_create and pcw_step_noalloc (PCW_ERR_ARG):
 * every kernel's grid then stays within 65,535 workgroups in y and z. */
#define PCW_MAX_SIDE 262136     /* nx, ny of one domain */
#define PCW_MAX_DOMAINS 65535   /* instances of one handle */

typedef struct pcw_handle pcw_handle;

/* Per-instance numerical settings; field meaning follows control.in
 * (src/utility/MOD_IO.f90:61-94) and MOD_solver_parameters. */
typedef struct pcw_domain {
  int32_t nx, ny;              /* m_xi, m_eta (interior cells)                 */
  int32_t bc[4];               /* bc_xi_lower, bc_xi_upper, bc_eta_lower, bc_eta_upper */
  int32_t source;              /* PCW_SRC_*                                     */
  int32_t pad_;
  double xi_min, xi_max, eta_min, eta_max;
  double manning_n;            /* per-instance Manning coefficient             */
} pcw_domain;

/* Settings shared by all instances of a handle. */
typedef struct pcw_params {
  int32_t precision;           /* PCW_FP64 / PCW_FP32                           */
  int32_t engine;              /* PCW_ENGINE_FUSED / PCW_ENGINE_STAGED          */
  int32_t higher_order;        /* add_higher_order_flux_correction              */
  int32_t use_limiter;         /* use_limiters                                  */
  int32_t limiter;             /* 1 minmod, 2 superbee, 3 van Leer, 4 MC        */
  int32_t tile_rows;           /* rows per tile (0 = autotuned/default)         */
  int32_t waves_per_block;     /* waves per workgroup (0 = default)             */
  int32_t pad_;
  double grav;                 /* g (the reference stores REAL*4 9.81 widened)  */
  double eps;                  /* ud_precision                                  */
  double dry_tol;              /* wet/dry threshold, 0 = off (reference)        */
} pcw_params;

/* time_control settings (MOD_time_control.f90:70-151). */
typedef struct pcw_time {
  double t0, t_final, dt0, cfl_target, dt_max, dt_min, growth, shrink;
  int32_t n_snapshots;
  int32_t adaptive;
} pcw_time;

/* One attempted step, as recorded by the device controller. */
typedef struct pcw_step_record {
  int32_t accepted;
  int32_t snapshot;            /* snapshot index reached at this step, 0 if none */
  double dt_used, cfl, time, dt_next;
} pcw_step_record;

/* Per-instance status after pcw_run. */
typedef struct pcw_status {
  int32_t status;              /* PCW_STATUS_*                                  */
  int32_t error;               /* PCW_ERR_* when status is ERROR                */
  int64_t accepted, rejected;
  double time, dt, cfl_last;
} pcw_status;

/* ---- lifetime ----------------------------------------------------------
 * pcw_create: PCW_ERR_ARG, with a message, for sizes beyond the limits
 * above, nx or ny below 2 and bad settings; a handle is returned then too,
 * for pcw_last_error, and is released with pcw_destroy.  FP32 tiles span at
 * most the rows whose three planes' 32-bit buffer offsets stay below 2^31
 * bytes per plane: a larger tile_rows is lowered to that for the instance.
 * Threads: a handle is used by one host thread at a time; different handles
 * may be driven from different threads at once.  Each handle has its own
 * non-blocking HIP stream, and its transfers are ordered on that stream. */
int pcw_create(const pcw_params* p, const pcw_domain* domains, int32_t n_domains,
               pcw_handle** out);
int pcw_destroy(pcw_handle* h);
const char* pcw_last_error(const pcw_handle* h);
int pcw_device_bytes(const pcw_handle* h, int64_t* bytes);

/* ---- state transfer (Fortran AoS frame <-> device SoA) ------------------
 * vars points at solution%vars(1,-1,-1): 3*(nx+4)*(ny+4) doubles, equation
 * index fastest, then i, then j.  FP32 instances round on upload. */
int pcw_set_state(pcw_handle* h, int32_t inst, const double* vars);
int pcw_get_state(pcw_handle* h, int32_t inst, double* vars);
/* bathymetry frame b(-1:nx+2, -1:ny+2), doubles */
int pcw_set_bathymetry(pcw_handle* h, int32_t inst, const double* b);
/* Mesh metrics of a mapped grid, frames (-1:nx+2, -1:ny+2) of doubles:
 * mesh%quadrilateral_mapping%capacity, length_ratio_xi(1,:,:) and
 * length_ratio_eta(1,:,:) (src/domain/SMOD_mesh_manipulation.f90:147-173).
 * Without this call an instance uses the identity map's metrics, computed in
 * the kernels with the reference's rounding. */
int pcw_set_metrics(pcw_handle* h, int32_t inst, const double* kappa, const double* gamma_xi,
                    const double* gamma_eta);

/* ---- host-controlled stepping (the Fortran host's path) -----------------
 * pcw_step advances every instance by one attempt with its own dt into the
 * pending buffer (fused: both sweeps + source + ghost fill; staged: the
 * reference stages in place after a backup copy).  smax[2*k], smax[2*k+1]
 * receive instance k's max wave speed of the xi and eta sweep; the CFL of a
 * sweep is (smax*dt)/delta, exactly max(lambda*dt/delta) of MOD_solve.f90:81.
 * The host then calls pcw_accept (keep) or pcw_reject (roll back). */
int pcw_step(pcw_handle* h, const double* dt, double* smax);
/* pcw_accept / pcw_reject apply to the instances pcw_step advanced: those
 * whose status is RUNNING (every instance after pcw_create; a device-
 * controlled run leaves finished ones DONE until pcw_time_init).  They only
 * enqueue work on the handle's stream; the entry points that read results
 * back (pcw_step, pcw_get_state, pcw_totals, ...) synchronise it. */
int pcw_accept(pcw_handle* h);
int pcw_reject(pcw_handle* h);
/* Single sweeps in place on the current state (staged engine semantics). */
int pcw_sweep(pcw_handle* h, int32_t ixy, const double* dt, double* smax);
int pcw_source(pcw_handle* h, const double* dt);
int pcw_fill_ghosts(pcw_handle* h);

/* One attempt of one domain on device frames the caller owns, allocating
 * nothing.  It has no counterpart in the reference: it is pcw_step for
 * callers that cannot allocate device memory (serving loops, harnesses that
 * forbid hipMalloc).  It waits on `stream` to return smax, so it cannot be
 * recorded under stream capture.
 *   state_in, state_out: device frames (256-byte aligned, not overlapping) of
 *     three planes h, hu, hv, each ny+4 rows of `pitch` elements of the
 *     handle's layout: FP64 rows hold column i (-1 .. nx+2) at element i+1,
 *     pitch = nx+4 rounded up to 64; FP32 rows are 64-element slots (slot 0
 *     holds columns -1, 0 in its last two elements, slot s+1 columns
 *     60s+1 .. 60s+60 then 4 unused elements, the last slot also nx+1, nx+2
 *     after column nx), pitch = 64 * (ceil(nx/60) + 1).  state_out receives
 *     the new interior and physical ghost cells (not the ghost corners).
 *   dt: the attempt's step; smax[0], smax[1]: the xi / eta sweep maxima, as
 *     pcw_step returns them.  stream: a hipStream_t, NULL for the default.
 *   Limits: the fused engine, source kinds NONE or FRICTION, and at most
 *     130,560 tiles (ceil(nx/60) * ceil(ny/tile_rows); tile_rows 0 means 64
 *     rows in FP32, 128 in FP64).
 * Descriptor, controller record and tile list live in a 2 MiB device array
 * per GPU that each call rewrites; calls that overlap on one GPU race.
 * Returns as pcw_step; pcw_last_error(NULL) describes a failure. */
int pcw_step_noalloc(const pcw_params* p, const pcw_domain* d, const void* state_in,
                     void* state_out, double dt, double* smax, void* stream);

/* ---- device-controlled run (batched / persistent path) ------------------
 * Initialises the device replica of time_control for every instance. */
int pcw_time_init(pcw_handle* h, const pcw_time* t);
/* Optional per-instance step log (capacity records per instance). */
int pcw_enable_log(pcw_handle* h, int32_t capacity);
int pcw_read_log(pcw_handle* h, int32_t inst, pcw_step_record* out, int32_t cap,
                 int32_t* n);
/* Runs graph-captured blocks of steps_per_graph attempts until every
 * instance is done or max_attempts attempts were made (0 = unbounded).
 * max_accepted > 0 also stops an instance after that many accepted steps;
 * a block is then shortened to the accepted steps the running instances
 * still need, so that no launch of it only finds finished instances.  An
 * instance that an earlier run stopped at its max_accepted continues when
 * the limit is raised (or 0).  A fused batch runs as two lanes, halves of
 * its instances whose step graphs run on two streams at once; the call
 * returns after both. */
int pcw_run(pcw_handle* h, int32_t steps_per_graph, int64_t max_attempts,
            int64_t max_accepted);
/* on = 1: device-controlled runs stop each instance after the accepted step
 * that reaches one of its snapshot times, so that the host can write the
 * snapshot there (the snapshot branch of the time loop, src/main.f90:256-268);
 * the next pcw_run continues the instance.  on = 0 (the default): runs go
 * on through snapshot times. */
int pcw_stop_at_snapshots(pcw_handle* h, int32_t on);
int pcw_get_status(pcw_handle* h, int32_t inst, pcw_status* out);

/* ---- measurement --------------------------------------------------------- */
/* FP64 sums over the interior of instance inst: h, hu, hv. */
int pcw_totals(pcw_handle* h, int32_t inst, double out[3]);
/* Device time (ms) of the last pcw_run / pcw_step, measured with HIP events
 * on the handle's stream. */
int pcw_last_elapsed_ms(pcw_handle* h, double* ms);
/* Average duration (ms) of one attempt over the last pcw_run: HIP events
 * bracket each graph launch on the handle's stream; the average is taken over
 * the graphs whose every launch made an attempt (else over all attempts).
 * launches: step launches of the run. */
int pcw_kernel_stats(pcw_handle* h, double* avg_step_ms, int64_t* launches);
/* Attainable bandwidth of the step's compulsory traffic on the current GPU:
 * a stream of three planes of plane_bytes in and three out (16-byte
 * non-temporal loads and stores, no stencil), timed with HIP events over
 * reps launches after two warm-up launches; GB/s of the fastest and the
 * median launch (bytes moved = 6 * plane_bytes).  Allocates and frees its
 * six planes. */
int pcw_stream_bandwidth(int64_t plane_bytes, int32_t reps, double* gbs_best, double* gbs_median);

/* The tile list pcw_create would step these instances with, for a given
 * band height (p->tile_rows; 0 means 64 rows in FP32, 128 in FP64, without
 * the shortening pcw_create applies where tiles would not fill the GPU) and
 * p->waves_per_block, computed on the host: it touches no device.  Tile t is
 * out[4t .. 4t+3] = (instance, x0, y0, y1): output columns x0 .. x0+59
 * (within 1 .. nx; the last strip is shifted to end at column nx), rows
 * y0 .. y1, each instance's tiles contiguous.  lanes != 0: the list of a
 * fused batch is split for pcw_run's two lanes, *split receiving the
 * list position where the second starts (0: one lane).  Writes at most cap
 * tiles (out may be NULL when cap is 0) and returns the number of tiles, or
 * PCW_ERR_ARG.  It has no counterpart in the reference: it exposes the
 * engine's scheduling to tests and tools. */
int64_t pcw_plan_tiles(const pcw_params* p, const pcw_domain* d, int32_t n, int32_t lanes,
                       int32_t* out, int64_t cap, int32_t* split);
/* Autotuner: times candidate (tile_rows, waves_per_block) on the handle's
 * current state (`steps` attempts each, a batch's two lanes at once) and
 * keeps the fastest; for a batch, the tallest height within 1 % of the
 * fastest, at that height's fastest wave count.  Writes the choice.  No
 * counterpart in the reference (the reference has no tiles). */
int pcw_autotune(pcw_handle* h, int32_t steps, int32_t* tile_rows,
                 int32_t* waves_per_block);

#ifdef __cplusplus
}
#endif
#endif
