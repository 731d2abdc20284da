// pcw_kernels.hpp -- launch interfaces between the C ABI and the kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "pcw_engine.hpp"

namespace pcw {

// work-array planes of the staged engine (each one frame of doubles)
enum {
  W_LAM = 0,    // lambda(2) = s2 (lambda(1) = -s2)
  W_W1 = 1,     // 3 planes: wave 1
  W_W2 = 4,     // 3 planes: wave 2
  W_AM = 7,     // 3 planes: A-dQ
  W_AP = 10,    // 3 planes: A+dQ
  W_T1 = 13,    // 3 planes: limited wave 1
  W_T2 = 16,    // 3 planes: limited wave 2
  W_F = 19,     // 3 planes: correction flux
  W_COUNT = 22
};

struct StagedArgs {
  const KInst* inst;   // this instance's descriptor (device)
  InstState* st;       // this instance's state (device)
  double* q;           // current buffer, h plane frame origin
  double* w;           // work planes
  const double* m;     // metric planes of a mapped grid, or nullptr (identity map)
  double g, eps, dry;
  int limiter;         // 0 = unlimited, 1..4
  int pad;
};

template <typename T>
struct FusedArgs {
  T* pool;
  const T* bpool;
  const T* mpool;      // metric planes (instances with moff >= 0)
  const KInst* inst;
  InstState* st;
  const int4* tiles;   // (instance, x0, y0, y1): output columns x0..x0+59, rows y0..y1
  const int* live;     // live[0]: tiles this launch takes from the list, live[1]: the list
                       // position of the first (device memory: the host may change both
                       // between launches of a captured graph; a batch may run as two
                       // launches over the two halves of one list)
  int ntiles;          // tiles the launch's grid was sized for
  int ho;              // add_higher_order_flux_correction
  int wetdry;          // dry_tol > 0
  int pad;
  double g, eps, dry;
  TileSlot* slots;     // FP32: kSlots completion accumulators per instance
};

// phys: 0 instances without source terms, 1 with source terms and without the
// wet/dry cut-off, 2 with the wet/dry cut-off (every instance then)
hipError_t launch_fused_f32(int lim, bool met, int phys, const FusedArgs<float>& a, int blocks,
                            int threads, hipStream_t s);
// resident blocks per CU of the FP32 fused kernel variant (0 if unknown)
int fused_f32_blocks_per_cu(int lim, bool ho, bool met, int phys, int threads);
// Mesh metrics in an FP64 fused launch: identity-map instances with the
// reference's rounded metrics computed in the kernel, instances with metric
// arrays, and identity-map instances whose rounded metrics are all exactly 1
// (multiplying and dividing by them changes no bit, so the kernel skips them).
// Each launch runs the instances of one kind.
constexpr int kMetComputed = 0, kMetArrays = 1, kMetUnit = 2;
__host__ __device__ inline int metric_mode(const KInst& I) {
  return I.moff >= 0 ? kMetArrays : (I.unit_met ? kMetUnit : kMetComputed);
}
hipError_t launch_fused_f64(int lim, int met_mode, const FusedArgs<double>& a, int blocks,
                            int threads, hipStream_t s);
hipError_t launch_staged_sweep(const StagedArgs& a, const KInst& hI, int dir, int ho,
                               hipStream_t s);
// buf_sel of launch_ghosts: the buffer whose ghost cells are filled
constexpr int kGhostsCur = -1;         // the current buffer of every instance
constexpr int kGhostsAfterFused = 2;   // after an FP64 fused attempt: under host control the
                                       // pending buffer of a running instance, under device
                                       // control the current buffer if the attempt was accepted
template <typename T>
hipError_t launch_ghosts(T* pool, const KInst* dinst, InstState* dst, int n, int max_side_cells,
                         int buf_sel, hipStream_t s);
template <typename T>
hipError_t launch_source(T* pool, const T* bpool, const KInst* dinst, InstState* dst, int n,
                         int max_nx, int max_ny, int wetdry, double g, double dry, hipStream_t s);
template <typename T>
hipError_t launch_copy(T* pool, const KInst* dinst, InstState* dst, int n, int to_backup,
                       int only_if_rejected, hipStream_t s);
hipError_t launch_st_finish(const KInst* dinst, InstState* dst, int n, hipStream_t s);
template <typename T>
hipError_t launch_totals(const T* pool, const KInst* dinst, const InstState* dst, double* part,
                         int nparts, double* out, hipStream_t s);
hipError_t launch_set_dt(InstState* dst, const double* ddt, int n, int controlled, hipStream_t s);
hipError_t launch_flip(InstState* dst, int n, hipStream_t s);

// Instance descriptors on the host (pcw_api.hip); neither allocates.
bool metrics_are_unit(const KInst& I);
void describe_instance(const pcw_params& p, const pcw_domain& d, KInst& I);
// error message of the last failed call without a handle, on this thread
void set_loose_error(const char* msg);

}  // namespace pcw
