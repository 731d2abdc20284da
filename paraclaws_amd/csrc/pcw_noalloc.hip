// pcw_noalloc.hip -- pcw_step_noalloc: one attempt of one domain on frames
// the caller owns, with no device allocation, for callers that cannot
// allocate (serving loops, harnesses that forbid hipMalloc).  It waits on
// its stream to return the sweeps' speed maxima, so it is not for code under
// stream capture.
//
// The step kernels of pcw_step need, besides the state, a descriptor, the
// attempt's controller record, the completion accumulators (FP32) and the
// tile list.  Here they live in `scratch`, a file-scope device array of
// kScratchBytes per GPU, held until the process ends.  Each call rewrites
// all of it from the host's settings with one set-up kernel (no memset, no
// copy from the host), then launches the same kernels as pcw_step.  Calls
// that overlap on one GPU race on that GPU's scratch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <mutex>

#include "../../include/pcw.h"
#include "pcw_engine.hpp"
#include "pcw_kernels.hpp"

namespace pcw {
namespace {

// scratch layout (bytes): descriptor, controller record, kSlots completion
// accumulators, tile count and first list position, tile list
constexpr size_t kScratchBytes = size_t(2) << 20;
constexpr size_t kOffInst = 0, kOffState = 512, kOffSlots = 1024, kOffLive = 6144, kOffTiles = 8192;
constexpr long long kMaxTiles = (long long)((kScratchBytes - kOffTiles) / sizeof(int4));
static_assert(sizeof(KInst) <= kOffState - kOffInst, "descriptor does not fit");
static_assert(sizeof(InstState) <= kOffSlots - kOffState, "controller record does not fit");
static_assert(sizeof(TileSlot) * kSlots <= kOffLive - kOffSlots, "accumulators do not fit");

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// scratch's device address on each GPU (each holds its own copy); the array's
// name in host code is only a stand-in and is never passed to a kernel
constexpr int kMaxDevices = 64;
unsigned char* g_scratch[kMaxDevices] = {};
std::mutex g_scratch_mu;

const char* scratch_of(int dev, unsigned char** out) {
  if (dev < 0 || dev >= kMaxDevices) return "device index out of range";
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return hipGetErrorString(e);
    g_scratch[dev] = static_cast<unsigned char*>(p);
  }
  *out = g_scratch[dev];
  return nullptr;
}

// Writes the descriptor, a fresh host-controlled controller record with dt,
// zeroed accumulators and the tile list: strips of kStripOut columns (the
// last shifted to end at column nx) times nb row bands of equal heights,
// ordered by band then strip.
__global__ void k_noalloc_setup(unsigned char* scr, KInst I, double dt, int nsx, int nb) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t == 0) {
    *reinterpret_cast<KInst*>(scr + kOffInst) = I;
    InstState S{};
    S.status = PCW_STATUS_RUNNING;
    S.last_ok = 1;
    S.tc.dt = dt;
    *reinterpret_cast<InstState*>(scr + kOffState) = S;
    reinterpret_cast<int*>(scr + kOffLive)[0] = I.ntiles;  // tiles of the launch
    reinterpret_cast<int*>(scr + kOffLive)[1] = 0;         // from list position 0
  }
  constexpr int kSlotWords = (int)(sizeof(TileSlot) * kSlots / sizeof(unsigned));
  if (t < kSlotWords) reinterpret_cast<unsigned*>(scr + kOffSlots)[t] = 0u;
  if (t < I.ntiles) {
    const int s = t % nsx, b = t / nsx;
    int x0 = 1 + kStripOut * s;
    if (x0 + kStripOut - 1 > I.nx) x0 = max(1, I.nx - kStripOut + 1);
    const int y0 = 1 + (int)((long long)I.ny * b / nb);
    const int y1 = (int)((long long)I.ny * (b + 1) / nb);
    reinterpret_cast<int4*>(scr + kOffTiles)[t] = make_int4(0, x0, y0, y1);
  }
}

int fail(int code, const char* msg) {
  set_loose_error(msg);
  return code;
}

}  // namespace
}  // namespace pcw

using namespace pcw;

extern "C" int pcw_step_noalloc(const pcw_params* p, const pcw_domain* d, const void* state_in,
                                void* state_out, double dt, double* smax, void* stream) {
  char msg[200];
  if (!p || !d || !state_in || !state_out || !smax) return fail(PCW_ERR_ARG, "null argument");
  if (p->precision != PCW_FP32 && p->precision != PCW_FP64)
    return fail(PCW_ERR_ARG, "precision must be PCW_FP64 or PCW_FP32");
  if (p->engine != PCW_ENGINE_FUSED) return fail(PCW_ERR_ARG, "pcw_step_noalloc runs the fused engine");
  if (p->higher_order && p->use_limiter && (p->limiter < 1 || p->limiter > 4))
    return fail(PCW_ERR_ARG, "invalid limiter choice => ABORTING");
  if (d->nx < 2 || d->ny < 2) return fail(PCW_ERR_ARG, "the domain needs nx >= 2 and ny >= 2");
  if (d->nx > PCW_MAX_SIDE || d->ny > PCW_MAX_SIDE)
    return fail(PCW_ERR_ARG, "the domain needs nx, ny <= PCW_MAX_SIDE (262,136)");
  for (int s = 0; s < 4; ++s)
    if (d->bc[s] < PCW_BC_ZERO_GRADIENT || d->bc[s] > PCW_BC_REFLECTIVE)
      return fail(PCW_ERR_BC, d->bc[s] == 1 ? "User-defined boundary condition not implemented yet!"
                                             : "boundary condition not implemented yet!");
  if (d->source != PCW_SRC_NONE && d->source != PCW_SRC_FRICTION)
    return fail(PCW_ERR_ARG, "pcw_step_noalloc takes no bathymetry (source NONE or FRICTION)");
  const int esize = p->precision == PCW_FP32 ? 4 : 8;
  const uintptr_t in = reinterpret_cast<uintptr_t>(state_in), out = reinterpret_cast<uintptr_t>(state_out);
  if (in % 256 || out % 256) return fail(PCW_ERR_ARG, "state frames must be 256-byte aligned");

  KInst I;
  describe_instance(*p, *d, I);
  const long long frame_bytes = 3 * I.plane * esize;
  const long long diff = (long long)(out - in);  // two's complement: out below in is negative
  if (diff > -frame_bytes && diff < frame_bytes) return fail(PCW_ERR_ARG, "state_in and state_out overlap");
  I.off[0] = 0;
  I.off[1] = diff / esize;
  // bands of at least two rows (the FP32 step finishes the eta ghost rows from
  // the two rows next to the boundary within one tile)
  const int ty = std::max(kMinBandRows,
                          tile_rows_for(I, p->tile_rows > 1 ? p->tile_rows : (p->precision == PCW_FP32 ? 64 : 128)));
  const int nsx = (d->nx + kStripOut - 1) / kStripOut;
  const int nb = (d->ny + ty - 1) / ty;
  const long long ntiles = (long long)nsx * nb;
  if (ntiles > kMaxTiles) {
    std::snprintf(msg, sizeof msg, "%lld tiles exceed the scratch's %lld: raise tile_rows", ntiles,
                  kMaxTiles);
    return fail(PCW_ERR_ARG, msg);
  }
  I.ntiles = (int)ntiles;
  I.tile0 = 0;

  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return fail(PCW_ERR_HIP, hipGetErrorString(e));
  unsigned char* scr = nullptr;
  if (const char* why = scratch_of(dev, &scr)) {
    std::snprintf(msg, sizeof msg, "hipGetSymbolAddress(scratch): %s", why);
    return fail(PCW_ERR_HIP, msg);
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int setup_threads = (int)std::max<long long>(ntiles, (long long)(sizeof(TileSlot) * kSlots / 4));
  k_noalloc_setup<<<(setup_threads + 255) / 256, 256, 0, s>>>(scr, I, dt, nsx, nb);
  e = hipGetLastError();
  if (e != hipSuccess) return fail(PCW_ERR_HIP, hipGetErrorString(e));

  KInst* dinst = reinterpret_cast<KInst*>(scr + kOffInst);
  InstState* dst = reinterpret_cast<InstState*>(scr + kOffState);
  const int4* dtiles = reinterpret_cast<const int4*>(scr + kOffTiles);
  TileSlot* dslots = reinterpret_cast<TileSlot*>(scr + kOffSlots);
  const int* dlive = reinterpret_cast<const int*>(scr + kOffLive);
  const int wpb = (p->waves_per_block >= 1 && p->waves_per_block <= 4) ? p->waves_per_block : 2;
  const int blocks = (int)((ntiles + wpb - 1) / wpb + 7) / 8 * 8;
  const int lim = (p->higher_order && p->use_limiter) ? p->limiter : 0;
  const bool wetdry = p->dry_tol > 0.0;
  if (p->precision == PCW_FP32) {
    FusedArgs<float> a{static_cast<float*>(const_cast<void*>(state_in)), nullptr, nullptr, dinst, dst,
                       dtiles, dlive, (int)ntiles, p->higher_order, wetdry, 0, p->grav, p->eps,
                       p->dry_tol, dslots};
    e = launch_fused_f32(lim, false, wetdry ? 2 : (d->source != PCW_SRC_NONE ? 1 : 0), a, blocks,
                         64 * wpb, s);
  } else {
    FusedArgs<double> a{static_cast<double*>(const_cast<void*>(state_in)), nullptr, nullptr, dinst,
                        dst, dtiles, dlive, (int)ntiles, p->higher_order, wetdry, 0, p->grav, p->eps,
                        p->dry_tol, nullptr};
    e = launch_fused_f64(lim, metric_mode(I), a, blocks, 64 * wpb, s);
    if (e == hipSuccess)  // ghost cells of the new state (the pending buffer)
      e = launch_ghosts<double>(a.pool, dinst, dst, 1, d->nx + d->ny, kGhostsAfterFused, s);
  }
  if (e != hipSuccess) return fail(PCW_ERR_HIP, hipGetErrorString(e));

  InstState S;
  e = hipMemcpyAsync(&S, dst, sizeof S, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return fail(PCW_ERR_HIP, hipGetErrorString(e));
  smax[0] = S.last_smax_x;
  smax[1] = S.last_smax_y;
  if (S.last_bad & 1)
    return fail(PCW_ERR_STATE, "NaN in the Jacobian or complex eigenvalues in the xi sweep");
  if (S.last_bad & 2)
    return fail(PCW_ERR_STATE_ETA,
                "NaN in the Jacobian or complex eigenvalues in the eta sweep (fatal unless the xi "
                "sweep alone rejects the step)");
  return PCW_OK;
}
