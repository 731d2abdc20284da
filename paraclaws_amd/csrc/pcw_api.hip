// pcw_api.hip -- the extern "C" entry points declared in include/pcw.h.
//
// Memory: every instance owns two frames (ping-pong) of three SoA planes
// h, hu, hv of (pitch x (ny+4)) elements, all in one HBM pool (FP64: plain
// rows, pitch = nx+4 rounded up to 64 elements; FP32: rows of 64-element
// slots, col_pos in pcw_engine.hpp); descriptors (KInst) and the per-instance
// controller state (InstState) are device resident.  One HIP stream per
// handle, default priority.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/pcw.h"
#include "pcw_device.hpp"
#include "pcw_engine.hpp"
#include "pcw_kernels.hpp"

using namespace pcw;

struct pcw_handle {
  pcw_params p{};
  int n = 0;
  int esize = 8;
  std::vector<pcw_domain> dom;
  std::vector<KInst> hinst;
  std::vector<InstState> hst;
  std::vector<int4> htiles;
  KInst* dinst = nullptr;
  InstState* dst = nullptr;
  int4* dtiles = nullptr;
  int dtiles_cap = 0;             // tiles dtiles holds room for
  int* dlive = nullptr;           // {tiles, first list position} of a launch, read by the
                                  // step kernels: the whole list, lane 1, lane 2
  TileSlot* dslots = nullptr;     // FP32 completion accumulators, kSlots per instance
  int ntiles = 0;                 // tiles in the list now
  int tile_cap = 0;               // tiles the step launches' grids are sized for
  bool tiles_subset = false;      // the list holds the running instances only
  int tiled_running = 0;          // running instances when the list was built
  void* pool = nullptr;
  long long pool_elems = 0;
  void* bpool = nullptr;
  long long bpool_elems = 0;
  void* mpool = nullptr;          // metric planes of mapped-grid instances
  long long mpool_elems = 0;
  bool any_met = false, any_plain = true;
  double* dsnaps = nullptr;
  pcw_step_record* dlog = nullptr;
  int log_cap = 0;
  double* work = nullptr;
  double* dscratch = nullptr;     // 4096 doubles: dt upload, totals partials
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  int tile_rows = 64, wpb = 2;
  // Lanes.  A device-controlled run of a fused batch splits the tile list at
  // an instance boundary into [0, split) and [split, ntiles) and launches each
  // half's steps on a stream of its own, so that one lane's end of launch (its
  // last tiles finishing while others wait) overlaps the other lane's step.
  int split = 0;                  // lane boundary in the list (0: one lane)
  bool two_lanes = false;         // a fused handle of two or more instances
  int lane_k = 0;                 // first instance of lane 2 in the captured lane graphs
  int lane_cap[2] = {0, 0};       // tiles the lanes' captured grids are sized for
  hipStream_t stream2 = nullptr;  // lane 2 (lane 1 and the whole list: stream)
  hipEvent_t evf = nullptr, evj = nullptr;
  // captured step graphs by length (attempts per launch) and lane (0: the
  // whole list); a run uses its full length and, for the attempts left
  // before max_accepted, a shorter one
  struct Captured {
    int len, lane;
    hipGraphExec_t exec;
  };
  std::vector<Captured> graphs;
  std::string err;
  double last_ms = 0.0, avg_step_ms = 0.0;
  long long launches = 0;
  int max_nx = 0, max_ny = 0, max_side = 0;
};

// message of the last failed entry point that has no handle (pcw_step_noalloc)
thread_local char g_loose_error[256] = {0};

namespace pcw {

void set_loose_error(const char* msg) {
  std::snprintf(g_loose_error, sizeof g_loose_error, "%s", msg);
}

// True when the identity map's metrics, rounded as the FP64 kernel computes
// them (kappa_exact, gamma_xi_exact, gamma_eta_exact), are exactly 1 on every
// frame cell: the FP64 kernel then skips them (kMetUnit), with the same bits.
// kappa depends on the column only through the rounded cell width x2 - x1
// (its other two x-terms are x2 - x2 and x1 - x1, i.e. zero), so it is
// evaluated on one column per distinct width.  Allocates nothing.
bool metrics_are_unit(const KInst& I) {
  const Mesh M{I.xi_min, I.eta_min, I.dxi, I.deta};
  constexpr int kMaxWidths = 16;
  double widths[kMaxWidths];
  int cols[kMaxWidths];
  int nw = 0;
  for (int i = -1; i <= I.nx + 2; ++i) {
    if (gamma_eta_exact(M, i) != 1.0) return false;
    const double x1 = corner_x1(M, i), w = (x1 + M.dxi) - x1;
    bool seen = false;
    for (int k = 0; k < nw; ++k) seen = seen || widths[k] == w;
    if (!seen) {
      // a uniform grid rounds to one or a few widths; many widths mean metrics
      // that are not all 1 anyway (gamma_eta = width / dxi), so stop early
      if (nw == kMaxWidths) return false;
      widths[nw] = w;
      cols[nw++] = i;
    }
  }
  for (int j = -1; j <= I.ny + 2; ++j) {
    if (gamma_xi_exact(M, j) != 1.0) return false;
    for (int k = 0; k < nw; ++k)
      if (kappa_exact(M, cols[k], j) != 1.0) return false;
  }
  return true;
}

// Descriptor of one instance from its settings: frame layout (FP32 slot rows,
// FP64 plain rows), spacings, BCs, extrapolation ratios and metric kind.
// Offsets of the two buffers, bathymetry, metrics and tiles are the caller's.
// Allocates nothing.
void describe_instance(const pcw_params& p, const pcw_domain& d, KInst& I) {
  std::memset(&I, 0, sizeof I);
  I.nx = d.nx;
  I.ny = d.ny;
  if (p.precision == PCW_FP32) {
    I.slot_rows = 1;
    I.nslot = (d.nx + kStripOut - 1) / kStripOut;
    I.pitch = 64 * (I.nslot + 1);
  } else {
    I.pitch = (int)(((long long)d.nx + 4 + 63) / 64 * 64);
  }
  I.plane = (long long)I.pitch * (d.ny + 4);
  for (int s = 0; s < 4; ++s) I.bc[s] = d.bc[s];
  I.source = d.source;
  I.dxi = (d.xi_max - d.xi_min) / d.nx;   // SMOD_mesh_manipulation.f90:28-29
  I.deta = (d.eta_max - d.eta_min) / d.ny;
  I.xi_min = d.xi_min;
  I.eta_min = d.eta_min;
  I.manning = d.manning_n;
  I.unit_met = p.precision == PCW_FP64 && metrics_are_unit(I);
  auto c = [](double lo, double dd, int i) { return lo + (double)(i - 1) * dd + dd / 2; };
  for (int kk = 1; kk <= 2; ++kk) {
    const double xl = c(I.xi_min, I.dxi, 1), xl2 = c(I.xi_min, I.dxi, 2);
    const double xr = c(I.xi_min, I.dxi, d.nx), xr2 = c(I.xi_min, I.dxi, d.nx - 1);
    const double yb = c(I.eta_min, I.deta, 1), yb2 = c(I.eta_min, I.deta, 2);
    const double yt = c(I.eta_min, I.deta, d.ny), yt2 = c(I.eta_min, I.deta, d.ny - 1);
    I.gr[kk - 1] = (float)((c(I.xi_min, I.dxi, 1 - kk) - xl) / (xl2 - xl));
    I.gr[1 + kk] = (float)((c(I.xi_min, I.dxi, d.nx + kk) - xr) / (xr2 - xr));
    I.gr[3 + kk] = (float)((c(I.eta_min, I.deta, 1 - kk) - yb) / (yb2 - yb));
    I.gr[5 + kk] = (float)((c(I.eta_min, I.deta, d.ny + kk) - yt) / (yt2 - yt));
  }
  I.boff = -1;
  I.moff = -1;
}

}  // namespace pcw

namespace {

constexpr int kScratch = 4096;

int fail(pcw_handle* h, int code, const std::string& msg) {
  if (h) h->err = msg;
  return code;
}

#define HIPCHK(x)                                                                        \
  do {                                                                                   \
    hipError_t e_ = (x);                                                                 \
    if (e_ != hipSuccess) return fail(h, PCW_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

long long round_up(long long v, long long m) { return (v + m - 1) / m * m; }

// Transfers and fills of a handle's memory go through the handle's own
// stream, and the call waits for it.  That stream is created non-blocking, so
// nothing here is ordered with the legacy (null) stream: handles used from
// several host threads at once, one thread capturing a step graph while
// another uploads a state, do not wait on or break each other.
hipError_t copy_wait(pcw_handle* h, void* dst, const void* src, size_t n, hipMemcpyKind kind) {
  const hipError_t e = hipMemcpyAsync(dst, src, n, kind, h->stream);
  return e == hipSuccess ? hipStreamSynchronize(h->stream) : e;
}
hipError_t zero_wait(pcw_handle* h, void* p, size_t n) {
  const hipError_t e = hipMemsetAsync(p, 0, n, h->stream);
  return e == hipSuccess ? hipStreamSynchronize(h->stream) : e;
}

// A tile that touches a domain boundary (the first strip, the strip ending at
// column nx, the first two or last two rows) runs the kernels' general body
// and takes longer than an interior tile.
bool boundary_tile(const KInst& I, const int4& t) {
  return t.y == 1 || t.y + kStripOut - 1 >= I.nx || t.z < 3 || t.w + 2 > I.ny;
}

// The fused kernels give XCD x the blocks x*nb/8 .. (x+1)*nb/8-1 of the grid
// (nb blocks of wpb tiles), in dispatch order.  An instance's boundary tiles
// are spread over the XCD ranges its tiles cover, in proportion to each
// range's share of them, and each range starts with its share: boundary tiles
// then start in the first round of resident waves on every XCD instead of
// finishing last on some.  Interior tiles keep their order, and every
// instance keeps its contiguous range [tile0, tile0+ntiles).
// lo, end: the part of the list one launch takes (a lane, or all of it).
void order_tiles(pcw_handle* h, long long lo, long long end) {
  const long long T = end - lo;
  if (T <= 0) return;
  const long long nb = round_up((T + h->wpb - 1) / h->wpb, 8);
  const long long span = nb / 8 * h->wpb;
  long long i0 = lo;
  while (i0 < end) {
    const int k = h->htiles[i0].x;
    long long i1 = i0;
    while (i1 < end && h->htiles[i1].x == k) ++i1;
    const KInst& I = h->hinst[k];
    std::vector<int4> bnd, inner;
    for (long long i = i0; i < i1; ++i)
      (boundary_tile(I, h->htiles[i]) ? bnd : inner).push_back(h->htiles[i]);
    const long long n = i1 - i0, nbnd = (long long)bnd.size();
    long long out = i0, bi = 0, ii = 0;
    for (long long a = lo + (i0 - lo) / span * span; a < i1; a += span) {
      const long long hi = std::min(a + span, i1);
      // boundary tiles due up to the end of this range, in proportion
      const long long due = (hi == i1) ? nbnd : nbnd * (hi - i0) / n;
      for (; bi < due && out < hi; ++bi) h->htiles[out++] = bnd[bi];
      while (out < hi) {
        if (ii < (long long)inner.size()) h->htiles[out++] = inner[ii++];
        else h->htiles[out++] = bnd[bi++];
      }
    }
    i0 = i1;
  }
}

// Row bands [y0, y1] covering 1..ny: bands of ty rows, or, when short_ends,
// first and last bands of about four fifths of ty (short_all: every band).
// Heights below 3 are raised to 3, so that every band, and in particular the
// first and the last, holds at least two rows (ny >= 2): the FP32 step
// writes the eta ghost rows from the two rows next to the boundary, which
// must be finished by the same tile.
void row_bands(int ny, int ty, bool short_ends, bool short_all, std::vector<int2>& out) {
  out.clear();
  ty = std::max(ty, kMinBandRows);
  const int te = std::max(kMinBandRows, (4 * ty + 2) / 5);
  if (short_all) ty = te;
  if (!short_ends || short_all || ny <= 2 * te + ty) {
    const int nb = (ny + ty - 1) / ty;
    for (int b = 0; b < nb; ++b) {  // equal heights within one row
      const int y0 = 1 + (int)((long long)ny * b / nb), y1 = (int)((long long)ny * (b + 1) / nb);
      out.push_back(make_int2(y0, y1));
    }
    return;
  }
  out.push_back(make_int2(1, te));
  const int mid = ny - 2 * te, nb = (mid + ty - 1) / ty;
  for (int b = 0; b < nb; ++b)
    out.push_back(make_int2(te + 1 + (int)((long long)mid * b / nb), te + (int)((long long)mid * (b + 1) / nb)));
  out.push_back(make_int2(ny - te + 1, ny));
}

// Lane boundary of a fused handle's list: the instance boundary nearest
// its middle (0, one lane, when the list holds one instance).  Called while
// every instance's tiles are contiguous and in instance order.
void choose_split(pcw_handle* h) {
  h->split = 0;
  if (!h->two_lanes) return;
  long long best = -1;
  for (int i = 1; i < h->ntiles; ++i)
    if (h->htiles[i].x != h->htiles[i - 1].x) {
      const long long d = std::llabs(2LL * i - h->ntiles);
      if (best < 0 || d < best) {
        best = d;
        h->split = i;
      }
    }
}

// Tiles of every instance: strips of kStripOut columns (the last one shifted
// to end at column nx) times row bands, ordered by band start then column.
// FP32: tiles on a domain boundary run the kernel's general body, about 1.2x
// an interior tile's time per row, so they get about four fifths of the rows
// (every band of the two edge strips, the first and last band of the others):
// a launch then does not wait on its boundary tiles.
// number of tiles build_tiles makes for tile height ty
long long count_tiles(const pcw_handle* h, int ty, const std::vector<char>* only = nullptr) {
  const bool shorten = h->p.precision == PCW_FP32;
  std::vector<int2> inner, edge;
  long long t = 0;
  for (int k = 0; k < h->n; ++k) {
    if (only && !(*only)[k]) continue;
    const KInst& I = h->hinst[k];
    const int nsx = (I.nx + kStripOut - 1) / kStripOut;
    row_bands(I.ny, tile_rows_for(I, ty), shorten, false, inner);
    row_bands(I.ny, tile_rows_for(I, ty), shorten, shorten, edge);
    const int nedge = std::min(nsx, 2);
    t += (long long)(nsx - nedge) * (long long)inner.size() + (long long)nedge * (long long)edge.size();
  }
  return t;
}

// only: the instances to tile (all when null); the others get no tiles
void build_tiles(pcw_handle* h, int ty, const std::vector<char>* only = nullptr) {
  h->htiles.clear();
  const bool shorten = h->p.precision == PCW_FP32;
  std::vector<int2> inner, edge;
  for (int k = 0; k < h->n; ++k) {
    KInst& I = h->hinst[k];
    I.tile0 = (int)h->htiles.size();
    if (only && !(*only)[k]) {
      I.ntiles = 0;
      continue;
    }
    std::vector<int> xs;
    const int nsx = (I.nx + kStripOut - 1) / kStripOut;
    for (int s = 0; s < nsx; ++s) {
      int x0 = 1 + kStripOut * s;
      if (x0 + kStripOut - 1 > I.nx) x0 = std::max(1, I.nx - kStripOut + 1);
      xs.push_back(x0);
    }
    row_bands(I.ny, tile_rows_for(I, ty), shorten, false, inner);
    row_bands(I.ny, tile_rows_for(I, ty), shorten, shorten, edge);
    std::vector<int4> t;
    for (int s = 0; s < nsx; ++s) {
      const bool is_edge = s == 0 || s == nsx - 1;
      for (const int2& b : is_edge ? edge : inner) t.push_back(make_int4(k, xs[s], b.x, b.y));
    }
    std::stable_sort(t.begin(), t.end(), [](const int4& a, const int4& b) {
      return a.z != b.z ? a.z < b.z : a.y < b.y;
    });
    h->htiles.insert(h->htiles.end(), t.begin(), t.end());
    I.ntiles = (int)t.size();
  }
  h->ntiles = (int)h->htiles.size();
  choose_split(h);
  order_tiles(h, 0, h->split > 0 ? h->split : h->ntiles);
  if (h->split > 0) order_tiles(h, h->split, h->ntiles);
}

// tile list, descriptors and live tile count to the device (between launches)
void drop_graph(pcw_handle* h);

int upload_tiles(pcw_handle* h) {
  if (!h->dtiles || h->ntiles > h->dtiles_cap) {
    drop_graph(h);  // captured launches hold the old list's address
    HIPCHK(hipStreamSynchronize(h->stream));  // and launched ones may still read it
    if (h->dtiles) HIPCHK(hipFree(h->dtiles));
    h->dtiles = nullptr;
    h->dtiles_cap = std::max(1, h->ntiles);
    HIPCHK(hipMalloc(&h->dtiles, sizeof(int4) * h->dtiles_cap));
  }
  HIPCHK(copy_wait(h, h->dtiles, h->htiles.data(), sizeof(int4) * h->ntiles, hipMemcpyHostToDevice));
  HIPCHK(copy_wait(h, h->dinst, h->hinst.data(), sizeof(KInst) * h->n, hipMemcpyHostToDevice));
  const int lv[6] = {h->ntiles, 0, h->split, 0, h->ntiles - h->split, h->split};
  HIPCHK(copy_wait(h, h->dlive, lv, sizeof lv, hipMemcpyHostToDevice));
  return PCW_OK;
}

int pull_states(pcw_handle* h) {
  HIPCHK(hipMemcpyAsync(h->hst.data(), h->dst, sizeof(InstState) * h->n, hipMemcpyDeviceToHost,
                        h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  return PCW_OK;
}

int push_states(pcw_handle* h) {
  HIPCHK(hipMemcpyAsync(h->dst, h->hst.data(), sizeof(InstState) * h->n, hipMemcpyHostToDevice,
                        h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  return PCW_OK;
}

void drop_graph(pcw_handle* h) {
  for (auto& g : h->graphs) (void)hipGraphExecDestroy(g.exec);
  h->graphs.clear();
}

int limiter_code(const pcw_params& p) { return (p.higher_order && p.use_limiter) ? p.limiter : 0; }

// enqueue one fused attempt for every instance.  Instances with metric arrays
// and those on the identity map run in separate launches of the kernel (FP64:
// identity-map instances whose metrics are exactly 1 in a third one).
// lane 0: the whole list on the handle's stream; lanes 1 and 2: the two parts
// of a split list (FP32), each on its own stream with a grid of its own.
int enqueue_fused(pcw_handle* h, int lane = 0) {
  const int threads = 64 * h->wpb;
  const int cap = lane == 0 ? h->tile_cap : h->lane_cap[lane - 1];
  int blocks = (std::max(1, cap) + h->wpb - 1) / h->wpb;
  blocks = (int)round_up(blocks, 8);
  const int* live = h->dlive + 2 * lane;
  hipStream_t stream = lane == 2 ? h->stream2 : h->stream;
  if (h->p.precision == PCW_FP64) {
    bool has[3] = {false, false, false};
    for (const auto& I : h->hinst) has[metric_mode(I)] = true;
    FusedArgs<double> a{(double*)h->pool, (const double*)h->bpool, (const double*)h->mpool,
                        h->dinst, h->dst, h->dtiles, live, cap, h->p.higher_order,
                        h->p.dry_tol > 0.0, 0, h->p.grav, h->p.eps, h->p.dry_tol, nullptr};
    for (int mode = 0; mode < 3; ++mode) {
      if (!has[mode]) continue;
      hipError_t e = launch_fused_f64(limiter_code(h->p), mode, a, blocks, threads, stream);
      if (e != hipSuccess)
        return fail(h, PCW_ERR_HIP, std::string("fused launch: ") + hipGetErrorString(e));
    }
    // ghost cells of the new state, of the lane's instances: lane 1 holds the
    // instances below lane_k, lane 2 the others (a captured lane graph is
    // captured again when the split moves to another instance)
    const int k0 = lane == 2 ? h->lane_k : 0, k1 = lane == 1 ? h->lane_k : h->n;
    hipError_t e = launch_ghosts<double>((double*)h->pool, h->dinst + k0, h->dst + k0, k1 - k0,
                                         h->max_side, kGhostsAfterFused, stream);
    if (e != hipSuccess)
      return fail(h, PCW_ERR_HIP, std::string("ghost launch: ") + hipGetErrorString(e));
    return PCW_OK;
  }
  // FP32: instances with and without source terms run in separate launches
  const bool wetdry = h->p.dry_tol > 0.0;
  bool any_src = wetdry, any_nosrc = !wetdry;
  for (const auto& d : h->dom) {
    any_src = any_src || d.source != PCW_SRC_NONE;
    any_nosrc = any_nosrc || (d.source == PCW_SRC_NONE && !wetdry);
  }
  FusedArgs<float> a{(float*)h->pool, (const float*)h->bpool, (const float*)h->mpool,
                     h->dinst, h->dst, h->dtiles, live, cap, h->p.higher_order,
                     wetdry, 0, h->p.grav, h->p.eps, h->p.dry_tol, h->dslots};
  for (int met = 0; met <= 1; ++met) {
    if (met ? !h->any_met : !h->any_plain) continue;
    hipError_t e = hipSuccess;
    for (int phys = 0; phys <= 1 && e == hipSuccess; ++phys) {
      if (phys ? !any_src : !any_nosrc) continue;
      e = launch_fused_f32(limiter_code(h->p), met, phys ? (wetdry ? 2 : 1) : 0, a, blocks, threads,
                           stream);
    }
    if (e != hipSuccess)
      return fail(h, PCW_ERR_HIP, std::string("fused launch: ") + hipGetErrorString(e));
  }
  return PCW_OK;
}

// enqueue the staged stage kernels of one attempt (without the finish)
int enqueue_staged_sweeps(pcw_handle* h, int backup) {
  const size_t wplane_total = 0;
  (void)wplane_total;
  if (backup) {
    hipError_t e = launch_copy<double>((double*)h->pool, h->dinst, h->dst, h->n, 1, 0, h->stream);
    if (e != hipSuccess) return fail(h, PCW_ERR_HIP, hipGetErrorString(e));
  }
  long long woff = 0;
  for (int k = 0; k < h->n; ++k) {
    const KInst& I = h->hinst[k];
    StagedArgs a{h->dinst + k, h->dst + k, (double*)h->pool + I.off[0], h->work + woff,
                 I.moff >= 0 ? (const double*)h->mpool + I.moff : nullptr, h->p.grav, h->p.eps,
                 h->p.dry_tol, limiter_code(h->p), 0};
    for (int dir = 1; dir <= 2; ++dir) {
      hipError_t e = launch_staged_sweep(a, I, dir, h->p.higher_order, h->stream);
      if (e != hipSuccess) return fail(h, PCW_ERR_HIP, hipGetErrorString(e));
    }
    woff += (long long)W_COUNT * I.plane;
  }
  return PCW_OK;
}

// source term (every instance: without a source kind it is the reference's
// q + 0*dt) and ghost fill of the staged engine's attempt
template <typename T>
int enqueue_source_ghosts(pcw_handle* h) {
  hipError_t e = launch_source<T>((T*)h->pool, (const T*)h->bpool, h->dinst, h->dst, h->n,
                                  h->max_nx, h->max_ny, h->p.dry_tol > 0.0, h->p.grav,
                                  h->p.dry_tol, h->stream);
  if (e != hipSuccess) return fail(h, PCW_ERR_HIP, hipGetErrorString(e));
  e = launch_ghosts<T>((T*)h->pool, h->dinst, h->dst, h->n, h->max_side, kGhostsCur, h->stream);
  if (e != hipSuccess) return fail(h, PCW_ERR_HIP, hipGetErrorString(e));
  return PCW_OK;
}

// one device-controlled attempt (graph body) of the whole list or of a lane
int enqueue_attempt(pcw_handle* h, int lane) {
  if (h->p.engine == PCW_ENGINE_FUSED) return enqueue_fused(h, lane);
  int rc = enqueue_staged_sweeps(h, 1);
  if (rc) return rc;
  rc = enqueue_source_ghosts<double>(h);
  if (rc) return rc;
  hipError_t e = launch_st_finish(h->dinst, h->dst, h->n, h->stream);
  if (e != hipSuccess) return fail(h, PCW_ERR_HIP, hipGetErrorString(e));
  e = launch_copy<double>((double*)h->pool, h->dinst, h->dst, h->n, 0, 1, h->stream);
  if (e != hipSuccess) return fail(h, PCW_ERR_HIP, hipGetErrorString(e));
  return PCW_OK;
}

int set_dts(pcw_handle* h, const double* dt, int controlled) {
  if (h->n > kScratch) return fail(h, PCW_ERR_ARG, "too many instances for host control");
  HIPCHK(hipMemcpyAsync(h->dscratch, dt, sizeof(double) * h->n, hipMemcpyHostToDevice, h->stream));
  HIPCHK(launch_set_dt(h->dst, h->dscratch, h->n, controlled, h->stream));
  return PCW_OK;
}

int check_errors(pcw_handle* h) {
  for (int k = 0; k < h->n; ++k)
    if (h->hst[k].status == PCW_STATUS_ERROR) {
      int e = h->hst[k].error ? h->hst[k].error : PCW_ERR_STATE;
      return fail(h, e,
                  e == PCW_ERR_STATE
                      ? "instance " + std::to_string(k) +
                            ": NaN in the Jacobian or complex eigenvalues (h <= 0)"
                      : "instance " + std::to_string(k) + ": time controller reached end_program");
    }
  return PCW_OK;
}

// host control: NaN/complex eigenvalues of the last attempt, by sweep
int step_bad_code(pcw_handle* h) {
  int eta_only = 0;
  for (int k = 0; k < h->n; ++k) {
    if (h->hst[k].last_bad & 1)
      return fail(h, PCW_ERR_STATE, "instance " + std::to_string(k) +
                                        ": NaN in the Jacobian or complex eigenvalues in the xi sweep");
    if (h->hst[k].last_bad & 2) eta_only = 1;
  }
  if (eta_only)
    return fail(h, PCW_ERR_STATE_ETA,
                "NaN in the Jacobian or complex eigenvalues in the eta sweep (fatal unless the xi "
                "sweep alone rejects the step)");
  return PCW_OK;
}

// The instantiated graph of `len` device-controlled attempts of the whole
// list (lane 0) or of one lane, captured on first use on the stream it runs
// on.  At most a few are kept (a run uses two lengths, of one or two lanes).
int graph_of(pcw_handle* h, int len, int lane, hipGraphExec_t* out) {
  for (auto& g : h->graphs)
    if (g.len == len && g.lane == lane) {
      *out = g.exec;
      return PCW_OK;
    }
  if (h->graphs.size() >= 6) {
    (void)hipGraphExecDestroy(h->graphs.front().exec);
    h->graphs.erase(h->graphs.begin());
  }
  hipStream_t stream = lane == 2 ? h->stream2 : h->stream;
  HIPCHK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
  for (int s = 0; s < len; ++s) {
    int rc = enqueue_attempt(h, lane);
    if (rc) {
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(stream, &g);
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
  }
  hipGraph_t graph = nullptr;
  HIPCHK(hipStreamEndCapture(stream, &graph));
  hipGraphExec_t gx = nullptr;
  hipError_t e = hipGraphInstantiate(&gx, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess) return fail(h, PCW_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
  h->graphs.push_back({len, lane, gx});
  *out = gx;
  return PCW_OK;
}

// The tile list of every instance at the tuned height, after a run left a
// list of its running instances only.
int ensure_full_tiles(pcw_handle* h) {
  if (!h->tiles_subset) return PCW_OK;
  build_tiles(h, h->tile_rows);
  h->tiles_subset = false;
  if (h->ntiles != h->tile_cap) {
    h->tile_cap = h->ntiles;
    drop_graph(h);
  }
  return upload_tiles(h);
}

// Between the launches of a device-controlled run of several instances: when
// instances have finished, the tile list is rebuilt for the running ones, so
// that no launch dispatches tiles that exit at once, and with shorter tiles
// (shrink_rows) while the running instances' tiles would not fill the GPU.
// The step launches read the list's length from device memory, so a captured
// graph stays valid; it is captured again, for a smaller grid, once the list
// is below a quarter of the grid.  Cell values do not depend on the tiling.
constexpr int kMinRetileRows = 16;
constexpr int kRebuiltRows = 64;
constexpr int kResidentWavesPerCU = 16;
constexpr int kSimdsPerCU = 4;

// compute units of the current device (the one the handle's memory is on)
int cu_count() {
  int dev = 0, cu = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 1;
  if (hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 1;
  return cu > 0 ? cu : 1;
}

// Tile height for ty requested: bands shortened by quarters down to
// kMinRetileRows while the tiles would not fill the GPU's resident waves
// once (kResidentWavesPerCU per CU), then further, down to 3 rows (FP32) or
// 4 (FP64), while they would not give every SIMD one wave.  Such a small
// domain's step is bound by one wave's march down its tile (about 0.7 us a
// row in FP32), which shorter bands cut by more than the halo rows they
// recompute cost (256^2: FP32 20.5 -> 14 us, FP64 52.2 -> 31.3 us per step);
// FP64 stops at 4 rows, where its arithmetic for the recomputed rows starts
// to dominate.  only: the instances to count (all when null).
int shrink_rows(const pcw_handle* h, int ty, const std::vector<char>* only = nullptr) {
  const long long cu = cu_count();
  while (ty > kMinRetileRows && count_tiles(h, ty, only) < cu * kResidentWavesPerCU)
    ty = std::max(kMinRetileRows, ty * 3 / 4);
  const int floor_rows = h->p.precision == PCW_FP32 ? kMinBandRows : 4;
  while (ty > floor_rows && count_tiles(h, ty, only) < cu * kSimdsPerCU)
    ty = std::max(floor_rows, ty * 3 / 4);
  return ty;
}

int maybe_retile(pcw_handle* h) {
  if (h->p.engine != PCW_ENGINE_FUSED || h->n < 2) return PCW_OK;
  std::vector<char> run(h->n, 0);
  int nrun = 0;
  for (int k = 0; k < h->n; ++k) {
    run[k] = h->hst[k].status == PCW_STATUS_RUNNING;
    nrun += run[k];
  }
  if (nrun == 0 || nrun == h->tiled_running) return PCW_OK;
  // keep the list while at least three quarters of it belongs to running
  // instances and those tiles alone fill the resident waves: a finished
  // instance's tiles exit at once, which costs less than a rebuild (about
  // 1 ms for 60,000 tiles; the 512-instance ensemble ran 0.5 % faster)
  long long live = 0;
  for (int k = 0; k < h->n; ++k)
    if (run[k]) live += h->hinst[k].ntiles;
  if (live * 4 >= (long long)h->ntiles * 3 && live >= (long long)cu_count() * kResidentWavesPerCU)
    return PCW_OK;
  // A rebuilt list starts from bands of at least kRebuiltRows rows, shortened
  // while the running instances' tiles would not fill the GPU.  The instances
  // still running when the smaller ones have finished are the large ones, on
  // which short bands (the autotuner's pick while every instance runs, when
  // small instances make 32 and 64 rows time alike) recompute more halo rows:
  // the 512-instance ensemble with 32-row tiles ran 0.417 s instead of
  // 0.423-0.443 s.
  int ty = shrink_rows(h, std::max(h->tile_rows, kRebuiltRows), &run);
  while (count_tiles(h, ty, &run) > h->tile_cap) ty += std::max(1, ty / 4);
  build_tiles(h, ty, &run);
  h->tiles_subset = true;
  h->tiled_running = nrun;
  // an instance without tiles takes no attempt, so the FP64 ghost fill after
  // each step must not refill it (its first tile would have cleared this)
  for (int k = 0; k < h->n; ++k)
    if (!run[k]) h->hst[k].ghosts_due = 0;
  int rc = push_states(h);
  if (rc) return rc;
  if ((long long)h->ntiles * 4 < h->tile_cap) {
    h->tile_cap = std::max(1, h->ntiles);
    drop_graph(h);
  }
  return upload_tiles(h);
}

}  // namespace

extern "C" {

const char* pcw_last_error(const pcw_handle* h) {
  if (h) return h->err.c_str();
  return g_loose_error[0] ? g_loose_error : "null handle";
}

int pcw_create(const pcw_params* p, const pcw_domain* domains, int32_t n_domains,
               pcw_handle** out) {
  if (!p || !domains || n_domains < 1 || !out) return PCW_ERR_ARG;
  *out = nullptr;
  pcw_handle* h = new pcw_handle();
  h->p = *p;
  h->n = n_domains;
  h->esize = (p->precision == PCW_FP32) ? 4 : 8;
  if (n_domains > PCW_MAX_DOMAINS) {
    fail(h, PCW_ERR_ARG, "a handle holds at most PCW_MAX_DOMAINS (65,535) domains");
    *out = h;
    return PCW_ERR_ARG;
  }
  if (p->precision != PCW_FP32 && p->precision != PCW_FP64) {
    fail(h, PCW_ERR_ARG, "precision must be PCW_FP64 or PCW_FP32");
    *out = h;
    return PCW_ERR_ARG;
  }
  if (p->engine == PCW_ENGINE_STAGED && p->precision != PCW_FP64) {
    fail(h, PCW_ERR_ARG, "the staged engine is FP64 only");
    *out = h;
    return PCW_ERR_ARG;
  }
  if (p->higher_order && p->use_limiter && (p->limiter < 1 || p->limiter > 4)) {
    fail(h, PCW_ERR_ARG, "invalid limiter choice => ABORTING");
    *out = h;
    return PCW_ERR_ARG;
  }
  h->dom.assign(domains, domains + n_domains);
  h->hinst.resize(n_domains);
  h->hst.assign(n_domains, InstState{});
  bool any_bathy = false;
  long long base = 0, bbase = 0;
  for (int k = 0; k < n_domains; ++k) {
    const pcw_domain& d = domains[k];
    if (d.nx < 2 || d.ny < 2) {
      fail(h, PCW_ERR_ARG, "each domain needs nx >= 2 and ny >= 2");
      *out = h;
      return PCW_ERR_ARG;
    }
    if (d.nx > PCW_MAX_SIDE || d.ny > PCW_MAX_SIDE) {
      fail(h, PCW_ERR_ARG, "each domain needs nx, ny <= PCW_MAX_SIDE (262,136)");
      *out = h;
      return PCW_ERR_ARG;
    }
    for (int s = 0; s < 4; ++s)
      if (d.bc[s] < PCW_BC_ZERO_GRADIENT || d.bc[s] > PCW_BC_REFLECTIVE) {
        fail(h, PCW_ERR_BC,
             d.bc[s] == 1 ? "User-defined boundary condition not implemented yet!"
                          : "boundary condition not implemented yet!");
        *out = h;
        return PCW_ERR_BC;
      }
    KInst& I = h->hinst[k];
    describe_instance(*p, d, I);
    I.off[0] = base;
    I.off[1] = base + 3 * I.plane;
    base += 6 * I.plane;
    if (d.source == PCW_SRC_BATHY || d.source == PCW_SRC_BATHY_FRICTION) {
      I.boff = bbase;
      bbase += I.plane;
      any_bathy = true;
    }
    h->max_nx = std::max(h->max_nx, d.nx);
    h->max_ny = std::max(h->max_ny, d.ny);
    h->max_side = std::max(h->max_side, d.nx + d.ny);
    InstState& S = h->hst[k];
    std::memset(&S, 0, sizeof S);
    S.status = PCW_STATUS_RUNNING;
    S.last_ok = 1;
  }
  *out = h;
  h->pool_elems = base;
  h->bpool_elems = bbase;
  // default tile heights: FP32 64 rows (memory-bound; the autotuner refines it), FP64
  // 128 rows (arithmetic-bound: fewer recomputed halo rows per tile).  Without a
  // height from the caller, the bands are shortened (shrink_rows) while the tiles
  // would not fill the GPU: a small domain stepped without the autotuner (the
  // Fortran host, pcw_step) then still spreads over the CUs (256^2 in FP32: 430
  // tiles of 3 rows or fewer instead of 10 of 64).  Cell values do not depend on it.
  int ty = p->tile_rows > 1 ? p->tile_rows : (p->precision == PCW_FP32 ? 64 : 128);
  if (p->tile_rows <= 1) ty = shrink_rows(h, ty);
  h->tile_rows = ty;
  h->wpb = (p->waves_per_block >= 1 && p->waves_per_block <= 4) ? p->waves_per_block : 2;
  HIPCHK(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
  HIPCHK(hipEventCreate(&h->ev0));
  HIPCHK(hipEventCreate(&h->ev1));
  if (p->engine == PCW_ENGINE_FUSED && n_domains >= 2) {
    // the second lane's stream (default priority, like the first)
    HIPCHK(hipStreamCreateWithFlags(&h->stream2, hipStreamNonBlocking));
    h->two_lanes = true;
    HIPCHK(hipEventCreateWithFlags(&h->evf, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&h->evj, hipEventDisableTiming));
  }
  HIPCHK(hipMalloc(&h->pool, (size_t)base * h->esize));
  HIPCHK(zero_wait(h, h->pool, (size_t)base * h->esize));
  if (any_bathy) {
    HIPCHK(hipMalloc(&h->bpool, (size_t)bbase * h->esize));
    HIPCHK(zero_wait(h, h->bpool, (size_t)bbase * h->esize));
  }
  HIPCHK(hipMalloc(&h->dinst, sizeof(KInst) * n_domains));
  HIPCHK(hipMalloc(&h->dst, sizeof(InstState) * n_domains));
  HIPCHK(hipMalloc(&h->dscratch, sizeof(double) * kScratch));
  if (p->precision == PCW_FP32) {
    HIPCHK(hipMalloc(&h->dslots, sizeof(TileSlot) * kSlots * n_domains));
    HIPCHK(zero_wait(h, h->dslots, sizeof(TileSlot) * kSlots * n_domains));
  }
  if (p->engine == PCW_ENGINE_STAGED) {
    long long welems = 0;
    for (const auto& I : h->hinst) welems += (long long)W_COUNT * I.plane;
    HIPCHK(hipMalloc(&h->work, (size_t)welems * sizeof(double)));
    HIPCHK(zero_wait(h, h->work, (size_t)welems * sizeof(double)));
  }
  HIPCHK(hipMalloc(&h->dlive, 6 * sizeof(int)));
  build_tiles(h, h->tile_rows);
  h->tile_cap = h->ntiles;
  int rc = upload_tiles(h);
  if (rc) return rc;
  return push_states(h);
}

int pcw_destroy(pcw_handle* h) {
  if (!h) return PCW_ERR_ARG;
  drop_graph(h);
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  if (h->stream2) (void)hipStreamSynchronize(h->stream2);
  void* ptrs[] = {h->pool, h->bpool, h->mpool, h->dinst, h->dst, h->dtiles, h->dslots, h->dsnaps, h->dlog,
                  h->work, h->dscratch, h->dlive};
  for (void* q : ptrs)
    if (q) (void)hipFree(q);
  if (h->ev0) (void)hipEventDestroy(h->ev0);
  if (h->ev1) (void)hipEventDestroy(h->ev1);
  if (h->evf) (void)hipEventDestroy(h->evf);
  if (h->evj) (void)hipEventDestroy(h->evj);
  if (h->stream2) (void)hipStreamDestroy(h->stream2);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return PCW_OK;
}

int pcw_device_bytes(const pcw_handle* h, int64_t* bytes) {
  if (!h || !bytes) return PCW_ERR_ARG;
  long long w = 0;
  if (h->work)
    for (const auto& I : h->hinst) w += (long long)W_COUNT * I.plane * 8;
  *bytes = (h->pool_elems + h->bpool_elems + h->mpool_elems) * h->esize + w;
  return PCW_OK;
}

// ------------------------------------------------------------ state I/O
int pcw_set_state(pcw_handle* h, int32_t inst, const double* vars) {
  if (!h || !vars || inst < 0 || inst >= h->n) return fail(h, PCW_ERR_ARG, "bad argument");
  const KInst& I = h->hinst[inst];
  const int nxf = I.nx + 4, nyf = I.ny + 4;
  std::vector<char> buf((size_t)3 * I.plane * h->esize, 0);
  for (int j = 0; j < nyf; ++j)
    for (int i = 0; i < nxf; ++i)
      for (int me = 0; me < 3; ++me) {
        double v = vars[me + 3 * ((size_t)i + (size_t)nxf * j)];
        size_t o = (size_t)me * I.plane + (size_t)j * I.pitch + col_pos(I, i - 1);
        if (h->esize == 4) ((float*)buf.data())[o] = (float)v;
        else ((double*)buf.data())[o] = v;
      }
  for (int b = 0; b < 2; ++b)
    HIPCHK(copy_wait(h, (char*)h->pool + (size_t)I.off[b] * h->esize, buf.data(), buf.size(),
                     hipMemcpyHostToDevice));
  return PCW_OK;
}

int pcw_get_state(pcw_handle* h, int32_t inst, double* vars) {
  if (!h || !vars || inst < 0 || inst >= h->n) return fail(h, PCW_ERR_ARG, "bad argument");
  int rc = pull_states(h);
  if (rc) return rc;
  const KInst& I = h->hinst[inst];
  const int cur = h->hst[inst].cur;
  const int nxf = I.nx + 4, nyf = I.ny + 4;
  std::vector<char> buf((size_t)3 * I.plane * h->esize);
  HIPCHK(copy_wait(h, buf.data(), (char*)h->pool + (size_t)I.off[cur] * h->esize, buf.size(),
                   hipMemcpyDeviceToHost));
  for (int j = 0; j < nyf; ++j)
    for (int i = 0; i < nxf; ++i)
      for (int me = 0; me < 3; ++me) {
        size_t o = (size_t)me * I.plane + (size_t)j * I.pitch + col_pos(I, i - 1);
        vars[me + 3 * ((size_t)i + (size_t)nxf * j)] =
            h->esize == 4 ? (double)((float*)buf.data())[o] : ((double*)buf.data())[o];
      }
  return PCW_OK;
}

int pcw_set_bathymetry(pcw_handle* h, int32_t inst, const double* b) {
  if (!h || !b || inst < 0 || inst >= h->n) return fail(h, PCW_ERR_ARG, "bad argument");
  const KInst& I = h->hinst[inst];
  if (I.boff < 0) return fail(h, PCW_ERR_ARG, "instance has no bathymetry source");
  const int nxf = I.nx + 4, nyf = I.ny + 4;
  std::vector<char> buf((size_t)I.plane * h->esize, 0);
  for (int j = 0; j < nyf; ++j)
    for (int i = 0; i < nxf; ++i) {
      double v = b[(size_t)i + (size_t)nxf * j];
      size_t o = (size_t)j * I.pitch + col_pos(I, i - 1);
      if (h->esize == 4) ((float*)buf.data())[o] = (float)v;
      else ((double*)buf.data())[o] = v;
    }
  HIPCHK(copy_wait(h, (char*)h->bpool + (size_t)I.boff * h->esize, buf.data(), buf.size(),
                   hipMemcpyHostToDevice));
  return PCW_OK;
}

int pcw_set_metrics(pcw_handle* h, int32_t inst, const double* kappa, const double* gamma_xi,
                    const double* gamma_eta) {
  if (!h || !kappa || !gamma_xi || !gamma_eta || inst < 0 || inst >= h->n)
    return fail(h, PCW_ERR_ARG, "bad argument");
  if (!h->mpool) {  // one metric region per instance, allocated on first use
    long long off = 0;
    for (auto& I : h->hinst) off += 3 * I.plane;
    HIPCHK(hipMalloc(&h->mpool, (size_t)off * h->esize));
    HIPCHK(zero_wait(h, h->mpool, (size_t)off * h->esize));
    h->mpool_elems = off;
  }
  long long off = 0;
  for (int k = 0; k < inst; ++k) off += 3 * h->hinst[k].plane;
  KInst& I = h->hinst[inst];
  I.moff = off;
  const int nxf = I.nx + 4, nyf = I.ny + 4;
  std::vector<char> buf((size_t)3 * I.plane * h->esize, 0);
  const double* src[3] = {kappa, gamma_xi, gamma_eta};
  for (int c = 0; c < 3; ++c)
    for (int j = 0; j < nyf; ++j)
      for (int i = 0; i < nxf; ++i) {
        double v = src[c][(size_t)i + (size_t)nxf * j];
        size_t o = (size_t)c * I.plane + (size_t)j * I.pitch + col_pos(I, i - 1);
        if (h->esize == 4) ((float*)buf.data())[o] = (float)v;
        else ((double*)buf.data())[o] = v;
      }
  HIPCHK(copy_wait(h, (char*)h->mpool + (size_t)off * h->esize, buf.data(), buf.size(),
                   hipMemcpyHostToDevice));
  HIPCHK(copy_wait(h, h->dinst + inst, &I, sizeof(KInst), hipMemcpyHostToDevice));
  h->any_met = false;
  h->any_plain = false;
  for (const auto& J : h->hinst) {
    h->any_met = h->any_met || J.moff >= 0;
    h->any_plain = h->any_plain || J.moff < 0;
  }
  drop_graph(h);  // the set of launches per step changed
  return PCW_OK;
}

// ------------------------------------------------- host-controlled stepping
int pcw_step(pcw_handle* h, const double* dt, double* smax) {
  if (!h || !dt) return fail(h, PCW_ERR_ARG, "bad argument");
  int rc = ensure_full_tiles(h);
  if (rc) return rc;
  rc = set_dts(h, dt, 0);
  if (rc) return rc;
  if (h->p.engine == PCW_ENGINE_FUSED) {
    rc = enqueue_fused(h);
  } else {
    rc = enqueue_staged_sweeps(h, 1);
    if (!rc) rc = enqueue_source_ghosts<double>(h);
    if (!rc) HIPCHK(launch_st_finish(h->dinst, h->dst, h->n, h->stream));
  }
  if (rc) return rc;
  rc = pull_states(h);
  if (rc) return rc;
  rc = check_errors(h);
  if (rc) return rc;
  if (smax)
    for (int k = 0; k < h->n; ++k) {
      smax[2 * k] = h->hst[k].last_smax_x;
      smax[2 * k + 1] = h->hst[k].last_smax_y;
    }
  return step_bad_code(h);
}

// Accept and reject only enqueue their work on the handle's stream: every
// entry point that reads results back synchronises the stream first.
int pcw_accept(pcw_handle* h) {
  if (!h) return PCW_ERR_ARG;
  if (h->p.engine == PCW_ENGINE_FUSED) HIPCHK(launch_flip(h->dst, h->n, h->stream));
  return PCW_OK;
}

int pcw_reject(pcw_handle* h) {
  if (!h) return PCW_ERR_ARG;
  if (h->p.engine == PCW_ENGINE_STAGED)
    HIPCHK(launch_copy<double>((double*)h->pool, h->dinst, h->dst, h->n, 0, 0, h->stream));
  return PCW_OK;
}

int pcw_sweep(pcw_handle* h, int32_t ixy, const double* dt, double* smax) {
  if (!h || !dt || (ixy != 1 && ixy != 2)) return fail(h, PCW_ERR_ARG, "bad argument");
  if (h->p.engine != PCW_ENGINE_STAGED)
    return fail(h, PCW_ERR_ARG, "single sweeps need the staged engine");
  int rc = set_dts(h, dt, 0);
  if (rc) return rc;
  long long woff = 0;
  for (int k = 0; k < h->n; ++k) {
    const KInst& I = h->hinst[k];
    StagedArgs a{h->dinst + k, h->dst + k, (double*)h->pool + I.off[0], h->work + woff,
                 I.moff >= 0 ? (const double*)h->mpool + I.moff : nullptr, h->p.grav, h->p.eps,
                 h->p.dry_tol, limiter_code(h->p), 0};
    HIPCHK(launch_staged_sweep(a, I, ixy, h->p.higher_order, h->stream));
    woff += (long long)W_COUNT * I.plane;
  }
  HIPCHK(launch_st_finish(h->dinst, h->dst, h->n, h->stream));
  rc = pull_states(h);
  if (rc) return rc;
  rc = check_errors(h);
  if (rc) return rc;
  if (smax)
    for (int k = 0; k < h->n; ++k) smax[k] = ixy == 1 ? h->hst[k].last_smax_x : h->hst[k].last_smax_y;
  for (int k = 0; k < h->n; ++k)
    if (h->hst[k].last_bad)
      return fail(h, PCW_ERR_STATE, "instance " + std::to_string(k) +
                                        ": NaN in the Jacobian or complex eigenvalues");
  return PCW_OK;
}

int pcw_source(pcw_handle* h, const double* dt) {
  if (!h || !dt) return fail(h, PCW_ERR_ARG, "bad argument");
  int rc = set_dts(h, dt, 0);
  if (rc) return rc;
  hipError_t e;
  if (h->esize == 4)
    e = launch_source<float>((float*)h->pool, (const float*)h->bpool, h->dinst, h->dst, h->n,
                             h->max_nx, h->max_ny, h->p.dry_tol > 0.0, h->p.grav, h->p.dry_tol,
                             h->stream);
  else
    e = launch_source<double>((double*)h->pool, (const double*)h->bpool, h->dinst, h->dst, h->n,
                              h->max_nx, h->max_ny, h->p.dry_tol > 0.0, h->p.grav, h->p.dry_tol,
                              h->stream);
  if (e != hipSuccess) return fail(h, PCW_ERR_HIP, hipGetErrorString(e));
  HIPCHK(hipStreamSynchronize(h->stream));
  return PCW_OK;
}

int pcw_fill_ghosts(pcw_handle* h) {
  if (!h) return PCW_ERR_ARG;
  hipError_t e = h->esize == 4
                     ? launch_ghosts<float>((float*)h->pool, h->dinst, h->dst, h->n, h->max_side, kGhostsCur, h->stream)
                     : launch_ghosts<double>((double*)h->pool, h->dinst, h->dst, h->n, h->max_side, kGhostsCur, h->stream);
  if (e != hipSuccess) return fail(h, PCW_ERR_HIP, hipGetErrorString(e));
  HIPCHK(hipStreamSynchronize(h->stream));
  return PCW_OK;
}

// -------------------------------------------------- device-controlled run
int pcw_time_init(pcw_handle* h, const pcw_time* t) {
  if (!h || !t) return PCW_ERR_ARG;
  if (t->dt_min > t->dt_max) return fail(h, PCW_ERR_CONTROL, "ABORT: Invalid timestep bounds!");
  int rc = pull_states(h);
  if (rc) return rc;
  const int ns = t->n_snapshots < 1 ? 1 : t->n_snapshots;
  if (h->dsnaps) HIPCHK(hipFree(h->dsnaps));
  h->dsnaps = nullptr;
  HIPCHK(hipMalloc(&h->dsnaps, sizeof(double) * ns));
  std::vector<double> snaps(ns);
  tc_snapshots(*t, snaps.data());
  HIPCHK(copy_wait(h, h->dsnaps, snaps.data(), sizeof(double) * ns, hipMemcpyHostToDevice));
  for (int k = 0; k < h->n; ++k) {
    InstState& S = h->hst[k];
    tc_init(S.tc, *t);
    S.tc.snaps = h->dsnaps;
    S.smax_x = S.smax_y = 0ull;
    S.tiles_done = 0u;
    S.accepted = S.rejected = 0;
    S.max_accepted = 0;
    S.status = PCW_STATUS_RUNNING;
    S.error = 0;
    S.bad = 0;
    S.controlled = 1;
    S.last_ok = 1;
    S.ghosts_due = 0;
    S.log_n = 0;
  }
  return push_states(h);
}

int pcw_enable_log(pcw_handle* h, int32_t capacity) {
  if (!h || capacity < 0) return PCW_ERR_ARG;
  int rc = pull_states(h);
  if (rc) return rc;
  if (h->dlog) HIPCHK(hipFree(h->dlog));
  h->dlog = nullptr;
  h->log_cap = capacity;
  if (capacity > 0)
    HIPCHK(hipMalloc(&h->dlog, sizeof(pcw_step_record) * (size_t)capacity * h->n));
  for (int k = 0; k < h->n; ++k) {
    h->hst[k].log = capacity > 0 ? h->dlog + (size_t)k * capacity : nullptr;
    h->hst[k].log_cap = capacity;
    h->hst[k].log_n = 0;
  }
  return push_states(h);
}

int pcw_read_log(pcw_handle* h, int32_t inst, pcw_step_record* out, int32_t cap, int32_t* n) {
  if (!h || !out || !n || inst < 0 || inst >= h->n) return PCW_ERR_ARG;
  int rc = pull_states(h);
  if (rc) return rc;
  int m = std::min(cap, h->hst[inst].log_n);
  if (m > 0)
    HIPCHK(copy_wait(h, out, h->dlog + (size_t)inst * h->log_cap, sizeof(pcw_step_record) * m,
                     hipMemcpyDeviceToHost));
  *n = m;
  return PCW_OK;
}

int pcw_run(pcw_handle* h, int32_t steps_per_graph, int64_t max_attempts, int64_t max_accepted) {
  if (!h || steps_per_graph < 1) return PCW_ERR_ARG;
  int rc = pull_states(h);
  if (rc) return rc;
  for (int k = 0; k < h->n; ++k) {
    InstState& S = h->hst[k];
    if (!S.controlled) return fail(h, PCW_ERR_ARG, "call pcw_time_init before pcw_run");
    S.max_accepted = max_accepted;
    // an instance that an earlier run stopped at its accepted-step limit
    // (not at the final time, not by an error) continues under a higher limit
    if (S.status == PCW_STATUS_DONE && !S.tc.ended && S.tc.time < S.tc.time_final &&
        (max_accepted == 0 || S.accepted < max_accepted))
      S.status = PCW_STATUS_RUNNING;
  }
  rc = push_states(h);
  if (rc) return rc;
  // a list of an earlier run's running instances serves while every running
  // instance has tiles in it (pcw_time_init or a raised max_accepted can
  // restart others)
  bool covered = true;
  for (int k = 0; k < h->n; ++k)
    covered = covered && (h->hst[k].status != PCW_STATUS_RUNNING || h->hinst[k].ntiles > 0);
  if (!covered) {
    rc = ensure_full_tiles(h);
    if (rc) return rc;
    h->tiled_running = 0;
    for (int k = 0; k < h->n; ++k) h->tiled_running += h->hst[k].status == PCW_STATUS_RUNNING;
  }
  long long before = 0;
  for (int k = 0; k < h->n; ++k)
    before = std::max(before, h->hst[k].accepted + h->hst[k].rejected);
  long long attempts = 0, work_launches = 0;
  double graph_ms = 0.0, work_ms = 0.0;
  h->launches = 0;
  for (;;) {
    bool running = false;
    for (int k = 0; k < h->n; ++k) running = running || h->hst[k].status == PCW_STATUS_RUNNING;
    if (!running) break;
    if (max_attempts > 0 && attempts >= max_attempts) break;
    // attempts in this launch: the full graph, or fewer when every running
    // instance needs fewer accepted steps to reach max_accepted (a launch
    // past that point would only dispatch tiles that exit at once)
    int len = steps_per_graph;
    if (max_accepted > 0) {
      long long need = 0;
      for (int k = 0; k < h->n; ++k)
        if (h->hst[k].status == PCW_STATUS_RUNNING)
          need = std::max(need, max_accepted - h->hst[k].accepted);
      len = (int)std::max(1ll, std::min<long long>(len, need));
    }
    if (max_attempts > 0) len = (int)std::max(1ll, std::min<long long>(len, max_attempts - attempts));
    // two lanes while both halves of the list hold running instances
    bool run_lo = false, run_hi = false;
    for (int k = 0; k < h->n; ++k)
      if (h->hst[k].status == PCW_STATUS_RUNNING && h->hinst[k].ntiles > 0)
        (h->hinst[k].tile0 < h->split ? run_lo : run_hi) = true;
    const bool lanes = h->split > 0 && run_lo && run_hi;
    long long a0 = 0;
    for (int k = 0; k < h->n; ++k) a0 = std::max(a0, h->hst[k].accepted + h->hst[k].rejected);
    if (lanes) {
      // lane grids are captured for the lanes' sizes; a lane that outgrew
      // its grid or shrank below a quarter of it is captured again
      const int want[2] = {h->split, h->ntiles - h->split};
      const int k_split = h->htiles[h->split].x;
      bool fits = true;
      for (int l = 0; l < 2; ++l)
        fits = fits && want[l] <= h->lane_cap[l] && (long long)want[l] * 4 >= h->lane_cap[l];
      // the FP64 ghost launches of a lane graph name its instances
      if (h->p.precision == PCW_FP64) fits = fits && k_split == h->lane_k;
      if (!fits) {
        drop_graph(h);
        h->lane_cap[0] = want[0];
        h->lane_cap[1] = want[1];
        h->lane_k = k_split;
      }
      hipGraphExec_t g1 = nullptr, g2 = nullptr;
      rc = graph_of(h, len, 1, &g1);
      if (!rc) rc = graph_of(h, len, 2, &g2);
      if (rc) return rc;
      HIPCHK(hipEventRecord(h->ev0, h->stream));
      HIPCHK(hipEventRecord(h->evf, h->stream));  // lane 2 starts after the uploads on stream
      HIPCHK(hipStreamWaitEvent(h->stream2, h->evf, 0));
      HIPCHK(hipGraphLaunch(g1, h->stream));
      HIPCHK(hipGraphLaunch(g2, h->stream2));
      HIPCHK(hipEventRecord(h->evj, h->stream2));  // and stream waits for lane 2
      HIPCHK(hipStreamWaitEvent(h->stream, h->evj, 0));
      HIPCHK(hipEventRecord(h->ev1, h->stream));
    } else {
      hipGraphExec_t gx = nullptr;
      rc = graph_of(h, len, 0, &gx);
      if (rc) return rc;
      HIPCHK(hipEventRecord(h->ev0, h->stream));
      HIPCHK(hipGraphLaunch(gx, h->stream));
      HIPCHK(hipEventRecord(h->ev1, h->stream));
    }
    rc = pull_states(h);
    if (rc) return rc;
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, h->ev0, h->ev1));
    graph_ms += ms;
    h->launches += len;
    attempts += len;
    long long a1 = 0;
    for (int k = 0; k < h->n; ++k) a1 = std::max(a1, h->hst[k].accepted + h->hst[k].rejected);
    if (a1 - a0 == len) {  // every launch of this graph did an attempt
      work_ms += ms;
      work_launches += len;
    }
    rc = maybe_retile(h);
    if (rc) return rc;
  }
  long long after = 0;
  for (int k = 0; k < h->n; ++k)
    after = std::max(after, h->hst[k].accepted + h->hst[k].rejected);
  h->last_ms = graph_ms;
  // average launch time over the graphs whose launches all did work; when a
  // run ends inside every graph, over the attempts made
  h->avg_step_ms = work_launches > 0 ? work_ms / (double)work_launches
                   : (after > before) ? graph_ms / (double)(after - before) : 0.0;
  return check_errors(h);
}

int pcw_stop_at_snapshots(pcw_handle* h, int32_t on) {
  if (!h) return PCW_ERR_ARG;
  int rc = pull_states(h);
  if (rc) return rc;
  for (int k = 0; k < h->n; ++k) h->hst[k].stop_snap = on ? 1 : 0;
  return push_states(h);
}

int pcw_get_status(pcw_handle* h, int32_t inst, pcw_status* out) {
  if (!h || !out || inst < 0 || inst >= h->n) return PCW_ERR_ARG;
  int rc = pull_states(h);
  if (rc) return rc;
  const InstState& S = h->hst[inst];
  out->status = S.status;
  out->error = S.error;
  out->accepted = S.accepted;
  out->rejected = S.rejected;
  out->time = S.tc.time;
  out->dt = S.tc.dt;
  out->cfl_last = S.cfl_last;
  return PCW_OK;
}

// ------------------------------------------------------------ measurement
int pcw_totals(pcw_handle* h, int32_t inst, double out[3]) {
  if (!h || !out || inst < 0 || inst >= h->n) return PCW_ERR_ARG;
  const int nparts = std::min(1024, h->hinst[inst].ny);
  hipError_t e = h->esize == 4
                     ? launch_totals<float>((const float*)h->pool, h->dinst + inst, h->dst + inst,
                                            h->dscratch, nparts, h->dscratch + 3 * 1024, h->stream)
                     : launch_totals<double>((const double*)h->pool, h->dinst + inst, h->dst + inst,
                                             h->dscratch, nparts, h->dscratch + 3 * 1024, h->stream);
  if (e != hipSuccess) return fail(h, PCW_ERR_HIP, hipGetErrorString(e));
  HIPCHK(hipMemcpyAsync(out, h->dscratch + 3 * 1024, 3 * sizeof(double), hipMemcpyDeviceToHost,
                        h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  return PCW_OK;
}

int pcw_last_elapsed_ms(pcw_handle* h, double* ms) {
  if (!h || !ms) return PCW_ERR_ARG;
  *ms = h->last_ms;
  return PCW_OK;
}

constexpr long long kMaxPlannedTiles = 1LL << 28;

int64_t pcw_plan_tiles(const pcw_params* p, const pcw_domain* d, int32_t n, int32_t lanes,
                       int32_t* out, int64_t cap, int32_t* split) {
  if (!p || !d || n < 1 || n > PCW_MAX_DOMAINS || cap < 0 || (cap > 0 && !out)) return PCW_ERR_ARG;
  if (p->precision != PCW_FP32 && p->precision != PCW_FP64) return PCW_ERR_ARG;
  for (int k = 0; k < n; ++k)
    if (d[k].nx < 2 || d[k].ny < 2 || d[k].nx > PCW_MAX_SIDE || d[k].ny > PCW_MAX_SIDE)
      return PCW_ERR_ARG;
  pcw_handle h;  // host side only: no stream, no device memory
  h.p = *p;
  h.n = n;
  h.esize = p->precision == PCW_FP32 ? 4 : 8;
  h.dom.assign(d, d + n);
  h.hinst.resize(n);
  for (int k = 0; k < n; ++k) describe_instance(*p, d[k], h.hinst[k]);
  h.wpb = (p->waves_per_block >= 1 && p->waves_per_block <= 4) ? p->waves_per_block : 2;
  h.two_lanes = lanes != 0 && p->engine == PCW_ENGINE_FUSED && n >= 2;
  const int ty = p->tile_rows > 1 ? p->tile_rows : (p->precision == PCW_FP32 ? 64 : 128);
  if (count_tiles(&h, ty) > kMaxPlannedTiles) return PCW_ERR_ARG;  // list positions are ints
  build_tiles(&h, ty);
  if (split) *split = h.split;
  for (long long t = 0; t < std::min<long long>(cap, h.ntiles); ++t) {
    const int4 v = h.htiles[t];
    out[4 * t] = v.x;
    out[4 * t + 1] = v.y;
    out[4 * t + 2] = v.z;
    out[4 * t + 3] = v.w;
  }
  return h.ntiles;
}

int pcw_kernel_stats(pcw_handle* h, double* avg_step_ms, int64_t* launches) {
  if (!h) return PCW_ERR_ARG;
  if (avg_step_ms) *avg_step_ms = h->avg_step_ms;
  if (launches) *launches = h->launches;
  return PCW_OK;
}

int pcw_autotune(pcw_handle* h, int32_t steps, int32_t* tile_rows, int32_t* waves_per_block) {
  if (!h || steps < 1) return PCW_ERR_ARG;
  if (h->p.engine != PCW_ENGINE_FUSED) return fail(h, PCW_ERR_ARG, "autotuning applies to the fused engine");
  drop_graph(h);
  int rc = pull_states(h);
  if (rc) return rc;
  std::vector<InstState> saved = h->hst;
  for (auto& S : h->hst) {
    S.controlled = 0;
    S.status = PCW_STATUS_RUNNING;
  }
  rc = push_states(h);
  if (rc) return rc;
  // Candidates: powers of two, and for FP32 the tile heights whose tile
  // count just fits k rounds of the resident waves (k = 1..6), so that the
  // last round is full.
  const int wpb_c[] = {1, 2, 4};
  std::vector<std::pair<int, int>> cand;
  const int cu = cu_count();
  bool any_src = h->p.dry_tol > 0.0, any_nosrc = !(h->p.dry_tol > 0.0);
  for (const auto& d : h->dom) {
    any_src = any_src || d.source != PCW_SRC_NONE;
    any_nosrc = any_nosrc || (d.source == PCW_SRC_NONE && !(h->p.dry_tol > 0.0));
  }
  // heights below 16 rows where even 16-row tiles would not give every SIMD a
  // wave (a small domain, whose step one wave's march bounds)
  const bool small = count_tiles(h, 16) < (long long)cu * kSimdsPerCU;
  for (int wv : wpb_c) {
    for (int tr : {16, 32, 64, 128, 256}) cand.push_back({tr, wv});
    if (small)
      for (int tr : {kMinBandRows, 4, 6, 8}) cand.push_back({tr, wv});
    if (h->p.precision != PCW_FP32 || cu <= 0 || (any_src && any_nosrc) || (h->any_met && h->any_plain))
      continue;
    const int bpc = fused_f32_blocks_per_cu(limiter_code(h->p), h->p.higher_order != 0, h->any_met,
                                            any_src ? (h->p.dry_tol > 0.0 ? 2 : 1) : 0, 64 * wv);
    if (bpc <= 0) continue;
    const long long slots = (long long)bpc * wv * cu;  // resident waves
    auto ntiles = [&](int tr) { return count_tiles(h, tr); };
    for (int k = 1; k <= 6; ++k) {
      int lo = 8, hi = std::max(8, h->max_ny);
      if (ntiles(hi) > k * slots) continue;
      while (lo < hi) {  // smallest height whose tile count fits k rounds
        const int mid = (lo + hi) / 2;
        if (ntiles(mid) <= k * slots) hi = mid; else lo = mid + 1;
      }
      cand.push_back({lo, wv});
    }
  }
  std::sort(cand.begin(), cand.end());
  cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
  double best = 1e30;
  int br = h->tile_rows, bw = h->wpb;
  struct Timed {
    float ms;
    int rows, waves;
  };
  std::vector<Timed> timed;
  for (const auto& c : cand) {
    const int tr = c.first, wv = c.second;
    h->wpb = wv;
    build_tiles(h, tr);
    h->tile_cap = h->ntiles;
    rc = upload_tiles(h);
    if (rc) return rc;
    // one warm-up attempt, then `steps` timed ones, launched as pcw_run
    // launches them: the whole list on one stream, or a batch's two lanes
    // at once on their two streams
    const bool lanes = h->split > 0;
    h->lane_cap[0] = h->split;
    h->lane_cap[1] = h->ntiles - h->split;
    h->lane_k = lanes ? h->htiles[h->split].x : h->n;
    for (int pass = 0; pass < 2; ++pass) {
      if (pass == 1) HIPCHK(hipEventRecord(h->ev0, h->stream));
      if (lanes) {
        HIPCHK(hipEventRecord(h->evf, h->stream));
        HIPCHK(hipStreamWaitEvent(h->stream2, h->evf, 0));
      }
      for (int s = 0; s < (pass == 0 ? 1 : steps); ++s) {
        rc = enqueue_fused(h, lanes ? 1 : 0);
        if (!rc && lanes) rc = enqueue_fused(h, 2);
        if (rc) return rc;
      }
      if (lanes) {
        HIPCHK(hipEventRecord(h->evj, h->stream2));
        HIPCHK(hipStreamWaitEvent(h->stream, h->evj, 0));
      }
    }
    HIPCHK(hipEventRecord(h->ev1, h->stream));
    HIPCHK(hipEventSynchronize(h->ev1));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, h->ev0, h->ev1));
    timed.push_back({ms, tr, wv});
    if (ms < best) {
      best = ms;
      br = tr;
      bw = wv;
    }
  }
  if (h->split > 0) {
    // A batch: the candidates are timed at the run's current state, where
    // small instances can make short and tall bands time alike; the large
    // instances that run on after them take fewer halo rows with tall ones
    // (the 512-instance ensemble: 0.412 s with 64-row tiles against 0.417 s
    // with 32-row ones, which timed alike at its start).  So the tallest
    // of the candidates within 1 % of the fastest gives the height, and the
    // fastest of that height is taken.
    int tallest = br;
    for (const auto& c : timed)
      if (c.ms <= 1.01 * best) tallest = std::max(tallest, c.rows);
    float fastest = 1e30f;
    for (const auto& c : timed)
      if (c.rows == tallest && c.ms < fastest) {
        fastest = c.ms;
        br = c.rows;
        bw = c.waves;
      }
  }
  h->tile_rows = br;
  h->wpb = bw;
  build_tiles(h, br);
  h->tile_cap = h->ntiles;
  h->tiles_subset = false;
  rc = upload_tiles(h);
  if (rc) return rc;
  h->hst = saved;
  for (auto& S : h->hst) {
    S.smax_x = S.smax_y = 0ull;
    S.tiles_done = 0u;
  }
  rc = push_states(h);
  if (rc) return rc;
  if (tile_rows) *tile_rows = br;
  if (waves_per_block) *waves_per_block = bw;
  return PCW_OK;
}

}  // extern "C"
