// pcw_staged.hip -- the plain FP64 path: one kernel per reference stage with
// HBM work arrays, in place on the current buffer after a backup copy
// (src/main.f90:201 solution_old = solution).  Also the standalone ghost
// fill, source, copy and totals kernels shared by both engines.
//
// Stage kernels (src/solver/MOD_solve.f90:71-157 and callees):
//   k_st_waves      wave_decomposition / HLL + CFL max   (:72-86, :126-139)
//   k_st_split      split_flux                           (MOD_flux_calculation.f90:82-121)
//   k_st_integrate  integrate_flux_difference            (MOD_solve.f90:177-238)
//   k_st_limit      limiter (unlimited dot products, into separate arrays)
//                                                        (MOD_flux_correction.f90:17-125)
//   k_st_second     second_order_flux_correction         (MOD_flux_correction.f90:127-187)
//   k_st_increment  integrator_flux_correction_increment (MOD_solve.f90:240-293)
#include "pcw_device.hpp"
#include "pcw_engine.hpp"
#include "pcw_kernels.hpp"

namespace pcw {

__device__ __forceinline__ long long fidx(int pitch, int i, int j) {
  return (long long)(j + 1) * pitch + (i + 1);
}
// element of frame cell (i, j) in either row layout (the kernels shared with
// the FP32 engine: ghosts, source, totals)
__device__ __forceinline__ long long fidx(const KInst& I, int i, int j) {
  return (long long)(j + 1) * I.pitch + col_pos(I, i);
}

// metrics of cell (i, j): metric planes of a mapped grid, else the identity map's
__device__ __forceinline__ double st_kappa(const StagedArgs& a, const KInst& I, int i, int j) {
  if (a.m) return a.m[fidx(I.pitch, i, j)];
  Mesh M{I.xi_min, I.eta_min, I.dxi, I.deta};
  return kappa_exact(M, i, j);
}
__device__ __forceinline__ double st_gamma(const StagedArgs& a, const KInst& I, int dir, int i,
                                           int j) {
  if (a.m) return a.m[(dir == 1 ? 1 : 2) * I.plane + fidx(I.pitch, i, j)];
  Mesh M{I.xi_min, I.eta_min, I.dxi, I.deta};
  return dir == 1 ? gamma_xi_exact(M, j) : gamma_eta_exact(M, i);
}

// interface (i, j) of sweep DIR, threads over the sweep's interface range
template <int DIR>
__device__ __forceinline__ bool iface_of(const KInst& I, int& i, int& j) {
  const int tx = blockIdx.x * blockDim.x + threadIdx.x;
  const int ty = blockIdx.y * blockDim.y + threadIdx.y;
  if (DIR == 1) { i = tx; j = ty + 1; return i <= I.nx + 1 && j <= I.ny; }
  i = tx + 1; j = ty;
  return i <= I.nx && j <= I.ny + 1;
}

__global__ void k_st_waves(StagedArgs a, int dir) {
  const KInst& I = *a.inst;
  InstState* S = a.st;
  if (S->status != PCW_STATUS_RUNNING) return;
  int i, j;
  bool ok = (dir == 1) ? iface_of<1>(I, i, j) : iface_of<2>(I, i, j);
  const double* q = a.q;
  const long long pl = I.plane;
  double s2 = 0.0;
  int bad = 0;
  if (ok) {
    const int p = I.pitch;
    long long kR = fidx(p, i, j), kL = (dir == 1) ? fidx(p, i - 1, j) : fidx(p, i, j - 1);
    double qL[3] = {q[kL], q[pl + kL], q[2 * pl + kL]};
    double qR[3] = {q[kR], q[pl + kR], q[2 * pl + kR]};
    Phys P{a.g, a.eps, a.dry};
    CellQ<double> cL = (dir == 1) ? cell_exact<1>(qL[0], qL[1], qL[2], P) : cell_exact<2>(qL[0], qL[1], qL[2], P);
    CellQ<double> cR = (dir == 1) ? cell_exact<1>(qR[0], qR[1], qR[2], P) : cell_exact<2>(qR[0], qR[1], qR[2], P);
    bad = cL.bad | cR.bad;
    Waves<double> w = hll_exact(qL, cL, qR, cR, a.eps);
    a.w[W_LAM * pl + kR] = w.s2;
    for (int k = 0; k < 3; ++k) {
      a.w[(W_W1 + k) * pl + kR] = w.w1[k];
      a.w[(W_W2 + k) * pl + kR] = w.w2[k];
    }
    s2 = w.s2;
  }
  const double m = wave_max(s2);
  const int anybad = __any(bad);
  if ((threadIdx.x & 63) == 0 && (threadIdx.y * blockDim.x + threadIdx.x) % 64 == 0) {
    atomicMax(dir == 1 ? &S->smax_x : &S->smax_y, (unsigned long long)__double_as_longlong(m));
    if (anybad) atomicOr(&S->bad, dir == 1 ? 1 : 2);
  }
}

// split_flux: A-dQ += min(lambda,0) W, A+dQ += max(lambda,0) W (mask form)
__global__ void k_st_split(StagedArgs a, int dir) {
  const KInst& I = *a.inst;
  if (a.st->status != PCW_STATUS_RUNNING) return;
  int i, j;
  bool ok = (dir == 1) ? iface_of<1>(I, i, j) : iface_of<2>(I, i, j);
  if (!ok) return;
  const long long pl = I.plane, k = fidx(I.pitch, i, j);
  const double s2 = a.w[W_LAM * pl + k];
  const double lam[2] = {-s2, s2};
  for (int me = 0; me < 3; ++me) {
    double am = 0.0, ap = 0.0;
    for (int mw = 0; mw < 2; ++mw) {
      double s = lam[mw];
      double wv = a.w[((mw == 0 ? W_W1 : W_W2) + me) * pl + k];
      double mask = (s < 0.0) ? 1.0 : 0.0;
      am = am + mask * s * wv;
      ap = ap + (1.0 - mask) * s * wv;
    }
    a.w[(W_AM + me) * pl + k] = am;
    a.w[(W_AP + me) * pl + k] = ap;
  }
}

__device__ __forceinline__ bool cell_of(const KInst& I, int& i, int& j) {
  i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  j = blockIdx.y * blockDim.y + threadIdx.y + 1;
  return i <= I.nx && j <= I.ny;
}

__global__ void k_st_integrate(StagedArgs a, int dir) {
  const KInst& I = *a.inst;
  InstState* S = a.st;
  if (S->status != PCW_STATUS_RUNNING) return;
  int i, j;
  if (!cell_of(I, i, j)) return;
  const double dt = S->tc.dt;
  const double dtod = dt / (dir == 1 ? I.dxi : I.deta);
  const double kap = st_kappa(a, I, i, j);
  double gl, gr;
  if (dir == 1) { gl = st_gamma(a, I, 1, i, j); gr = st_gamma(a, I, 1, i + 1, j); }
  else { gl = st_gamma(a, I, 2, i, j); gr = st_gamma(a, I, 2, i, j + 1); }
  const long long pl = I.plane, k = fidx(I.pitch, i, j);
  const long long kr = (dir == 1) ? fidx(I.pitch, i + 1, j) : fidx(I.pitch, i, j + 1);
  for (int me = 0; me < 3; ++me) {
    double q = a.q[me * pl + k];
    a.q[me * pl + k] = q - dtod / kap * (gl * a.w[(W_AP + me) * pl + k] + gr * a.w[(W_AM + me) * pl + kr]);
  }
}

template <int LIM>
__device__ void st_limit_body(StagedArgs& a, int dir) {
  const KInst& I = *a.inst;
  int i, j;
  // range of second_order_flux_correction
  bool ok = (dir == 1) ? iface_of<1>(I, i, j) : iface_of<2>(I, i, j);
  if (!ok) return;
  const int p = I.pitch;
  const long long pl = I.plane, k = fidx(p, i, j);
  const long long kn = (dir == 1) ? fidx(p, i + 1, j) : fidx(p, i, j + 1);
  const long long kp = (dir == 1) ? fidx(p, i - 1, j) : fidx(p, i, j - 1);
  double w1[3], w2[3];
  for (int me = 0; me < 3; ++me) {
    w1[me] = a.w[(W_W1 + me) * pl + k];
    w2[me] = a.w[(W_W2 + me) * pl + k];
  }
  double p1 = 1.0, p2 = 1.0;
  const bool limited = (dir == 1) ? (i >= 1 && i <= I.nx) : (j >= 1 && j <= I.ny);
  if (LIM > 0 && limited) {
    double n1[3], n2[3];
    for (int me = 0; me < 3; ++me) {
      n1[me] = a.w[(W_W1 + me) * pl + kn];
      n2[me] = a.w[(W_W2 + me) * pl + kp];
    }
    p1 = limit_exact<LIM>(w1, n1, a.eps);
    p2 = limit_exact<LIM>(w2, n2, a.eps);
  }
  for (int me = 0; me < 3; ++me) {
    a.w[(W_T1 + me) * pl + k] = p1 * w1[me];
    a.w[(W_T2 + me) * pl + k] = p2 * w2[me];
  }
}

__global__ void k_st_limit(StagedArgs a, int dir) {
  if (a.st->status != PCW_STATUS_RUNNING) return;
  switch (a.limiter) {
    case 0: st_limit_body<0>(a, dir); break;
    case 1: st_limit_body<1>(a, dir); break;
    case 2: st_limit_body<2>(a, dir); break;
    case 3: st_limit_body<3>(a, dir); break;
    default: st_limit_body<4>(a, dir); break;
  }
}

__global__ void k_st_second(StagedArgs a, int dir) {
  const KInst& I = *a.inst;
  InstState* S = a.st;
  if (S->status != PCW_STATUS_RUNNING) return;
  int i, j;
  bool ok = (dir == 1) ? iface_of<1>(I, i, j) : iface_of<2>(I, i, j);
  if (!ok) return;
  const double dtod = S->tc.dt / (dir == 1 ? I.dxi : I.deta);
  const double ka = (dir == 1) ? st_kappa(a, I, i - 1, j) : st_kappa(a, I, i, j - 1);
  const double kb = st_kappa(a, I, i, j);
  const long long pl = I.plane, k = fidx(I.pitch, i, j);
  const double al = a.w[W_LAM * pl + k];
  const double coef = 0.5 * al * (1.0 - al * dtod / (0.5 * (ka + kb)));
  for (int me = 0; me < 3; ++me) {
    double t1 = coef * a.w[(W_T1 + me) * pl + k];
    double t2 = coef * a.w[(W_T2 + me) * pl + k];
    a.w[(W_F + me) * pl + k] = t1 + t2;
  }
}

__global__ void k_st_increment(StagedArgs a, int dir) {
  const KInst& I = *a.inst;
  InstState* S = a.st;
  if (S->status != PCW_STATUS_RUNNING) return;
  int i, j;
  if (!cell_of(I, i, j)) return;
  const double dtod = S->tc.dt / (dir == 1 ? I.dxi : I.deta);
  const double kap = st_kappa(a, I, i, j);
  const long long pl = I.plane, k = fidx(I.pitch, i, j);
  const long long kr = (dir == 1) ? fidx(I.pitch, i + 1, j) : fidx(I.pitch, i, j + 1);
  for (int me = 0; me < 3; ++me) {
    double q = a.q[me * pl + k];
    a.q[me * pl + k] = q - dtod / kap * (a.w[(W_F + me) * pl + kr] - a.w[(W_F + me) * pl + k]);
  }
}

// ------------------------------------------------------- ghost fill
// apply_physical_boundaries (SMOD_boundary_conditions.f90:58-213) on the
// buffer buf_sel selects (kGhosts*) of every instance; one thread per ghost
// cell of the four sides.  Ghost cells are computed from interior cells only.
template <typename T>
__global__ void k_ghosts(T* pool, const KInst* inst, InstState* st, int buf_sel) {
  const KInst& I = inst[blockIdx.y];
  InstState* S = st + blockIdx.y;
  const int nx = I.nx, ny = I.ny;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int nlr = 2 * ny, nbt = 2 * nx;
  if (t >= 2 * nlr + 2 * nbt) return;
  int b = S->cur;
  if (buf_sel == kGhostsAfterFused) {
    if (S->controlled) {
      if (!S->ghosts_due) return;      // rejected, or no attempt: the current state stays as it is
    } else {
      if (S->status != PCW_STATUS_RUNNING) return;
      b ^= 1;                          // the pending state of pcw_step
    }
  }
  T* q = pool + I.off[b];
  const long long pl = I.plane;
  int side, pos, kk;
  if (t < 2 * nlr) { side = t / nlr; pos = (t % nlr) / 2 + 1; kk = (t % 2) + 1; }
  else { int u = t - 2 * nlr; side = 2 + u / nbt; pos = (u % nbt) / 2 + 1; kk = (u % 2) + 1; }
  Mesh M{I.xi_min, I.eta_min, I.dxi, I.deta};
  long long kn, kf, kg;
  double xn, xf, xg;
  int normal;
  if (side == 0) {
    kn = fidx(I, 1, pos); kf = fidx(I, 2, pos); kg = fidx(I, 1 - kk, pos);
    xn = centre(M.xi_min, M.dxi, 1); xf = centre(M.xi_min, M.dxi, 2); xg = centre(M.xi_min, M.dxi, 1 - kk);
    normal = 1;
  } else if (side == 1) {
    kn = fidx(I, nx, pos); kf = fidx(I, nx - 1, pos); kg = fidx(I, nx + kk, pos);
    xn = centre(M.xi_min, M.dxi, nx); xf = centre(M.xi_min, M.dxi, nx - 1); xg = centre(M.xi_min, M.dxi, nx + kk);
    normal = 1;
  } else if (side == 2) {
    kn = fidx(I, pos, 1); kf = fidx(I, pos, 2); kg = fidx(I, pos, 1 - kk);
    xn = centre(M.eta_min, M.deta, 1); xf = centre(M.eta_min, M.deta, 2); xg = centre(M.eta_min, M.deta, 1 - kk);
    normal = 2;
  } else {
    kn = fidx(I, pos, ny); kf = fidx(I, pos, ny - 1); kg = fidx(I, pos, ny + kk);
    xn = centre(M.eta_min, M.deta, ny); xf = centre(M.eta_min, M.deta, ny - 1); xg = centre(M.eta_min, M.deta, ny + kk);
    normal = 2;
  }
  const int bc = I.bc[side];
  for (int me = 0; me < 3; ++me)
    q[me * pl + kg] = ghost_value<T>(bc, q[me * pl + kn], q[me * pl + kf], kk, me == normal, xn, xf, xg);
}

// integrate_source_term with the built-in source kinds on buffer cur
template <typename T>
__global__ void k_source(T* pool, const T* bpool, const KInst* inst, InstState* st, int wetdry,
                         double g, double dry) {
  const KInst& I = inst[blockIdx.z];
  InstState* S = st + blockIdx.z;
  if (S->status != PCW_STATUS_RUNNING) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
  if (i > I.nx || j > I.ny) return;
  T* q = pool + I.off[S->cur];
  const long long pl = I.plane, k = fidx(I, i, j);
  if (I.source == PCW_SRC_NONE && !wetdry) {  // the reference's q + 0*dt
    T h = q[k], hu = q[pl + k], hv = q[2 * pl + k];
    reference_source<T>(h, hu, hv);
    q[k] = h;
    q[pl + k] = hu;
    q[2 * pl + k] = hv;
    return;
  }
  T bxm = 0, bxp = 0, bym = 0, byp = 0;
  if (I.boff >= 0) {
    const T* b = bpool + I.boff;
    bxm = b[fidx(I, i - 1, j)]; bxp = b[fidx(I, i + 1, j)];
    bym = b[fidx(I, i, j - 1)]; byp = b[fidx(I, i, j + 1)];
  }
  T h = q[k], hu = q[pl + k], hv = q[2 * pl + k];
  apply_source<T>(h, hu, hv, I.source, (T)dry, (T)g, (T)I.dxi, (T)I.deta, (T)S->tc.dt,
                  (T)(I.manning * I.manning), bxm, bxp, bym, byp);
  q[k] = h;
  q[pl + k] = hu;
  q[2 * pl + k] = hv;
}

// copy the three planes of one buffer to the other (backup / restore)
template <typename T>
__global__ void k_copy_frame(T* pool, const KInst* inst, InstState* st, int to_backup,
                             int only_if_rejected) {
  const KInst& I = inst[blockIdx.y];
  InstState* S = st + blockIdx.y;
  if (only_if_rejected) {
    if (S->last_ok) return;
  } else if (S->status != PCW_STATUS_RUNNING) {
    return;
  }
  const int c = S->cur;
  const T* src = pool + I.off[to_backup ? c : c ^ 1];
  T* dst = pool + I.off[to_backup ? c ^ 1 : c];
  const long long n = 3 * I.plane;
  for (long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x; k < n;
       k += (long long)gridDim.x * blockDim.x)
    dst[k] = src[k];
}

// end of a staged attempt: publish the sweep maxima, run the controller
// when device-controlled, and mark acceptance for the restore kernel
__global__ void k_st_finish(const KInst* inst, InstState* st, int n) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const KInst& I = inst[k];
  InstState* S = st + k;
  if (S->status != PCW_STATUS_RUNNING) { S->last_ok = 1; return; }
  const double sx = __longlong_as_double((long long)S->smax_x);
  const double sy = __longlong_as_double((long long)S->smax_y);
  S->smax_x = 0ull;
  S->smax_y = 0ull;
  S->last_smax_x = sx;
  S->last_smax_y = sy;
  S->last_ok = 1;
  const int badbits = S->bad;
  S->last_bad = badbits;
  S->bad = 0;
  if (!S->controlled) return;
  pcw_step_record rec;
  int ok;
  if (finish_checks(S, sx, sy, I.dxi, I.deta, badbits, &rec, &ok)) return;
  S->cfl_last = rec.cfl;
  S->last_ok = ok;
  if (ok) S->accepted += 1;
  else S->rejected += 1;
  after_attempt(S, rec);
}

// deterministic FP64 sums over the interior of buffer cur: pass 1
template <typename T>
__global__ void k_totals_partial(const T* pool, const KInst* inst, const InstState* st,
                                 double* part) {
  const KInst& I = *inst;
  const T* q = pool + I.off[st->cur];
  const long long pl = I.plane;
  double s[3] = {0.0, 0.0, 0.0};
  for (int j = 1 + blockIdx.x; j <= I.ny; j += gridDim.x)
    for (int i = 1 + threadIdx.x; i <= I.nx; i += blockDim.x) {
      long long k = fidx(I, i, j);
      s[0] += (double)q[k];
      s[1] += (double)q[pl + k];
      s[2] += (double)q[2 * pl + k];
    }
  __shared__ double red[3][256];
  for (int c = 0; c < 3; ++c) red[c][threadIdx.x] = s[c];
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o)
      for (int c = 0; c < 3; ++c) red[c][threadIdx.x] += red[c][threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int c = 0; c < 3; ++c) part[3 * blockIdx.x + c] = red[c][0];
}

__global__ void k_totals_final(const double* part, int nparts, double* out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double s[3] = {0.0, 0.0, 0.0};
  for (int b = 0; b < nparts; ++b)
    for (int c = 0; c < 3; ++c) s[c] += part[3 * b + c];
  for (int c = 0; c < 3; ++c) out[c] = s[c];
}

__global__ void k_set_dt(InstState* st, const double* dt, int n, int controlled) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  st[k].tc.dt = dt[k];
  st[k].controlled = controlled;
}

// pcw_accept of the fused engines: the step kernels wrote the pending buffer
// of the running instances only, so only those switch to it
__global__ void k_flip(InstState* st, int n) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  if (st[k].status == PCW_STATUS_RUNNING) st[k].cur ^= 1;
}

// ------------------------------------------------------------- launchers
static dim3 iface_grid(const KInst& I, int dir, dim3 blk) {
  int nxr = (dir == 1) ? I.nx + 2 : I.nx;
  int nyr = (dir == 1) ? I.ny : I.ny + 2;
  return dim3((nxr + blk.x - 1) / blk.x, (nyr + blk.y - 1) / blk.y);
}
static dim3 cell_grid(const KInst& I, dim3 blk) {
  return dim3((I.nx + blk.x - 1) / blk.x, (I.ny + blk.y - 1) / blk.y);
}

hipError_t launch_staged_sweep(const StagedArgs& a, const KInst& hI, int dir, int ho,
                               hipStream_t s) {
  const dim3 blk(64, 4);
  const dim3 gi = iface_grid(hI, dir, blk), gc = cell_grid(hI, blk);
  k_st_waves<<<gi, blk, 0, s>>>(a, dir);
  k_st_split<<<gi, blk, 0, s>>>(a, dir);
  k_st_integrate<<<gc, blk, 0, s>>>(a, dir);
  if (ho) {
    k_st_limit<<<gi, blk, 0, s>>>(a, dir);
    k_st_second<<<gi, blk, 0, s>>>(a, dir);
    k_st_increment<<<gc, blk, 0, s>>>(a, dir);
  }
  return hipGetLastError();
}

template <typename T>
hipError_t launch_ghosts(T* pool, const KInst* dinst, InstState* dst, int n, int max_side_cells,
                         int buf_sel, hipStream_t s) {
  const int tot = 4 * max_side_cells;
  k_ghosts<T><<<dim3((tot + 255) / 256, n), 256, 0, s>>>(pool, dinst, dst, buf_sel);
  return hipGetLastError();
}
template hipError_t launch_ghosts<float>(float*, const KInst*, InstState*, int, int, int, hipStream_t);
template hipError_t launch_ghosts<double>(double*, const KInst*, InstState*, int, int, int, hipStream_t);

template <typename T>
hipError_t launch_source(T* pool, const T* bpool, const KInst* dinst, InstState* dst, int n,
                         int max_nx, int max_ny, int wetdry, double g, double dry, hipStream_t s) {
  const dim3 blk(64, 4);
  k_source<T><<<dim3((max_nx + 63) / 64, (max_ny + 3) / 4, n), blk, 0, s>>>(pool, bpool, dinst,
                                                                              dst, wetdry, g, dry);
  return hipGetLastError();
}
template hipError_t launch_source<float>(float*, const float*, const KInst*, InstState*, int, int, int, int, double, double, hipStream_t);
template hipError_t launch_source<double>(double*, const double*, const KInst*, InstState*, int, int, int, int, double, double, hipStream_t);

template <typename T>
hipError_t launch_copy(T* pool, const KInst* dinst, InstState* dst, int n, int to_backup,
                       int only_if_rejected, hipStream_t s) {
  k_copy_frame<T><<<dim3(1024, n), 256, 0, s>>>(pool, dinst, dst, to_backup, only_if_rejected);
  return hipGetLastError();
}
template hipError_t launch_copy<float>(float*, const KInst*, InstState*, int, int, int, hipStream_t);
template hipError_t launch_copy<double>(double*, const KInst*, InstState*, int, int, int, hipStream_t);

hipError_t launch_st_finish(const KInst* dinst, InstState* dst, int n, hipStream_t s) {
  k_st_finish<<<(n + 63) / 64, 64, 0, s>>>(dinst, dst, n);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_totals(const T* pool, const KInst* dinst, const InstState* dst, double* part,
                         int nparts, double* out, hipStream_t s) {
  k_totals_partial<T><<<nparts, 256, 0, s>>>(pool, dinst, dst, part);
  k_totals_final<<<1, 64, 0, s>>>(part, nparts, out);
  return hipGetLastError();
}
template hipError_t launch_totals<float>(const float*, const KInst*, const InstState*, double*, int, double*, hipStream_t);
template hipError_t launch_totals<double>(const double*, const KInst*, const InstState*, double*, int, double*, hipStream_t);

hipError_t launch_set_dt(InstState* dst, const double* ddt, int n, int controlled, hipStream_t s) {
  k_set_dt<<<(n + 63) / 64, 64, 0, s>>>(dst, ddt, n, controlled);
  return hipGetLastError();
}
hipError_t launch_flip(InstState* dst, int n, hipStream_t s) {
  k_flip<<<(n + 63) / 64, 64, 0, s>>>(dst, n);
  return hipGetLastError();
}

}  // namespace pcw
