// pcw_stream.hip -- the attainable-bandwidth reference for the step kernels:
// a stream of three planes in and three planes out (the state access of one
// step, without stencil, halos or arithmetic), 16-byte loads and stores with
// the non-temporal cache policy, one 16-byte element per thread.  Its rate is
// what HBM sustains for the step's compulsory traffic on this GPU.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "../../include/pcw.h"

namespace pcw {
namespace {

typedef float v4f __attribute__((ext_vector_type(4)));

__global__ void __launch_bounds__(256) k_stream3(const v4f* __restrict__ a0, const v4f* __restrict__ a1,
                                                 const v4f* __restrict__ a2, v4f* __restrict__ b0,
                                                 v4f* __restrict__ b1, v4f* __restrict__ b2,
                                                 long long n4) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n4) return;
  const v4f x0 = __builtin_nontemporal_load(a0 + i);
  const v4f x1 = __builtin_nontemporal_load(a1 + i);
  const v4f x2 = __builtin_nontemporal_load(a2 + i);
  __builtin_nontemporal_store(x0, b0 + i);
  __builtin_nontemporal_store(x1, b1 + i);
  __builtin_nontemporal_store(x2, b2 + i);
}

}  // namespace
}  // namespace pcw

extern "C" int pcw_stream_bandwidth(int64_t plane_bytes, int32_t reps, double* gbs_best,
                                    double* gbs_median) {
  using namespace pcw;
  if (plane_bytes < 16 || reps < 1 || !gbs_best || !gbs_median) return PCW_ERR_ARG;
  const long long n4 = plane_bytes / 16;
  const size_t bytes = (size_t)n4 * 16;
  v4f* p[6] = {};
  hipEvent_t e0 = nullptr, e1 = nullptr;
  std::vector<float> ms;
  int rc = PCW_OK;
  for (int k = 0; k < 6 && rc == PCW_OK; ++k) {
    if (hipMalloc(&p[k], bytes) != hipSuccess || hipMemset(p[k], 0, bytes) != hipSuccess) rc = PCW_ERR_HIP;
  }
  if (rc == PCW_OK && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
    rc = PCW_ERR_HIP;
  const unsigned blocks = (unsigned)((n4 + 255) / 256);
  for (int r = 0; r < reps + 2 && rc == PCW_OK; ++r) {  // two untimed warm-up launches
    float t = 0.f;
    if (hipEventRecord(e0, nullptr) != hipSuccess) { rc = PCW_ERR_HIP; break; }
    k_stream3<<<blocks, 256, 0, nullptr>>>(p[0], p[1], p[2], p[3], p[4], p[5], n4);
    if (hipGetLastError() != hipSuccess || hipEventRecord(e1, nullptr) != hipSuccess ||
        hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&t, e0, e1) != hipSuccess) {
      rc = PCW_ERR_HIP;
      break;
    }
    if (r >= 2) ms.push_back(t);
  }
  if (rc == PCW_OK) {
    std::sort(ms.begin(), ms.end());
    const double moved = 6.0 * (double)bytes;
    *gbs_best = moved / (ms.front() * 1e-3) / 1e9;
    *gbs_median = moved / (ms[ms.size() / 2] * 1e-3) / 1e9;
  }
  (void)hipDeviceSynchronize();
  for (auto* q : p)
    if (q) (void)hipFree(q);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  return rc;
}
