// Copy-kernel bandwidth: the measured HBM ceiling reported beside the
// 8 TB/s specification, and the known-size kernel that calibrates the
// FETCH_SIZE / WRITE_SIZE counters of a profiled run (k_copy reads and
// writes exactly `bytes` each).
#include <hip/hip_runtime.h>

#include "sbfs.h"
#include "sbfs_internal.hpp"

namespace sbfs {

// one 16-byte element per thread, non-temporal loads and stores: the
// fastest plain copy pattern measured on the device (more waves in flight
// beat loops over several elements per thread)
__global__ __launch_bounds__(256) void k_copy(const double2* __restrict__ src,
                                              double2* __restrict__ dst, size_t n) {
  const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const double x = __builtin_nontemporal_load(&src[k].x), y = __builtin_nontemporal_load(&src[k].y);
  __builtin_nontemporal_store(x, &dst[k].x);
  __builtin_nontemporal_store(y, &dst[k].y);
}

}  // namespace sbfs

using namespace sbfs;

extern "C" int sbfs_copy_bandwidth(size_t bytes, int reps, double* gbs) {
  if (init_runtime()) return 1;
  SBFS_REQUIRE(gbs != nullptr, "gbs is null");
  SBFS_REQUIRE(reps >= 1, "reps must be >= 1");
  SBFS_REQUIRE(bytes >= 16 && bytes % 16 == 0, "bytes must be a positive multiple of 16");
  const size_t n = bytes / 16;
  double2 *src = nullptr, *dst = nullptr;
  SBFS_HIP(hipMalloc(&src, bytes));
  SBFS_HIP(hipMalloc(&dst, bytes));
  int rc = 0;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  do {
    if (hipMemsetAsync(src, 0, bytes, stream()) != hipSuccess) { rc = 1; break; }
    const size_t blocks = (n + 255) / 256;
    hipLaunchKernelGGL(k_copy, dim3((unsigned)blocks), dim3(256), 0, stream(), src, dst, n);
    if (check_launch("copy warm-up")) { rc = 1; break; }
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) { rc = 1; break; }
    (void)hipEventRecord(e0, stream());
    for (int r = 0; r < reps; r++)
      SBFS_TIMED("k_copy", hipLaunchKernelGGL(k_copy, dim3((unsigned)blocks), dim3(256), 0, stream(), src, dst, n));
    (void)hipEventRecord(e1, stream());
    if (hipEventSynchronize(e1) != hipSuccess) { rc = 1; break; }
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    *gbs = 2.0 * (double)bytes * reps / ((double)ms * 1e-3) / 1e9;
  } while (0);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipFree(src);
  (void)hipFree(dst);
  if (rc) set_error("sbfs_copy_bandwidth: HIP failure");
  return rc;
}
