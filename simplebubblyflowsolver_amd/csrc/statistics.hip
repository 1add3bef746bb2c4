// Running sums of ux, uy and vof for temporal means (src/statistics.c):
// the accumulators stay in device memory and are added to in place, so a
// collection costs one pass over three fields and no host transfer; they
// are copied out only when written.  One set per process, as in the
// reference.
#include <hip/hip_runtime.h>

#include <vector>

#include "sbfs.h"
#include "sbfs_internal.hpp"

namespace sbfs {

// acc[(j-1) * w + c] += src[(j - 1 + yoff) * w + c] for rows j = 1..ny
__global__ void k_accumulate(double* __restrict__ acc, const double* __restrict__ src, int w, int ny,
                             int yoff) {
  const long n = (long)w * ny;
  for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < n; k += (long)gridDim.x * blockDim.x)
    acc[k] += src[k + (long)yoff * w];
}

struct Stats {
  size_t nx = 0, ny = 0;
  size_t num = 0;
  double* ux1 = nullptr;   // [1 : nx+1] x [1 : ny]   (statistics/ux1.h)
  double* uy1 = nullptr;   // [0 : nx+1] x [1 : ny]
  double* vof1 = nullptr;  // [0 : nx+1] x [1 : ny]
};
static Stats g_stats;

static int stats_prepare(const domain_t* d) {
  const size_t nx = d->mysizes[0], ny = d->mysizes[1];
  if (g_stats.ux1 && g_stats.nx == nx && g_stats.ny == ny) return 0;
  (void)hipFree(g_stats.ux1);
  (void)hipFree(g_stats.uy1);
  (void)hipFree(g_stats.vof1);
  g_stats = Stats();
  SBFS_HIP(hipMalloc(&g_stats.ux1, sizeof(double) * (nx + 1) * ny));
  SBFS_HIP(hipMalloc(&g_stats.uy1, sizeof(double) * (nx + 2) * ny));
  SBFS_HIP(hipMalloc(&g_stats.vof1, sizeof(double) * (nx + 2) * ny));
  SBFS_HIP(hipMemsetAsync(g_stats.ux1, 0, sizeof(double) * (nx + 1) * ny, stream()));
  SBFS_HIP(hipMemsetAsync(g_stats.uy1, 0, sizeof(double) * (nx + 2) * ny, stream()));
  SBFS_HIP(hipMemsetAsync(g_stats.vof1, 0, sizeof(double) * (nx + 2) * ny, stream()));
  g_stats.nx = nx;
  g_stats.ny = ny;
  return 0;
}

}  // namespace sbfs

using namespace sbfs;

extern "C" int sbfs_statistics_collect(const domain_t* domain, const fluid_t* fluid,
                                       const interface_t* interface) {
  SBFS_REQUIRE(domain && fluid && interface, "sbfs_statistics_collect: null argument");
  if (init_runtime()) return 1;
  if (check_array(domain, &fluid->ux, NA_UX, 8, "ux")) return 1;
  if (check_array(domain, &fluid->uy, NA_UY, 8, "uy")) return 1;
  if (check_array(domain, &interface->vof, NA_VOF, 8, "vof")) return 1;
  if (stats_prepare(domain)) return 1;
  const int nx = (int)g_stats.nx, ny = (int)g_stats.ny;
  const dim3 blk(256), grd(2048);
  hipLaunchKernelGGL(k_accumulate, grd, blk, 0, stream(), g_stats.ux1, (const double*)fluid->ux.data,
                     nx + 1, ny, 1);
  hipLaunchKernelGGL(k_accumulate, grd, blk, 0, stream(), g_stats.uy1, (const double*)fluid->uy.data,
                     nx + 2, ny, 1);
  hipLaunchKernelGGL(k_accumulate, grd, blk, 0, stream(), g_stats.vof1,
                     (const double*)interface->vof.data, nx + 2, ny, 2);
  if (check_launch("sbfs_statistics_collect")) return 1;
  g_stats.num += 1;
  return 0;
}

extern "C" int sbfs_statistics_num(size_t* num) {
  SBFS_REQUIRE(num, "sbfs_statistics_num: null argument");
  *num = g_stats.num;
  return 0;
}

extern "C" int sbfs_statistics_save(const char dirname[], const domain_t* domain) {
  SBFS_REQUIRE(dirname && domain, "sbfs_statistics_save: null argument");
  SBFS_REQUIRE(g_stats.ux1 && g_stats.nx == domain->mysizes[0] && g_stats.ny == domain->mysizes[1],
               "sbfs_statistics_save: nothing collected for this domain");
  const size_t nx = g_stats.nx, ny = g_stats.ny;
  if (npy_write(dirname, "num", "<u8", {}, sizeof(size_t), &g_stats.num)) return 1;
  struct {
    const char* name;
    const double* dev;
    size_t w;
  } vars[] = {{"ux1", g_stats.ux1, nx + 1}, {"uy1", g_stats.uy1, nx + 2}, {"vof1", g_stats.vof1, nx + 2}};
  std::vector<double> buf;
  for (const auto& v : vars) {
    buf.resize(v.w * ny);
    SBFS_HIP(hipMemcpyAsync(buf.data(), v.dev, sizeof(double) * buf.size(), hipMemcpyDeviceToHost, stream()));
    SBFS_HIP(hipStreamSynchronize(stream()));
    if (npy_write(dirname, v.name, "<f8", {ny, v.w}, sizeof(double), buf.data())) return 1;
  }
  return 0;
}
