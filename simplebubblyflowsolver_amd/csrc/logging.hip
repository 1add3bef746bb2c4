// Every quantity the reference logs (src/logging/*.c), plus the bubble's
// volume, centroid and rise velocity, as deterministic FP64 reductions on
// the device: one pass over the x faces x rows, per-workgroup partials, then
// one workgroup combines them in a fixed order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "reduce.hpp"
#include "sbfs.h"
#include "sbfs_internal.hpp"

namespace sbfs {

// quantity slots of the partials; kind: 0 min, 1 max, 2 sum
enum {
  Q_DIVMAX, Q_MOMX, Q_MOMY, Q_EXTX, Q_EXTY, Q_ENGX, Q_ENGY, Q_STRM, Q_STRP,
  Q_VMIN, Q_VMAX, Q_VOL, Q_AREA_DS, Q_AREA, Q_XVOL, Q_UVOL, Q_N
};
__constant__ int kQKind[Q_N] = {1, 2, 2, 1, 1, 2, 2, 2, 2, 0, 1, 2, 2, 2, 2, 2};

struct LogArgs {
  const double* vof;
  const double* ux;
  const double* uy;
  const double* den;  // den[1]
  const double* txy;
  const double* xc;
  const double* hxxf;
  const double* jdxf;
  const double* jdxc;
  double hy, diffusivity;
  int nx, ny;
};

__global__ __launch_bounds__(kRedThreads) void k_log_partial(LogArgs a, double* __restrict__ part) {
  __shared__ double sh[kRedThreads / 64];
  const int nx = a.nx;
  double q[Q_N];
  for (int k = 0; k < Q_N; k++) q[k] = 0.;
  q[Q_VMIN] = 1.;
  const long total = (long)(nx + 1) * a.ny;
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < total;
       idx += (long)gridDim.x * blockDim.x) {
    const int i = (int)(idx % (nx + 1)) + 1, j = (int)(idx / (nx + 1)) + 1;
    const double jf = a.jdxf[i - 1], hf = a.hxxf[i - 1];
    const double vof = a.vof[L::vof(i, j, nx)];
    const double vofm = a.vof[L::vof(i - 1, j, nx)];
    // surface area, x faces (interface.c:119-127)
    q[Q_AREA] += fabs(vof - vofm) * (jf / hf);
    // wall shear stress (stress.c:30-36)
    if (i == 1) q[Q_STRM] += a.diffusivity * jf / hf * a.txy[L::xf(1, j, nx)];
    if (i == nx + 1) {
      q[Q_STRP] += a.diffusivity * jf / hf * a.txy[L::xf(nx + 1, j, nx)];
      continue;
    }
    if (i >= 2) {
      // x momentum, energy, extremum (momentum.c:33-39, energy.c:33-39, extrema.c:27-31)
      const double lden = 0.5 * a.den[L::cc(i - 1, j, nx)] + 0.5 * a.den[L::cc(i, j, nx)];
      const double lvel = a.ux[L::xf(i, j, nx)];
      q[Q_MOMX] += lden * lvel * jf;
      q[Q_ENGX] += 0.5 * lden * (lvel * lvel) * jf;
      q[Q_EXTX] = fmax(q[Q_EXTX], fabs(lvel));
    }
    const double jc = a.jdxc[i - 1];
    {
      // y momentum, energy, extremum
      const double lden = 0.5 * a.den[L::cc(i, j - 1, nx)] + 0.5 * a.den[L::cc(i, j, nx)];
      const double lvel = a.uy[L::cc(i, j, nx)];
      q[Q_MOMY] += lden * lvel * jc;
      q[Q_ENGY] += 0.5 * lden * (lvel * lvel) * jc;
      q[Q_EXTY] = fmax(q[Q_EXTY], fabs(lvel));
    }
    // vof extrema, volume, y-face area (interface.c:104-135)
    q[Q_VMIN] = fmin(q[Q_VMIN], vof);
    q[Q_VMAX] = fmax(q[Q_VMAX], vof);
    q[Q_VOL] += vof * jc;
    q[Q_AREA_DS] += jc;
    q[Q_AREA] += fabs(vof - a.vof[L::vof(i, j - 1, nx)]) * (jc / a.hy);
    // bubble centroid and rise velocity (+x is up)
    const double uxm = a.ux[L::xf(i, j, nx)], uxp = a.ux[L::xf(i + 1, j, nx)];
    q[Q_XVOL] += vof * a.xc[i] * jc;
    q[Q_UVOL] += vof * (0.5 * uxm + 0.5 * uxp) * jc;
    // divergence (divergence.c:44-58)
    const double div = 1. / jc *
                       (-jf / hf * uxm + a.jdxf[i] / a.hxxf[i] * uxp
                        - jc / a.hy * a.uy[L::cc(i, j, nx)] + jc / a.hy * a.uy[L::cc(i, j + 1, nx)]);
    q[Q_DIVMAX] = fmax(q[Q_DIVMAX], fabs(div));
  }
  double r[Q_N];
#pragma unroll
  for (int k = 0; k < Q_N; k++) {
    const int kind = kQKind[k];
    r[k] = kind == 0 ? block_reduce<0>(q[k], sh) : kind == 1 ? block_reduce<1>(q[k], sh)
                                                            : block_reduce<2>(q[k], sh);
  }
  if (threadIdx.x == 0)
    for (int k = 0; k < Q_N; k++) part[(size_t)Q_N * blockIdx.x + k] = r[k];
}

__global__ __launch_bounds__(kRedThreads) void k_log_final(const double* __restrict__ part, int n,
                                                           double* __restrict__ out) {
  __shared__ double sh[kRedThreads / 64];
  for (int k = 0; k < Q_N; k++) {
    const int kind = kQKind[k];
    double v = kind == 0 ? 1. : 0.;
    for (int b = threadIdx.x; b < n; b += blockDim.x) {
      const double x = part[(size_t)Q_N * b + k];
      v = kind == 0 ? fmin(v, x) : kind == 1 ? fmax(v, x) : v + x;
    }
    const double r = kind == 0 ? block_reduce<0>(v, sh) : kind == 1 ? block_reduce<1>(v, sh)
                                                                    : block_reduce<2>(v, sh);
    if (threadIdx.x == 0) out[k] = r;
  }
}

}  // namespace sbfs

using namespace sbfs;

extern "C" int sbfs_logging_values(const domain_t* domain, const fluid_t* fluid,
                                   const interface_t* interface, double out[SBFS_NLOG]) {
  SBFS_REQUIRE(domain && fluid && interface && out, "sbfs_logging_values: null argument");
  DomainCtx* c = domain_ctx(domain);
  if (!c) return 1;
  if (check_array(domain, &interface->vof, NA_VOF, 8, "vof")) return 1;
  if (check_array(domain, &fluid->ux, NA_UX, 8, "ux")) return 1;
  if (check_array(domain, &fluid->uy, NA_UY, 8, "uy")) return 1;
  if (check_array(domain, &fluid->den[1], NA_DEN, 8, "den[1]")) return 1;
  if (check_array(domain, &fluid->txy, NA_TXY, 8, "txy")) return 1;
  const int nx = (int)c->nx, ny = (int)c->ny;
  static double* buf = nullptr;  // partials + result, grown once
  constexpr int kBlocks = 1024;
  if (!buf) SBFS_HIP(hipMalloc(&buf, sizeof(double) * (size_t)Q_N * (kBlocks + 1)));
  LogArgs a;
  a.vof = (const double*)interface->vof.data;
  a.ux = (const double*)fluid->ux.data;
  a.uy = (const double*)fluid->uy.data;
  a.den = (const double*)fluid->den[1].data;
  a.txy = (const double*)fluid->txy.data;
  a.xc = c->xc;
  a.hxxf = c->hxxf;
  a.jdxf = c->jdxf;
  a.jdxc = c->jdxc;
  a.hy = c->hy;
  a.diffusivity = 1. / fluid->Re;
  a.nx = nx;
  a.ny = ny;
  const long total = (long)(nx + 1) * ny;
  const int nblk = (int)std::min<long>(kBlocks, (total + kRedThreads - 1) / kRedThreads);
  hipLaunchKernelGGL(k_log_partial, dim3(nblk), dim3(kRedThreads), 0, stream(), a, buf);
  hipLaunchKernelGGL(k_log_final, dim3(1), dim3(kRedThreads), 0, stream(), (const double*)buf, nblk,
                     buf + (size_t)Q_N * kBlocks);
  if (check_launch("sbfs_logging_values")) return 1;
  double q[Q_N];
  SBFS_HIP(hipMemcpyAsync(q, buf + (size_t)Q_N * kBlocks, sizeof(q), hipMemcpyDeviceToHost, stream()));
  SBFS_HIP(hipStreamSynchronize(stream()));
  out[SBFS_LOG_DIVMAX] = q[Q_DIVMAX];
  out[SBFS_LOG_MOMX] = q[Q_MOMX];
  out[SBFS_LOG_MOMY] = q[Q_MOMY];
  out[SBFS_LOG_EXTX] = q[Q_EXTX];
  out[SBFS_LOG_EXTY] = q[Q_EXTY];
  out[SBFS_LOG_ENGX] = q[Q_ENGX];
  out[SBFS_LOG_ENGY] = q[Q_ENGY];
  out[SBFS_LOG_STRESS_XM] = q[Q_STRM];
  out[SBFS_LOG_STRESS_XP] = q[Q_STRP];
  out[SBFS_LOG_VOFMIN] = q[Q_VMIN];
  out[SBFS_LOG_VOFMAX] = q[Q_VMAX];
  out[SBFS_LOG_VOFMEAN] = q[Q_VOL] / q[Q_AREA_DS];
  out[SBFS_LOG_AREA] = q[Q_AREA];
  out[SBFS_LOG_VOLUME] = q[Q_VOL];
  out[SBFS_LOG_CENTROID] = q[Q_VOL] > 0. ? q[Q_XVOL] / q[Q_VOL] : 0.;
  out[SBFS_LOG_RISE] = q[Q_VOL] > 0. ? q[Q_UVOL] / q[Q_VOL] : 0.;
  return 0;
}
