// Pointwise and stencil arithmetic of the time step, as device functions
// on values.  Each follows the reference expression it cites with the same
// operation order (the library is built with -ffp-contract=off, and FP64
// division and sqrt are correctly rounded on gfx950), so the staged and the
// fused kernels produce bitwise the same numbers as the reference's loops.
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>

#include "sbfs_internal.hpp"

namespace sbfs {

struct V2 {
  double x, y;
};
struct V3 {
  double x, y, z;
};

// density / viscosity of a volume fraction (src/fluid/density.c:25-27,
// src/fluid/viscosity.c:13-23)
__device__ __forceinline__ double mix_clamp(double ratio, double lo, double hi,
                                            double v) {
  double r = 1. + (ratio - 1.) * v;
  r = fmax(lo, r);
  r = fmin(hi, r);
  return r;
}

// corner unit gradient of vof (src/interface/curvature_tensor.c:26-47);
// mm = VOF(i-1,j-1), pm = VOF(i,j-1), mp = VOF(i-1,j), pp = VOF(i,j)
__device__ __forceinline__ V2 corner_gradient(double mm, double pm, double mp,
                                              double pp, double hxxf_i,
                                              double hy) {
  const double gx = 1. / hxxf_i * (-mm + pm - mp + pp);
  const double gy = 1. / hy * (-mm - pm + mp + pp);
  const double norm = sqrt(gx * gx + gy * gy);
  const double ninv = 1. / fmax(norm, DBL_EPSILON);
  return {gx * ninv, gy * ninv};
}

// cells where the surface is reconstructed (curvature_tensor.c:67-69)
__device__ __forceinline__ bool in_band(double v) {
  return !(v < kVofMin || 1. - kVofMin < v);
}

// cell normal + THINC intercept from the four corner gradients
// (curvature_tensor.c:71-93); c00 = DVOF(i,j), c10 = DVOF(i+1,j),
// c01 = DVOF(i,j+1), c11 = DVOF(i+1,j+1)
__device__ __forceinline__ V3 cell_normal(V2 c00, V2 c10, V2 c01, V2 c11,
                                          double hxxc_i, double hy, double v) {
  double nx = c00.x + c10.x + c01.x + c11.x;
  double ny = c00.y + c10.y + c01.y + c11.y;
  nx /= hxxc_i;
  ny /= hy;
  const double norm = sqrt(nx * nx + ny * ny);
  const double ninv = 1. / fmax(norm, DBL_EPSILON);
  nx *= ninv;
  ny *= ninv;
  return {nx, ny, -0.5 / kVofBeta * log(1. / v - 1.)};
}

// curvature = -div of the corner normals (curvature_tensor.c:114-131)
__device__ __forceinline__ double curvature(V2 c00, V2 c10, V2 c01, V2 c11,
                                            double hm, double hp, double jm,
                                            double j0, double jp, double hy) {
  const double nxm = 0.5 * c00.x + 0.5 * c01.x;
  const double nxp = 0.5 * c10.x + 0.5 * c11.x;
  const double nym = 0.5 * c00.y + 0.5 * c10.y;
  const double nyp = 0.5 * c01.y + 0.5 * c11.y;
  const double div = 1. / j0 * (-jm / hm * nxm + jp / hp * nxp - j0 / hy * nym + j0 / hy * nyp);
  return -1. * div;
}

// continuum surface force at a face (src/interface/force.c:37-49, :72-84);
// dm/dp, km/kp, vm/vp: density, curvature, vof on the two sides
__device__ __forceinline__ double surface_force(double dm, double dp, double km,
                                                double kp, double vm, double vp,
                                                double h, double refdeninv,
                                                double tension) {
  const double d = 0.5 * dm + 0.5 * dp;
  const double k = 0.5 * km + 0.5 * kp;
  const double delta = 1. / h * (-vm + vp);
  return d * refdeninv * tension * k * delta;
}

// diffused indicator 0.5 (1 + tanh(beta (n.x + d))) (update/main.c:17-35)
__device__ __forceinline__ double thinc(V3 n, double x, double y) {
  const double sf = n.x * x + n.y * y + n.z;
  return 0.5 * (1. + tanh(kVofBeta * sf));
}

// upwind THINC flux through an x face (update/fluxx.c:53-68): `v` and `n`
// are the vof and normal of the upwind cell
__device__ __forceinline__ double thinc_flux_x(double vel, double v, V3 n) {
  const double x = vel < 0. ? -0.5 : +0.5;
  if (!in_band(v)) return vel * v;
  double fl = 0.;
  fl += kGaussW * thinc(n, x, -kGaussP);
  fl += kGaussW * thinc(n, x, +kGaussP);
  return vel * fl;
}

// same through a y face (update/fluxy.c:53-68)
__device__ __forceinline__ double thinc_flux_y(double vel, double v, V3 n) {
  const double y = vel < 0. ? -0.5 : +0.5;
  if (!in_band(v)) return vel * v;
  double fl = 0.;
  fl += kGaussW * thinc(n, -kGaussP, y);
  fl += kGaussW * thinc(n, +kGaussP, y);
  return vel * fl;
}

// write a value of row j and its periodic image when row j lies within `nh`
// rows of either end (src/halo.c:51-74 on one rank)
template <typename IndexFn>
__device__ __forceinline__ void store_with_halo(double* a, int i, int j, int ny,
                                                int nh, double v, IndexFn idx) {
  a[idx(i, j)] = v;
  if (j > ny - nh) a[idx(i, j - ny)] = v;
  if (j <= nh) a[idx(i, j + ny)] = v;
}

}  // namespace sbfs
