// Per-stage kernels of the time step and the reference-named entry points
// (include/fluid_solver.h, include/interface_solver.h).
//
// One thread per output element; a thread that writes a boundary or halo
// value writes it from the value it computed, so no stage needs a separate
// halo pass (the reference's halo_communicate_in_y, src/halo.c, becomes a
// periodic index image on one GPU).
#include <cmath>

#include "fluid_solver.h"
#include "interface_solver.h"
#include "sbfs_internal.hpp"
#include "stencils.hpp"

namespace sbfs {

#define KRANGE int i0, int i1, int j0, int j1
#define KINDEX                                                     \
  const int i = i0 + (int)(blockIdx.x * blockDim.x + threadIdx.x); \
  const int j = j0 + (int)(blockIdx.y * blockDim.y + threadIdx.y); \
  if (i > i1 || j > j1) return

template <typename K, typename... A>
static void launch2n(const char* name, K kernel, int i0, int i1, int j0, int j1, A... args) {
  const dim3 block(64, 4);
  const dim3 grid(cdiv(i1 - i0 + 1, 64), cdiv(j1 - j0 + 1, 4));
  SBFS_TIMED(name, hipLaunchKernelGGL(kernel, grid, block, 0, stream(), i0, i1, j0, j1, args...));
}
#define launch2(k, ...) launch2n(#k, k, __VA_ARGS__)

// index helpers bound to an interior width
#define IXF(i, j) L::xf(i, j, nx)
#define ICC(i, j) L::cc(i, j, nx)
#define IVF(i, j) L::vof(i, j, nx)
#define ICX(i, j) L::cx(i, j, nx)
#define IFX(i, j) L::fx(i, j, nx)
#define ICI(i, j) L::ci(i, j, nx)

__device__ __forceinline__ void store_cc_neumann(double* a, int i, int j, int nx,
                                                 int ny, double v) {
  auto idx = [nx](int ii, int jj) { return L::cc(ii, jj, nx); };
  store_with_halo(a, i, j, ny, 1, v, idx);
  if (i == 1) store_with_halo(a, 0, j, ny, 1, v, idx);
  if (i == nx) store_with_halo(a, nx + 1, j, ny, 1, v, idx);
}

// ---------------------------------------------------------------------------
// density / viscosity
// ---------------------------------------------------------------------------
__global__ void k_density(KRANGE, const double* __restrict__ vof,
                          double* __restrict__ den, int nx, double denr,
                          double lo, double hi) {
  KINDEX;
  den[ICC(i, j)] = mix_clamp(denr, lo, hi, vof[IVF(i, j)]);
}

// src/fluid/viscosity.c:37-53
__global__ void k_visux(KRANGE, const double* __restrict__ vof,
                        double* __restrict__ visux, int nx, int ny, double visr,
                        double lo, double hi) {
  KINDEX;
  double v;
  if (i == 1) v = vof[IVF(0, j)];
  else if (i == nx + 1) v = vof[IVF(nx + 1, j)];
  else v = 0.5 * vof[IVF(i - 1, j)] + 0.5 * vof[IVF(i, j)];
  store_with_halo(visux, i, j, ny, 1, mix_clamp(visr, lo, hi, v),
                  [nx](int a, int b) { return L::xf(a, b, nx); });
}

// src/fluid/viscosity.c:75-83
__global__ void k_visuy(KRANGE, const double* __restrict__ vof,
                        double* __restrict__ visuy, int nx, int ny, double visr,
                        double lo, double hi) {
  KINDEX;
  const double v = 0.5 * vof[IVF(i, j - 1)] + 0.5 * vof[IVF(i, j)];
  store_with_halo(visuy, i, j, ny, 1, mix_clamp(visr, lo, hi, v),
                  [nx](int a, int b) { return L::cc(a, b, nx); });
}

// ---------------------------------------------------------------------------
// interface geometry
// ---------------------------------------------------------------------------
__global__ void k_dvof(KRANGE, const double* __restrict__ vof,
                       double* __restrict__ dvof, const double* __restrict__ hxxf,
                       double hy, int nx) {
  KINDEX;
  const V2 g = corner_gradient(vof[IVF(i - 1, j - 1)], vof[IVF(i, j - 1)],
                               vof[IVF(i - 1, j)], vof[IVF(i, j)], hxxf[i - 1], hy);
  const size_t k = IXF(i, j);
  dvof[2 * k] = g.x;
  dvof[2 * k + 1] = g.y;
}

__device__ __forceinline__ V2 ld2(const double* __restrict__ a, size_t k) {
  return {a[2 * k], a[2 * k + 1]};
}

__global__ void k_normal_curv(KRANGE, const double* __restrict__ vof,
                              const double* __restrict__ dvof,
                              double* __restrict__ normal, double* __restrict__ curv,
                              const double* __restrict__ hxxf,
                              const double* __restrict__ hxxc,
                              const double* __restrict__ jdxf,
                              const double* __restrict__ jdxc, double hy, int nx) {
  KINDEX;
  const V2 c00 = ld2(dvof, IXF(i, j)), c10 = ld2(dvof, IXF(i + 1, j));
  const V2 c01 = ld2(dvof, IXF(i, j + 1)), c11 = ld2(dvof, IXF(i + 1, j + 1));
  const double v = vof[IVF(i, j)];
  if (in_band(v)) {
    const V3 n = cell_normal(c00, c10, c01, c11, hxxc[i - 1], hy, v);
    const size_t k = ICX(i, j);
    normal[3 * k] = n.x;
    normal[3 * k + 1] = n.y;
    normal[3 * k + 2] = n.z;
  }
  curv[ICC(i, j)] = curvature(c00, c10, c01, c11, hxxf[i - 1], hxxf[i], jdxf[i - 1],
                              jdxc[i - 1], jdxf[i], hy);
}

// src/interface/force.c:20-98
__global__ void k_force(KRANGE, const double* __restrict__ den,
                        const double* __restrict__ curv,
                        const double* __restrict__ vof, double* __restrict__ ifrcx,
                        double* __restrict__ ifrcy, const double* __restrict__ hxxf,
                        double hy, double refdeninv, double tension, int nx) {
  KINDEX;
  if (i >= 2)
    ifrcx[IFX(i, j)] = surface_force(den[ICC(i - 1, j)], den[ICC(i, j)], curv[ICC(i - 1, j)],
                                     curv[ICC(i, j)], vof[IVF(i - 1, j)], vof[IVF(i, j)],
                                     hxxf[i - 1], refdeninv, tension);
  ifrcy[ICI(i, j)] = surface_force(den[ICC(i, j - 1)], den[ICC(i, j)], curv[ICC(i, j - 1)],
                                   curv[ICC(i, j)], vof[IVF(i, j - 1)], vof[IVF(i, j)], hy,
                                   refdeninv, tension);
}

// ---------------------------------------------------------------------------
// vof transport
// ---------------------------------------------------------------------------
__device__ __forceinline__ V3 ld3(const double* __restrict__ a, size_t k) {
  return {a[3 * k], a[3 * k + 1], a[3 * k + 2]};
}

// src/interface/update/fluxx.c:14-73
__global__ void k_fluxx(KRANGE, const double* __restrict__ ux,
                        const double* __restrict__ vof,
                        const double* __restrict__ normal, double* __restrict__ fluxx,
                        int nx, int ny) {
  KINDEX;
  double f = 0.;
  if (i != 1 && i != nx + 1) {
    const double vel = ux[IXF(i, j)];
    const int ii = vel < 0. ? i : i - 1;
    const double v = vof[IVF(ii, j)];
    const V3 n = in_band(v) ? ld3(normal, ICX(ii, j)) : V3{0., 0., 0.};
    f = thinc_flux_x(vel, v, n);
  }
  store_with_halo(fluxx, i, j, ny, 1, f, [nx](int a, int b) { return L::xf(a, b, nx); });
}

// src/interface/update/fluxy.c:14-73
__global__ void k_fluxy(KRANGE, const double* __restrict__ uy,
                        const double* __restrict__ vof,
                        const double* __restrict__ normal, double* __restrict__ fluxy,
                        int nx, int ny) {
  KINDEX;
  double f = 0.;
  if (i != 0 && i != nx + 1) {
    const double vel = uy[ICC(i, j)];
    const int jj = vel < 0. ? j : j - 1;
    const double v = vof[IVF(i, jj)];
    const V3 n = in_band(v) ? ld3(normal, ICX(i, jj)) : V3{0., 0., 0.};
    f = thinc_flux_y(vel, v, n);
  }
  store_with_halo(fluxy, i, j, ny, 1, f, [nx](int a, int b) { return L::cc(a, b, nx); });
}

// flux divergence, Runge-Kutta update and vof boundaries
// (src/interface/update/main.c:54-120, src/interface/boundary.c:15-42)
__global__ void k_vof_advance(KRANGE, const double* __restrict__ fluxx,
                              const double* __restrict__ fluxy,
                              double* __restrict__ srca, const double* __restrict__ srcb,
                              double* __restrict__ vof, const double* __restrict__ hxxf,
                              const double* __restrict__ jdxf,
                              const double* __restrict__ jdxc, double hy, double dtca,
                              double dtcb, int has_b, int nx, int ny) {
  KINDEX;
  double v = 0.;
  if (i != 0 && i != nx + 1) {
    const double hm = hxxf[i - 1], hp = hxxf[i];
    const double jm = jdxf[i - 1], j0 = jdxc[i - 1], jp = jdxf[i];
    const double s = 1. / j0 * (+jm / hm * fluxx[IXF(i, j)] - jp / hp * fluxx[IXF(i + 1, j)]
                                + j0 / hy * fluxy[ICC(i, j)] - j0 / hy * fluxy[ICC(i, j + 1)]);
    srca[ICI(i, j)] = s;
    v = vof[IVF(i, j)];
    v += dtca * s;
    if (has_b) v += dtcb * srcb[ICI(i, j)];
  }
  store_with_halo(vof, i, j, ny, 2, v, [nx](int a, int b) { return L::vof(a, b, nx); });
}

// src/interface/mass_flux.c:10-58
__global__ void k_mass_flux(KRANGE, const double* __restrict__ ux,
                            const double* __restrict__ uy, double* __restrict__ fluxx,
                            double* __restrict__ fluxy, double denr, int nx) {
  KINDEX;
  if (i >= 1) fluxx[IXF(i, j)] = ux[IXF(i, j)] + (denr - 1.) * fluxx[IXF(i, j)];
  fluxy[ICC(i, j)] = uy[ICC(i, j)] + (denr - 1.) * fluxy[ICC(i, j)];
}

// ---------------------------------------------------------------------------
// momentum predictor
// ---------------------------------------------------------------------------
// src/fluid/predict/txx.c:21-31, txy.c:26-41, tyy.c:20-29
__global__ void k_stress(KRANGE, const double* __restrict__ ux,
                         const double* __restrict__ uy,
                         const double* __restrict__ visux,
                         const double* __restrict__ visuy, double* __restrict__ txx,
                         double* __restrict__ txy, double* __restrict__ tyy,
                         const double* __restrict__ hxxf,
                         const double* __restrict__ hxxc, double hy, int nx, int ny) {
  KINDEX;
  if (i <= nx) {
    const double vx = 0.5 * visux[IXF(i, j)] + 0.5 * visux[IXF(i + 1, j)];
    const double lxx = 1. / hxxc[i - 1] * (-ux[IXF(i, j)] + ux[IXF(i + 1, j)]);
    store_with_halo(txx, i, j, ny, 1, vx * lxx + vx * lxx,
                    [nx](int a, int b) { return L::cx(a, b, nx); });
    const double vy = 0.5 * visuy[ICC(i, j)] + 0.5 * visuy[ICC(i, j + 1)];
    const double lyy = 1. / hy * (-uy[ICC(i, j)] + uy[ICC(i, j + 1)]);
    store_with_halo(tyy, i, j, ny, 1, vy * lyy + vy * lyy,
                    [nx](int a, int b) { return L::cx(a, b, nx); });
  }
  const double vx = 0.5 * visux[IXF(i, j - 1)] + 0.5 * visux[IXF(i, j)];
  const double vy = 0.5 * visuy[ICC(i - 1, j)] + 0.5 * visuy[ICC(i, j)];
  const double lxy = 1. / hy * (-ux[IXF(i, j - 1)] + ux[IXF(i, j)]);
  const double lyx = 1. / hxxf[i - 1] * (-uy[ICC(i - 1, j)] + uy[ICC(i, j)]);
  store_with_halo(txy, i, j, ny, 1, vx * lxy + vy * lyx,
                  [nx](int a, int b) { return L::xf(a, b, nx); });
}

// explicit sources of ux (src/fluid/predict/ux.c:29-226); every term is
// accumulated in the reference's order starting from the cleared source
__global__ void k_rhs_ux(KRANGE, const double* __restrict__ ux,
                         const double* __restrict__ fluxx,
                         const double* __restrict__ fluxy,
                         const double* __restrict__ txx, const double* __restrict__ txy,
                         const double* __restrict__ p, const double* __restrict__ ifrcx,
                         const double* __restrict__ den0, double* __restrict__ srca,
                         double* __restrict__ srcg, const double* __restrict__ hxxf,
                         const double* __restrict__ hxxc,
                         const double* __restrict__ jdxf,
                         const double* __restrict__ jdxc, double hy,
                         double diffusivity, double grav, int nx) {
  KINDEX;
  double s = 0.;
  {
    const double mm = 0.5 * jdxf[i - 2] / hxxf[i - 2] * fluxx[IXF(i - 1, j)] + 0.5 * jdxf[i - 1] / hxxf[i - 1] * fluxx[IXF(i, j)];
    const double mp = 0.5 * jdxf[i - 1] / hxxf[i - 1] * fluxx[IXF(i, j)] + 0.5 * jdxf[i] / hxxf[i] * fluxx[IXF(i + 1, j)];
    const double um = 0.5 * ux[IXF(i - 1, j)] + 0.5 * ux[IXF(i, j)];
    const double up = 0.5 * ux[IXF(i, j)] + 0.5 * ux[IXF(i + 1, j)];
    s -= 1. / jdxf[i - 1] * (-mm * um + mp * up);
  }
  {
    const double mm = 0.5 * jdxc[i - 2] / hy * fluxy[ICC(i - 1, j)] + 0.5 * jdxc[i - 1] / hy * fluxy[ICC(i, j)];
    const double mp = 0.5 * jdxc[i - 2] / hy * fluxy[ICC(i - 1, j + 1)] + 0.5 * jdxc[i - 1] / hy * fluxy[ICC(i, j + 1)];
    const double um = 0.5 * ux[IXF(i, j - 1)] + 0.5 * ux[IXF(i, j)];
    const double up = 0.5 * ux[IXF(i, j)] + 0.5 * ux[IXF(i, j + 1)];
    s -= 1. / jdxf[i - 1] * (-mm * um + mp * up);
  }
  s += diffusivity / jdxf[i - 1] * (-jdxc[i - 2] / hxxc[i - 2] * txx[ICX(i - 1, j)] + jdxc[i - 1] / hxxc[i - 1] * txx[ICX(i, j)]);
  s += diffusivity / jdxf[i - 1] * (-jdxf[i - 1] / hy * txy[IXF(i, j)] + jdxf[i - 1] / hy * txy[IXF(i, j + 1)]);
  double g = 0.;
  g -= 1. / hxxf[i - 1] * (-p[ICC(i - 1, j)] + p[ICC(i, j)]);
  s += ifrcx[IFX(i, j)];
  s += grav * (0.5 * den0[ICC(i - 1, j)] + 0.5 * den0[ICC(i, j)]);
  srca[IFX(i, j)] = s;
  srcg[IFX(i, j)] = g;
}

// explicit sources of uy (src/fluid/predict/uy.c:29-204)
__global__ void k_rhs_uy(KRANGE, const double* __restrict__ uy,
                         const double* __restrict__ fluxx,
                         const double* __restrict__ fluxy,
                         const double* __restrict__ txy, const double* __restrict__ tyy,
                         const double* __restrict__ p, const double* __restrict__ ifrcy,
                         double* __restrict__ srca, double* __restrict__ srcg,
                         const double* __restrict__ hxxf,
                         const double* __restrict__ jdxf,
                         const double* __restrict__ jdxc, double hy,
                         double diffusivity, int nx) {
  KINDEX;
  double s = 0.;
  {
    const double mm = 0.5 * jdxf[i - 1] / hxxf[i - 1] * fluxx[IXF(i, j - 1)] + 0.5 * jdxf[i - 1] / hxxf[i - 1] * fluxx[IXF(i, j)];
    const double mp = 0.5 * jdxf[i] / hxxf[i] * fluxx[IXF(i + 1, j - 1)] + 0.5 * jdxf[i] / hxxf[i] * fluxx[IXF(i + 1, j)];
    const double um = 0.5 * uy[ICC(i - 1, j)] + 0.5 * uy[ICC(i, j)];
    const double up = 0.5 * uy[ICC(i, j)] + 0.5 * uy[ICC(i + 1, j)];
    s -= 1. / jdxc[i - 1] * (-mm * um + mp * up);
  }
  {
    const double mm = 0.5 * jdxc[i - 1] / hy * fluxy[ICC(i, j - 1)] + 0.5 * jdxc[i - 1] / hy * fluxy[ICC(i, j)];
    const double mp = 0.5 * jdxc[i - 1] / hy * fluxy[ICC(i, j)] + 0.5 * jdxc[i - 1] / hy * fluxy[ICC(i, j + 1)];
    const double um = 0.5 * uy[ICC(i, j - 1)] + 0.5 * uy[ICC(i, j)];
    const double up = 0.5 * uy[ICC(i, j)] + 0.5 * uy[ICC(i, j + 1)];
    s -= 1. / jdxc[i - 1] * (-mm * um + mp * up);
  }
  s += diffusivity / jdxc[i - 1] * (-jdxf[i - 1] / hxxf[i - 1] * txy[IXF(i, j)] + jdxf[i] / hxxf[i] * txy[IXF(i + 1, j)]);
  s += diffusivity / jdxc[i - 1] * (-jdxc[i - 1] / hy * tyy[ICX(i, j - 1)] + jdxc[i - 1] / hy * tyy[ICX(i, j)]);
  double g = 0.;
  g -= 1. / hy * (-p[ICC(i, j - 1)] + p[ICC(i, j)]);
  s += ifrcy[ICI(i, j)];
  srca[ICI(i, j)] = s;
  srcg[ICI(i, j)] = g;
}

// predicted ux + walls + halo (src/fluid/predict/ux.c:236-273)
__global__ void k_predict_ux(KRANGE, double* __restrict__ ux,
                             const double* __restrict__ sa, const double* __restrict__ sb,
                             const double* __restrict__ sg,
                             const double* __restrict__ den0,
                             const double* __restrict__ den1, double ca, double cb,
                             double cg, double dt, int nx, int ny) {
  KINDEX;
  double u = 0.;
  if (i != 1 && i != nx + 1) {
    const double d0 = 0.5 * den0[ICC(i - 1, j)] + 0.5 * den0[ICC(i, j)];
    u = d0 * ux[IXF(i, j)] + ca * dt * sa[IFX(i, j)] + cb * dt * sb[IFX(i, j)] + cg * dt * sg[IFX(i, j)];
    const double d1 = 0.5 * den1[ICC(i - 1, j)] + 0.5 * den1[ICC(i, j)];
    u /= d1;
  }
  store_with_halo(ux, i, j, ny, 1, u, [nx](int a, int b) { return L::xf(a, b, nx); });
}

// predicted uy + wall ghosts + halo (src/fluid/predict/uy.c:214-251)
__global__ void k_predict_uy(KRANGE, double* __restrict__ uy,
                             const double* __restrict__ sa, const double* __restrict__ sb,
                             const double* __restrict__ sg,
                             const double* __restrict__ den0,
                             const double* __restrict__ den1, double ca, double cb,
                             double cg, double dt, int nx, int ny) {
  KINDEX;
  double u;
  if (i == 0) u = kUyXm;
  else if (i == nx + 1) u = kUyXp;
  else {
    const double d0 = 0.5 * den0[ICC(i, j - 1)] + 0.5 * den0[ICC(i, j)];
    u = d0 * uy[ICC(i, j)] + ca * dt * sa[ICI(i, j)] + cb * dt * sb[ICI(i, j)] + cg * dt * sg[ICI(i, j)];
    const double d1 = 0.5 * den1[ICC(i, j - 1)] + 0.5 * den1[ICC(i, j)];
    u /= d1;
  }
  store_with_halo(uy, i, j, ny, 1, u, [nx](int a, int b) { return L::cc(a, b, nx); });
}

// ---------------------------------------------------------------------------
// projection and pressure
// ---------------------------------------------------------------------------
// src/fluid/correct_velocity/ux.c:54-89
__global__ void k_correct_ux(KRANGE, double* __restrict__ ux,
                             const double* __restrict__ psi1,
                             const double* __restrict__ psi0,
                             const double* __restrict__ den1,
                             const double* __restrict__ hxxf, double dtn,
                             double refden, double coef, int nx, int ny) {
  KINDEX;
  double u = 0.;
  if (i != 1 && i != nx + 1) {
    const double h = hxxf[i - 1];
    u = ux[IXF(i, j)];
    u -= dtn / refden / h * (-psi1[ICC(i - 1, j)] + psi1[ICC(i, j)]);
    const double d = 0.5 * den1[ICC(i - 1, j)] + 0.5 * den1[ICC(i, j)];
    u += dtn * coef * (1. / d - 1. / refden) / h * (-psi0[ICC(i - 1, j)] + psi0[ICC(i, j)]);
  }
  store_with_halo(ux, i, j, ny, 1, u, [nx](int a, int b) { return L::xf(a, b, nx); });
}

// src/fluid/correct_velocity/uy.c:120-154
__global__ void k_correct_uy(KRANGE, double* __restrict__ uy,
                             const double* __restrict__ psi1,
                             const double* __restrict__ psi0,
                             const double* __restrict__ den1, double hy, double dtn,
                             double refden, double coef, int nx, int ny) {
  KINDEX;
  double u;
  if (i == 0) u = kUyXm;
  else if (i == nx + 1) u = kUyXp;
  else {
    u = uy[ICC(i, j)];
    u -= dtn / refden / hy * (-psi1[ICC(i, j - 1)] + psi1[ICC(i, j)]);
    const double d = 0.5 * den1[ICC(i, j - 1)] + 0.5 * den1[ICC(i, j)];
    u += dtn * coef * (1. / d - 1. / refden) / hy * (-psi0[ICC(i, j - 1)] + psi0[ICC(i, j)]);
  }
  store_with_halo(uy, i, j, ny, 1, u, [nx](int a, int b) { return L::cc(a, b, nx); });
}

// src/fluid/update_pressure.c:7-19 + p boundaries
__global__ void k_update_p(KRANGE, double* __restrict__ p,
                           const double* __restrict__ psi1, int nx, int ny) {
  KINDEX;
  store_cc_neumann(p, i, j, nx, ny, p[ICC(i, j)] + psi1[ICC(i, j)]);
}

// ---------------------------------------------------------------------------
// boundary conditions alone
// ---------------------------------------------------------------------------
__global__ void k_bc_ux(KRANGE, double* __restrict__ ux, int nx, int ny) {
  KINDEX;
  const double u = (i == 1 || i == nx + 1) ? 0. : ux[IXF(i, j)];
  store_with_halo(ux, i, j, ny, 1, u, [nx](int a, int b) { return L::xf(a, b, nx); });
}

__global__ void k_bc_uy(KRANGE, double* __restrict__ uy, int nx, int ny) {
  KINDEX;
  const double u = (i == 0) ? kUyXm : (i == nx + 1) ? kUyXp : uy[ICC(i, j)];
  store_with_halo(uy, i, j, ny, 1, u, [nx](int a, int b) { return L::cc(a, b, nx); });
}

__global__ void k_bc_neumann(KRANGE, double* __restrict__ a, int nx, int ny) {
  KINDEX;
  store_cc_neumann(a, i, j, nx, ny, a[ICC(i, j)]);
}

__global__ void k_bc_vof(KRANGE, double* __restrict__ vof, int nx, int ny) {
  KINDEX;
  const double v = (i == 0 || i == nx + 1) ? 0. : vof[IVF(i, j)];
  store_with_halo(vof, i, j, ny, 2, v, [nx](int a, int b) { return L::vof(a, b, nx); });
}

// ---------------------------------------------------------------------------
// launchers shared with the integrator
// ---------------------------------------------------------------------------
int launch_density(const domain_t* d, DomainCtx* c, const double* vof, double* den,
                   double denr) {
  const int nx = (int)c->nx, ny = (int)c->ny;
  (void)d;
  launch2(k_density, 0, nx + 1, 0, ny + 1, vof, den, nx, denr, fmin(1., denr), fmax(1., denr));
  return check_launch("fluid_compute_density");
}

}  // namespace sbfs

using namespace sbfs;

// ---------------------------------------------------------------------------
// entry points
// ---------------------------------------------------------------------------
#define GET_CTX(d)                     \
  DomainCtx* c = domain_ctx(d);        \
  if (!c) return 1;                    \
  const int nx = (int)c->nx;           \
  const int ny = (int)c->ny;           \
  (void)nx;                            \
  (void)ny

#define D(a) ((double*)(a).data)
#define CHK(a, na, elem, name) \
  if (check_array(domain, &(a), na, elem, name)) return 1

extern "C" {

int fluid_compute_density(const domain_t* domain, fluid_t* fluid,
                          const interface_t* interface, const size_t index) {
  GET_CTX(domain);
  SBFS_REQUIRE(index < 2, "fluid_compute_density: index %zu out of range", index);
  CHK(fluid->den[index], NA_DEN, 8, "den");
  CHK(interface->vof, NA_VOF, 8, "vof");
  return launch_density(domain, c, D(interface->vof), D(fluid->den[index]), fluid->denr);
}

int fluid_compute_viscosity(const domain_t* domain, fluid_t* fluid,
                            const interface_t* interface) {
  GET_CTX(domain);
  CHK(fluid->visux, NA_VISUX, 8, "visux");
  CHK(fluid->visuy, NA_VISUY, 8, "visuy");
  CHK(interface->vof, NA_VOF, 8, "vof");
  const double visr = fluid->visr, lo = fmin(1., visr), hi = fmax(1., visr);
  launch2(k_visux, 1, nx + 1, 1, ny, D(interface->vof), D(fluid->visux), nx, ny, visr, lo, hi);
  launch2(k_visuy, 0, nx + 1, 1, ny, D(interface->vof), D(fluid->visuy), nx, ny, visr, lo, hi);
  return check_launch("fluid_compute_viscosity");
}

int interface_compute_curvature_tensor(const domain_t* domain, interface_t* interface) {
  GET_CTX(domain);
  CHK(interface->vof, NA_VOF, 8, "vof");
  CHK(interface->dvof, NA_DVOF, 16, "dvof");
  CHK(interface->normal, NA_NORMAL, 24, "normal");
  CHK(interface->curv, NA_CURV, 8, "curv");
  launch2(k_dvof, 1, nx + 1, 0, ny + 2, D(interface->vof), D(interface->dvof), c->hxxf, c->hy, nx);
  launch2(k_normal_curv, 1, nx, 0, ny + 1, D(interface->vof), D(interface->dvof),
          D(interface->normal), D(interface->curv), c->hxxf, c->hxxc, c->jdxf, c->jdxc, c->hy, nx);
  return check_launch("interface_compute_curvature_tensor");
}

int interface_compute_force(const domain_t* domain, const fluid_t* fluid,
                            interface_t* interface) {
  GET_CTX(domain);
  CHK(fluid->den[0], NA_DEN, 8, "den[0]");
  CHK(interface->curv, NA_CURV, 8, "curv");
  CHK(interface->vof, NA_VOF, 8, "vof");
  CHK(interface->ifrcx, NA_IFRCX, 8, "ifrcx");
  CHK(interface->ifrcy, NA_IFRCY, 8, "ifrcy");
  const double tension = 1. / interface->We;
  const double refdeninv = 2. / (1. + fluid->denr);  // force.c:14-18
  launch2(k_force, 1, nx, 1, ny, D(fluid->den[0]), D(interface->curv), D(interface->vof),
          D(interface->ifrcx), D(interface->ifrcy), c->hxxf, c->hy, refdeninv, tension, nx);
  return check_launch("interface_compute_force");
}

int interface_update_vof(const domain_t* domain, const size_t rkstep, const double dt,
                         const fluid_t* fluid, interface_t* interface) {
  GET_CTX(domain);
  SBFS_REQUIRE(rkstep < 3, "interface_update_vof: rkstep %zu out of range", rkstep);
  CHK(fluid->ux, NA_UX, 8, "ux");
  CHK(fluid->uy, NA_UY, 8, "uy");
  CHK(interface->vof, NA_VOF, 8, "vof");
  CHK(interface->normal, NA_NORMAL, 24, "normal");
  CHK(interface->fluxx, NA_FLUXX, 8, "fluxx");
  CHK(interface->fluxy, NA_FLUXY, 8, "fluxy");
  CHK(interface->src[0], NA_SRC, 8, "src[0]");
  CHK(interface->src[1], NA_SRC, 8, "src[1]");
  // stash the previous alpha source as beta (update/main.c:37-52)
  if (rkstep != 0) {
    void* t = interface->src[0].data;
    interface->src[0].data = interface->src[1].data;
    interface->src[1].data = t;
  }
  launch2(k_fluxx, 1, nx + 1, 1, ny, D(fluid->ux), D(interface->vof), D(interface->normal),
          D(interface->fluxx), nx, ny);
  launch2(k_fluxy, 0, nx + 1, 1, ny, D(fluid->uy), D(interface->vof), D(interface->normal),
          D(interface->fluxy), nx, ny);
  const RK r = rk(rkstep);
  launch2(k_vof_advance, 0, nx + 1, 1, ny, D(interface->fluxx), D(interface->fluxy),
          D(interface->src[0]), D(interface->src[1]), D(interface->vof), c->hxxf, c->jdxf,
          c->jdxc, c->hy, dt * r.a, dt * r.b, rkstep != 0 ? 1 : 0, nx, ny);
  return check_launch("interface_update_vof");
}

int interface_compute_mass_flux(const domain_t* domain, const fluid_t* fluid,
                                interface_t* interface) {
  GET_CTX(domain);
  CHK(fluid->ux, NA_UX, 8, "ux");
  CHK(fluid->uy, NA_UY, 8, "uy");
  CHK(interface->fluxx, NA_FLUXX, 8, "fluxx");
  CHK(interface->fluxy, NA_FLUXY, 8, "fluxy");
  launch2(k_mass_flux, 0, nx + 1, 0, ny + 1, D(fluid->ux), D(fluid->uy), D(interface->fluxx),
          D(interface->fluxy), fluid->denr, nx);
  return check_launch("interface_compute_mass_flux");
}

int fluid_predict_field(const domain_t* domain, const size_t rkstep, const double dt,
                        fluid_t* fluid, const interface_t* interface) {
  GET_CTX(domain);
  SBFS_REQUIRE(rkstep < 3, "fluid_predict_field: rkstep %zu out of range", rkstep);
  CHK(fluid->ux, NA_UX, 8, "ux");
  CHK(fluid->uy, NA_UY, 8, "uy");
  CHK(fluid->p, NA_P, 8, "p");
  CHK(fluid->den[0], NA_DEN, 8, "den[0]");
  CHK(fluid->den[1], NA_DEN, 8, "den[1]");
  CHK(fluid->visux, NA_VISUX, 8, "visux");
  CHK(fluid->visuy, NA_VISUY, 8, "visuy");
  CHK(fluid->txx, NA_TXX, 8, "txx");
  CHK(fluid->txy, NA_TXY, 8, "txy");
  CHK(fluid->tyy, NA_TYY, 8, "tyy");
  for (int n = 0; n < 3; n++) {
    CHK(fluid->srcux[n], NA_SRCUX, 8, "srcux");
    CHK(fluid->srcuy[n], NA_SRCUY, 8, "srcuy");
  }
  CHK(interface->fluxx, NA_FLUXX, 8, "fluxx");
  CHK(interface->fluxy, NA_FLUXY, 8, "fluxy");
  CHK(interface->ifrcx, NA_IFRCX, 8, "ifrcx");
  CHK(interface->ifrcy, NA_IFRCY, 8, "ifrcy");
  // stash the previous alpha sources as beta (predict/main.c:9-29); the
  // alpha and gamma sources are overwritten completely below
  if (rkstep != 0) {
    void* t = fluid->srcux[0].data;
    fluid->srcux[0].data = fluid->srcux[1].data;
    fluid->srcux[1].data = t;
    t = fluid->srcuy[0].data;
    fluid->srcuy[0].data = fluid->srcuy[1].data;
    fluid->srcuy[1].data = t;
  }
  launch2(k_stress, 1, nx + 1, 1, ny, D(fluid->ux), D(fluid->uy), D(fluid->visux), D(fluid->visuy),
          D(fluid->txx), D(fluid->txy), D(fluid->tyy), c->hxxf, c->hxxc, c->hy, nx, ny);
  const double diffusivity = 1. / fluid->Re;
  const double grav = -1. / fluid->Fr / fluid->Fr;
  launch2(k_rhs_ux, 2, nx, 1, ny, D(fluid->ux), D(interface->fluxx), D(interface->fluxy),
          D(fluid->txx), D(fluid->txy), D(fluid->p), D(interface->ifrcx), D(fluid->den[0]),
          D(fluid->srcux[0]), D(fluid->srcux[2]), c->hxxf, c->hxxc, c->jdxf, c->jdxc, c->hy,
          diffusivity, grav, nx);
  launch2(k_rhs_uy, 1, nx, 1, ny, D(fluid->uy), D(interface->fluxx), D(interface->fluxy),
          D(fluid->txy), D(fluid->tyy), D(fluid->p), D(interface->ifrcy), D(fluid->srcuy[0]),
          D(fluid->srcuy[2]), c->hxxf, c->jdxf, c->jdxc, c->hy, diffusivity, nx);
  const RK r = rk(rkstep);
  launch2(k_predict_ux, 1, nx + 1, 1, ny, D(fluid->ux), D(fluid->srcux[0]), D(fluid->srcux[1]),
          D(fluid->srcux[2]), D(fluid->den[0]), D(fluid->den[1]), r.a, r.b, r.g, dt, nx, ny);
  launch2(k_predict_uy, 0, nx + 1, 1, ny, D(fluid->uy), D(fluid->srcuy[0]), D(fluid->srcuy[1]),
          D(fluid->srcuy[2]), D(fluid->den[0]), D(fluid->den[1]), r.a, r.b, r.g, dt, nx, ny);
  return check_launch("fluid_predict_field");
}

int fluid_compute_potential(const domain_t* domain, const double dt_old, const double dt_new,
                            fluid_t* fluid) {
  GET_CTX(domain);
  CHK(fluid->ux, NA_UX, 8, "ux");
  CHK(fluid->uy, NA_UY, 8, "uy");
  CHK(fluid->psi[0], NA_P, 8, "psi[0]");
  CHK(fluid->psi[1], NA_P, 8, "psi[1]");
  CHK(fluid->den[1], NA_DEN, 8, "den[1]");
  return poisson_solve(domain, c, dt_old, dt_new, fluid);
}

int fluid_correct_velocity(const domain_t* domain, const double dt_old, const double dt_new,
                           fluid_t* fluid) {
  GET_CTX(domain);
  CHK(fluid->ux, NA_UX, 8, "ux");
  CHK(fluid->uy, NA_UY, 8, "uy");
  CHK(fluid->psi[0], NA_P, 8, "psi[0]");
  CHK(fluid->psi[1], NA_P, 8, "psi[1]");
  CHK(fluid->den[1], NA_DEN, 8, "den[1]");
  const double coef = -1. / dt_new * dt_old;  // correct_velocity/ux.c:72
  launch2(k_correct_ux, 1, nx + 1, 1, ny, D(fluid->ux), D(fluid->psi[1]), D(fluid->psi[0]),
          D(fluid->den[1]), c->hxxf, dt_new, fluid->refden, coef, nx, ny);
  launch2(k_correct_uy, 0, nx + 1, 1, ny, D(fluid->uy), D(fluid->psi[1]), D(fluid->psi[0]),
          D(fluid->den[1]), c->hy, dt_new, fluid->refden, coef, nx, ny);
  return check_launch("fluid_correct_velocity");
}

int fluid_update_pressure(const domain_t* domain, fluid_t* fluid) {
  GET_CTX(domain);
  CHK(fluid->p, NA_P, 8, "p");
  CHK(fluid->psi[1], NA_P, 8, "psi[1]");
  launch2(k_update_p, 1, nx, 1, ny, D(fluid->p), D(fluid->psi[1]), nx, ny);
  // update_pressure.c:181-190
  void* t = fluid->psi[0].data;
  fluid->psi[0].data = fluid->psi[1].data;
  fluid->psi[1].data = t;
  return check_launch("fluid_update_pressure");
}

int fluid_update_boundaries_ux(const domain_t* domain, array_t* array) {
  GET_CTX(domain);
  CHK(*array, NA_UX, 8, "ux");
  launch2(k_bc_ux, 1, nx + 1, 1, ny, D(*array), nx, ny);
  return check_launch("fluid_update_boundaries_ux");
}

int fluid_update_boundaries_uy(const domain_t* domain, array_t* array) {
  GET_CTX(domain);
  CHK(*array, NA_UY, 8, "uy");
  launch2(k_bc_uy, 0, nx + 1, 1, ny, D(*array), nx, ny);
  return check_launch("fluid_update_boundaries_uy");
}

int fluid_update_boundaries_p(const domain_t* domain, array_t* array) {
  GET_CTX(domain);
  CHK(*array, NA_P, 8, "p");
  launch2(k_bc_neumann, 1, nx, 1, ny, D(*array), nx, ny);
  return check_launch("fluid_update_boundaries_p");
}

int fluid_update_boundaries_psi(const domain_t* domain, array_t* array) {
  GET_CTX(domain);
  CHK(*array, NA_P, 8, "psi");
  launch2(k_bc_neumann, 1, nx, 1, ny, D(*array), nx, ny);
  return check_launch("fluid_update_boundaries_psi");
}

int interface_update_boundaries_vof(const domain_t* domain, array_t* vof) {
  GET_CTX(domain);
  CHK(*vof, NA_VOF, 8, "vof");
  launch2(k_bc_vof, 0, nx + 1, 1, ny, D(*vof), nx, ny);
  return check_launch("interface_update_boundaries_vof");
}

}  // extern "C"
