// Time-step control (reference src/decide_dt.c), device diagnostics and the
// integrator (src/integrate.c).
#include <cfloat>
#include <cmath>

#include "decide_dt.h"
#include "fluid_solver.h"
#include "integrate.h"
#include "interface_solver.h"
#include "sbfs.h"
#include "reduce.hpp"
#include "sbfs_internal.hpp"

namespace sbfs {

static const double kPiDt = 3.1415926535897932384626433832;  // decide_dt.c:16
static const double kDtMax = 1.;                              // decide_dt.c:25


// advective limit min over faces of h / (|u| + small) (decide_dt.c:49-64)
// advective limit min over faces of h / (|u| + small) (decide_dt.c:49-64):
// a workgroup takes kRedThreads consecutive columns of `rows` rows (16, or
// more when the partial minima would not fit the reduction buffer), each
// thread one column, so the loads of a thread are independent and in flight
// together; a min does not depend on the combination order

template <bool NT>
__device__ __forceinline__ void dt_partial_body(const double* __restrict__ ux, const double* __restrict__ uy,
                                                const double* __restrict__ hxxf, double hy, int nx, int ny,
                                                int rows, double* __restrict__ partial, int bx, int by, int gx) {
  __shared__ double sh[kRedThreads / 64];
  const double small = 1.e-8;
  double m = kDtMax;
  const int i = 1 + (int)(bx * kRedThreads + threadIdx.x);
  const int j0 = 1 + by * rows;
  if (i <= nx) {
    const double hx = hxxf[i - 1];
    // The column's minimum of h / (|u| + small) is h over its largest
    // |u| + small, bit for bit: a correctly rounded quotient does not grow
    // with its divisor, so the largest divisor gives the smallest quotient
    // (and fmax passes over a NaN as fmin does).  One division per field
    // and thread instead of one per cell.
    double mx = 0., my = 0.;
    // rows past ny repeat row ny (a repeated value leaves a max unchanged),
    // so the loads carry no branch and are all in flight together; with NT
    // they are streaming loads (read once; 124 -> 88 us at 4096 x 8192, see
    // stream_fields)
    auto ld = [](const double* p) {
      if constexpr (NT) return __builtin_nontemporal_load(p);
      else return *p;
    };
#pragma unroll 16
    for (int r = 0; r < rows; r++) {
      const int j = min(j0 + r, ny);
      if (i >= 2) mx = fmax(mx, fabs(ld(ux + L::xf(i, j, nx))) + small);
      my = fmax(my, fabs(ld(uy + L::cc(i, j, nx))) + small);
    }
    if (i >= 2) m = fmin(m, hx / mx);
    m = fmin(m, hy / my);
  }
  m = block_reduce<0>(m, sh);
  if (threadIdx.x == 0) partial[by * gx + bx] = m;
}

template <bool NT>
__global__ __launch_bounds__(kRedThreads) void k_dt_partial(const double* __restrict__ ux,
                                                           const double* __restrict__ uy,
                                                           const double* __restrict__ hxxf,
                                                           double hy, int nx, int ny, int rows,
                                                           double* __restrict__ partial) {
  dt_partial_body<NT>(ux, uy, hxxf, hy, nx, ny, rows, partial, blockIdx.x, blockIdx.y, gridDim.x);
}

// combine the limits (decide_dt.c:157-167)
__device__ __forceinline__ void dt_final_body(const double* __restrict__ partial, int n, double cadv, double dif0,
                                              double dif1, double dint, double* __restrict__ dt) {
  __shared__ double sh[kRedThreads / 64];
  double m = kDtMax;
  // four running minima, so that the loads are in flight together (a
  // minimum does not depend on the combination order)
  double m1 = kDtMax, m2 = kDtMax, m3 = kDtMax;
  const int s = (int)blockDim.x;
  int k = threadIdx.x;
  for (; k + 3 * s < n; k += 4 * s) {
    m = fmin(m, partial[k]);
    m1 = fmin(m1, partial[k + s]);
    m2 = fmin(m2, partial[k + 2 * s]);
    m3 = fmin(m3, partial[k + 3 * s]);
  }
  for (; k < n; k += s) m = fmin(m, partial[k]);
  m = fmin(fmin(m, m1), fmin(m2, m3));
  m = block_reduce<0>(m, sh);
  if (threadIdx.x == 0) {
    double r = m * cadv;
    r = fmin(r, dif0);
    r = fmin(r, dif1);
    r = fmin(r, dint);
    dt[0] = r;
  }
}

__global__ void k_dt_final(const double* __restrict__ partial, int n, double cadv,
                           double dif0, double dif1, double dint, double* __restrict__ dt) {
  dt_final_body(partial, n, cadv, dif0, dif1, dint, dt);
}

// ensembles: one map entry (case, local block) per workgroup / one
// workgroup per case
__global__ __launch_bounds__(kRedThreads) void k_dt_partial_batch(const DtArgs* __restrict__ tab,
                                                                 const uint2* __restrict__ map) {
  const uint2 m = map[blockIdx.x];
  const DtArgs& a = tab[m.x];
  dt_partial_body<false>(a.ux, a.uy, a.hxxf, a.hy, a.nx, a.ny, a.rows, a.partial, (int)(m.y % a.gx), (int)(m.y / a.gx), a.gx);
}

__global__ __launch_bounds__(kRedThreads) void k_dt_final_batch(const DtArgs* __restrict__ tab) {
  const DtArgs& a = tab[blockIdx.x];
  dt_final_body(a.partial, a.nblk, a.cadv, a.dif0, a.dif1, a.dint, a.dt);
}

// the arguments of one case's decide_dt (dt limits from the host, as
// compute_dt_device forms them)
int dt_args(const domain_t* d, DomainCtx* c, const fluid_t* f, const interface_t* s, DtArgs* a) {
  if (load_dt_coefs()) return 1;
  const int nx = (int)c->nx, ny = (int)c->ny;
  dt_limits(d, f, s, &a->dif0, &a->dif1, &a->dint);
  int rows = 0, nblk = 0;
  if (dt_grid(nx, ny, kRedSlots - 64, &rows, &nblk)) return 1;
  a->ux = (const double*)f->ux.data;
  a->uy = (const double*)f->uy.data;
  a->hxxf = c->hxxf;
  a->hy = c->hy;
  a->cadv = options().cadv;
  a->partial = c->red;
  a->dt = c->dt_dev;
  a->nx = nx;
  a->ny = ny;
  a->rows = rows;
  a->gx = (int)cdiv(nx, kRedThreads);
  a->nblk = a->gx * (int)cdiv(ny, rows);
  return 0;
}

// grid-only limits (decide_dt.c:79-130); host arithmetic identical to the
// reference, including pow(h, 3.) from libm
void dt_limits(const domain_t* d, const fluid_t* f, const interface_t* s, double* dif0,
               double* dif1, double* dint) {
  const Options& o = options();
  const size_t nx = d->mysizes[0];
  double gs[2] = {DBL_MAX, d->hy};
  for (size_t i = 2; i <= nx; i++) gs[0] = fmin(gs[0], d->hxxf[i - 1]);
  const double diffusivity = 1. / f->Re;
  double dif[2];
  for (int k = 0; k < 2; k++) {
    dif[k] = fmin(1., f->denr / f->visr) / diffusivity * 0.5 / 2. * (gs[k] * gs[k]);
    dif[k] *= o.cdif;
  }
  const double tension = 1. / s->We;
  double di = kDtMax;
  for (int k = 0; k < 2; k++) di = fmin(di, sqrt(1. / tension / 4. / kPiDt * pow(gs[k], 3.)));
  di *= o.cint;
  *dif0 = dif[0];
  *dif1 = dif[1];
  *dint = di;
}

// Rows per partial-reduction workgroup: 16, doubled until the partial
// minima fit `capacity` slots.  Fails (instead of looping on) a domain so
// wide that one workgroup row alone needs more slots.
int dt_grid(int nx, int ny, int capacity, int* rows, int* nblk) {
  const long gx = cdiv(nx, kRedThreads);
  int r = 16;
  while (gx * cdiv(ny, r) > capacity && r < ny) r *= 2;
  const long n = gx * cdiv(ny, r);
  SBFS_REQUIRE(n <= capacity, "decide_dt: %d columns need %ld partial minima, more than the %d slots", nx, n,
               capacity);
  *rows = r;
  *nblk = (int)n;
  return 0;
}

// The dt reduction of one case into `partial` (at least dt_grid's nblk
// doubles) and `dt_out` (one double), on the library stream.
int dt_launch(const domain_t* d, const DomainCtx* c, const fluid_t* f, const interface_t* s,
              int capacity, double* partial, double* dt_out) {
  if (load_dt_coefs()) return 1;
  const int nx = (int)c->nx, ny = (int)c->ny;
  double dif0, dif1, dint;
  dt_limits(d, f, s, &dif0, &dif1, &dint);
  int rows = 0, nblk = 0;
  if (dt_grid(nx, ny, capacity, &rows, &nblk)) return 1;
  const dim3 grd(cdiv(nx, kRedThreads), cdiv(ny, rows));
  SBFS_TIMED("k_dt_partial",
             hipLaunchKernelGGL((stream_fields(nx, ny) ? k_dt_partial<true> : k_dt_partial<false>), grd,
                                dim3(kRedThreads), 0, stream(),
                                (const double*)f->ux.data, (const double*)f->uy.data, c->hxxf,
                                c->hy, nx, ny, rows, partial));
  SBFS_TIMED("k_dt_final",
             hipLaunchKernelGGL(k_dt_final, dim3(1), dim3(kRedThreads), 0, stream(),
                                (const double*)partial, nblk, options().cadv, dif0, dif1, dint,
                                dt_out));
  return check_launch("decide_dt");
}

int compute_dt_device(const domain_t* d, DomainCtx* c, const fluid_t* f,
                      const interface_t* s) {
  return dt_launch(d, c, f, s, kRedSlots - 64, c->red, c->dt_dev);
}

// ---------------------------------------------------------------------------
// diagnostics: volume, centroid, rise velocity, max |div|, vof extrema, area
// ---------------------------------------------------------------------------
__global__ void k_diag_partial(const double* __restrict__ vof, const double* __restrict__ ux,
                               const double* __restrict__ uy, const double* __restrict__ xc,
                               const double* __restrict__ hxxf,
                               const double* __restrict__ jdxf,
                               const double* __restrict__ jdxc, double hy, int nx, int ny,
                               double* __restrict__ part) {
  __shared__ double sh[kRedThreads / 64];
  double vol = 0., sds = 0., xm = 0., um = 0., area = 0.;
  double vmin = 1., vmax = 0., dmax = 0.;
  const long total = (long)(nx + 1) * ny;
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < total;
       idx += (long)gridDim.x * blockDim.x) {
    const int i = (int)(idx % (nx + 1)) + 1, j = (int)(idx / (nx + 1)) + 1;
    // x-face area term, src/logging/interface.c:121-127
    area += fabs(vof[L::vof(i, j, nx)] - vof[L::vof(i - 1, j, nx)]) * (jdxf[i - 1] / hxxf[i - 1]);
    if (i > nx) continue;
    const double ds = jdxc[i - 1];
    const double lv = vof[L::vof(i, j, nx)];
    vmin = fmin(vmin, lv);
    vmax = fmax(vmax, lv);
    vol += lv * ds;
    sds += ds;
    xm += lv * xc[i] * ds;
    um += lv * (0.5 * ux[L::xf(i, j, nx)] + 0.5 * ux[L::xf(i + 1, j, nx)]) * ds;
    area += fabs(lv - vof[L::vof(i, j - 1, nx)]) * (jdxc[i - 1] / hy);
    // src/logging/divergence.c:44-58
    const double div = 1. / jdxc[i - 1] *
                       (-jdxf[i - 1] / hxxf[i - 1] * ux[L::xf(i, j, nx)] + jdxf[i] / hxxf[i] * ux[L::xf(i + 1, j, nx)]
                        - jdxc[i - 1] / hy * uy[L::cc(i, j, nx)] + jdxc[i - 1] / hy * uy[L::cc(i, j + 1, nx)]);
    dmax = fmax(dmax, fabs(div));
  }
  const double r0 = block_reduce<2>(vol, sh);
  const double r1 = block_reduce<2>(sds, sh);
  const double r2 = block_reduce<2>(xm, sh);
  const double r3 = block_reduce<2>(um, sh);
  const double r4 = block_reduce<2>(area, sh);
  const double r5 = block_reduce<0>(vmin, sh);
  const double r6 = block_reduce<1>(vmax, sh);
  const double r7 = block_reduce<1>(dmax, sh);
  if (threadIdx.x == 0) {
    double* o = part + 8 * blockIdx.x;
    o[0] = r0; o[1] = r1; o[2] = r2; o[3] = r3; o[4] = r4; o[5] = r5; o[6] = r6; o[7] = r7;
  }
}

__global__ void k_diag_final(const double* __restrict__ part, int n, double* __restrict__ out) {
  __shared__ double sh[kRedThreads / 64];
  double s[5] = {0., 0., 0., 0., 0.};
  double vmin = 1., vmax = 0., dmax = 0.;
  for (int k = threadIdx.x; k < n; k += blockDim.x) {
    for (int q = 0; q < 5; q++) s[q] += part[8 * k + q];
    vmin = fmin(vmin, part[8 * k + 5]);
    vmax = fmax(vmax, part[8 * k + 6]);
    dmax = fmax(dmax, part[8 * k + 7]);
  }
  double r[5];
  for (int q = 0; q < 5; q++) r[q] = block_reduce<2>(s[q], sh);
  vmin = block_reduce<0>(vmin, sh);
  vmax = block_reduce<1>(vmax, sh);
  dmax = block_reduce<1>(dmax, sh);
  if (threadIdx.x == 0) {
    const double vol = r[0];
    out[0] = vol;
    out[1] = vol != 0. ? r[2] / vol : 0.;
    out[2] = vol != 0. ? r[3] / vol : 0.;
    out[3] = dmax;
    out[4] = vmin;
    out[5] = vmax;
    out[6] = vol / r[1];
    out[7] = r[4];
  }
}

// ---------------------------------------------------------------------------
// staged integrator: the reference's composition, stage by stage
// ---------------------------------------------------------------------------
static int integrate_staged(const domain_t* domain, fluid_t* fluid, interface_t* interface,
                            double* dt) {
  if (decide_dt(domain, fluid, interface, dt)) return 1;
  for (size_t k = 0; k < 3; k++) {
    if (fluid_compute_density(domain, fluid, interface, 0)) return 1;
    if (fluid_compute_viscosity(domain, fluid, interface)) return 1;
    if (interface_compute_curvature_tensor(domain, interface)) return 1;
    if (interface_compute_force(domain, fluid, interface)) return 1;
    if (interface_update_vof(domain, k, *dt, fluid, interface)) return 1;
    if (fluid_compute_density(domain, fluid, interface, 1)) return 1;
    if (interface_compute_mass_flux(domain, fluid, interface)) return 1;
    if (fluid_predict_field(domain, k, *dt, fluid, interface)) return 1;
    const double dt_old = rk((k + 2) % 3).g * (*dt);
    const double dt_new = rk(k).g * (*dt);
    if (fluid_compute_potential(domain, dt_old, dt_new, fluid)) return 1;
    if (fluid_correct_velocity(domain, dt_old, dt_new, fluid)) return 1;
    if (fluid_update_pressure(domain, fluid)) return 1;
  }
  return 0;
}

int integrate_fused(const domain_t* domain, fluid_t* fluid, interface_t* interface, double* dt);

}  // namespace sbfs

using namespace sbfs;

extern "C" {

int decide_dt(const domain_t* domain, const fluid_t* fluid, const interface_t* interface,
              double* dt) {
  DomainCtx* c = domain_ctx(domain);
  if (!c) return 1;
  SBFS_REQUIRE(dt != nullptr, "decide_dt: null dt");
  if (check_array(domain, &fluid->ux, NA_UX, 8, "ux")) return 1;
  if (check_array(domain, &fluid->uy, NA_UY, 8, "uy")) return 1;
  if (compute_dt_device(domain, c, fluid, interface)) return 1;
  SBFS_HIP(hipMemcpyAsync(c->host_pin, c->dt_dev, sizeof(double), hipMemcpyDeviceToHost, stream()));
  SBFS_HIP(hipStreamSynchronize(stream()));
  *dt = c->host_pin[0];
  return 0;
}

int integrate(const domain_t* domain, fluid_t* fluid, interface_t* interface, double* dt) {
  if (init_runtime()) return 1;
  SBFS_REQUIRE(domain && fluid && interface && dt, "integrate: null argument");
  if (options().path == SBFS_PATH_FUSED) {
    if (options().precision == SBFS_PRECISION_MIXED ||
        fp64_march_addressable(domain->mysizes[0], domain->mysizes[1]))
      return integrate_fused(domain, fluid, interface, dt);
    // fields of 4 GiB or more: the FP64 step runs on the staged path
    SBFS_REQUIRE(options().curvature != SBFS_CURVATURE_LEARNED,
                 "integrate: the learned curvature runs on the fused path, whose FP64 step needs fields below 4 GiB "
                 "(%zu x %zu cells); use mixed precision", domain->mysizes[0], domain->mysizes[1]);
  }
  return integrate_staged(domain, fluid, interface, dt);
}

int sbfs_diagnostics(const domain_t* domain, const fluid_t* fluid, const interface_t* interface,
                     double out[SBFS_NDIAG]) {
  DomainCtx* c = domain_ctx(domain);
  if (!c) return 1;
  if (check_array(domain, &fluid->ux, NA_UX, 8, "ux")) return 1;
  if (check_array(domain, &fluid->uy, NA_UY, 8, "uy")) return 1;
  if (check_array(domain, &interface->vof, NA_VOF, 8, "vof")) return 1;
  const int nx = (int)c->nx, ny = (int)c->ny;
  const long total = (long)(nx + 1) * ny;
  const int nblk = (int)std::min<long>(kRedBlocks, (total + kRedThreads - 1) / kRedThreads);
  hipLaunchKernelGGL(k_diag_partial, dim3(nblk), dim3(kRedThreads), 0, stream(),
                     (const double*)interface->vof.data, (const double*)fluid->ux.data,
                     (const double*)fluid->uy.data, c->xc, c->hxxf, c->jdxf, c->jdxc, c->hy, nx,
                     ny, c->red);
  hipLaunchKernelGGL(k_diag_final, dim3(1), dim3(kRedThreads), 0, stream(), (const double*)c->red,
                     nblk, c->red + 8 * kRedBlocks);
  if (check_launch("sbfs_diagnostics")) return 1;
  SBFS_HIP(hipMemcpyAsync(c->host_pin, c->red + 8 * kRedBlocks, sizeof(double) * SBFS_NDIAG,
                          hipMemcpyDeviceToHost, stream()));
  SBFS_HIP(hipStreamSynchronize(stream()));
  for (int k = 0; k < SBFS_NDIAG; k++) out[k] = c->host_pin[k];
  return 0;
}

}  // extern "C"
