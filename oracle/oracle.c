/*
 * TEST INFRASTRUCTURE ONLY -- the CPU oracle.
 *
 * FP64 restatement of the reference's 2D two-phase time step
 * (NaokiHori/SimpleBubblyFlowSolver, src/integrate.c and everything it
 * calls).  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load this code; the product library never links it.
 *
 * Every function cites the reference file:line it follows.  Evaluation
 * order of every floating-point expression matches the reference's C
 * source (left-to-right, no FMA contraction: build with
 * -ffp-contract=off), so on the same input the oracle reproduces the
 * reference's rounding except where the reference calls FFTW (absent
 * here; replaced by the DCT definitions below) and where libm's tanh / log
 * / sqrt are involved (the oracle calls the same glibc functions).
 *
 * Two documented choices differ from a literal reading of the reference:
 *  - Poisson DCTs: FFTW's REDFT10 / REDFT01 are computed either by the
 *    O(N^2) definition (ORACLE_DCT_DIRECT) or by the N/2-point complex FFT
 *    algorithm the GPU uses (ORACLE_DCT_FFT).  FFTW itself chooses plans by
 *    timing (FFTW_PATIENT, src/fluid/compute_potential.c:200), so the
 *    reference's own rounding here is machine dependent.
 *  - Gauge of the singular x-mode 0 (src/tdm.c:92-94): by default mode 0
 *    takes the reference's singular branch x_{n-1} = 0 deterministically
 *    (ORACLE_GAUGE_PINNED); ORACLE_GAUGE_REFERENCE follows the num/den
 *    closure literally.  The two differ by a constant in psi only.
 */
#define _GNU_SOURCE
#include <math.h>
#include <float.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#if defined(_OPENMP)
#include <omp.h>
#endif
#include "oracle.h"

/* ------------------------------------------------------------------ */
/* constants                                                           */
/* ------------------------------------------------------------------ */

/* src/decide_dt.c:16 */
static const double k_pi_dt = 3.1415926535897932384626433832;
/* src/fluid/compute_potential.c:21 */
static const double k_pi_poisson = 3.14159265358979324;
/* src/decide_dt.c:25 */
static const double k_dt_max = 1.;
/* src/interface/common.c:4-14 */
static const double k_gauss_p[2] = {-0.2886751345948129, +0.2886751345948129};
static const double k_gauss_w[2] = {+0.5, +0.5};
static const double k_vofbeta = 1.;
static const double k_vofmin = 1.e-8;
/* src/param/boundary-condition.c:3-4 */
static const double k_uy_xm = 0.;
static const double k_uy_xp = 0.;

/* src/runge_kutta.c:8-15: {alpha, beta, gamma = alpha + beta} per stage */
static const double k_rk[3][3] = {
  {32. / 60., 0. / 60., 32. / 60. + 0. / 60.},
  {25. / 60., -17. / 60., 25. / 60. + -17. / 60.},
  {45. / 60., -25. / 60., 45. / 60. + -25. / 60.},
};
enum { RKA = 0, RKB = 1, RKG = 2 };

/* the table above, row by row: {alpha, beta, gamma} of stages 0, 1, 2 */
int oracle_rkcoefs(double out[9]) {
  for (int k = 0; k < 3; k++)
    for (int c = 0; c < 3; c++) out[3 * k + c] = k_rk[k][c];
  return 0;
}

/* ------------------------------------------------------------------ */
/* options                                                             */
/* ------------------------------------------------------------------ */

static struct {
  int dct;
  int tdm;
  int chunk;
  int gauge;
  int coefs_set;
  double cadv, cdif, cint;
} g_opt = {ORACLE_DCT_FFT, ORACLE_TDM_SEQUENTIAL, 32, ORACLE_GAUGE_PINNED,
           0, 0., 0., 0.};

int oracle_set_poisson(int dct_kind, int tdm_kind, int chunk, int gauge) {
  g_opt.dct = dct_kind;
  g_opt.tdm = tdm_kind;
  g_opt.chunk = chunk;
  g_opt.gauge = gauge;
  return 0;
}

int oracle_set_dt_coefs(double adv, double dif, double intf) {
  g_opt.cadv = adv;
  g_opt.cdif = dif;
  g_opt.cint = intf;
  g_opt.coefs_set = 1;
  return 0;
}

int oracle_set_threads(int nthreads) {
#if defined(_OPENMP)
  omp_set_num_threads(nthreads > 0 ? nthreads : 1);
#else
  (void)nthreads;
#endif
  return 0;
}

/* ------------------------------------------------------------------ */
/* layouts (include/array_macros/): x fastest, per-field ghost cells */
/* ------------------------------------------------------------------ */

/* nadds {x:{lo,hi}, y:{lo,hi}} of each field */
static const int NA_UX[2][2] = {{0, 1}, {1, 1}};
static const int NA_UY[2][2] = {{1, 1}, {1, 1}};
static const int NA_P[2][2] = {{1, 1}, {1, 1}};
static const int NA_DEN[2][2] = {{1, 1}, {1, 1}};
static const int NA_VISUX[2][2] = {{0, 1}, {1, 1}};
static const int NA_VISUY[2][2] = {{1, 1}, {1, 1}};
static const int NA_TXX[2][2] = {{0, 0}, {1, 1}};
static const int NA_TXY[2][2] = {{0, 1}, {1, 1}};
static const int NA_TYY[2][2] = {{0, 0}, {1, 1}};
static const int NA_SRCUX[2][2] = {{-1, 0}, {0, 0}};
static const int NA_SRCUY[2][2] = {{0, 0}, {0, 0}};
static const int NA_VOF[2][2] = {{1, 1}, {2, 2}};
static const int NA_IFRCX[2][2] = {{-1, 0}, {0, 0}};
static const int NA_IFRCY[2][2] = {{0, 0}, {0, 0}};
static const int NA_DVOF[2][2] = {{0, 1}, {1, 2}};
static const int NA_NORMAL[2][2] = {{0, 0}, {1, 1}};
static const int NA_CURV[2][2] = {{1, 1}, {1, 1}};
static const int NA_FLUXX[2][2] = {{0, 1}, {1, 1}};
static const int NA_FLUXY[2][2] = {{1, 1}, {1, 1}};
static const int NA_SRC[2][2] = {{0, 0}, {0, 0}};

/* element (i, j) with 1-based interior indices; lo = nadds lower bounds */
#define IDX(i, j, nxl, nxtot, nyl) ((size_t)((i) - 1 + (nxl)) + (size_t)(nxtot) * (size_t)((j) - 1 + (nyl)))
#define UXa(i, j)    ux[IDX(i, j, 0, nx + 1, 1)]
#define UYa(i, j)    uy[IDX(i, j, 1, nx + 2, 1)]
#define Pa(i, j)     p[IDX(i, j, 1, nx + 2, 1)]
#define PSIa(i, j)   psi[IDX(i, j, 1, nx + 2, 1)]
#define PSI0a(i, j)  psi0[IDX(i, j, 1, nx + 2, 1)]
#define PSI1a(i, j)  psi1[IDX(i, j, 1, nx + 2, 1)]
#define DENa(i, j)   den[IDX(i, j, 1, nx + 2, 1)]
#define DEN0a(i, j)  den0[IDX(i, j, 1, nx + 2, 1)]
#define DEN1a(i, j)  den1[IDX(i, j, 1, nx + 2, 1)]
#define VISUXa(i, j) visux[IDX(i, j, 0, nx + 1, 1)]
#define VISUYa(i, j) visuy[IDX(i, j, 1, nx + 2, 1)]
#define TXXa(i, j)   txx[IDX(i, j, 0, nx, 1)]
#define TXYa(i, j)   txy[IDX(i, j, 0, nx + 1, 1)]
#define TYYa(i, j)   tyy[IDX(i, j, 0, nx, 1)]
#define SRCUXa(arr, i, j) arr[IDX(i, j, -1, nx - 1, 0)]
#define SRCUYa(arr, i, j) arr[IDX(i, j, 0, nx, 0)]
#define VOFa(i, j)   vof[IDX(i, j, 1, nx + 2, 2)]
#define IFRCXa(i, j) ifrcx[IDX(i, j, -1, nx - 1, 0)]
#define IFRCYa(i, j) ifrcy[IDX(i, j, 0, nx, 0)]
#define DVOFa(i, j)  dvof[IDX(i, j, 0, nx + 1, 1)]
#define NORMALa(i, j) normal[IDX(i, j, 0, nx, 1)]
#define CURVa(i, j)  curv[IDX(i, j, 1, nx + 2, 1)]
#define FLUXXa(i, j) fluxx[IDX(i, j, 0, nx + 1, 1)]
#define FLUXYa(i, j) fluxy[IDX(i, j, 1, nx + 2, 1)]
#define SRCa(arr, i, j) arr[IDX(i, j, 0, nx, 0)]
/* 1D metrics (include/array_macros/domain/) */
#define HXXF(i) hxxf[(i) - 1]
#define HXXC(i) hxxc[(i) - 1]
#define JDXF(i) jdxf[(i) - 1]
#define JDXC(i) jdxc[(i) - 1]
#define XC(i)   xc[(i)]

#define DOMAIN_VARS                                   \
  const int nx = (int)domain->mysizes[0];             \
  const int ny = (int)domain->mysizes[1];             \
  const double * hxxf = domain->hxxf;                 \
  const double * hxxc = domain->hxxc;                 \
  const double * jdxf = domain->jdxf;                 \
  const double * jdxc = domain->jdxc;                 \
  const double hy = domain->hy;                       \
  (void)nx; (void)ny; (void)hxxf; (void)hxxc;         \
  (void)jdxf; (void)jdxc; (void)hy

/* ------------------------------------------------------------------ */
/* allocation (src/array.c:10-54)                                      */
/* ------------------------------------------------------------------ */

static int prepare(const domain_t * domain, const int nadds[2][2], size_t size,
                   array_t * a) {
  const size_t nxt = domain->mysizes[0] + nadds[0][0] + nadds[0][1];
  const size_t nyt = domain->mysizes[1] + nadds[1][0] + nadds[1][1];
  a->size = size;
  a->nadds = calloc(2, sizeof(int[2]));
  if (!a->nadds) return 1;
  for (int d = 0; d < 2; d++) {
    a->nadds[d][0] = nadds[d][0];
    a->nadds[d][1] = nadds[d][1];
  }
  a->datasize = nxt * nyt * size;
  a->data = calloc(nxt * nyt, size);
  return a->data ? 0 : 1;
}

static void release(array_t * a) {
  free(a->nadds);
  free(a->data);
  a->nadds = NULL;
  a->data = NULL;
}

/* uniform grid as initial_condition/main.py:15-24 builds it; metrics as
 * src/domain.c:98-146 */
int oracle_domain_create(size_t nxs, size_t nys, double lx, double ly,
                         domain_t * domain) {
  memset(domain, 0, sizeof(*domain));
  const int nx = (int)nxs;
  domain->glsizes[0] = domain->mysizes[0] = nxs;
  domain->glsizes[1] = domain->mysizes[1] = nys;
  domain->lengths[0] = lx;
  domain->lengths[1] = ly;
  double * xf = calloc(nxs + 1, sizeof(double));
  double * xc = calloc(nxs + 2, sizeof(double));
  double * hxxf = calloc(nxs + 1, sizeof(double));
  double * hxxc = calloc(nxs, sizeof(double));
  double * jdxf = calloc(nxs + 1, sizeof(double));
  double * jdxc = calloc(nxs, sizeof(double));
  if (!xf || !xc || !hxxf || !hxxc || !jdxf || !jdxc) return 1;
  /* numpy.linspace(0, lx, nx+1): step = lx / nx, x_k = k * step, last = lx */
  const double step = lx / (double)nx;
  for (int k = 0; k <= nx; k++) xf[k] = (double)k * step;
  xf[nx] = lx;
  /* xc = [0, 0.5 xf[:-1] + 0.5 xf[1:], lx] */
  xc[0] = 0.;
  for (int k = 0; k < nx; k++) xc[k + 1] = 0.5 * xf[k] + 0.5 * xf[k + 1];
  xc[nx + 1] = lx;
  const double hy = ly / (double)nys; /* src/domain.c:196 */
  for (int i = 1; i <= nx + 1; i++) HXXF(i) = XC(i) - XC(i - 1);   /* :104-106 */
  for (int i = 1; i <= nx; i++) HXXC(i) = xf[i] - xf[i - 1];        /* :116-118 */
  for (int i = 1; i <= nx + 1; i++) JDXF(i) = HXXF(i) * hy;         /* :129-131 */
  for (int i = 1; i <= nx; i++) JDXC(i) = HXXC(i) * hy;             /* :142-144 */
  domain->xf = xf;
  domain->xc = xc;
  domain->hxxf = hxxf;
  domain->hxxc = hxxc;
  domain->jdxf = jdxf;
  domain->jdxc = jdxc;
  domain->hy = hy;
  return 0;
}

int oracle_domain_destroy(domain_t * domain) {
  free(domain->xf);
  free(domain->xc);
  free(domain->hxxf);
  free(domain->hxxc);
  free(domain->jdxf);
  free(domain->jdxc);
  memset(domain, 0, sizeof(*domain));
  return 0;
}

/* src/fluid/init.c:28-57, :102-106 (without the NPY load) */
int oracle_fluid_create(const domain_t * domain, double Re, double Fr,
                        double denr, double visr, fluid_t * f) {
  memset(f, 0, sizeof(*f));
  int e = 0;
  e |= prepare(domain, NA_UX, sizeof(double), &f->ux);
  e |= prepare(domain, NA_UY, sizeof(double), &f->uy);
  e |= prepare(domain, NA_P, sizeof(double), &f->p);
  for (int n = 0; n < 2; n++) e |= prepare(domain, NA_P, sizeof(double), &f->psi[n]);
  for (int n = 0; n < 2; n++) e |= prepare(domain, NA_DEN, sizeof(double), &f->den[n]);
  e |= prepare(domain, NA_VISUX, sizeof(double), &f->visux);
  e |= prepare(domain, NA_VISUY, sizeof(double), &f->visuy);
  e |= prepare(domain, NA_TXX, sizeof(double), &f->txx);
  e |= prepare(domain, NA_TXY, sizeof(double), &f->txy);
  e |= prepare(domain, NA_TYY, sizeof(double), &f->tyy);
  for (int n = 0; n < 3; n++) {
    e |= prepare(domain, NA_SRCUX, sizeof(double), &f->srcux[n]);
    e |= prepare(domain, NA_SRCUY, sizeof(double), &f->srcuy[n]);
  }
  f->Re = Re;
  f->Fr = Fr;
  f->denr = denr;
  f->visr = visr;
  f->refden = fmin(1., denr);
  return e;
}

int oracle_fluid_destroy(fluid_t * f) {
  release(&f->ux);
  release(&f->uy);
  release(&f->p);
  for (int n = 0; n < 2; n++) release(&f->psi[n]);
  for (int n = 0; n < 2; n++) release(&f->den[n]);
  release(&f->visux);
  release(&f->visuy);
  release(&f->txx);
  release(&f->txy);
  release(&f->tyy);
  for (int n = 0; n < 3; n++) {
    release(&f->srcux[n]);
    release(&f->srcuy[n]);
  }
  return 0;
}

/* src/interface/init.c:22-38 (without the NPY load) */
int oracle_interface_create(const domain_t * domain, double We,
                            interface_t * s) {
  memset(s, 0, sizeof(*s));
  int e = 0;
  e |= prepare(domain, NA_VOF, sizeof(double), &s->vof);
  e |= prepare(domain, NA_IFRCX, sizeof(double), &s->ifrcx);
  e |= prepare(domain, NA_IFRCY, sizeof(double), &s->ifrcy);
  e |= prepare(domain, NA_DVOF, sizeof(vector_t), &s->dvof);
  e |= prepare(domain, NA_NORMAL, sizeof(normal_t), &s->normal);
  e |= prepare(domain, NA_CURV, sizeof(double), &s->curv);
  e |= prepare(domain, NA_FLUXX, sizeof(double), &s->fluxx);
  e |= prepare(domain, NA_FLUXY, sizeof(double), &s->fluxy);
  for (int n = 0; n < 2; n++) e |= prepare(domain, NA_SRC, sizeof(double), &s->src[n]);
  s->We = We;
  return e;
}

int oracle_interface_destroy(interface_t * s) {
  release(&s->vof);
  release(&s->ifrcx);
  release(&s->ifrcy);
  release(&s->dvof);
  release(&s->normal);
  release(&s->curv);
  release(&s->fluxx);
  release(&s->fluxy);
  for (int n = 0; n < 2; n++) release(&s->src[n]);
  return 0;
}

/* ------------------------------------------------------------------ */
/* periodic y halo (src/halo.c:18-76 on a single rank)                 */
/* ------------------------------------------------------------------ */

static void halo_y(const domain_t * domain, array_t * a) {
  const size_t w = (domain->mysizes[0] + a->nadds[0][0] + a->nadds[0][1]) * a->size;
  const size_t rows = domain->mysizes[1] + a->nadds[1][0] + a->nadds[1][1];
  const size_t nh = (size_t)a->nadds[1][0];
  char * d = a->data;
  /* bottom halo <- last interior rows, then top halo <- first interior rows */
  memmove(d, d + w * (rows - 2 * nh), w * nh);
  memmove(d + w * (rows - nh), d + w * nh, w * nh);
}

/* ------------------------------------------------------------------ */
/* boundary conditions                                                 */
/* ------------------------------------------------------------------ */

/* src/fluid/boundary/ux.c:15-42 */
int oracle_fluid_update_boundaries_ux(const domain_t * domain, array_t * a) {
  DOMAIN_VARS;
  double * ux = a->data;
  for (int j = 1; j <= ny; j++) {
    UXa(1, j) = 0.;
    UXa(nx + 1, j) = 0.;
  }
  halo_y(domain, a);
  return 0;
}

/* src/fluid/boundary/uy.c:16-43 */
int oracle_fluid_update_boundaries_uy(const domain_t * domain, array_t * a) {
  DOMAIN_VARS;
  double * uy = a->data;
  for (int j = 1; j <= ny; j++) {
    UYa(0, j) = k_uy_xm;
    UYa(nx + 1, j) = k_uy_xp;
  }
  halo_y(domain, a);
  return 0;
}

/* src/fluid/boundary/p.c:15-42 */
int oracle_fluid_update_boundaries_p(const domain_t * domain, array_t * a) {
  DOMAIN_VARS;
  double * p = a->data;
  for (int j = 1; j <= ny; j++) {
    Pa(0, j) = Pa(1, j);
    Pa(nx + 1, j) = Pa(nx, j);
  }
  halo_y(domain, a);
  return 0;
}

/* src/fluid/boundary/psi.c:15-42 */
int oracle_fluid_update_boundaries_psi(const domain_t * domain, array_t * a) {
  DOMAIN_VARS;
  double * psi = a->data;
  for (int j = 1; j <= ny; j++) {
    PSIa(0, j) = PSIa(1, j);
    PSIa(nx + 1, j) = PSIa(nx, j);
  }
  halo_y(domain, a);
  return 0;
}

/* src/interface/boundary.c:15-42 */
int oracle_interface_update_boundaries_vof(const domain_t * domain, array_t * a) {
  DOMAIN_VARS;
  double * vof = a->data;
  for (int j = 1; j <= ny; j++) {
    VOFa(0, j) = 0.;
    VOFa(nx + 1, j) = 0.;
  }
  halo_y(domain, a);
  return 0;
}

/* ------------------------------------------------------------------ */
/* decide_dt (src/decide_dt.c:34-169)                                  */
/* ------------------------------------------------------------------ */

static int load_coef(const char * name, double * v) {
  const char * s = getenv(name);
  if (!s) {
    fprintf(stderr, "%s not found\n", name);
    return 1;
  }
  *v = strtod(s, NULL);
  return 0;
}

int oracle_decide_dt(const domain_t * domain, const fluid_t * fluid,
                     const interface_t * interface, double * dt) {
  DOMAIN_VARS;
  if (!g_opt.coefs_set) {
    if (load_coef("coef_dt_adv", &g_opt.cadv)) return 1;
    if (load_coef("coef_dt_dif", &g_opt.cdif)) return 1;
    if (load_coef("coef_dt_int", &g_opt.cint)) return 1;
    g_opt.coefs_set = 1;
  }
  const double * ux = fluid->ux.data;
  const double * uy = fluid->uy.data;
  /* advective limit, :34-70; fmin is exact, so any order gives the same */
  const double small = 1.e-8;
  double dt_adv = k_dt_max;
  for (int j = 1; j <= ny; j++)
    for (int i = 2; i <= nx; i++) {
      const double vel = fabs(UXa(i, j)) + small;
      dt_adv = fmin(dt_adv, HXXF(i) / vel);
    }
  for (int j = 1; j <= ny; j++)
    for (int i = 1; i <= nx; i++) {
      const double vel = fabs(UYa(i, j)) + small;
      dt_adv = fmin(dt_adv, hy / vel);
    }
  dt_adv *= g_opt.cadv;
  /* grid sizes, :89-96 */
  double gs[2] = {DBL_MAX, hy};
  for (int i = 2; i <= nx; i++) gs[0] = fmin(gs[0], HXXF(i));
  /* diffusive limit, :98-103.  pow(h, 2.) is folded to h * h by GCC. */
  const double diffusivity = 1. / fluid->Re;
  double dt_dif[2];
  for (int d = 0; d < 2; d++) {
    dt_dif[d] = fmin(1., fluid->denr / fluid->visr) / diffusivity * 0.5 / 2. * (gs[d] * gs[d]);
    dt_dif[d] *= g_opt.cdif;
  }
  /* interfacial limit, :123-128 */
  const double tension = 1. / interface->We;
  double dt_int = k_dt_max;
  for (int d = 0; d < 2; d++)
    dt_int = fmin(dt_int, sqrt(1. / tension / 4. / k_pi_dt * pow(gs[d], 3.)));
  dt_int *= g_opt.cint;
  /* :162-167 */
  double r = dt_adv;
  r = fmin(r, dt_dif[0]);
  r = fmin(r, dt_dif[1]);
  r = fmin(r, dt_int);
  *dt = r;
  return 0;
}

/* ------------------------------------------------------------------ */
/* density and viscosity                                               */
/* ------------------------------------------------------------------ */

/* src/fluid/density.c:9-31 */
int oracle_fluid_compute_density(const domain_t * domain, fluid_t * fluid,
                                 const interface_t * interface, size_t index) {
  DOMAIN_VARS;
  double * den = fluid->den[index].data;
  const double * vof = interface->vof.data;
  const double denr = fluid->denr;
  const double lo = fmin(1., denr);
  const double hi = fmax(1., denr);
#pragma omp parallel for schedule(static)
  for (int j = 0; j <= ny + 1; j++)
    for (int i = 0; i <= nx + 1; i++) {
      double d = 1. + (denr - 1.) * VOFa(i, j);
      d = fmax(lo, d);
      d = fmin(hi, d);
      DENa(i, j) = d;
    }
  return 0;
}

/* src/fluid/viscosity.c:13-23 */
static inline double vis_of(double visr, double lo, double hi, double v) {
  double r = 1. + (visr - 1.) * v;
  r = fmax(lo, r);
  r = fmin(hi, r);
  return r;
}

/* src/fluid/viscosity.c:25-108 */
int oracle_fluid_compute_viscosity(const domain_t * domain, fluid_t * fluid,
                                   const interface_t * interface) {
  DOMAIN_VARS;
  const double visr = fluid->visr;
  const double lo = fmin(1., visr);
  const double hi = fmax(1., visr);
  const double * vof = interface->vof.data;
  double * visux = fluid->visux.data;
  double * visuy = fluid->visuy.data;
#pragma omp parallel for schedule(static)
  for (int j = 1; j <= ny; j++) {
    VISUXa(1, j) = vis_of(visr, lo, hi, VOFa(0, j));
    for (int i = 2; i <= nx; i++)
      VISUXa(i, j) = vis_of(visr, lo, hi, 0.5 * VOFa(i - 1, j) + 0.5 * VOFa(i, j));
    VISUXa(nx + 1, j) = vis_of(visr, lo, hi, VOFa(nx + 1, j));
  }
  halo_y(domain, &fluid->visux);
#pragma omp parallel for schedule(static)
  for (int j = 1; j <= ny; j++)
    for (int i = 0; i <= nx + 1; i++)
      VISUYa(i, j) = vis_of(visr, lo, hi, 0.5 * VOFa(i, j - 1) + 0.5 * VOFa(i, j));
  halo_y(domain, &fluid->visuy);
  return 0;
}

/* ------------------------------------------------------------------ */
/* interface geometry (src/interface/curvature_tensor.c)               */
/* ------------------------------------------------------------------ */

int oracle_interface_compute_curvature_tensor(const domain_t * domain,
                                              interface_t * interface) {
  DOMAIN_VARS;
  const double * vof = interface->vof.data;
  vector_t * dvof = interface->dvof.data;
  normal_t * normal = interface->normal.data;
  double * curv = interface->curv.data;
  /* corner unit gradients, :16-49 (pow(g, 2.) folded to g * g) */
#pragma omp parallel for schedule(static)
  for (int j = 0; j <= ny + 2; j++)
    for (int i = 1; i <= nx + 1; i++) {
      const double gx = 1. / HXXF(i) * (-VOFa(i - 1, j - 1) + VOFa(i, j - 1) - VOFa(i - 1, j) + VOFa(i, j));
      const double gy = 1. / hy * (-VOFa(i - 1, j - 1) - VOFa(i, j - 1) + VOFa(i - 1, j) + VOFa(i, j));
      const double norm = sqrt(gx * gx + gy * gy);
      const double ninv = 1. / fmax(norm, DBL_EPSILON);
      DVOFa(i, j)[0] = gx * ninv;
      DVOFa(i, j)[1] = gy * ninv;
    }
  /* cell normals and THINC intercepts inside the band, :51-97 */
#pragma omp parallel for schedule(static)
  for (int j = 0; j <= ny + 1; j++)
    for (int i = 1; i <= nx; i++) {
      const double lv = VOFa(i, j);
      if (lv < k_vofmin || 1. - k_vofmin < lv) continue;
      double nxv = DVOFa(i, j)[0] + DVOFa(i + 1, j)[0] + DVOFa(i, j + 1)[0] + DVOFa(i + 1, j + 1)[0];
      double nyv = DVOFa(i, j)[1] + DVOFa(i + 1, j)[1] + DVOFa(i, j + 1)[1] + DVOFa(i + 1, j + 1)[1];
      nxv /= HXXC(i);
      nyv /= hy;
      const double norm = sqrt(nxv * nxv + nyv * nyv);
      const double ninv = 1. / fmax(norm, DBL_EPSILON);
      nxv *= ninv;
      nyv *= ninv;
      NORMALa(i, j)[0] = nxv;
      NORMALa(i, j)[1] = nyv;
      NORMALa(i, j)[2] = -0.5 / k_vofbeta * log(1. / lv - 1.);
    }
  /* curvature = -div(corner normals), :100-135 */
#pragma omp parallel for schedule(static)
  for (int j = 0; j <= ny + 1; j++)
    for (int i = 1; i <= nx; i++) {
      const double hm = HXXF(i), hp = HXXF(i + 1);
      const double jm = JDXF(i), j0 = JDXC(i), jp = JDXF(i + 1);
      const double nxm = 0.5 * DVOFa(i, j)[0] + 0.5 * DVOFa(i, j + 1)[0];
      const double nxp = 0.5 * DVOFa(i + 1, j)[0] + 0.5 * DVOFa(i + 1, j + 1)[0];
      const double nym = 0.5 * DVOFa(i, j)[1] + 0.5 * DVOFa(i + 1, j)[1];
      const double nyp = 0.5 * DVOFa(i, j + 1)[1] + 0.5 * DVOFa(i + 1, j + 1)[1];
      const double div = 1. / j0 * (-jm / hm * nxm + jp / hp * nxp - j0 / hy * nym + j0 / hy * nyp);
      CURVa(i, j) = -1. * div;
    }
  return 0;
}

/* ------------------------------------------------------------------ */
/* surface tension (src/interface/force.c:13-98)                       */
/* ------------------------------------------------------------------ */

int oracle_interface_compute_force(const domain_t * domain, const fluid_t * fluid,
                                   interface_t * interface) {
  DOMAIN_VARS;
  const double tension = 1. / interface->We;
  const double refdeninv = 2. / (1. + fluid->denr);
  const double * den = fluid->den[0].data;
  const double * vof = interface->vof.data;
  const double * curv = interface->curv.data;
  double * ifrcx = interface->ifrcx.data;
  double * ifrcy = interface->ifrcy.data;
#pragma omp parallel for schedule(static)
  for (int j = 1; j <= ny; j++)
    for (int i = 2; i <= nx; i++) {
      const double d = 0.5 * DENa(i - 1, j) + 0.5 * DENa(i, j);
      const double k = 0.5 * CURVa(i - 1, j) + 0.5 * CURVa(i, j);
      const double delta = 1. / HXXF(i) * (-VOFa(i - 1, j) + VOFa(i, j));
      IFRCXa(i, j) = d * refdeninv * tension * k * delta;
    }
#pragma omp parallel for schedule(static)
  for (int j = 1; j <= ny; j++)
    for (int i = 1; i <= nx; i++) {
      const double d = 0.5 * DENa(i, j - 1) + 0.5 * DENa(i, j);
      const double k = 0.5 * CURVa(i, j - 1) + 0.5 * CURVa(i, j);
      const double delta = 1. / hy * (-VOFa(i, j - 1) + VOFa(i, j));
      IFRCYa(i, j) = d * refdeninv * tension * k * delta;
    }
  return 0;
}

/* ------------------------------------------------------------------ */
/* vof transport (src/interface/update/)                              */
/* ------------------------------------------------------------------ */

/* src/interface/update/main.c:17-35 */
static inline double thinc_indicator(const double * n, double x, double y) {
  const double sf = n[0] * x + n[1] * y + n[2];
  return 0.5 * (1. + tanh(k_vofbeta * sf));
}

int oracle_interface_update_vof(const domain_t * domain, size_t rkstep, double dt,
                                const fluid_t * fluid, interface_t * interface) {
  DOMAIN_VARS;
  /* stash the previous alpha source as beta, main.c:37-52 */
  if (0 != rkstep) {
    void * t = interface->src[RKA].data;
    interface->src[RKA].data = interface->src[RKB].data;
    interface->src[RKB].data = t;
  }
  const double * ux = fluid->ux.data;
  const double * uy = fluid->uy.data;
  double * vof = interface->vof.data;
  const normal_t * normal = interface->normal.data;
  double * fluxx = interface->fluxx.data;
  double * fluxy = interface->fluxy.data;
  /* x fluxes, fluxx.c:39-73 */
#pragma omp parallel for schedule(static)
  for (int j = 1; j <= ny; j++)
    for (int i = 2; i <= nx; i++) {
      const double vel = UXa(i, j);
      const int ii = vel < 0. ? i : i - 1;
      const double x = vel < 0. ? -0.5 : +0.5;
      const double lv = VOFa(ii, j);
      if (lv < k_vofmin || 1. - k_vofmin < lv) {
        FLUXXa(i, j) = vel * lv;
        continue;
      }
      double fl = 0.;
      for (int g = 0; g < 2; g++)
        fl += k_gauss_w[g] * thinc_indicator(NORMALa(ii, j), x, k_gauss_p[g]);
      FLUXXa(i, j) = vel * fl;
    }
  for (int j = 1; j <= ny; j++) {
    FLUXXa(1, j) = 0.;
    FLUXXa(nx + 1, j) = 0.;
  }
  halo_y(domain, &interface->fluxx);
  /* y fluxes, fluxy.c:39-73 */
#pragma omp parallel for schedule(static)
  for (int j = 1; j <= ny; j++)
    for (int i = 1; i <= nx; i++) {
      const double vel = UYa(i, j);
      const int jj = vel < 0. ? j : j - 1;
      const double y = vel < 0. ? -0.5 : +0.5;
      const double lv = VOFa(i, jj);
      if (lv < k_vofmin || 1. - k_vofmin < lv) {
        FLUXYa(i, j) = vel * lv;
        continue;
      }
      double fl = 0.;
      for (int g = 0; g < 2; g++)
        fl += k_gauss_w[g] * thinc_indicator(NORMALa(i, jj), k_gauss_p[g], y);
      FLUXYa(i, j) = vel * fl;
    }
  for (int j = 1; j <= ny; j++) {
    FLUXYa(0, j) = 0.;
    FLUXYa(nx + 1, j) = 0.;
  }
  halo_y(domain, &interface->fluxy);
  /* flux divergence, main.c:54-88 */
  double * srca = interface->src[RKA].data;
  const double * srcb = interface->src[RKB].data;
#pragma omp parallel for schedule(static)
  for (int j = 1; j <= ny; j++)
    for (int i = 1; i <= nx; i++) {
      const double hm = HXXF(i), hp = HXXF(i + 1);
      const double jm = JDXF(i), j0 = JDXC(i), jp = JDXF(i + 1);
      SRCa(srca, i, j) = 1. / j0 * (+jm / hm * FLUXXa(i, j) - jp / hp * FLUXXa(i + 1, j)
                                    + j0 / hy * FLUXYa(i, j) - j0 / hy * FLUXYa(i, j + 1));
    }
  /* vof update, main.c:90-120 */
  {
    const double ca = k_rk[rkstep][RKA];
#pragma omp parallel for schedule(static)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++) VOFa(i, j) += dt * ca * SRCa(srca, i, j);
  }
  if (0 != rkstep) {
    const double cb = k_rk[rkstep][RKB];
#pragma omp parallel for schedule(static)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++) VOFa(i, j) += dt * cb * SRCa(srcb, i, j);
  }
  oracle_interface_update_boundaries_vof(domain, &interface->vof);
  return 0;
}

/* src/interface/mass_flux.c:10-58 */
int oracle_interface_compute_mass_flux(const domain_t * domain, const fluid_t * fluid,
                                       interface_t * interface) {
  DOMAIN_VARS;
  const double * ux = fluid->ux.data;
  const double * uy = fluid->uy.data;
  double * fluxx = interface->fluxx.data;
  double * fluxy = interface->fluxy.data;
  const double denr = fluid->denr;
#pragma omp parallel for schedule(static)
  for (int j = 0; j <= ny + 1; j++)
    for (int i = 1; i <= nx + 1; i++) FLUXXa(i, j) = UXa(i, j) + (denr - 1.) * FLUXXa(i, j);
#pragma omp parallel for schedule(static)
  for (int j = 0; j <= ny + 1; j++)
    for (int i = 0; i <= nx + 1; i++) FLUXYa(i, j) = UYa(i, j) + (denr - 1.) * FLUXYa(i, j);
  return 0;
}

/* ------------------------------------------------------------------ */
/* momentum predictor (src/fluid/predict/)                            */
/* ------------------------------------------------------------------ */

static void reset_fluid_srcs(size_t rkstep, array_t * a, array_t * b, array_t * g) {
  /* src/fluid/predict/main.c:9-29 */
  if (0 != rkstep) {
    void * t = a->data;
    a->data = b->data;
    b->data = t;
  }
  memset(a->data, 0, a->datasize);
  memset(g->data, 0, g->datasize);
}

int oracle_fluid_predict_field(const domain_t * domain, size_t rkstep, double dt,
                               fluid_t * fluid, const interface_t * interface) {
  DOMAIN_VARS;
  reset_fluid_srcs(rkstep, &fluid->srcux[RKA], &fluid->srcux[RKB], &fluid->srcux[RKG]);
  reset_fluid_srcs(rkstep, &fluid->srcuy[RKA], &fluid->srcuy[RKB], &fluid->srcuy[RKG]);
  double * ux = fluid->ux.data;
  double * uy = fluid->uy.data;
  const double * p = fluid->p.data;
  const double * visux = fluid->visux.data;
  const double * visuy = fluid->visuy.data;
  double * txx = fluid->txx.data;
  double * txy = fluid->txy.data;
  double * tyy = fluid->tyy.data;
  const double * fluxx = interface->fluxx.data;
  const double * fluxy = interface->fluxy.data;
  const double * ifrcx = interface->ifrcx.data;
  const double * ifrcy = interface->ifrcy.data;
  const double * den0 = fluid->den[0].data;
  const double * den1 = fluid->den[1].data;
  /* txx at centres, txx.c:11-40 */
#pragma omp parallel for schedule(static)
  for (int j = 1; j <= ny; j++)
    for (int i = 1; i <= nx; i++) {
      const double vis = 0.5 * VISUXa(i, j) + 0.5 * VISUXa(i + 1, j);
      const double l = 1. / HXXC(i) * (-UXa(i, j) + UXa(i + 1, j));
      TXXa(i, j) = vis * l + vis * l;
    }
  halo_y(domain, &fluid->txx);
  /* txy at corners, txy.c:13-50 */
#pragma omp parallel for schedule(static)
  for (int j = 1; j <= ny; j++)
    for (int i = 1; i <= nx + 1; i++) {
      const double vx = 0.5 * VISUXa(i, j - 1) + 0.5 * VISUXa(i, j);
      const double vy = 0.5 * VISUYa(i - 1, j) + 0.5 * VISUYa(i, j);
      const double lxy = 1. / hy * (-UXa(i, j - 1) + UXa(i, j));
      const double lyx = 1. / HXXF(i) * (-UYa(i - 1, j) + UYa(i, j));
      TXYa(i, j) = vx * lxy + vy * lyx;
    }
  halo_y(domain, &fluid->txy);
  /* tyy at centres, tyy.c:10-38 */
#pragma omp parallel for schedule(static)
  for (int j = 1; j <= ny; j++)
    for (int i = 1; i <= nx; i++) {
      const double vis = 0.5 * VISUYa(i, j) + 0.5 * VISUYa(i, j + 1);
      const double l = 1. / hy * (-UYa(i, j) + UYa(i, j + 1));
      TYYa(i, j) = vis * l + vis * l;
    }
  halo_y(domain, &fluid->tyy);
  const double diffusivity = 1. / fluid->Re;
  const double grav = -1. / fluid->Fr / fluid->Fr;
  /* ux sources, predict/ux.c:29-226; each term is its own read-modify-write */
  {
    double * sa = fluid->srcux[RKA].data;
    double * sg = fluid->srcux[RKG].data;
#pragma omp parallel for schedule(static)
    for (int j = 1; j <= ny; j++)
      for (int i = 2; i <= nx; i++) {
        double s = SRCUXa(sa, i, j);
        /* advection in x, :29-61 */
        {
          const double mm = 0.5 * JDXF(i - 1) / HXXF(i - 1) * FLUXXa(i - 1, j) + 0.5 * JDXF(i) / HXXF(i) * FLUXXa(i, j);
          const double mp = 0.5 * JDXF(i) / HXXF(i) * FLUXXa(i, j) + 0.5 * JDXF(i + 1) / HXXF(i + 1) * FLUXXa(i + 1, j);
          const double um = 0.5 * UXa(i - 1, j) + 0.5 * UXa(i, j);
          const double up = 0.5 * UXa(i, j) + 0.5 * UXa(i + 1, j);
          s -= 1. / JDXF(i) * (-mm * um + mp * up);
        }
        /* advection in y, :63-93 */
        {
          const double mm = 0.5 * JDXC(i - 1) / hy * FLUXYa(i - 1, j) + 0.5 * JDXC(i) / hy * FLUXYa(i, j);
          const double mp = 0.5 * JDXC(i - 1) / hy * FLUXYa(i - 1, j + 1) + 0.5 * JDXC(i) / hy * FLUXYa(i, j + 1);
          const double um = 0.5 * UXa(i, j - 1) + 0.5 * UXa(i, j);
          const double up = 0.5 * UXa(i, j) + 0.5 * UXa(i, j + 1);
          s -= 1. / JDXF(i) * (-mm * um + mp * up);
        }
        /* diffusion in x, :95-121 */
        s += diffusivity / JDXF(i) * (-JDXC(i - 1) / HXXC(i - 1) * TXXa(i - 1, j) + JDXC(i) / HXXC(i) * TXXa(i, j));
        /* diffusion in y, :123-144 */
        s += diffusivity / JDXF(i) * (-JDXF(i) / hy * TXYa(i, j) + JDXF(i) / hy * TXYa(i, j + 1));
        /* pressure gradient -> gamma source, :146-162 */
        SRCUXa(sg, i, j) -= 1. / HXXF(i) * (-Pa(i - 1, j) + Pa(i, j));
        /* surface tension, :164-175 */
        s += IFRCXa(i, j);
        /* gravity along -x, :177-192 */
        s += grav * (0.5 * DEN0a(i - 1, j) + 0.5 * DEN0a(i, j));
        SRCUXa(sa, i, j) = s;
      }
  }
  /* uy sources, predict/uy.c:29-204 */
  {
    double * sa = fluid->srcuy[RKA].data;
    double * sg = fluid->srcuy[RKG].data;
#pragma omp parallel for schedule(static)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++) {
        double s = SRCUYa(sa, i, j);
        /* advection in x, :29-61 */
        {
          const double mm = 0.5 * JDXF(i) / HXXF(i) * FLUXXa(i, j - 1) + 0.5 * JDXF(i) / HXXF(i) * FLUXXa(i, j);
          const double mp = 0.5 * JDXF(i + 1) / HXXF(i + 1) * FLUXXa(i + 1, j - 1) + 0.5 * JDXF(i + 1) / HXXF(i + 1) * FLUXXa(i + 1, j);
          const double um = 0.5 * UYa(i - 1, j) + 0.5 * UYa(i, j);
          const double up = 0.5 * UYa(i, j) + 0.5 * UYa(i + 1, j);
          s -= 1. / JDXC(i) * (-mm * um + mp * up);
        }
        /* advection in y, :63-90 */
        {
          const double mm = 0.5 * JDXC(i) / hy * FLUXYa(i, j - 1) + 0.5 * JDXC(i) / hy * FLUXYa(i, j);
          const double mp = 0.5 * JDXC(i) / hy * FLUXYa(i, j) + 0.5 * JDXC(i) / hy * FLUXYa(i, j + 1);
          const double um = 0.5 * UYa(i, j - 1) + 0.5 * UYa(i, j);
          const double up = 0.5 * UYa(i, j) + 0.5 * UYa(i, j + 1);
          s -= 1. / JDXC(i) * (-mm * um + mp * up);
        }
        /* diffusion in x, :92-118 */
        s += diffusivity / JDXC(i) * (-JDXF(i) / HXXF(i) * TXYa(i, j) + JDXF(i + 1) / HXXF(i + 1) * TXYa(i + 1, j));
        /* diffusion in y, :120-141 */
        s += diffusivity / JDXC(i) * (-JDXC(i) / hy * TYYa(i, j - 1) + JDXC(i) / hy * TYYa(i, j));
        /* pressure gradient, :143-159 */
        SRCUYa(sg, i, j) -= 1. / hy * (-Pa(i, j - 1) + Pa(i, j));
        /* surface tension, :161-172 */
        s += IFRCYa(i, j);
        SRCUYa(sa, i, j) = s;
      }
  }
  /* predict ux, predict/ux.c:236-273 */
  {
    const double ca = k_rk[rkstep][RKA], cb = k_rk[rkstep][RKB], cg = k_rk[rkstep][RKG];
    const double * sa = fluid->srcux[RKA].data;
    const double * sb = fluid->srcux[RKB].data;
    const double * sg = fluid->srcux[RKG].data;
#pragma omp parallel for schedule(static)
    for (int j = 1; j <= ny; j++)
      for (int i = 2; i <= nx; i++) {
        const double d0 = 0.5 * DEN0a(i - 1, j) + 0.5 * DEN0a(i, j);
        UXa(i, j) = d0 * UXa(i, j) + ca * dt * SRCUXa(sa, i, j) + cb * dt * SRCUXa(sb, i, j) + cg * dt * SRCUXa(sg, i, j);
      }
#pragma omp parallel for schedule(static)
    for (int j = 1; j <= ny; j++)
      for (int i = 2; i <= nx; i++) {
        const double d1 = 0.5 * DEN1a(i - 1, j) + 0.5 * DEN1a(i, j);
        UXa(i, j) /= d1;
      }
    oracle_fluid_update_boundaries_ux(domain, &fluid->ux);
  }
  /* predict uy, predict/uy.c:214-251 */
  {
    const double ca = k_rk[rkstep][RKA], cb = k_rk[rkstep][RKB], cg = k_rk[rkstep][RKG];
    const double * sa = fluid->srcuy[RKA].data;
    const double * sb = fluid->srcuy[RKB].data;
    const double * sg = fluid->srcuy[RKG].data;
#pragma omp parallel for schedule(static)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++) {
        const double d0 = 0.5 * DEN0a(i, j - 1) + 0.5 * DEN0a(i, j);
        UYa(i, j) = d0 * UYa(i, j) + ca * dt * SRCUYa(sa, i, j) + cb * dt * SRCUYa(sb, i, j) + cg * dt * SRCUYa(sg, i, j);
      }
#pragma omp parallel for schedule(static)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++) {
        const double d1 = 0.5 * DEN1a(i, j - 1) + 0.5 * DEN1a(i, j);
        UYa(i, j) /= d1;
      }
    oracle_fluid_update_boundaries_uy(domain, &fluid->uy);
  }
  return 0;
}

/* ------------------------------------------------------------------ */
/* DCTs                                                                */
/* ------------------------------------------------------------------ */

static const long double k_pi_l = 3.141592653589793238462643383279502884L;

/* FFTW REDFT10 by definition: Y_k = 2 sum_j X_j cos(pi (2j+1) k / (2N)) */
static void dct2_direct(size_t n, const double * x, double * y) {
  for (size_t k = 0; k < n; k++) {
    long double s = 0.L;
    for (size_t j = 0; j < n; j++) {
      const size_t e = ((2 * j + 1) * k) % (4 * n);
      s += (long double)x[j] * cosl(k_pi_l * (long double)e / (long double)(2 * n));
    }
    y[k] = (double)(2.L * s);
  }
}

/* FFTW REDFT01 by definition: Y_k = X_0 + 2 sum_{j>=1} X_j cos(pi j (2k+1) / (2N)) */
static void dct3_direct(size_t n, const double * x, double * y) {
  for (size_t k = 0; k < n; k++) {
    long double s = 0.L;
    for (size_t j = 1; j < n; j++) {
      const size_t e = (j * (2 * k + 1)) % (4 * n);
      s += (long double)x[j] * cosl(k_pi_l * (long double)e / (long double)(2 * n));
    }
    y[k] = (double)((long double)x[0] + 2.L * s);
  }
}

/* twiddle e^{-2 pi i e / m} rounded from long double; the library builds its
 * tables with the same expression so both sides hold identical values */
static void twiddle(size_t e, size_t m, double * re, double * im) {
  const long double a = 2.L * k_pi_l * (long double)e / (long double)m;
  *re = (double)cosl(a);
  *im = (double)(-sinl(a));
}

static inline void cmul(double ar, double ai, double br, double bi, double * cr, double * ci) {
  *cr = ar * br - ai * bi;
  *ci = ar * bi + ai * br;
}

/*
 * Forward complex FFT of length m (power of two), Stockham autosort,
 * decimation in frequency: radix-4 stages while the sub-length is >= 4,
 * then one radix-2 stage if log2(m) is odd.  x, y: interleaved (re, im)
 * buffers of 2m doubles; tw: twiddles e^{-2 pi i e/m}, e < m.  Returns the
 * buffer holding the result.
 */
static double * fft_stockham(size_t m, double * x, double * y, const double * tw) {
  size_t n = m, s = 1;
  while (n >= 4) {
    const size_t q4 = n / 4;
    for (size_t p = 0; p < q4; p++)
      for (size_t q = 0; q < s; q++) {
        const double * a = x + 2 * (q + s * p);
        const double * b = x + 2 * (q + s * (p + q4));
        const double * c = x + 2 * (q + s * (p + 2 * q4));
        const double * d = x + 2 * (q + s * (p + 3 * q4));
        const double t0r = a[0] + c[0], t0i = a[1] + c[1];
        const double t1r = a[0] - c[0], t1i = a[1] - c[1];
        const double t2r = b[0] + d[0], t2i = b[1] + d[1];
        const double t3r = b[0] - d[0], t3i = b[1] - d[1];
        const double x0r = t0r + t2r, x0i = t0i + t2i;
        const double x2r = t0r - t2r, x2i = t0i - t2i;
        const double x1r = t1r + t3i, x1i = t1i - t3r;
        const double x3r = t1r - t3i, x3i = t1i + t3r;
        const size_t e1 = p * s, e2 = 2 * p * s, e3 = 3 * p * s;
        double * o = y + 2 * (q + s * 4 * p);
        o[0] = x0r;
        o[1] = x0i;
        cmul(x1r, x1i, tw[2 * e1], tw[2 * e1 + 1], o + 2 * s, o + 2 * s + 1);
        cmul(x2r, x2i, tw[2 * e2], tw[2 * e2 + 1], o + 4 * s, o + 4 * s + 1);
        cmul(x3r, x3i, tw[2 * e3], tw[2 * e3 + 1], o + 6 * s, o + 6 * s + 1);
      }
    double * t = x;
    x = y;
    y = t;
    n /= 4;
    s *= 4;
  }
  if (n == 2) {
    for (size_t q = 0; q < s; q++) {
      const double * a = x + 2 * q;
      const double * b = x + 2 * (q + s);
      double * o = y + 2 * q;
      o[0] = a[0] + b[0];
      o[1] = a[1] + b[1];
      cmul(a[0] - b[0], a[1] - b[1], tw[0], tw[1], o + 2 * s, o + 2 * s + 1);
    }
    double * t = x;
    x = y;
    y = t;
  }
  return x;
}

int oracle_fft(size_t m, const double * in, double * out) {
  double * tw = malloc(sizeof(double) * 2 * m);
  double * a = malloc(sizeof(double) * 2 * m);
  double * b = malloc(sizeof(double) * 2 * m);
  for (size_t e = 0; e < m; e++) twiddle(e, m, tw + 2 * e, tw + 2 * e + 1);
  memcpy(a, in, sizeof(double) * 2 * m);
  const double * r = fft_stockham(m, a, b, tw);
  memcpy(out, r, sizeof(double) * 2 * m);
  free(tw);
  free(a);
  free(b);
  return 0;
}

/* tables for the FFT-based DCTs of length n = 2m */
typedef struct {
  size_t n, m;
  double * tw;  /* FFT twiddles e^{-2 pi i e/m}, e < m */
  double * t1;  /* e^{-2 pi i k/n}, k <= m */
  double * t2;  /* e^{-i pi k/(2n)}, k < n */
} dct_plan_t;

static void dct_plan_init(dct_plan_t * pl, size_t n) {
  pl->n = n;
  pl->m = n / 2;
  pl->tw = malloc(sizeof(double) * 2 * (pl->m > 0 ? pl->m : 1));
  pl->t1 = malloc(sizeof(double) * 2 * (pl->m + 1));
  pl->t2 = malloc(sizeof(double) * 2 * n);
  for (size_t e = 0; e < pl->m; e++) twiddle(e, pl->m, pl->tw + 2 * e, pl->tw + 2 * e + 1);
  for (size_t k = 0; k <= pl->m; k++) twiddle(k, n, pl->t1 + 2 * k, pl->t1 + 2 * k + 1);
  for (size_t k = 0; k < n; k++) twiddle(k, 4 * n, pl->t2 + 2 * k, pl->t2 + 2 * k + 1);
}

static void dct_plan_free(dct_plan_t * pl) {
  free(pl->tw);
  free(pl->t1);
  free(pl->t2);
}

/* REDFT10 via an m-point complex FFT (Makhoul's reordering).  work: 4m doubles x 2 */
static void dct2_fft(const dct_plan_t * pl, const double * x, double * y, double * wa, double * wb) {
  const size_t n = pl->n, m = pl->m;
  /* v[k] = x[2k], v[n-1-k] = x[2k+1]; z[k] = v[2k] + i v[2k+1] */
  for (size_t k = 0; k < m; k++) {
    const size_t e = 2 * k, o = 2 * k + 1;
    const double ve = (e < m) ? x[2 * e] : x[2 * (n - 1 - e) + 1];
    const double vo = (o < m) ? x[2 * o] : x[2 * (n - 1 - o) + 1];
    wa[2 * k] = ve;
    wa[2 * k + 1] = vo;
  }
  const double * z = fft_stockham(m, wa, wb, pl->tw);
  for (size_t k = 0; k <= m; k++) {
    const size_t kk = (k == m) ? 0 : k;
    const size_t kr = (k == 0) ? 0 : m - k;
    const double ar = z[2 * kk], ai = z[2 * kk + 1];
    const double br = z[2 * kr], bi = -z[2 * kr + 1];
    const double er = (ar + br) * 0.5, ei = (ai + bi) * 0.5;
    const double dr = ar - br, di = ai - bi;
    const double orr = di * 0.5, oi = -(dr * 0.5);
    double tr, ti;
    cmul(orr, oi, pl->t1[2 * k], pl->t1[2 * k + 1], &tr, &ti);
    const double vr = er + tr, vi = ei + ti;
    double wr, wi;
    cmul(vr, vi, pl->t2[2 * k], pl->t2[2 * k + 1], &wr, &wi);
    y[k] = 2. * wr;
    if (k != 0 && k != m) y[n - k] = -2. * wi;
  }
}

/* REDFT01 via an m-point inverse complex FFT (inverse of the above) */
static void dct3_fft(const dct_plan_t * pl, const double * x, double * y, double * wa, double * wb) {
  const size_t n = pl->n, m = pl->m;
  for (size_t k = 0; k < m; k++) {
    const double pr = x[k], pim = -((k == 0) ? 0. : x[n - k]);
    const double cr = pl->t2[2 * k], ci = -pl->t2[2 * k + 1];
    double vkr, vki;
    cmul(pr, pim, cr, ci, &vkr, &vki);
    const double qr = x[k + m], qim = -x[m - k];
    const double dr = pl->t2[2 * (k + m)], di = -pl->t2[2 * (k + m) + 1];
    double vmr, vmi;
    cmul(qr, qim, dr, di, &vmr, &vmi);
    const double sr = vkr + vmr, si = vki + vmi;
    const double fr0 = vkr - vmr, fi0 = vki - vmi;
    double fr, fi;
    cmul(fr0, fi0, pl->t1[2 * k], -pl->t1[2 * k + 1], &fr, &fi);
    /* Z[k] = S + i F, conjugated for the forward transform */
    wa[2 * k] = sr - fi;
    wa[2 * k + 1] = -(si + fr);
  }
  const double * z = fft_stockham(m, wa, wb, pl->tw);
  for (size_t k = 0; k < m; k++) {
    const double v0 = z[2 * k], v1 = -z[2 * k + 1];
    /* v[2k] = v0, v[2k+1] = v1; y[2t] = v[t], y[2t+1] = v[n-1-t] */
    const size_t e = 2 * k, o = 2 * k + 1;
    if (e < m) y[2 * e] = v0; else y[2 * (n - 1 - e) + 1] = v0;
    if (o < m) y[2 * o] = v1; else y[2 * (n - 1 - o) + 1] = v1;
  }
}

static int dct_rows(int inverse, int kind, size_t n, size_t howmany, const double * in, double * out) {
  if (kind == ORACLE_DCT_DIRECT || n < 2 || (n & (n - 1)) != 0) {
#pragma omp parallel for schedule(static)
    for (size_t r = 0; r < howmany; r++) {
      if (inverse) dct3_direct(n, in + r * n, out + r * n);
      else dct2_direct(n, in + r * n, out + r * n);
    }
    return 0;
  }
  dct_plan_t pl;
  dct_plan_init(&pl, n);
#pragma omp parallel
  {
    double * wa = malloc(sizeof(double) * 2 * pl.m);
    double * wb = malloc(sizeof(double) * 2 * pl.m);
#pragma omp for schedule(static)
    for (size_t r = 0; r < howmany; r++) {
      if (inverse) dct3_fft(&pl, in + r * n, out + r * n, wa, wb);
      else dct2_fft(&pl, in + r * n, out + r * n, wa, wb);
    }
    free(wa);
    free(wb);
  }
  dct_plan_free(&pl);
  return 0;
}

int oracle_dct2(int kind, size_t n, size_t howmany, const double * in, double * out) {
  return dct_rows(0, kind, n, howmany, in, out);
}

int oracle_dct3(int kind, size_t n, size_t howmany, const double * in, double * out) {
  return dct_rows(1, kind, n, howmany, in, out);
}

/* ------------------------------------------------------------------ */
/* tridiagonal solvers (src/tdm.c)                                     */
/* ------------------------------------------------------------------ */

/* src/tdm.c:17-50, constant coefficients per call */
static void gtsv_const(int n, double l, double c, double u, double * v, double * q) {
  v[0] = u / c;
  q[0] = q[0] / c;
  for (int i = 1; i < n - 1; i++) {
    const double val = 1. / (c - l * v[i - 1]);
    v[i] = val * u;
    q[i] = val * (q[i] - l * q[i - 1]);
  }
  const double val = c - l * v[n - 2];
  if (fabs(val) > DBL_EPSILON) q[n - 1] = 1. / val * (q[n - 1] - l * q[n - 2]);
  else q[n - 1] = 0.;
  for (int i = n - 2; i >= 0; i--) q[i] -= v[i] * q[i + 1];
}

/* src/tdm.c:66-99, periodic branch, nrhs = 1; `pin` forces the singular branch */
static void tdm_periodic_work(int n, double l, double c, double u, double * q,
                              int pin, double * v, double * q1) {
  for (int i = 0; i < n - 1; i++) q1[i] = (i == 0) ? -l : (i == n - 2) ? -u : 0.;
  gtsv_const(n - 1, l, c, u, v, q1);
  gtsv_const(n - 1, l, c, u, v, q);
  const double num = q[n - 1] - u * q[0] - l * q[n - 2];
  const double den = c + u * q1[0] + l * q1[n - 2];
  q[n - 1] = (pin || fabs(den) < DBL_EPSILON) ? 0. : num / den;
  for (int i = 0; i < n - 1; i++) q[i] = q[i] + q[n - 1] * q1[i];
}

int oracle_tdm_periodic(int n, double l, double c, double u, double * q, int gauge) {
  if (n < 3) return 1;
  double * v = calloc((size_t)n, sizeof(double));
  double * q1 = calloc((size_t)n, sizeof(double));
  tdm_periodic_work(n, l, c, u, q, gauge == ORACLE_GAUGE_PINNED ? 1 : 0, v, q1);
  free(v);
  free(q1);
  return 0;
}

/*
 * Chunked solve of the same periodic system (the GPU's algorithm).
 * Chunk c covers rows [cL, cL+L); its last row s_c is a separator.  The
 * L-1 interior rows are solved with zero ends (g), plus the two unit spikes
 * w (driven by the lower separator) and z (upper separator); the separators
 * then satisfy a periodic tridiagonal system of size P = n / L with
 * constant coefficients (a circulant), solved as the circular convolution
 * of its right-hand side with the first column of the inverse, which the
 * reference algorithm (src/tdm.c) finds from e_0.  The singular mode pins
 * the last separator (x_{n-1} = 0), the same gauge as the reference's
 * singular branch.
 */
typedef struct {
  int L, K, P;
  double a, b;
  double * inv; /* K */
  double * v;   /* K */
  double * w;   /* K */
  double * z;   /* K */
  double lo, di, up;
} part_tables_t;

static void part_thomas(const part_tables_t * t, const double * f, double * g) {
  /* forward: g_k = inv_k (f_k - a g_{k-1}); backward: g_k -= v_k g_{k+1} */
  double prev = 0.;
  for (int k = 0; k < t->K; k++) {
    g[k] = t->inv[k] * (f[k] - t->a * prev);
    prev = g[k];
  }
  for (int k = t->K - 2; k >= 0; k--) g[k] = g[k] - t->v[k] * g[k + 1];
}

static void part_tables_build(part_tables_t * t, int L, int n, double a, double b) {
  t->L = L;
  t->K = L - 1;
  t->P = n / L;
  t->a = a;
  t->b = b;
  t->inv = calloc((size_t)t->K, sizeof(double));
  t->v = calloc((size_t)t->K, sizeof(double));
  t->w = calloc((size_t)t->K, sizeof(double));
  t->z = calloc((size_t)t->K, sizeof(double));
  double vp = 0.;
  for (int k = 0; k < t->K; k++) {
    t->inv[k] = 1. / (b - a * vp);
    t->v[k] = t->inv[k] * a;
    vp = t->v[k];
  }
  double * f = calloc((size_t)t->K, sizeof(double));
  f[0] = -a;
  part_thomas(t, f, t->w);
  memset(f, 0, sizeof(double) * (size_t)t->K);
  f[t->K - 1] = -a;
  part_thomas(t, f, t->z);
  free(f);
  t->lo = a * t->w[t->K - 1];
  t->up = a * t->z[0];
  t->di = (b + a * t->z[t->K - 1]) + a * t->w[0];
}

static void part_tables_free(part_tables_t * t) {
  free(t->inv);
  free(t->v);
  free(t->w);
  free(t->z);
}

static void part_solve(const part_tables_t * t, double * q, int pin, double * g,
                       double * gf, double * gl, double * X, double * v, double * q1) {
  const int L = t->L, K = t->K, P = t->P;
  for (int c = 0; c < P; c++) {
    part_thomas(t, q + (size_t)c * L, g);
    gf[c] = g[0];
    gl[c] = g[K - 1];
  }
  /* reduced right-hand side in gl (its chunk values are no longer needed) */
  for (int c = 0; c < P; c++) gl[c] = (q[(size_t)c * L + K] - t->a * gl[c]) - t->a * gf[(c + 1) % P];
  /* The reduced matrix is circulant (every chunk is the same), so its
   * inverse is the circulant of h = C^-1 e_0, found with the reference
   * algorithm; the singular mode uses h = C^+ (e_0 - 1/P) with h[P-1] = 0
   * and then pins X[P-1] = 0.  X = h (*) R, summed in ascending k. */
  for (int c = 0; c < P; c++) gf[c] = pin ? (c == 0 ? 1. - 1. / (double)P : -1. / (double)P) : (c == 0 ? 1. : 0.);
  tdm_periodic_work(P, t->lo, t->di, t->up, gf, pin, v, q1);
  for (int i = 0; i < P; i++) {
    double acc = 0.;
    for (int k = 0; k < P; k++) acc += gf[(i - k + P) % P] * gl[k];
    X[i] = acc;
  }
  if (pin) {
    const double xl = X[P - 1];
    for (int i = 0; i < P; i++) X[i] = X[i] - xl;
  }
  for (int c = 0; c < P; c++) {
    part_thomas(t, q + (size_t)c * L, g);
    const double xl = X[(c + P - 1) % P], xu = X[c];
    double * o = q + (size_t)c * L;
    for (int k = 0; k < K; k++) o[k] = (g[k] + xl * t->w[k]) + xu * t->z[k];
    o[K] = xu;
  }
}

/* largest chunk <= requested that divides n and leaves >= 3 chunks; 0 if none */
static int part_chunk(int n, int requested) {
  int L = requested;
  while (L >= 2 && (n % L != 0 || n / L < 3)) L /= 2;
  return L >= 2 ? L : 0;
}

int oracle_tdm_partitioned(int n, int chunk, double a, double b, double * q) {
  const int L = part_chunk(n, chunk);
  if (L == 0) return 1;
  part_tables_t t;
  part_tables_build(&t, L, n, a, b);
  double * g = calloc((size_t)L, sizeof(double));
  double * gf = calloc((size_t)t.P, sizeof(double));
  double * gl = calloc((size_t)t.P, sizeof(double));
  double * X = calloc((size_t)t.P, sizeof(double));
  double * v = calloc((size_t)t.P, sizeof(double));
  double * q1 = calloc((size_t)t.P, sizeof(double));
  part_solve(&t, q, 0, g, gf, gl, X, v, q1);
  free(g);
  free(gf);
  free(gl);
  free(X);
  free(v);
  free(q1);
  part_tables_free(&t);
  return 0;
}

/* ------------------------------------------------------------------ */
/* Poisson solver (src/fluid/compute_potential.c)                      */
/* ------------------------------------------------------------------ */

/* eigenvalues of the x operator, :233-261 (pow(., 2.) folded to squares) */
int oracle_poisson_evals(const domain_t * domain, double * evals) {
  const size_t nx = domain->glsizes[0];
  const double sl = 2. * (double)nx;
  const double gsz = domain->lengths[0] / (double)nx;
  for (size_t i = 0; i < nx; i++) {
    const double s = sin(k_pi_poisson * (double)i / sl);
    evals[i] = -4. / (gsz * gsz) * (s * s);
  }
  return 0;
}

int oracle_fluid_compute_potential(const domain_t * domain, double dt_old,
                                   double dt_new, fluid_t * fluid) {
  DOMAIN_VARS;
  const size_t ncell = (size_t)nx * (size_t)ny;
  double * buf0 = malloc(sizeof(double) * ncell);
  double * buf1 = malloc(sizeof(double) * ncell);
  double * evals = malloc(sizeof(double) * (size_t)nx);
  if (!buf0 || !buf1 || !evals) return 1;
  oracle_poisson_evals(domain, evals);
  /* right-hand side, :285-363 */
  {
    const double * ux = fluid->ux.data;
    const double * uy = fluid->uy.data;
    const double * psi = fluid->psi[0].data;
    const double * den = fluid->den[1].data;
    const double norm = 2. * (double)domain->glsizes[0];
    const double refden = fluid->refden;
    const double c0 = 1. / norm / dt_new * refden;
    const double c1 = 1. / norm * -1. / dt_new * dt_old;
#pragma omp parallel for schedule(static)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++) {
        const double hm = HXXF(i), hp = HXXF(i + 1);
        const double jm = JDXF(i), j0 = JDXC(i), jp = JDXF(i + 1);
        const double div = 1. / j0 * (-jm / hm * UXa(i, j) + jp / hp * UXa(i + 1, j)
                                      - j0 / hy * UYa(i, j) + j0 / hy * UYa(i, j + 1));
        const double dxm = 0.5 * DENa(i - 1, j) + 0.5 * DENa(i, j);
        const double dxp = 0.5 * DENa(i, j) + 0.5 * DENa(i + 1, j);
        const double dym = 0.5 * DENa(i, j - 1) + 0.5 * DENa(i, j);
        const double dyp = 0.5 * DENa(i, j) + 0.5 * DENa(i, j + 1);
        const double pxm = -PSIa(i - 1, j) + PSIa(i, j);
        const double pxp = -PSIa(i, j) + PSIa(i + 1, j);
        const double pym = -PSIa(i, j - 1) + PSIa(i, j);
        const double pyp = -PSIa(i, j) + PSIa(i, j + 1);
        const double gxm = (refden / dxm - 1.) * pxm / hm;
        const double gxp = (refden / dxp - 1.) * pxp / hp;
        const double gym = (refden / dym - 1.) * pym / hy;
        const double gyp = (refden / dyp - 1.) * pyp / hy;
        const double add = 1. / j0 * (-jm / hm * gxm + jp / hp * gxp - j0 / hy * gym + j0 / hy * gyp);
        buf0[(size_t)(j - 1) * nx + (i - 1)] = c0 * div + c1 * add;
      }
  }
  /* DCT-II along x, :445 */
  dct_rows(0, g_opt.dct, (size_t)nx, (size_t)ny, buf0, buf1);
  /* transpose x1 -> y1 (single rank), :448-452: buf0[m][j] */
#pragma omp parallel for schedule(static)
  for (int m = 0; m < nx; m++)
    for (int j = 0; j < ny; j++) buf0[(size_t)m * ny + j] = buf1[(size_t)j * nx + m];
  /* periodic tridiagonal solves per x-mode, :382-410 */
  const double l = 1. / hy / hy; /* :173-176 */
  const double u = 1. / hy / hy;
  if (g_opt.tdm == ORACLE_TDM_PARTITIONED && part_chunk(ny, g_opt.chunk) != 0) {
    const int L = part_chunk(ny, g_opt.chunk);
#pragma omp parallel
    {
      const int P = ny / L;
      double * g = calloc((size_t)L, sizeof(double));
      double * gf = calloc((size_t)P, sizeof(double));
      double * gl = calloc((size_t)P, sizeof(double));
      double * X = calloc((size_t)P, sizeof(double));
      double * v = calloc((size_t)P, sizeof(double));
      double * q1 = calloc((size_t)P, sizeof(double));
#pragma omp for schedule(static)
      for (int m = 0; m < nx; m++) {
        const double c = -l - u + evals[m];
        part_tables_t t;
        part_tables_build(&t, L, ny, l, c);
        part_solve(&t, buf0 + (size_t)m * ny, m == 0, g, gf, gl, X, v, q1);
        part_tables_free(&t);
      }
      free(g);
      free(gf);
      free(gl);
      free(X);
      free(v);
      free(q1);
    }
  } else {
#pragma omp parallel
    {
      double * v = calloc((size_t)ny, sizeof(double));
      double * q1 = calloc((size_t)ny, sizeof(double));
#pragma omp for schedule(static)
      for (int m = 0; m < nx; m++) {
        const double c = -l - u + evals[m];
        const int pin = (g_opt.gauge == ORACLE_GAUGE_PINNED) && (m == 0);
        tdm_periodic_work(ny, l, c, u, buf0 + (size_t)m * ny, pin, v, q1);
      }
      free(v);
      free(q1);
    }
  }
  /* transpose y1 -> x1, :457-461 */
#pragma omp parallel for schedule(static)
  for (int j = 0; j < ny; j++)
    for (int m = 0; m < nx; m++) buf1[(size_t)j * nx + m] = buf0[(size_t)m * ny + j];
  /* DCT-III along x, :466 */
  dct_rows(1, g_opt.dct, (size_t)nx, (size_t)ny, buf1, buf0);
  /* psi[1], :365-380 */
  {
    double * psi = fluid->psi[1].data;
#pragma omp parallel for schedule(static)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++) PSIa(i, j) = buf0[(size_t)(j - 1) * nx + (i - 1)];
    oracle_fluid_update_boundaries_psi(domain, &fluid->psi[1]);
  }
  free(buf0);
  free(buf1);
  free(evals);
  return 0;
}

/* ------------------------------------------------------------------ */
/* projection and pressure                                             */
/* ------------------------------------------------------------------ */

/* src/fluid/correct_velocity/ux.c:18-66 and uy.c:17-64 */
int oracle_fluid_correct_velocity(const domain_t * domain, double dt_old,
                                  double dt_new, fluid_t * fluid) {
  DOMAIN_VARS;
  double * ux = fluid->ux.data;
  double * uy = fluid->uy.data;
  const double * psi1 = fluid->psi[1].data;
  const double * psi0 = fluid->psi[0].data;
  const double * den1 = fluid->den[1].data;
  const double refden = fluid->refden;
  const double coef = -1. / dt_new * dt_old;
#pragma omp parallel for schedule(static)
  for (int j = 1; j <= ny; j++)
    for (int i = 2; i <= nx; i++)
      UXa(i, j) -= dt_new / refden / HXXF(i) * (-PSI1a(i - 1, j) + PSI1a(i, j));
#pragma omp parallel for schedule(static)
  for (int j = 1; j <= ny; j++)
    for (int i = 2; i <= nx; i++) {
      const double d = 0.5 * DEN1a(i - 1, j) + 0.5 * DEN1a(i, j);
      UXa(i, j) += dt_new * coef * (1. / d - 1. / refden) / HXXF(i) * (-PSI0a(i - 1, j) + PSI0a(i, j));
    }
  oracle_fluid_update_boundaries_ux(domain, &fluid->ux);
#pragma omp parallel for schedule(static)
  for (int j = 1; j <= ny; j++)
    for (int i = 1; i <= nx; i++)
      UYa(i, j) -= dt_new / refden / hy * (-PSI1a(i, j - 1) + PSI1a(i, j));
#pragma omp parallel for schedule(static)
  for (int j = 1; j <= ny; j++)
    for (int i = 1; i <= nx; i++) {
      const double d = 0.5 * DEN1a(i, j - 1) + 0.5 * DEN1a(i, j);
      UYa(i, j) += dt_new * coef * (1. / d - 1. / refden) / hy * (-PSI0a(i, j - 1) + PSI0a(i, j));
    }
  oracle_fluid_update_boundaries_uy(domain, &fluid->uy);
  return 0;
}

/* src/fluid/update_pressure.c:7-53 */
int oracle_fluid_update_pressure(const domain_t * domain, fluid_t * fluid) {
  DOMAIN_VARS;
  double * p = fluid->p.data;
  const double * psi = fluid->psi[1].data;
#pragma omp parallel for schedule(static)
  for (int j = 1; j <= ny; j++)
    for (int i = 1; i <= nx; i++) Pa(i, j) += PSIa(i, j);
  oracle_fluid_update_boundaries_p(domain, &fluid->p);
  void * t = fluid->psi[0].data;
  fluid->psi[0].data = fluid->psi[1].data;
  fluid->psi[1].data = t;
  return 0;
}

/* ------------------------------------------------------------------ */
/* one time step (src/integrate.c:11-70)                               */
/* ------------------------------------------------------------------ */

int oracle_integrate(const domain_t * domain, fluid_t * fluid,
                     interface_t * interface, double * dt) {
  if (oracle_decide_dt(domain, fluid, interface, dt)) return 1;
  for (size_t k = 0; k < 3; k++) {
    if (oracle_fluid_compute_density(domain, fluid, interface, 0)) return 1;
    if (oracle_fluid_compute_viscosity(domain, fluid, interface)) return 1;
    if (oracle_interface_compute_curvature_tensor(domain, interface)) return 1;
    if (oracle_interface_compute_force(domain, fluid, interface)) return 1;
    if (oracle_interface_update_vof(domain, k, *dt, fluid, interface)) return 1;
    if (oracle_fluid_compute_density(domain, fluid, interface, 1)) return 1;
    if (oracle_interface_compute_mass_flux(domain, fluid, interface)) return 1;
    if (oracle_fluid_predict_field(domain, k, *dt, fluid, interface)) return 1;
    const double dt_old = k_rk[(k + 2) % 3][RKG] * (*dt);
    const double dt_new = k_rk[(k + 0) % 3][RKG] * (*dt);
    if (oracle_fluid_compute_potential(domain, dt_old, dt_new, fluid)) return 1;
    if (oracle_fluid_correct_velocity(domain, dt_old, dt_new, fluid)) return 1;
    if (oracle_fluid_update_pressure(domain, fluid)) return 1;
  }
  return 0;
}

/* ------------------------------------------------------------------ */
/* diagnostics                                                         */
/* ------------------------------------------------------------------ */

/*
 * Sums in the reference's loop order (j outer, i inner):
 * volume / mean / area as src/logging/interface.c:34-61, max |div| as
 * src/logging/divergence.c:41-60, centroid and rise velocity weighted by
 * vof * JDXC (not present in the reference).
 */
int oracle_diagnostics(const domain_t * domain, const fluid_t * fluid,
                       const interface_t * interface, double out[8]) {
  DOMAIN_VARS;
  const double * vof = interface->vof.data;
  const double * ux = fluid->ux.data;
  const double * uy = fluid->uy.data;
  const double * xc = domain->xc;
  double vol = 0., sds = 0., xm = 0., um = 0., area = 0.;
  double vmin = 1., vmax = 0., divmax = 0.;
  for (int j = 1; j <= ny; j++)
    for (int i = 1; i <= nx; i++) {
      const double ds = JDXC(i);
      const double lv = VOFa(i, j);
      vmin = fmin(vmin, lv);
      vmax = fmax(vmax, lv);
      vol += lv * ds;
      sds += ds;
      xm += lv * XC(i) * ds;
      um += lv * (0.5 * UXa(i, j) + 0.5 * UXa(i + 1, j)) * ds;
      const double div = 1. / JDXC(i) * (-JDXF(i) / HXXF(i) * UXa(i, j) + JDXF(i + 1) / HXXF(i + 1) * UXa(i + 1, j)
                                         - JDXC(i) / hy * UYa(i, j) + JDXC(i) / hy * UYa(i, j + 1));
      divmax = fmax(divmax, fabs(div));
    }
  for (int j = 1; j <= ny; j++)
    for (int i = 1; i <= nx + 1; i++)
      area += fabs(VOFa(i, j) - VOFa(i - 1, j)) * (JDXF(i) / HXXF(i));
  for (int j = 1; j <= ny; j++)
    for (int i = 1; i <= nx; i++)
      area += fabs(VOFa(i, j) - VOFa(i, j - 1)) * (JDXC(i) / hy);
  out[0] = vol;
  out[1] = vol != 0. ? xm / vol : 0.;
  out[2] = vol != 0. ? um / vol : 0.;
  out[3] = divmax;
  out[4] = vmin;
  out[5] = vmax;
  out[6] = vol / sds;
  out[7] = area;
  return 0;
}
