// Runtime of libsbfs: error reporting, the HIP stream, per-domain device
// contexts, allocation of the reference's structs on the device and the
// sbfs_* library calls (include/sbfs.h).
#include <cerrno>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>

#include "sbfs.h"
#include "sbfs_internal.hpp"

namespace sbfs {

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
static thread_local char g_err[1024] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  fprintf(stderr, "libsbfs: %s\n", g_err);
  fflush(stderr);
}

const char* last_error() { return g_err; }

int check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error("%s: kernel launch failed: %s", what, hipGetErrorString(e));
    return 1;
  }
  if (getenv("SBFS_SYNC_EACH_CALL")) {
    e = hipStreamSynchronize(stream());
    if (e != hipSuccess) {
      set_error("%s: %s", what, hipGetErrorString(e));
      return 1;
    }
  }
  return 0;
}

// ---------------------------------------------------------------------------
// layouts and constants
// ---------------------------------------------------------------------------
const Nadds NA_UX = {{{0, 1}, {1, 1}}};
const Nadds NA_UY = {{{1, 1}, {1, 1}}};
const Nadds NA_P = {{{1, 1}, {1, 1}}};
const Nadds NA_DEN = {{{1, 1}, {1, 1}}};
const Nadds NA_VISUX = {{{0, 1}, {1, 1}}};
const Nadds NA_VISUY = {{{1, 1}, {1, 1}}};
const Nadds NA_TXX = {{{0, 0}, {1, 1}}};
const Nadds NA_TXY = {{{0, 1}, {1, 1}}};
const Nadds NA_TYY = {{{0, 0}, {1, 1}}};
const Nadds NA_SRCUX = {{{-1, 0}, {0, 0}}};
const Nadds NA_SRCUY = {{{0, 0}, {0, 0}}};
const Nadds NA_VOF = {{{1, 1}, {2, 2}}};
const Nadds NA_IFRCX = {{{-1, 0}, {0, 0}}};
const Nadds NA_IFRCY = {{{0, 0}, {0, 0}}};
const Nadds NA_DVOF = {{{0, 1}, {1, 2}}};
const Nadds NA_NORMAL = {{{0, 0}, {1, 1}}};
const Nadds NA_CURV = {{{1, 1}, {1, 1}}};
const Nadds NA_FLUXX = {{{0, 1}, {1, 1}}};
const Nadds NA_FLUXY = {{{1, 1}, {1, 1}}};
const Nadds NA_SRC = {{{0, 0}, {0, 0}}};

int check_array(const domain_t* d, const array_t* a, const Nadds& na,
                size_t elem, const char* name) {
  SBFS_REQUIRE(a != nullptr, "%s: null array", name);
  SBFS_REQUIRE(a->data != nullptr, "%s: array has no data", name);
  SBFS_REQUIRE(a->size == elem, "%s: element size %zu, expected %zu", name,
               a->size, elem);
  if (a->nadds) {
    for (int k = 0; k < 2; k++)
      SBFS_REQUIRE(a->nadds[k][0] == na.v[k][0] && a->nadds[k][1] == na.v[k][1],
                   "%s: ghost widths differ from the layout", name);
  }
  const size_t nxt = d->mysizes[0] + na.v[0][0] + na.v[0][1];
  const size_t nyt = d->mysizes[1] + na.v[1][0] + na.v[1][1];
  SBFS_REQUIRE(a->datasize == nxt * nyt * elem,
               "%s: datasize %zu, expected %zu", name, a->datasize,
               nxt * nyt * elem);
  return 0;
}

// src/runge_kutta.c:8-15 {alpha, beta, alpha + beta}
RK rk(size_t stage) {
  static const RK t[3] = {
      {32. / 60., 0. / 60., 32. / 60. + 0. / 60.},
      {25. / 60., -17. / 60., 25. / 60. + -17. / 60.},
      {45. / 60., -25. / 60., 45. / 60. + -25. / 60.},
  };
  return t[stage % 3];
}

// ---------------------------------------------------------------------------
// runtime state
// ---------------------------------------------------------------------------
static Options g_opt;
static hipStream_t g_stream = nullptr;
static hipStream_t g_side = nullptr;
static hipEvent_t g_fork[3] = {}, g_join[3] = {};
static bool g_ready = false;
static std::mutex g_mutex;
static std::vector<std::pair<const domain_t*, DomainCtx*>> g_domains;

Options& options() { return g_opt; }
hipStream_t stream() { return g_stream; }
hipStream_t side_stream() { return g_side; }
hipEvent_t fork_event(int k) { return g_fork[k]; }
hipEvent_t join_event(int k) { return g_join[k]; }

int init_runtime() {
  if (g_ready) return 0;
  std::lock_guard<std::mutex> lock(g_mutex);
  if (g_ready) return 0;
  int n = 0;
  SBFS_HIP(hipGetDeviceCount(&n));
  SBFS_REQUIRE(n > 0, "no HIP device visible");
  // default-priority, non-blocking streams: the library's, and a side
  // stream for work that overlaps it within a step (forked and joined with
  // the events below, which graph capture records as dependencies)
  SBFS_HIP(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
  SBFS_HIP(hipStreamCreateWithFlags(&g_side, hipStreamNonBlocking));
  for (int k = 0; k < 3; k++) {
    SBFS_HIP(hipEventCreateWithFlags(&g_fork[k], hipEventDisableTiming));
    SBFS_HIP(hipEventCreateWithFlags(&g_join[k], hipEventDisableTiming));
  }
  const char* p = getenv("SBFS_PATH");
  if (p) g_opt.path = atoi(p);
  if (noalloc_setup()) return 1;
  g_ready = true;
  return 0;
}

static uint64_t hash_bytes(uint64_t h, const void* p, size_t n) {
  const unsigned char* b = (const unsigned char*)p;
  for (size_t k = 0; k < n; k++) {
    h ^= b[k];
    h *= 1099511628211ull;
  }
  return h;
}

// Identity of a host domain: sizes, lengths, hy and the addresses of its
// metric arrays.  As in the reference, the metrics are written once by
// domain_init and never modified in place afterwards, so the device mirror
// is refreshed only when one of these changes (e.g. a new domain_t reuses
// the same struct).
static uint64_t domain_hash(const domain_t* d) {
  uint64_t h = 1469598103934665603ull;
  h = hash_bytes(h, d->mysizes, sizeof(d->mysizes));
  h = hash_bytes(h, d->lengths, sizeof(d->lengths));
  h = hash_bytes(h, &d->hy, sizeof(double));
  const void* ptrs[5] = {d->hxxf, d->hxxc, d->jdxf, d->jdxc, d->xc};
  h = hash_bytes(h, ptrs, sizeof(ptrs));
  return h;
}

void poisson_destroy(PoissonPlan* p);  // poisson.hip

static uint64_t g_graph_epoch = 0;
void graph_epoch_bump() { g_graph_epoch++; }
uint64_t graph_epoch() { return g_graph_epoch; }

static void ctx_free(DomainCtx* c) {
  if (!c) return;
  (void)hipStreamSynchronize(g_stream);
  graph_epoch_bump();
  (void)hipFree(c->hxxf);
  (void)hipFree(c->hxxc);
  (void)hipFree(c->jdxf);
  (void)hipFree(c->jdxc);
  (void)hipFree(c->xc);
  (void)hipFree(c->red);
  (void)hipFree(c->dt_dev);
  (void)hipHostFree(c->host_pin);
  poisson_destroy(c->poisson);
  delete c;
}

static int upload_metrics(const domain_t* d, DomainCtx* c) {
  const size_t nx = d->mysizes[0];
  SBFS_HIP(hipMemcpy(c->hxxf, d->hxxf, sizeof(double) * (nx + 1), hipMemcpyHostToDevice));
  SBFS_HIP(hipMemcpy(c->hxxc, d->hxxc, sizeof(double) * nx, hipMemcpyHostToDevice));
  SBFS_HIP(hipMemcpy(c->jdxf, d->jdxf, sizeof(double) * (nx + 1), hipMemcpyHostToDevice));
  SBFS_HIP(hipMemcpy(c->jdxc, d->jdxc, sizeof(double) * nx, hipMemcpyHostToDevice));
  SBFS_HIP(hipMemcpy(c->xc, d->xc, sizeof(double) * (nx + 2), hipMemcpyHostToDevice));
  return 0;
}

DomainCtx* domain_ctx(const domain_t* d) {
  if (init_runtime()) return nullptr;
  if (!d || !d->hxxf || !d->hxxc || !d->jdxf || !d->jdxc || !d->xc) {
    set_error("domain is not initialised");
    return nullptr;
  }
  const size_t nx = d->mysizes[0], ny = d->mysizes[1];
  if (nx < 2 || ny < 4) {
    set_error("domain %zu x %zu is too small (need Nx >= 2, Ny >= 4)", nx, ny);
    return nullptr;
  }
  if (d->glsizes[0] != nx || d->glsizes[1] != ny) {
    set_error("one GPU holds the whole domain: mysizes must equal glsizes");
    return nullptr;
  }
  if ((nx + 2) * (ny + 4) > (size_t)INT32_MAX) {
    set_error("domain too large for 32-bit cell indices");
    return nullptr;
  }
  const uint64_t h = domain_hash(d);
  std::lock_guard<std::mutex> lock(g_mutex);
  for (auto& e : g_domains) {
    if (e.first == d) {
      DomainCtx* c = e.second;
      if (c->hash == h && c->nx == nx && c->ny == ny) return c;
      ctx_free(c);
      e.second = nullptr;
    }
  }
  DomainCtx* c = new DomainCtx();
  c->nx = nx;
  c->ny = ny;
  c->lx = d->lengths[0];
  c->ly = d->lengths[1];
  c->hy = d->hy;
  c->hash = h;
  bool ok = hipMalloc(&c->hxxf, sizeof(double) * (nx + 1)) == hipSuccess &&
            hipMalloc(&c->hxxc, sizeof(double) * nx) == hipSuccess &&
            hipMalloc(&c->jdxf, sizeof(double) * (nx + 1)) == hipSuccess &&
            hipMalloc(&c->jdxc, sizeof(double) * nx) == hipSuccess &&
            hipMalloc(&c->xc, sizeof(double) * (nx + 2)) == hipSuccess &&
            hipMalloc(&c->red, sizeof(double) * kRedSlots) == hipSuccess &&
            hipMalloc(&c->dt_dev, sizeof(double) * 8) == hipSuccess &&
            hipHostMalloc(&c->host_pin, sizeof(double) * 16, 0) == hipSuccess;
  if (!ok || upload_metrics(d, c)) {
    set_error("device allocation for the domain failed");
    ctx_free(c);
    return nullptr;
  }
  bool placed = false;
  for (auto& e : g_domains)
    if (e.first == d) {
      e.second = c;
      placed = true;
    }
  if (!placed) g_domains.push_back({d, c});
  return c;
}

DomainCtx* domain_ctx_find(const domain_t* d) {
  if (!g_ready || !d) return nullptr;
  const uint64_t h = domain_hash(d);
  std::lock_guard<std::mutex> lock(g_mutex);
  for (auto& e : g_domains)
    if (e.first == d && e.second && e.second->hash == h && e.second->nx == d->mysizes[0] &&
        e.second->ny == d->mysizes[1])
      return e.second;
  return nullptr;
}

static void forget_domain(const domain_t* d) {
  std::lock_guard<std::mutex> lock(g_mutex);
  for (auto it = g_domains.begin(); it != g_domains.end(); ++it)
    if (it->first == d) {
      ctx_free(it->second);
      g_domains.erase(it);
      return;
    }
}

// environment safety factors, as src/decide_dt.c:147-152
static int env_double(const char* name, double* v) {
  const char* s = getenv(name);
  SBFS_REQUIRE(s != nullptr, "%s not found", name);
  errno = 0;
  *v = strtod(s, nullptr);
  SBFS_REQUIRE(errno == 0, "%s: invalid value as double", name);
  return 0;
}

int load_dt_coefs() {
  if (g_opt.coefs_set) return 0;
  if (env_double("coef_dt_adv", &g_opt.cadv)) return 1;
  if (env_double("coef_dt_dif", &g_opt.cdif)) return 1;
  if (env_double("coef_dt_int", &g_opt.cint)) return 1;
  g_opt.coefs_set = true;
  return 0;
}

// ---------------------------------------------------------------------------
// allocation of the reference's structs (device arrays, zero-filled)
// ---------------------------------------------------------------------------
static int prepare(const domain_t* d, const Nadds& na, size_t elem, array_t* a) {
  const size_t nxt = d->mysizes[0] + na.v[0][0] + na.v[0][1];
  const size_t nyt = d->mysizes[1] + na.v[1][0] + na.v[1][1];
  a->size = elem;
  a->nadds = (int(*)[2])calloc(2, sizeof(int[2]));
  SBFS_REQUIRE(a->nadds != nullptr, "host allocation failed");
  for (int k = 0; k < 2; k++) {
    a->nadds[k][0] = na.v[k][0];
    a->nadds[k][1] = na.v[k][1];
  }
  a->datasize = nxt * nyt * elem;
  a->data = nullptr;
  SBFS_HIP(hipMalloc(&a->data, a->datasize));
  SBFS_HIP(hipMemsetAsync(a->data, 0, a->datasize, g_stream));
  return 0;
}

static void release(array_t* a) {
  if (a->data) (void)hipFree(a->data);
  free(a->nadds);
  a->data = nullptr;
  a->nadds = nullptr;
}

int allocate_fluid(const domain_t* d, fluid_t* f) {
  if (init_runtime()) return 1;
  int e = 0;
  e |= prepare(d, NA_UX, sizeof(double), &f->ux);
  e |= prepare(d, NA_UY, sizeof(double), &f->uy);
  e |= prepare(d, NA_P, sizeof(double), &f->p);
  for (int n = 0; n < 2; n++) e |= prepare(d, NA_P, sizeof(double), &f->psi[n]);
  for (int n = 0; n < 2; n++) e |= prepare(d, NA_DEN, sizeof(double), &f->den[n]);
  e |= prepare(d, NA_VISUX, sizeof(double), &f->visux);
  e |= prepare(d, NA_VISUY, sizeof(double), &f->visuy);
  e |= prepare(d, NA_TXX, sizeof(double), &f->txx);
  e |= prepare(d, NA_TXY, sizeof(double), &f->txy);
  e |= prepare(d, NA_TYY, sizeof(double), &f->tyy);
  for (int n = 0; n < 3; n++) {
    e |= prepare(d, NA_SRCUX, sizeof(double), &f->srcux[n]);
    e |= prepare(d, NA_SRCUY, sizeof(double), &f->srcuy[n]);
  }
  if (e) return 1;
  SBFS_HIP(hipStreamSynchronize(g_stream));
  return 0;
}

int allocate_interface(const domain_t* d, interface_t* s) {
  if (init_runtime()) return 1;
  int e = 0;
  e |= prepare(d, NA_VOF, sizeof(double), &s->vof);
  e |= prepare(d, NA_IFRCX, sizeof(double), &s->ifrcx);
  e |= prepare(d, NA_IFRCY, sizeof(double), &s->ifrcy);
  e |= prepare(d, NA_DVOF, sizeof(double) * 2, &s->dvof);
  e |= prepare(d, NA_NORMAL, sizeof(double) * 3, &s->normal);
  e |= prepare(d, NA_CURV, sizeof(double), &s->curv);
  e |= prepare(d, NA_FLUXX, sizeof(double), &s->fluxx);
  e |= prepare(d, NA_FLUXY, sizeof(double), &s->fluxy);
  for (int n = 0; n < 2; n++) e |= prepare(d, NA_SRC, sizeof(double), &s->src[n]);
  if (e) return 1;
  SBFS_HIP(hipStreamSynchronize(g_stream));
  return 0;
}

// uniform grid (initial_condition/main.py:15-24) and metrics
// (src/domain.c:98-146)
int build_metrics(domain_t* d) {
  const size_t nx = d->glsizes[0];
  d->hxxf = (double*)calloc(nx + 1, sizeof(double));
  d->hxxc = (double*)calloc(nx, sizeof(double));
  d->jdxf = (double*)calloc(nx + 1, sizeof(double));
  d->jdxc = (double*)calloc(nx, sizeof(double));
  SBFS_REQUIRE(d->hxxf && d->hxxc && d->jdxf && d->jdxc, "host allocation failed");
  d->hy = d->lengths[1] / (double)d->glsizes[1];
  for (size_t i = 1; i <= nx + 1; i++) d->hxxf[i - 1] = d->xc[i] - d->xc[i - 1];
  for (size_t i = 1; i <= nx; i++) d->hxxc[i - 1] = d->xf[i] - d->xf[i - 1];
  for (size_t i = 1; i <= nx + 1; i++) d->jdxf[i - 1] = d->hxxf[i - 1] * d->hy;
  for (size_t i = 1; i <= nx; i++) d->jdxc[i - 1] = d->hxxc[i - 1] * d->hy;
  // src/domain.c:77-96, :213-219: only uniform x grids are accepted
  double lo = 1.e300, hi = -1.e300;
  for (size_t i = 0; i < nx; i++) {
    lo = fmin(lo, d->hxxc[i]);
    hi = fmax(hi, d->hxxc[i]);
  }
  SBFS_REQUIRE(fabs(hi - lo) < 1.e-15, "x grid is not uniform, which is not allowed");
  return 0;
}

}  // namespace sbfs

using namespace sbfs;

// ---------------------------------------------------------------------------
// include/sbfs.h
// ---------------------------------------------------------------------------
extern "C" {

const char* sbfs_last_error(void) { return last_error(); }

int sbfs_synchronize(void) {
  if (init_runtime()) return 1;
  SBFS_HIP(hipStreamSynchronize(stream()));
  return 0;
}

int sbfs_device_name(char* buf, size_t len) {
  if (init_runtime()) return 1;
  int dev = 0;
  SBFS_HIP(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  SBFS_HIP(hipGetDeviceProperties(&prop, dev));
  snprintf(buf, len, "%s (%s)", prop.name, prop.gcnArchName);
  return 0;
}

int sbfs_set_dt_coefs(double adv, double dif, double intf) {
  Options& o = options();
  o.cadv = adv;
  o.cdif = dif;
  o.cint = intf;
  o.coefs_set = true;
  return 0;
}

int sbfs_set_path(int path) {
  SBFS_REQUIRE(path == SBFS_PATH_STAGED || path == SBFS_PATH_FUSED, "unknown path %d", path);
  options().path = path;
  return 0;
}

int sbfs_get_path(void) { return options().path; }

int sbfs_set_gauge(int gauge) {
  SBFS_REQUIRE(gauge == SBFS_GAUGE_PINNED || gauge == SBFS_GAUGE_REFERENCE, "unknown gauge %d", gauge);
  options().gauge = gauge;
  return 0;
}

int sbfs_set_precision(int precision) {
  SBFS_REQUIRE(precision == SBFS_PRECISION_FP64 || precision == SBFS_PRECISION_MIXED,
               "unknown precision %d", precision);
  options().precision = precision;
  return 0;
}

int sbfs_set_curvature(int mode) {
  SBFS_REQUIRE(mode == SBFS_CURVATURE_PLAIN || mode == SBFS_CURVATURE_LEARNED, "unknown curvature mode %d", mode);
  options().curvature = mode;
  return 0;
}

int sbfs_set_tdm(int kind, int chunk) {
  SBFS_REQUIRE(kind == SBFS_TDM_SEQUENTIAL || kind == SBFS_TDM_PARTITIONED, "unknown tdm kind %d", kind);
  SBFS_REQUIRE(chunk >= 2 && chunk <= 64 && (chunk & (chunk - 1)) == 0,
               "chunk must be a power of two in [2, 64]");
  options().tdm = kind;
  options().chunk = chunk;
  return 0;
}

int sbfs_domain_create(size_t nx, size_t ny, double lx, double ly, domain_t* d) {
  SBFS_REQUIRE(d != nullptr, "null domain");
  SBFS_REQUIRE(nx >= 2 && ny >= 4, "domain %zu x %zu too small", nx, ny);
  memset(d, 0, sizeof(*d));
  d->glsizes[0] = d->mysizes[0] = nx;
  d->glsizes[1] = d->mysizes[1] = ny;
  d->lengths[0] = lx;
  d->lengths[1] = ly;
  d->xf = (double*)calloc(nx + 1, sizeof(double));
  d->xc = (double*)calloc(nx + 2, sizeof(double));
  SBFS_REQUIRE(d->xf && d->xc, "host allocation failed");
  // numpy.linspace(0, lx, nx+1): k * (lx / nx), last point exactly lx
  const double step = lx / (double)nx;
  for (size_t k = 0; k <= nx; k++) d->xf[k] = (double)k * step;
  d->xf[nx] = lx;
  d->xc[0] = 0.;
  for (size_t k = 0; k < nx; k++) d->xc[k + 1] = 0.5 * d->xf[k] + 0.5 * d->xf[k + 1];
  d->xc[nx + 1] = lx;
  return build_metrics(d);
}

int sbfs_domain_destroy(domain_t* d) {
  if (!d) return 0;
  forget_domain(d);
  free(d->xf);
  free(d->xc);
  free(d->hxxf);
  free(d->hxxc);
  free(d->jdxf);
  free(d->jdxc);
  memset(d, 0, sizeof(*d));
  return 0;
}

int sbfs_fluid_create(const domain_t* d, double Re, double Fr, double denr,
                      double visr, fluid_t* f) {
  SBFS_REQUIRE(d && f, "null argument");
  memset(f, 0, sizeof(*f));
  if (allocate_fluid(d, f)) return 1;
  f->Re = Re;
  f->Fr = Fr;
  f->denr = denr;
  f->visr = visr;
  f->refden = fmin(1., denr);  // src/fluid/init.c:106
  return 0;
}

int sbfs_fluid_destroy(fluid_t* f) {
  if (!f) return 0;
  (void)hipStreamSynchronize(stream());
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

int sbfs_interface_create(const domain_t* d, double We, interface_t* s) {
  SBFS_REQUIRE(d && s, "null argument");
  memset(s, 0, sizeof(*s));
  if (allocate_interface(d, s)) return 1;
  s->We = We;
  return 0;
}

int sbfs_interface_destroy(interface_t* s) {
  if (!s) return 0;
  (void)hipStreamSynchronize(stream());
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

int sbfs_array_upload(array_t* a, const void* host, size_t bytes) {
  if (init_runtime()) return 1;
  SBFS_REQUIRE(a && a->data && host, "null argument");
  SBFS_REQUIRE(bytes == a->datasize, "upload of %zu bytes into an array of %zu", bytes, a->datasize);
  SBFS_HIP(hipMemcpyAsync(a->data, host, bytes, hipMemcpyHostToDevice, stream()));
  SBFS_HIP(hipStreamSynchronize(stream()));
  return 0;
}

int sbfs_array_download(const array_t* a, void* host, size_t bytes) {
  if (init_runtime()) return 1;
  SBFS_REQUIRE(a && a->data && host, "null argument");
  SBFS_REQUIRE(bytes == a->datasize, "download of %zu bytes from an array of %zu", bytes, a->datasize);
  SBFS_HIP(hipMemcpyAsync(host, a->data, bytes, hipMemcpyDeviceToHost, stream()));
  SBFS_HIP(hipStreamSynchronize(stream()));
  return 0;
}

}  // extern "C"
