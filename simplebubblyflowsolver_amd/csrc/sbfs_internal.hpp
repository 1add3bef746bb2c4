// Internal declarations of libsbfs: device context, layouts, error handling.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "domain.h"
#include "array.h"
#include "fluid.h"
#include "interface.h"

namespace sbfs {

// ---------------------------------------------------------------------------
// errors: every entry point returns non-zero and records a message
// ---------------------------------------------------------------------------
void set_error(const char* fmt, ...);
const char* last_error();

#define SBFS_HIP(expr)                                                        \
  do {                                                                        \
    hipError_t sbfs_e_ = (expr);                                              \
    if (sbfs_e_ != hipSuccess) {                                              \
      ::sbfs::set_error("%s:%d %s: %s", __FILE__, __LINE__, #expr,             \
                        hipGetErrorString(sbfs_e_));                          \
      return 1;                                                               \
    }                                                                         \
  } while (0)

#define SBFS_REQUIRE(cond, ...)                                               \
  do {                                                                        \
    if (!(cond)) {                                                            \
      ::sbfs::set_error(__VA_ARGS__);                                         \
      return 1;                                                               \
    }                                                                         \
  } while (0)

// check the launch status of the kernels just queued
int check_launch(const char* what);

// ---------------------------------------------------------------------------
// field layouts (reference include/array_macros/): x fastest, 1-based
// interior indices, per-field ghost widths.  `nx` is the interior width.
// ---------------------------------------------------------------------------
struct L {
  // ux, visux, fluxx, txy, dvof: i = 1..nx+1, j = 0..ny+1 (dvof: ..ny+2)
  __host__ __device__ static inline size_t xf(int i, int j, int nx) {
    return (size_t)(i - 1) + (size_t)(nx + 1) * (size_t)j;
  }
  // uy, p, psi, den, visuy, curv, fluxy: i = 0..nx+1, j = 0..ny+1
  __host__ __device__ static inline size_t cc(int i, int j, int nx) {
    return (size_t)i + (size_t)(nx + 2) * (size_t)j;
  }
  // vof: i = 0..nx+1, j = -1..ny+2
  __host__ __device__ static inline size_t vof(int i, int j, int nx) {
    return (size_t)i + (size_t)(nx + 2) * (size_t)(j + 1);
  }
  // txx, tyy, normal: i = 1..nx, j = 0..ny+1
  __host__ __device__ static inline size_t cx(int i, int j, int nx) {
    return (size_t)(i - 1) + (size_t)nx * (size_t)j;
  }
  // srcux, ifrcx: i = 2..nx, j = 1..ny
  __host__ __device__ static inline size_t fx(int i, int j, int nx) {
    return (size_t)(i - 2) + (size_t)(nx - 1) * (size_t)(j - 1);
  }
  // srcuy, ifrcy, src: i = 1..nx, j = 1..ny
  __host__ __device__ static inline size_t ci(int i, int j, int nx) {
    return (size_t)(i - 1) + (size_t)nx * (size_t)(j - 1);
  }
};

// ghost widths of each field: {{x lo, x hi}, {y lo, y hi}}
struct Nadds {
  int v[2][2];
};
extern const Nadds NA_UX, NA_UY, NA_P, NA_DEN, NA_VISUX, NA_VISUY, NA_TXX,
    NA_TXY, NA_TYY, NA_SRCUX, NA_SRCUY, NA_VOF, NA_IFRCX, NA_IFRCY, NA_DVOF,
    NA_NORMAL, NA_CURV, NA_FLUXX, NA_FLUXY, NA_SRC;

// verify that an array_t matches the layout the kernels assume
int check_array(const domain_t* d, const array_t* a, const Nadds& na,
                size_t elem, const char* name);

// ---------------------------------------------------------------------------
// constants (reference src/runge_kutta.c, src/interface/common.c)
// ---------------------------------------------------------------------------
struct RK {
  double a, b, g;
};
RK rk(size_t stage);

constexpr double kVofMin = 1.e-8;   // src/interface/common.c:14
constexpr double kVofBeta = 1.;     // src/interface/common.c:13
constexpr double kGaussP = 0.2886751345948129;  // src/interface/common.c:4-7
constexpr double kGaussW = 0.5;                  // src/interface/common.c:8-11
constexpr double kUyXm = 0.;        // src/param/boundary-condition.c:3
constexpr double kUyXp = 0.;        // src/param/boundary-condition.c:4

// ---------------------------------------------------------------------------
// device state per domain
// ---------------------------------------------------------------------------
struct PoissonPlan;  // poisson.hip

struct DomainCtx {
  // identity of the host domain this mirrors
  size_t nx = 0, ny = 0;
  double lx = 0., ly = 0., hy = 0.;
  uint64_t hash = 0;
  // device copies of the metrics (1-based through -1 offsets, as the host)
  double* hxxf = nullptr;  // nx+1
  double* hxxc = nullptr;  // nx
  double* jdxf = nullptr;  // nx+1
  double* jdxc = nullptr;  // nx
  double* xc = nullptr;    // nx+2
  // scratch for reductions
  double* red = nullptr;   // kRedSlots doubles
  double* dt_dev = nullptr;  // device-resident dt (1 double)
  double* host_pin = nullptr;  // pinned host staging (16 doubles)
  PoissonPlan* poisson = nullptr;
};

constexpr int kRedBlocks = 1024;          // partial-reduction blocks
constexpr int kRedSlots = kRedBlocks * 8 + 64;

struct Options {
  int path = 1;        // SBFS_PATH_FUSED
  int gauge = 0;       // SBFS_GAUGE_PINNED
  int tdm = 1;         // SBFS_TDM_PARTITIONED
  int chunk = 32;
  int precision = 0;   // SBFS_PRECISION_FP64
  int curvature = 0;   // SBFS_CURVATURE_PLAIN
  bool coefs_set = false;
  double cadv = 0., cdif = 0., cint = 0.;
};

Options& options();
hipStream_t stream();
// a second default-priority stream for work that overlaps the library
// stream within a step, and fork / join events (k = 0..2)
hipStream_t side_stream();
hipEvent_t fork_event(int k);
hipEvent_t join_event(int k);
int init_runtime();
// device context of a host domain (created / refreshed on demand)
DomainCtx* domain_ctx(const domain_t* d);

// NPY v1 file `dir/name.npy` of `shape` (C order) with dtype `descr`
// ("<f8", "<u8"); shape {} writes a 0-d array (npy_io.hip)
int npy_write(const char* dir, const char* name, const char* descr,
              const std::vector<size_t>& shape, size_t elem, const void* data);
int load_dt_coefs();

// The FP64 stage kernels address a field by its base and a 32-bit byte
// offset (march.hip): every field of the grid, the vof with its two halo
// rows on each side the largest, must stay below 4 GiB.
inline bool fp64_march_addressable(size_t nx, size_t ny) { return (ny + 4) * (nx + 2) * 8 < (1ull << 32); }
// Whether the single-case kernels stream the rows they touch once
// (non-temporal loads and stores): only when an FP64 field outgrows what the
// 256 MB Infinity Cache keeps between one kernel and the next; below that,
// ordinary accesses leave the producer's rows there for the consumer.
// SBFS_STREAM=0 / =1 forces either (the results are bitwise the same).
constexpr size_t kStreamFieldBytes = 128ull << 20;
inline bool stream_fields(size_t nx, size_t ny) {
  if (const char* e = getenv("SBFS_STREAM")) {
    if (e[0] == '0') return false;
    if (e[0] == '1') return true;
  }
  return (ny + 4) * (nx + 2) * 8 >= kStreamFieldBytes;
}
// the learned curvature's band: cells within one cell (3 x 3 block) of a
// cut cell, vof in [kCutLo, kCutHi]
constexpr double kCutLo = 0.02, kCutHi = 0.98;
// learned curvature (mlp.hip): kappa of the interior cells of `vof`, NaN
// where the plain curvature stays; h is the (uniform, square) cell size.
// The field keeps NaN outside the cells of the last launch's list (each
// list: nx ny + 1 entries, the count, then the listed cells).
// curv_field_init fills a new field with NaN and empties up to three lists;
// curv_mlp_launch lists the cells of `vof` into `list`, whose count must be
// zero, returns the cells of `prev` (the previous launch's list, null for
// none) to NaN and empties `next` (the list of the launch after this one,
// null for none), then fills the listed cells.  A step's stages cycle
// through three lists.
int curv_field_init(float* kappa, unsigned* const* lists, int nlists, long cells, hipStream_t st);
// `cut`, when not null, is the cut-cell bitmask that the stage kernel wrote
// with this vof (march.hpp), which the gate then reads instead of the vof.
int curv_mlp_launch(const double* vof, const unsigned long long* cut, float* kappa, unsigned* list,
                    const unsigned* prev, unsigned* next, int nx, int ny, double h, hipStream_t st);
// Captured step graphs hold device addresses that their keys do not list
// (Poisson plan buffers, domain metrics).
// Whatever frees such memory bumps this epoch, which is part of every key,
// so no graph replays against freed memory, even when a new allocation
// reuses the old addresses.
void graph_epoch_bump();
uint64_t graph_epoch();

// kernel timing (profiler.hip): events around launches while profiling
bool prof_active();  // kernel timing on (sbfs_profile_start)
void prof_begin(const char* name);
void prof_end();
#define SBFS_TIMED(name, ...) \
  do {                        \
    ::sbfs::prof_begin(name); \
    __VA_ARGS__;              \
    ::sbfs::prof_end();       \
  } while (0)

// arguments of one case's decide_dt kernels in an ensemble (step.hip)
struct DtArgs {
  const double* ux;
  const double* uy;
  const double* hxxf;
  double hy, cadv, dif0, dif1, dint;
  double* partial;
  double* dt;
  int nx, ny, rows, gx, nblk;
  int stream;  // streaming loads (stream_fields)
};
int dt_args(const domain_t* d, DomainCtx* c, const fluid_t* f, const interface_t* s, DtArgs* a);
__global__ void k_dt_partial_batch(const DtArgs* tab, const uint2* map);
__global__ void k_dt_final_batch(const DtArgs* tab);

// grid-dimension helpers
inline unsigned cdiv(long a, long b) { return (unsigned)((a + b - 1) / b); }

// ---------------------------------------------------------------------------
// stage launchers (stages.hip, poisson.hip, step.hip); all queue on stream()
// ---------------------------------------------------------------------------
int launch_density(const domain_t* d, DomainCtx* c, const double* vof,
                   double* den, double denr);
int poisson_solve(const domain_t* d, DomainCtx* c, double dt_old,
                  double dt_new, fluid_t* f);
int compute_dt_device(const domain_t* d, DomainCtx* c, const fluid_t* f,
                      const interface_t* s);
int dt_grid(int nx, int ny, int capacity, int* rows, int* nblk);
int dt_launch(const domain_t* d, const DomainCtx* c, const fluid_t* f, const interface_t* s,
              int capacity, double* partial, double* dt_out);
// the domain's device context if one exists and is current, without
// creating or refreshing one (nullptr otherwise)
DomainCtx* domain_ctx_find(const domain_t* d);
// the no-allocation variants' per-device set-up (noalloc.hip), run by
// init_runtime for the current device
int noalloc_setup();
void dt_limits(const domain_t* d, const fluid_t* f, const interface_t* s,
               double* dif0, double* dif1, double* dint);

}  // namespace sbfs
