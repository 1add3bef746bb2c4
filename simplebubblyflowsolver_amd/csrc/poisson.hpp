// Poisson solver state shared by poisson.hip and tdm_partition.hip.
#pragma once

#include <hip/hip_runtime.h>

#include "sbfs_internal.hpp"

namespace sbfs {

// work item of workgroup b among n: consecutive items on one XCD (the
// hardware deals workgroups to the 8 XCDs in turn) when n is a multiple of 8
__device__ __forceinline__ int xcd_group(int b, int n) {
  return (n & 7) ? b : (b & 7) * (n >> 3) + (b >> 3);
}

struct PoissonPlan {
  int nx = 0, ny = 0, m = 0;
  double lx = 0.;
  double* buf = nullptr;     // nx * ny, modes r[j][m]
  double* rhs = nullptr;     // nx * ny, right-hand side rows (marching path)
  double* v = nullptr;       // nx * ny, sequential-solver factors
  double* q1 = nullptr;      // nx * ny, sequential-solver spike
  double* evals = nullptr;   // nx
  double2* tw = nullptr;     // m      FFT twiddles e^{-2 pi i e/m}
  double2* t1 = nullptr;     // m + 1  e^{-2 pi i k/n}
  double2* t2 = nullptr;     // n      e^{-i pi k/(2n)}
  // partitioned solver
  int L = 0, P = 0;
  double a = 0.;             // off-diagonal 1/hy^2
  double* ptab = nullptr;    // [4][K][nx]: inv, v, w, z per (row-in-chunk, mode)
  double* htab = nullptr;    // [P][nx]: first column of the reduced circulant's inverse
  double* gfl = nullptr;     // [2P][nx]: first / last interior of each chunk
  double* X = nullptr;       // [P][nx]: separators
};

// one case's Poisson transforms in an ensemble (sequential tridiagonal)
struct RowArgs {
  const double* rhs;  // right-hand side rows (the stage kernel's output)
  double* buf;        // modes
  double* v;
  double* q1;
  double* psi;        // psi[1] of the stage
  const double* evals;
  const double2* tw;
  const double2* t1;
  const double2* t2;
  double l;
  int nx, ny, m, pin;
  // partitioned solve (L != 0; L == 0: sequential)
  const double* ptab;
  const double* htab;
  double* gfl;
  double* X;
  double pa;
  int L, P;
};
int poisson_row_args(const domain_t* d, DomainCtx* c, double* psi, RowArgs* a);
__global__ void k_dct2_rows_batch(const RowArgs* tab, const uint2* map);
__global__ void k_tdm_sequential_batch(const RowArgs* tab, const uint2* map);
__global__ void k_dct3_psi_batch(const RowArgs* tab, const uint2* map);
__global__ void k_part_reduce_batch(const RowArgs* tab, const uint2* map);
__global__ void k_part_conv_batch(const RowArgs* tab, const uint2* map);
constexpr int kConvModesBatch = 4;  // modes per workgroup of k_part_conv_batch

int partition_chunk(int n, int requested);
// largest reduced system of the partitioned solve: the convolution's
// rounding grows like P^2 for the lowest modes (at 8192 rows, relative
// error 6.7e-13 with P = 128 against 4.9e-12 with P = 256 and 7.0e-13 for
// the reference-order sequential solve), so long columns take longer chunks
constexpr int kMaxSeparators = 128;
int poisson_chunk(int ny);
int partition_setup(PoissonPlan* p, const domain_t* d, int L);
// row transforms for m complex points: rows, threads and LDS bytes of one
// unit (the threads that transform one group of rows together)
int wave_rows(int m);
int wave_unit(int m);
size_t wave_lds(int m);
int partition_solve(PoissonPlan* p, int pin_mode0);

}  // namespace sbfs
