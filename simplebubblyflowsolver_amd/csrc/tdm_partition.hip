// Partitioned periodic tridiagonal solve along y for every x-mode.
//
// The reference solves each mode with one sequential Thomas +
// Sherman-Morrison chain of length Ny (src/tdm.c:66-106); on a GPU that is
// Nx latency-bound chains.  Here each column is cut into P chunks of L rows
// whose last row is a separator:
//   T1  thread (mode, chunk): Thomas on the L-1 interior rows with zero ends
//       -> first / last interior value                       (reads r once)
//   T2  thread per mode: the separators satisfy a periodic tridiagonal
//       system of size P with constant coefficients, solved by the
//       reference algorithm; the singular mode 0 pins the last separator,
//       i.e. x_{Ny-1} = 0, the reference's singular branch (src/tdm.c:94)
//   T3  thread (mode, chunk): recompute the interior solution and add the
//       two unit spikes weighted by the bounding separators (writes x)
// The chunk matrices are identical for all chunks of a mode, so the Thomas
// factors and the spikes are per-(mode, row-in-chunk) tables built once.
// Data stay in the [j][m] layout: a wavefront handles 64 consecutive modes,
// so every row access is one contiguous 512-byte segment.
#include <cfloat>
#include <cmath>
#include <vector>

#include "poisson.hpp"
#include "sbfs_internal.hpp"

namespace sbfs {



// reference periodic algorithm on a strided column (src/tdm.c:17-99)
__device__ void periodic_tdm(double* __restrict__ q, double* __restrict__ v,
                             double* __restrict__ q1, size_t st, int n, double l, double c,
                             double u, bool pin) {
  const int N = n - 1;
  for (int r = 0; r < N; r++) q1[r * st] = (r == 0) ? -l : (r == n - 2) ? -u : 0.;
  v[0] = u / c;
  q1[0] = q1[0] / c;
  q[0] = q[0] / c;
  for (int r = 1; r < N - 1; r++) {
    const double val = 1. / (c - l * v[(r - 1) * st]);
    v[r * st] = val * u;
    q1[r * st] = val * (q1[r * st] - l * q1[(r - 1) * st]);
    q[r * st] = val * (q[r * st] - l * q[(r - 1) * st]);
  }
  {
    const double val = c - l * v[(N - 2) * st];
    if (fabs(val) > DBL_EPSILON) {
      q1[(N - 1) * st] = 1. / val * (q1[(N - 1) * st] - l * q1[(N - 2) * st]);
      q[(N - 1) * st] = 1. / val * (q[(N - 1) * st] - l * q[(N - 2) * st]);
    } else {
      q1[(N - 1) * st] = 0.;
      q[(N - 1) * st] = 0.;
    }
  }
  for (int r = N - 2; r >= 0; r--) {
    q1[r * st] -= v[r * st] * q1[(r + 1) * st];
    q[r * st] -= v[r * st] * q[(r + 1) * st];
  }
  const double num = q[(n - 1) * st] - u * q[0] - l * q[(n - 2) * st];
  const double den = c + u * q1[0] + l * q1[(n - 2) * st];
  const double xl = (pin || fabs(den) < DBL_EPSILON) ? 0. : num / den;
  q[(n - 1) * st] = xl;
  for (int r = 0; r < n - 1; r++) q[r * st] = q[r * st] + xl * q1[r * st];
}

// Thomas on one chunk interior with the per-mode tables:
// forward g_k = inv_k (f_k - a g_{k-1}); backward g_k -= v_k g_{k+1}
template <int K>
__device__ __forceinline__ void chunk_thomas(const double* __restrict__ q, int nx, int mo,
                                             int row0, const double* __restrict__ inv,
                                             const double* __restrict__ vv, double a,
                                             double (&g)[K]) {
  double prev = 0.;
#pragma unroll
  for (int k = 0; k < K; k++) {
    const double f = q[(size_t)(row0 + k) * nx + mo];
    g[k] = inv[(size_t)k * nx + mo] * (f - a * prev);
    prev = g[k];
  }
#pragma unroll
  for (int k = K - 2; k >= 0; k--) g[k] = g[k] - vv[(size_t)k * nx + mo] * g[k + 1];
}

template <int L>
__global__ void k_part_reduce(const double* __restrict__ q, const double* __restrict__ tab,
                              double* __restrict__ gfl, int nx, int P, double a) {
  constexpr int K = L - 1;
  const int mo = blockIdx.x * blockDim.x + threadIdx.x;
  const int ch = blockIdx.y * blockDim.y + threadIdx.y;
  if (mo >= nx || ch >= P) return;
  const double* inv = tab;
  const double* vv = tab + (size_t)K * nx;
  double g[K];
  chunk_thomas<K>(q, nx, mo, ch * L, inv, vv, a, g);
  gfl[(size_t)(2 * ch) * nx + mo] = g[0];
  gfl[(size_t)(2 * ch + 1) * nx + mo] = g[K - 1];
}

// Separators of every mode: assemble the reduced right-hand side and solve
// the periodic reduced system with the reference algorithm.  Its pivots,
// factors and spike depend only on the mode, so they come from tables
// (rtab, pred) built with the same arithmetic.  One wavefront handles 64
// modes; the separators live in LDS ([P][64], P <= kMaxP) and the table
// values are fetched in groups ahead of the dependent chain, whose only
// carried value stays in a register.
constexpr int kMaxP = 128;
constexpr int kSepGroup = 16;

template <int L>
__global__ __launch_bounds__(64) void k_part_separators(
    const double* __restrict__ q, const double* __restrict__ gfl, const double* __restrict__ pred,
    const double* __restrict__ rtab, double* __restrict__ X, int nx, int P, double a,
    int pin_mode0) {
  constexpr int K = L - 1;
  __shared__ double sx[kMaxP * 64];
  const int t = threadIdx.x;
  const int mo = blockIdx.x * 64 + t;
  const bool live = mo < nx;
  const int mm = live ? mo : nx - 1;
  for (int c = 0; c < P; c++) {
    const double gl = gfl[(size_t)(2 * c + 1) * nx + mm];
    const double gf = gfl[(size_t)(2 * ((c + 1) % P)) * nx + mm];
    sx[c * 64 + t] = (q[(size_t)(c * L + K) * nx + mm] - a * gl) - a * gf;
  }
  const double lo = pred[mm], di = pred[(size_t)nx + mm], up = pred[(size_t)2 * nx + mm];
  const double ilast = pred[(size_t)3 * nx + mm], ok = pred[(size_t)4 * nx + mm];
  const double den = pred[(size_t)5 * nx + mm];
  const double* val = rtab;
  const double* vv = rtab + (size_t)P * nx;
  const double* q1 = rtab + (size_t)2 * P * nx;
  const int N = P - 1;
  // gtsv on the first N rows (src/tdm.c:26-48)
  double prev = sx[t] / di;
  sx[t] = prev;
  for (int r0 = 1; r0 < N - 1; r0 += kSepGroup) {
    double tv[kSepGroup];
#pragma unroll
    for (int u = 0; u < kSepGroup; u++) tv[u] = (r0 + u < N - 1) ? val[(size_t)(r0 + u) * nx + mm] : 0.;
#pragma unroll
    for (int u = 0; u < kSepGroup; u++) {
      const int r = r0 + u;
      if (r < N - 1) {
        prev = tv[u] * (sx[r * 64 + t] - lo * prev);
        sx[r * 64 + t] = prev;
      }
    }
  }
  double nxt = ok != 0. ? ilast * (sx[(N - 1) * 64 + t] - lo * prev) : 0.;
  sx[(N - 1) * 64 + t] = nxt;
  for (int r0 = N - 2; r0 >= 0; r0 -= kSepGroup) {
    double tv[kSepGroup];
#pragma unroll
    for (int u = 0; u < kSepGroup; u++) tv[u] = (r0 - u >= 0) ? vv[(size_t)(r0 - u) * nx + mm] : 0.;
#pragma unroll
    for (int u = 0; u < kSepGroup; u++) {
      const int r = r0 - u;
      if (r >= 0) {
        nxt = sx[r * 64 + t] - tv[u] * nxt;
        sx[r * 64 + t] = nxt;
      }
    }
  }
  // closure (src/tdm.c:92-98); the singular mode pins the last separator
  const double num = sx[(P - 1) * 64 + t] - up * sx[t] - lo * sx[(P - 2) * 64 + t];
  const double xl = ((pin_mode0 && mo == 0) || fabs(den) < DBL_EPSILON) ? 0. : num / den;
  if (!live) return;
  X[(size_t)(P - 1) * nx + mo] = xl;
  for (int r = 0; r < P - 1; r++) X[(size_t)r * nx + mo] = sx[r * 64 + t] + xl * q1[(size_t)r * nx + mo];
}

template <int L>
__global__ void k_part_expand(double* __restrict__ q, const double* __restrict__ tab,
                              const double* __restrict__ X, int nx, int P, double a) {
  constexpr int K = L - 1;
  const int mo = blockIdx.x * blockDim.x + threadIdx.x;
  const int ch = blockIdx.y * blockDim.y + threadIdx.y;
  if (mo >= nx || ch >= P) return;
  const double* inv = tab;
  const double* vv = tab + (size_t)K * nx;
  const double* w = tab + (size_t)2 * K * nx;
  const double* z = tab + (size_t)3 * K * nx;
  double g[K];
  chunk_thomas<K>(q, nx, mo, ch * L, inv, vv, a, g);
  const double xl = X[(size_t)((ch + P - 1) % P) * nx + mo];
  const double xu = X[(size_t)ch * nx + mo];
#pragma unroll
  for (int k = 0; k < K; k++)
    q[(size_t)(ch * L + k) * nx + mo] = (g[k] + xl * w[(size_t)k * nx + mo]) + xu * z[(size_t)k * nx + mo];
  q[(size_t)(ch * L + K) * nx + mo] = xu;
}

// per-mode tables (same arithmetic as the oracle's chunk tables)
static void host_thomas(int K, const double* inv, const double* v, double a, const double* f,
                        double* g) {
  double prev = 0.;
  for (int k = 0; k < K; k++) {
    g[k] = inv[k] * (f[k] - a * prev);
    prev = g[k];
  }
  for (int k = K - 2; k >= 0; k--) g[k] = g[k] - v[k] * g[k + 1];
}

int partition_setup(PoissonPlan* p, const domain_t* d, int L) {
  const int nx = p->nx, ny = p->ny, K = L - 1, P = ny / L;
  (void)hipFree(p->ptab);
  (void)hipFree(p->pred);
  (void)hipFree(p->gfl);
  (void)hipFree(p->X);
  (void)hipFree(p->rtab);
  p->ptab = p->pred = p->gfl = p->X = p->rtab = nullptr;
  std::vector<double> ev(nx);
  SBFS_HIP(hipMemcpy(ev.data(), p->evals, sizeof(double) * nx, hipMemcpyDeviceToHost));
  const double l = 1. / d->hy / d->hy;
  const double u = l;
  std::vector<double> tab((size_t)4 * K * nx), red((size_t)6 * nx), rtab((size_t)3 * P * nx, 0.);
  std::vector<double> rv(P), rq(P), rval(P);
  std::vector<double> inv(K), vv(K), w(K), z(K), f(K);
  for (int mo = 0; mo < nx; mo++) {
    const double a = l;
    const double b = -l - u + ev[mo];
    double vp = 0.;
    for (int k = 0; k < K; k++) {
      inv[k] = 1. / (b - a * vp);
      vv[k] = inv[k] * a;
      vp = vv[k];
    }
    std::fill(f.begin(), f.end(), 0.);
    f[0] = -a;
    host_thomas(K, inv.data(), vv.data(), a, f.data(), w.data());
    std::fill(f.begin(), f.end(), 0.);
    f[K - 1] = -a;
    host_thomas(K, inv.data(), vv.data(), a, f.data(), z.data());
    for (int k = 0; k < K; k++) {
      tab[(size_t)(0 * K + k) * nx + mo] = inv[k];
      tab[(size_t)(1 * K + k) * nx + mo] = vv[k];
      tab[(size_t)(2 * K + k) * nx + mo] = w[k];
      tab[(size_t)(3 * K + k) * nx + mo] = z[k];
    }
    const double lo = a * w[K - 1];
    const double di = (b + a * z[K - 1]) + a * w[0];
    const double up = a * z[0];
    // data-independent part of the reduced periodic solve (src/tdm.c:78-94
    // with l = lo, c = di, u = up): pivots, factors, spike q1, denominator
    const int N = P - 1;
    std::fill(rv.begin(), rv.end(), 0.);
    std::fill(rval.begin(), rval.end(), 0.);
    for (int r = 0; r < N; r++) rq[r] = (r == 0) ? -lo : (r == P - 2) ? -up : 0.;
    rq[N] = 0.;
    rv[0] = up / di;
    rq[0] = rq[0] / di;
    for (int r = 1; r < N - 1; r++) {
      rval[r] = 1. / (di - lo * rv[r - 1]);
      rv[r] = rval[r] * up;
      rq[r] = rval[r] * (rq[r] - lo * rq[r - 1]);
    }
    const double lval = di - lo * rv[N - 2];
    double ilast = 0., ok = 0.;
    if (fabs(lval) > DBL_EPSILON) {
      ilast = 1. / lval;
      ok = 1.;
      rq[N - 1] = ilast * (rq[N - 1] - lo * rq[N - 2]);
    } else {
      rq[N - 1] = 0.;
    }
    for (int r = N - 2; r >= 0; r--) rq[r] -= rv[r] * rq[r + 1];
    const double den = di + up * rq[0] + lo * rq[P - 2];
    for (int r = 0; r < P; r++) {
      rtab[(size_t)(0 * P + r) * nx + mo] = rval[r];
      rtab[(size_t)(1 * P + r) * nx + mo] = rv[r];
      rtab[(size_t)(2 * P + r) * nx + mo] = rq[r];
    }
    red[mo] = lo;
    red[nx + mo] = di;
    red[2 * nx + mo] = up;
    red[3 * nx + mo] = ilast;
    red[4 * nx + mo] = ok;
    red[5 * nx + mo] = den;
  }
  SBFS_HIP(hipMalloc(&p->ptab, sizeof(double) * tab.size()));
  SBFS_HIP(hipMalloc(&p->pred, sizeof(double) * red.size()));
  SBFS_HIP(hipMalloc(&p->gfl, sizeof(double) * 2 * (size_t)P * nx));
  SBFS_HIP(hipMalloc(&p->X, sizeof(double) * (size_t)P * nx));
  SBFS_HIP(hipMalloc(&p->rtab, sizeof(double) * rtab.size()));
  SBFS_HIP(hipMemcpy(p->rtab, rtab.data(), sizeof(double) * rtab.size(), hipMemcpyHostToDevice));
  SBFS_HIP(hipMemcpy(p->ptab, tab.data(), sizeof(double) * tab.size(), hipMemcpyHostToDevice));
  SBFS_HIP(hipMemcpy(p->pred, red.data(), sizeof(double) * red.size(), hipMemcpyHostToDevice));
  p->L = L;
  p->P = P;
  p->a = l;
  return 0;
}

template <int L>
static void part_launch(PoissonPlan* p, double a, int pin) {
  const int nx = p->nx, P = p->P;
  const dim3 blk(64, 4), grd(cdiv(nx, 64), cdiv(P, 4));
  SBFS_TIMED("k_part_reduce",
             hipLaunchKernelGGL(k_part_reduce<L>, grd, blk, 0, stream(), p->buf, p->ptab, p->gfl,
                                nx, P, a));
  SBFS_TIMED("k_part_separators",
             hipLaunchKernelGGL(k_part_separators<L>, dim3(cdiv(nx, 64)), dim3(64), 0, stream(),
                                p->buf, p->gfl, p->pred, p->rtab, p->X, nx, P, a, pin));
  SBFS_TIMED("k_part_expand",
             hipLaunchKernelGGL(k_part_expand<L>, grd, blk, 0, stream(), p->buf, p->ptab, p->X,
                                nx, P, a));
}

int partition_solve(PoissonPlan* p, int pin) {
  switch (p->L) {
    case 2: part_launch<2>(p, p->a, pin); break;
    case 4: part_launch<4>(p, p->a, pin); break;
    case 8: part_launch<8>(p, p->a, pin); break;
    case 16: part_launch<16>(p, p->a, pin); break;
    case 32: part_launch<32>(p, p->a, pin); break;
    case 64: part_launch<64>(p, p->a, pin); break;
    default: set_error("partitioned tridiagonal: unsupported chunk %d", p->L); return 1;
  }
  return check_launch("poisson: partitioned tridiagonal");
}

}  // namespace sbfs
