// Partitioned periodic tridiagonal solve along y for every x-mode.
//
// The reference solves each mode with one sequential Thomas +
// Sherman-Morrison chain of length Ny (src/tdm.c:66-106); on a GPU that is
// Nx latency-bound chains.  Here each column is cut into P chunks of L rows
// whose last row is a separator:
//   T1  thread (mode, chunk): Thomas on the L-1 interior rows with zero ends
//       -> first / last interior value; the interior solution overwrites
//       the chunk's rows                                     (reads r once)
//   T2  the separators satisfy a periodic tridiagonal system of size P
//       with constant coefficients (a circulant), solved in parallel as a
//       circular convolution with the first column of its inverse; the
//       singular mode 0 pins the last separator, i.e. x_{Ny-1} = 0, the
//       reference's singular branch (src/tdm.c:94)
//   T3  add the two unit spikes weighted by the bounding separators to the
//       interior solution: T1 leaves that solution in place of the chunk's
//       rows, and the DCT-III kernel (poisson.hip) forms x = (g + xl w) +
//       xu z as it loads each row, so the expansion costs no pass of its own
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

// T1 for one (mode, chunk): Thomas on the chunk interior with the per-mode
// tables, forward g_k = inv_k (f_k - a g_{k-1}), backward g_k -= v_k g_{k+1};
// the interior solution replaces the chunk's rows, and its first and last
// values are returned.  Long chunks keep only their last KR forward values in
// registers (K = 63 in registers would allow one wavefront per SIMD); the
// first K - KR wait for the backward sweep in `spill`, a per-thread LDS
// column of stride kSpillStride when kLds, otherwise their own rows of q
// (read again by the backward sweep: a second pass over half the chunk in
// HBM).  Same operations either way.
constexpr int kSpillStride = 256;
template <int K, bool kLds>
__device__ __forceinline__ void chunk_reduce(double* __restrict__ q, int nx, int mo, int row0,
                                             const double* __restrict__ inv, const double* __restrict__ vv,
                                             double a, double& first, double& last, double* spill) {
  constexpr int KR = K > 40 ? K / 2 + 1 : K;
  constexpr int KM = K - KR;
  double g[KR];
  double prev = 0.;
#pragma unroll 8
  for (int k = 0; k < KM; k++) {
    const double f = q[(size_t)(row0 + k) * nx + mo];
    const double gk = inv[(size_t)k * nx + mo] * (f - a * prev);
    if constexpr (kLds) spill[k * kSpillStride] = gk;
    else q[(size_t)(row0 + k) * nx + mo] = gk;
    prev = gk;
  }
#pragma unroll
  for (int k = 0; k < KR; k++) {
    const double f = q[(size_t)(row0 + KM + k) * nx + mo];
    g[k] = inv[(size_t)(KM + k) * nx + mo] * (f - a * prev);
    prev = g[k];
  }
#pragma unroll
  for (int k = KR - 2; k >= 0; k--) g[k] = g[k] - vv[(size_t)(KM + k) * nx + mo] * g[k + 1];
#pragma unroll
  for (int k = 0; k < KR; k++) q[(size_t)(row0 + KM + k) * nx + mo] = g[k];
  double next = g[0];
#pragma unroll 8
  for (int k = KM - 1; k >= 0; k--) {
    const double fk = kLds ? spill[k * kSpillStride] : q[(size_t)(row0 + k) * nx + mo];
    const double gk = fk - vv[(size_t)k * nx + mo] * next;
    q[(size_t)(row0 + k) * nx + mo] = gk;
    next = gk;
  }
  first = next;
  last = g[KR - 1];
}

template <int K>
constexpr int spill_rows() {
  return K > 40 ? K - (K / 2 + 1) : 0;
}

// 64 modes x 4 chunks per workgroup (kSpillStride threads); the chunks of
// more than 40 rows keep their spilled forward values in LDS
template <int L>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_part_reduce(double* __restrict__ q, const double* __restrict__ tab,
                              double* __restrict__ gfl, int nx, int P, double a) {
  constexpr int K = L - 1;
  constexpr int KM = spill_rows<K>();
  __shared__ double spill[KM > 0 ? KM * kSpillStride : 1];
  const int mo = blockIdx.x * blockDim.x + threadIdx.x;
  const int ch = blockIdx.y * blockDim.y + threadIdx.y;
  if (mo >= nx || ch >= P) return;
  double first, last;
  chunk_reduce<K, (KM > 0)>(q, nx, mo, ch * L, tab, tab + (size_t)K * nx, a, first, last,
                            spill + threadIdx.y * blockDim.x + threadIdx.x);
  gfl[(size_t)(2 * ch) * nx + mo] = first;
  gfl[(size_t)(2 * ch + 1) * nx + mo] = last;
}

// Separators of every mode.  The reduced system is periodic tridiagonal with
// the same coefficients in every row (all chunks are alike), i.e. circulant,
// so its solution is the circular convolution X_i = sum_k h[(i-k) mod P] R_k
// of the reduced right-hand side R with h = C^-1 e_0 (one column per mode,
// built on the host with the reference algorithm).  For the singular mode
// h = C^+ (e_0 - 1/P), and the result is shifted so that X_{P-1} = 0, the
// reference's singular branch.  A workgroup holds R and h of kConvModes
// modes in LDS; each thread forms kConvS consecutive X_i of one mode, reading
// each R_k once and sliding a register window over h.  The sum runs over
// ascending k.  Few modes per workgroup give nx / 4 workgroups, enough to
// fill the chip; consecutive groups go to the same XCD (workgroups are
// dealt to the 8 XCDs in turn), so the 128-byte lines that 4 groups share
// are fetched into one L2.
constexpr int kConvModes = 4;
constexpr int kConvS = 2;


template <int L>
__global__ __launch_bounds__(256) void k_part_conv(const double* __restrict__ q,
                                                   const double* __restrict__ gfl,
                                                   const double* __restrict__ htab,
                                                   double* __restrict__ X, int nx, int P, double a,
                                                   int pin_mode0) {
  constexpr int K = L - 1;
  constexpr int G = 256 / kConvModes;
  extern __shared__ double lds[];
  double* R = lds;                               // [P][kConvModes] reduced right-hand side
  double* H = lds + (size_t)P * kConvModes;      // [P][kConvModes] inverse columns
  double* Y = lds + (size_t)2 * P * kConvModes;  // [P][kConvModes] result
  const int tm = threadIdx.x % kConvModes, g = threadIdx.x / kConvModes;
  const int mo = xcd_group((int)blockIdx.x, (int)gridDim.x) * kConvModes + tm;
  const bool live = mo < nx;
  const int mm = live ? mo : nx - 1;
  for (int c = g; c < P; c += G) {
    const double gl = gfl[(size_t)(2 * c + 1) * nx + mm];
    const double gf = gfl[(size_t)(2 * ((c + 1) % P)) * nx + mm];
    R[c * kConvModes + tm] = (q[(size_t)(c * L + K) * nx + mm] - a * gl) - a * gf;
    H[c * kConvModes + tm] = htab[(size_t)c * nx + mm];
  }
  __syncthreads();
  for (int i0 = g * kConvS; i0 < P; i0 += G * kConvS) {
    double acc[kConvS], w[kConvS];
#pragma unroll
    for (int r = 0; r < kConvS; r++) {
      acc[r] = 0.;
      w[r] = H[((i0 + r) % P) * kConvModes + tm];  // h[(i0 + r - k) mod P] at k = 0
    }
    int next = i0 > 0 ? i0 - 1 : P - 1;  // (i0 - k - 1) mod P
    for (int k = 0; k < P; k++) {
      const double rk = R[k * kConvModes + tm];
#pragma unroll
      for (int r = 0; r < kConvS; r++) acc[r] += w[r] * rk;
#pragma unroll
      for (int r = kConvS - 1; r > 0; r--) w[r] = w[r - 1];
      w[0] = H[next * kConvModes + tm];
      next = next > 0 ? next - 1 : P - 1;
    }
#pragma unroll
    for (int r = 0; r < kConvS; r++)
      if (i0 + r < P) Y[(i0 + r) * kConvModes + tm] = acc[r];
  }
  __syncthreads();
  const bool pin = pin_mode0 && mo == 0;
  const double xl = Y[(P - 1) * kConvModes + tm];
  if (!live) return;
  for (int i = g; i < P; i += G) {
    const double x = Y[i * kConvModes + tm];
    X[(size_t)i * nx + mo] = pin ? x - xl : x;
  }
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

// periodic tridiagonal with constant coefficients by the reference
// algorithm (src/tdm.c:17-99): Thomas on the first n-1 rows for the
// right-hand side and for the corner spike, then the closure for the last
// unknown; `pin` takes the singular branch (last unknown 0)
static void host_gtsv(int n, double l, double c, double u, double* v, double* q) {
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

static void host_periodic_tdm(int n, double l, double c, double u, bool pin, double* q, double* v,
                              double* q1) {
  for (int i = 0; i < n - 1; i++) q1[i] = (i == 0) ? -l : (i == n - 2) ? -u : 0.;
  host_gtsv(n - 1, l, c, u, v, q1);
  host_gtsv(n - 1, l, c, u, v, q);
  const double num = q[n - 1] - u * q[0] - l * q[n - 2];
  const double den = c + u * q1[0] + l * q1[n - 2];
  q[n - 1] = (pin || fabs(den) < DBL_EPSILON) ? 0. : num / den;
  for (int i = 0; i < n - 1; i++) q[i] = q[i] + q[n - 1] * q1[i];
}

int partition_setup(PoissonPlan* p, const domain_t* d, int L) {
  const int nx = p->nx, ny = p->ny, K = L - 1, P = ny / L;
  graph_epoch_bump();
  (void)hipFree(p->ptab);
  (void)hipFree(p->gfl);
  (void)hipFree(p->X);
  (void)hipFree(p->htab);
  p->ptab = p->gfl = p->X = p->htab = nullptr;
  std::vector<double> ev(nx);
  SBFS_HIP(hipMemcpy(ev.data(), p->evals, sizeof(double) * nx, hipMemcpyDeviceToHost));
  const double l = 1. / d->hy / d->hy;
  const double u = l;
  std::vector<double> tab((size_t)4 * K * nx), htab((size_t)P * nx);
  std::vector<double> hcol(P), hv(P), hq1(P);
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
    // first column of the reduced circulant's inverse (the singular mode 0:
    // of its pseudo-inverse, from e_0 - 1/P with the last entry pinned)
    const bool pin = mo == 0;
    for (int c = 0; c < P; c++)
      hcol[c] = pin ? (c == 0 ? 1. - 1. / (double)P : -1. / (double)P) : (c == 0 ? 1. : 0.);
    host_periodic_tdm(P, lo, di, up, pin, hcol.data(), hv.data(), hq1.data());
    for (int c = 0; c < P; c++) htab[(size_t)c * nx + mo] = hcol[c];
  }
  SBFS_HIP(hipMalloc(&p->ptab, sizeof(double) * tab.size()));
  SBFS_HIP(hipMalloc(&p->gfl, sizeof(double) * 2 * (size_t)P * nx));
  SBFS_HIP(hipMalloc(&p->X, sizeof(double) * (size_t)P * nx));
  SBFS_HIP(hipMalloc(&p->htab, sizeof(double) * htab.size()));
  SBFS_HIP(hipMemcpy(p->htab, htab.data(), sizeof(double) * htab.size(), hipMemcpyHostToDevice));
  SBFS_HIP(hipMemcpy(p->ptab, tab.data(), sizeof(double) * tab.size(), hipMemcpyHostToDevice));
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
  SBFS_TIMED("k_part_conv",
             hipLaunchKernelGGL(k_part_conv<L>, dim3(cdiv(nx, kConvModes)), dim3(256),
                                sizeof(double) * 3 * (size_t)P * kConvModes, stream(), p->buf,
                                p->gfl, p->htab, p->X, nx, P, a, pin));
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

// ---------------------------------------------------------------------------
// ensembles: the three kernels for every case in one launch each.  A case's
// chunk length L is uniform over its workgroups, so the switch over the
// instantiations does not diverge.
// ---------------------------------------------------------------------------
template <int L>
__device__ __forceinline__ void part_reduce_body(const RowArgs& a, int mo, int ch) {
  constexpr int K = L - 1;
  if (mo >= a.nx || ch >= a.P) return;
  double first, last;
  chunk_reduce<K, false>(a.buf, a.nx, mo, ch * L, a.ptab, a.ptab + (size_t)K * a.nx, a.pa, first, last, nullptr);
  a.gfl[(size_t)(2 * ch) * a.nx + mo] = first;
  a.gfl[(size_t)(2 * ch + 1) * a.nx + mo] = last;
}

#define SBFS_PART_SWITCH(BODY)                  \
  switch (a.L) {                                \
    case 2: BODY<2>(a, mo, ch); break;          \
    case 4: BODY<4>(a, mo, ch); break;          \
    case 8: BODY<8>(a, mo, ch); break;          \
    case 16: BODY<16>(a, mo, ch); break;        \
    case 32: BODY<32>(a, mo, ch); break;        \
    default: BODY<64>(a, mo, ch); break;        \
  }

// map entry (case, tile of 64 modes x 4 chunks)
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_part_reduce_batch(const RowArgs* __restrict__ tab,
                                                           const uint2* __restrict__ map) {
  const uint2 m = map[blockIdx.x];
  const RowArgs& a = tab[m.x];
  const int gx = (a.nx + 63) / 64;
  const int mo = (int)(m.y % gx) * 64 + threadIdx.x, ch = (int)(m.y / gx) * 4 + threadIdx.y;
  SBFS_PART_SWITCH(part_reduce_body)
}

#undef SBFS_PART_SWITCH

// map entry (case, group of kConvModes modes); LDS for the largest P
static_assert(kConvModes == kConvModesBatch, "the ensemble plans groups of kConvModesBatch modes");
__global__ __launch_bounds__(256) void k_part_conv_batch(const RowArgs* __restrict__ tab,
                                                         const uint2* __restrict__ map) {
  const uint2 mp = map[blockIdx.x];
  const RowArgs& a = tab[mp.x];
  const int nx = a.nx, P = a.P, L = a.L, K = L - 1;
  constexpr int G = 256 / kConvModes;
  extern __shared__ double lds[];
  double* R = lds;
  double* H = lds + (size_t)P * kConvModes;
  double* Y = lds + (size_t)2 * P * kConvModes;
  const int tm = threadIdx.x % kConvModes, g = threadIdx.x / kConvModes;
  const int mo = (int)mp.y * kConvModes + tm;
  const bool live = mo < nx;
  const int mm = live ? mo : nx - 1;
  for (int c = g; c < P; c += G) {
    const double gl = a.gfl[(size_t)(2 * c + 1) * nx + mm];
    const double gf = a.gfl[(size_t)(2 * ((c + 1) % P)) * nx + mm];
    R[c * kConvModes + tm] = (a.buf[(size_t)(c * L + K) * nx + mm] - a.pa * gl) - a.pa * gf;
    H[c * kConvModes + tm] = a.htab[(size_t)c * nx + mm];
  }
  __syncthreads();
  for (int i0 = g * kConvS; i0 < P; i0 += G * kConvS) {
    double acc[kConvS], w[kConvS];
#pragma unroll
    for (int r = 0; r < kConvS; r++) {
      acc[r] = 0.;
      w[r] = H[((i0 + r) % P) * kConvModes + tm];
    }
    int next = i0 > 0 ? i0 - 1 : P - 1;
    for (int k = 0; k < P; k++) {
      const double rk = R[k * kConvModes + tm];
#pragma unroll
      for (int r = 0; r < kConvS; r++) acc[r] += w[r] * rk;
#pragma unroll
      for (int r = kConvS - 1; r > 0; r--) w[r] = w[r - 1];
      w[0] = H[next * kConvModes + tm];
      next = next > 0 ? next - 1 : P - 1;
    }
#pragma unroll
    for (int r = 0; r < kConvS; r++)
      if (i0 + r < P) Y[(i0 + r) * kConvModes + tm] = acc[r];
  }
  __syncthreads();
  const bool pin = a.pin && mo == 0;
  const double xl = Y[(P - 1) * kConvModes + tm];
  if (!live) return;
  for (int i = g; i < P; i += G) {
    const double x = Y[i * kConvModes + tm];
    a.X[(size_t)i * nx + mo] = pin ? x - xl : x;
  }
}

}  // namespace sbfs
