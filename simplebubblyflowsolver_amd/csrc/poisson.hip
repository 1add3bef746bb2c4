// Split-pressure Poisson solver (reference src/fluid/compute_potential.c and
// src/tdm.c) on one GPU.
//
//  1. k_dct2_rhs   one workgroup per row j: builds the right-hand side of
//                  the row in LDS (assign_input, :285-363) and applies
//                  FFTW's REDFT10 along x through an N/2-point complex
//                  Stockham FFT in LDS; writes the row of modes r[j][m].
//                  The marching path's k_dct2_rows transforms rows that the
//                  stage kernel wrote, one wavefront per group of rows
//                  (rowfft.hpp), with the same arithmetic.
//  2. tridiagonal  per x-mode m along the periodic y direction, directly on
//                  the [j][m] layout (consecutive modes are consecutive
//                  addresses, so every sweep step is a coalesced row access;
//                  the reference's x1 <-> y1 pencil transposes disappear).
//  3. k_dct3_psi   one wavefront per group of rows: loads the solved modes
//                  (expanding the partitioned solve's chunks on the way),
//                  REDFT01 along x, then psi[1] with its Neumann x ghosts
//                  and periodic y halo (extract_output, :365-380,
//                  src/fluid/boundary/psi.c).
#include <cfloat>
#include <cmath>
#include <cstring>
#include <vector>

#include "sbfs.h"
#include "poisson.hpp"
#include "rowfft.hpp"
#include "sbfs_internal.hpp"
#include "stencils.hpp"

namespace sbfs {

static const long double kPiL = 3.141592653589793238462643383279502884L;
static const double kPiPoisson = 3.14159265358979324;  // compute_potential.c:21



void poisson_destroy(PoissonPlan* p) {
  if (!p) return;
  graph_epoch_bump();
  (void)hipFree(p->buf);
  (void)hipFree(p->v);
  (void)hipFree(p->q1);
  (void)hipFree(p->evals);
  (void)hipFree(p->tw);
  (void)hipFree(p->t1);
  (void)hipFree(p->t2);
  (void)hipFree(p->ptab);
  (void)hipFree(p->gfl);
  (void)hipFree(p->X);
  (void)hipFree(p->htab);
  (void)hipFree(p->rhs);
  delete p;
}

// twiddle e^{-2 pi i e/m}, rounded from long double (the oracle builds its
// tables with the same expression)
static void twiddle(size_t e, size_t m, double* re, double* im) {
  const long double a = 2.L * kPiL * (long double)e / (long double)m;
  *re = (double)cosl(a);
  *im = (double)(-sinl(a));
}

// largest chunk <= requested that divides n and leaves >= 3 chunks
int partition_chunk(int n, int requested) {
  int L = requested;
  while (L >= 2 && (n % L != 0 || n / L < 3)) L /= 2;
  return L >= 2 ? L : 0;
}

// the chunk of an ny-row solve: the requested length, doubled while the
// reduced system would exceed kMaxSeparators
int poisson_chunk(int ny) {
  int req = options().chunk;
  while (req < 64 && ny / req > kMaxSeparators) req *= 2;
  return partition_chunk(ny, req);
}

// ---------------------------------------------------------------------------
// FFT in LDS: forward, Stockham autosort, radix-4 stages then one radix-2
// ---------------------------------------------------------------------------
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

__device__ double2* fft_lds(double2* x, double2* y, int m, const double2* __restrict__ tw) {
  const int tid = threadIdx.x, nt = blockDim.x;
  int n = m, s = 1, ls = 0;
  while (n >= 4) {
    const int q4 = n >> 2;
    const int nb = m >> 2;
    for (int b = tid; b < nb; b += nt) {
      const int p = b >> ls, q = b & (s - 1);
      const double2 a = x[q + s * p];
      const double2 bb = x[q + s * (p + q4)];
      const double2 c = x[q + s * (p + 2 * q4)];
      const double2 d = x[q + s * (p + 3 * q4)];
      const double t0r = a.x + c.x, t0i = a.y + c.y;
      const double t1r = a.x - c.x, t1i = a.y - c.y;
      const double t2r = bb.x + d.x, t2i = bb.y + d.y;
      const double t3r = bb.x - d.x, t3i = bb.y - d.y;
      const double2 x0 = make_double2(t0r + t2r, t0i + t2i);
      const double2 x2 = make_double2(t0r - t2r, t0i - t2i);
      const double2 x1 = make_double2(t1r + t3i, t1i - t3r);
      const double2 x3 = make_double2(t1r - t3i, t1i + t3r);
      double2* o = y + q + s * 4 * p;
      o[0] = x0;
      o[s] = cmul(x1, tw[p * s]);
      o[2 * s] = cmul(x2, tw[2 * p * s]);
      o[3 * s] = cmul(x3, tw[3 * p * s]);
    }
    __syncthreads();
    double2* t = x;
    x = y;
    y = t;
    n >>= 2;
    s <<= 2;
    ls += 2;
  }
  if (n == 2) {
    for (int q = tid; q < s; q += nt) {
      const double2 a = x[q], b = x[q + s];
      y[q] = make_double2(a.x + b.x, a.y + b.y);
      y[q + s] = cmul(make_double2(a.x - b.x, a.y - b.y), tw[0]);
    }
    __syncthreads();
    double2* t = x;
    x = y;
    y = t;
  }
  return x;
}

// right-hand side of one cell (compute_potential.c:315-360); `den(i, j)`
// supplies den[1] (read from the array, or recomputed from the new vof)
template <typename Den>
__device__ __forceinline__ double poisson_rhs(
    int i, int j, int nx, const double* __restrict__ ux, const double* __restrict__ uy,
    const double* __restrict__ psi, Den den, const double* __restrict__ hxxf,
    const double* __restrict__ jdxf, const double* __restrict__ jdxc, double hy, double refden,
    double c0, double c1) {
  const double hm = hxxf[i - 1], hp = hxxf[i];
  const double jm = jdxf[i - 1], j0 = jdxc[i - 1], jp = jdxf[i];
  const double div = 1. / j0 * (-jm / hm * ux[L::xf(i, j, nx)] + jp / hp * ux[L::xf(i + 1, j, nx)]
                                - j0 / hy * uy[L::cc(i, j, nx)] + j0 / hy * uy[L::cc(i, j + 1, nx)]);
  const double d00 = den(i, j);
  const double dxm = 0.5 * den(i - 1, j) + 0.5 * d00;
  const double dxp = 0.5 * d00 + 0.5 * den(i + 1, j);
  const double dym = 0.5 * den(i, j - 1) + 0.5 * d00;
  const double dyp = 0.5 * d00 + 0.5 * den(i, j + 1);
  const double pxm = -psi[L::cc(i - 1, j, nx)] + psi[L::cc(i, j, nx)];
  const double pxp = -psi[L::cc(i, j, nx)] + psi[L::cc(i + 1, j, nx)];
  const double pym = -psi[L::cc(i, j - 1, nx)] + psi[L::cc(i, j, nx)];
  const double pyp = -psi[L::cc(i, j, nx)] + psi[L::cc(i, j + 1, nx)];
  const double gxm = (refden / dxm - 1.) * pxm / hm;
  const double gxp = (refden / dxp - 1.) * pxp / hp;
  const double gym = (refden / dym - 1.) * pym / hy;
  const double gyp = (refden / dyp - 1.) * pyp / hy;
  const double add = 1. / j0 * (-jm / hm * gxm + jp / hp * gxp - j0 / hy * gym + j0 / hy * gyp);
  return c0 * div + c1 * add;
}

// REDFT10 of a row held in LDS (`row`, n doubles aliasing B) -> out row
__device__ void dct2_row_lds(double2* A, double2* B, int n, int m,
                             const double2* __restrict__ tw, const double2* __restrict__ t1,
                             const double2* __restrict__ t2, double* __restrict__ out) {
  const int tid = threadIdx.x, nt = blockDim.x;
  const double* row = (const double*)B;
  // v[k] = x[2k], v[n-1-k] = x[2k+1]; z[k] = v[2k] + i v[2k+1]
  for (int k = tid; k < m; k += nt) {
    const int e = 2 * k, o = 2 * k + 1;
    const double ve = (e < m) ? row[2 * e] : row[2 * (n - 1 - e) + 1];
    const double vo = (o < m) ? row[2 * o] : row[2 * (n - 1 - o) + 1];
    A[k] = make_double2(ve, vo);
  }
  __syncthreads();
  const double2* z = fft_lds(A, B, m, tw);
  for (int k = tid; k <= m; k += nt) {
    const int kk = (k == m) ? 0 : k;
    const int kr = (k == 0) ? 0 : m - k;
    const double ar = z[kk].x, ai = z[kk].y;
    const double br = z[kr].x, bi = -z[kr].y;
    const double er = (ar + br) * 0.5, ei = (ai + bi) * 0.5;
    const double dr = ar - br, di = ai - bi;
    const double2 od = make_double2(di * 0.5, -(dr * 0.5));
    const double2 tt = cmul(od, t1[k]);
    const double2 vv = make_double2(er + tt.x, ei + tt.y);
    const double2 w = cmul(vv, t2[k]);
    out[k] = 2. * w.x;
    if (k != 0 && k != m) out[n - k] = -2. * w.y;
  }
}

__global__ void k_dct2_rhs(const double* __restrict__ ux, const double* __restrict__ uy,
                           const double* __restrict__ psi, const double* __restrict__ den,
                           const double* __restrict__ hxxf, const double* __restrict__ jdxf,
                           const double* __restrict__ jdxc, double hy, double refden,
                           double c0, double c1, int nx, int m,
                           const double2* __restrict__ tw, const double2* __restrict__ t1,
                           const double2* __restrict__ t2, double* __restrict__ out) {
  extern __shared__ double2 lds[];
  double2* A = lds;
  double2* B = lds + m;
  const int j = blockIdx.x + 1;
  double* row = (double*)B;
  auto dd = [den, nx](int p, int q) { return den[L::cc(p, q, nx)]; };
  for (int i = threadIdx.x + 1; i <= nx; i += blockDim.x)
    row[i - 1] = poisson_rhs(i, j, nx, ux, uy, psi, dd, hxxf, jdxf, jdxc, hy, refden, c0, c1);
  __syncthreads();
  dct2_row_lds(A, B, nx, m, tw, t1, t2, out + (size_t)(j - 1) * nx);
}

// ---------------------------------------------------------------------------
// Row transforms of the marching path, the partitioned solve's expansion and
// the tests: one wavefront per group of R rows (rowfft.hpp).  The arithmetic
// is dct2_row_lds / fft_lds's, operation for operation.
// ---------------------------------------------------------------------------

// REDFT10 of rows row0 .. row0 + R - 1 (of `rows`) of `in` into `out`
// With NT, streaming (non-temporal) loads and stores for the rows the
// transforms read and write once, so that they do not evict the twiddles, the
// spike tables and the separators, which every row reads again from L2: at
// 4096 x 8192, k_dct3_psi 226 -> 173 us and k_dct2_rows 128 -> 119 us.  The
// host picks NT for fields too large to stay in the Infinity Cache between
// kernels (stream_fields); the ensemble's batched kernels never stream.  (The
// same hints on the partitioned solve's rows slowed the DCT-III that follows
// it more than they sped the solve up, so it keeps ordinary accesses.)
template <bool NT>
__device__ __forceinline__ double2 ld_row(const double2* p) {
  if constexpr (NT) return make_double2(__builtin_nontemporal_load(&p->x), __builtin_nontemporal_load(&p->y));
  else return *p;
}
template <bool NT, typename T>
__device__ __forceinline__ void st_row(T* p, T v) {
  if constexpr (NT) {
    if constexpr (std::is_same<T, double2>::value) {
      __builtin_nontemporal_store(v.x, &p->x);
      __builtin_nontemporal_store(v.y, &p->y);
    } else {
      __builtin_nontemporal_store(v, p);
    }
  } else {
    *p = v;
  }
}

// the unit's rows in LDS: row rr at lds + rr PM, doubles through zd().
// Rows past the last (rr >= nrows >= 1) load the last row again, so every
// load is unconditional and all are in flight before the first LDS write.
template <int LOG, typename Src>
__device__ __forceinline__ void rows_to_lds(double2* lds, int lane, int nrows, Src src) {
  using F = WaveFft<LOG>;
  constexpr int H = F::N / 2;  // double2 per row
  if constexpr (F::R * H >= F::U) {
    constexpr int PER = F::R * H / F::U;
    double2 v[PER];
#pragma unroll
    for (int i = 0; i < PER; i++) {
      const int t = lane + i * F::U;
      v[i] = src(min(t / H, nrows - 1), t % H);
    }
#pragma unroll
    for (int i = 0; i < PER; i++) {
      const int t = lane + i * F::U;
      lds[(t / H) * F::PM + fpad(t % H)] = v[i];
    }
  } else {
    const int rr = lane / H, h = lane % H;
    if (lane < F::R * H) lds[rr * F::PM + fpad(h)] = src(min(rr, nrows - 1), h);
  }
}

// (the unit's LDS rows at `lds`, its thread `lane` in [0, U))
template <int LOG, bool NT>
__device__ __forceinline__ void dct2_wave(const double* __restrict__ in, double* __restrict__ out, int rows,
                                          int row0, const double2* __restrict__ tw,
                                          const double2* __restrict__ t1, const double2* __restrict__ t2,
                                          double2* lds, int lane) {
  using F = WaveFft<LOG>;
  constexpr int M = F::M, N = F::N, G = F::G;
  const int rg = lane / G, lt = lane % G;
  const int nrows = min(F::R, rows - row0);
  if (nrows <= 0) return;  // a second unit past the last row (one wavefront, no barrier)
  const auto tw0 = pass_tw<LOG, 0>(lt, tw);
  const double2* src = (const double2*)(in + (size_t)row0 * N);
  rows_to_lds<LOG>(lds, lane, nrows, [src](int rr, int h) { return ld_row<NT>(src + rr * (N / 2) + h); });
  F::sync();
  double2* Z = lds + rg * F::PM;
  // v[k] = x[2k], v[n-1-k] = x[2k+1]; z[k] = v[2k] + i v[2k+1]
  auto ld0 = [Z](int k) {
    const int e = 2 * k, o = 2 * k + 1;
    const double ve = (e < M) ? zd(Z, 2 * e) : zd(Z, 2 * (N - 1 - e) + 1);
    const double vo = (o < M) ? zd(Z, 2 * o) : zd(Z, 2 * (N - 1 - o) + 1);
    return make_double2(ve, vo);
  };
  wave_fft<LOG>(Z, lt, tw, ld0, tw0);
  if (rg >= nrows) return;
  double* dst = out + (size_t)(row0 + rg) * N;
#pragma unroll
  for (int i = 0; i <= M / G; i++) {
    const int k = lt + i * G;
    if (k > M) break;
    const int kk = (k == M) ? 0 : k;
    const int kr = (k == 0) ? 0 : M - k;
    const double2 za = Z[fpad(kk)], zb = Z[fpad(kr)];
    const double ar = za.x, ai = za.y;
    const double br = zb.x, bi = -zb.y;
    const double er = (ar + br) * 0.5, ei = (ai + bi) * 0.5;
    const double dr = ar - br, di = ai - bi;
    const double2 od = make_double2(di * 0.5, -(dr * 0.5));
    const double2 tt = wmul(od, t1[k]);
    const double2 vv = make_double2(er + tt.x, ei + tt.y);
    const double2 w = wmul(vv, t2[k]);
    st_row<NT>(dst + k, 2. * w.x);
    if (k != 0 && k != M) st_row<NT>(dst + (N - k), -2. * w.y);
  }
}

// where the DCT-III finds mode row jm: the solved buffer, or (L != 0) the
// partitioned solve's chunk solutions with zero ends plus the two unit
// spikes weighted by the bounding separators, x = (g + xl w) + xu z, and
// the separator itself on a chunk's last row (the expansion step of
// tdm_partition.hip, done as the rows are loaded by expanded_to_lds)
struct ModeSrc {
  const double* buf;
  const double* X;
  const double* w;  // spike tables [L - 1][nx]
  const double* z;
  int nx, L, P;  // L is a power of two
  // values c = 2h and 2h + 1 of mode row jm, L == 0
  __device__ __forceinline__ double2 solved(int jm, int h) const {
    return *(const double2*)(buf + (size_t)jm * nx + 2 * h);
  }
};

// rows_to_lds for the partitioned solve's modes (L != 0).  The chunk rows g
// of all the lane's points are loaded first (from HBM), then the separators
// and spikes in groups of points (from L2), so that the unit waits for one
// HBM latency and PER / GRP L2 latencies rather than one of each per point.
// GRP is 8 where the unit holds 2048 complex points (the group's 32 loads fit
// the registers the FFT leaves; 233 -> 226 us at 4096 x 8192), else 4.  With
// one row per unit (R = 1) the row, its chunk and whether it is the chunk's
// last row (the separator itself) are the same for every
// lane; with several, the loads are unconditional (a chunk's last row
// reads spike row 0, whose product it discards) and the value a select.
template <int LOG, bool NT>
__device__ __forceinline__ void expanded_to_lds(double2* lds, int lane, int nrows, const ModeSrc& src, int row0) {
  using F = WaveFft<LOG>;
  constexpr int H = F::N / 2;  // double2 per row
  static_assert(F::R * H >= F::U, "every lane loads at least one point");
  constexpr int PER = F::R * H / F::U, GMAX = LOG >= 11 ? 8 : 4, GRP = PER < GMAX ? PER : GMAX;
  static_assert(PER % GRP == 0, "groups of points");
  const int nx = src.nx, sh = 31 - __clz(src.L);
  auto row = [&](int i) { return F::R == 1 ? row0 : row0 + min((lane + i * F::U) / H, nrows - 1); };
  auto col = [&](int i) { return 2 * ((lane + i * F::U) % H); };
  auto put = [&](int i, double2 v) {
    const int t = lane + i * F::U;
    lds[(t / H) * F::PM + fpad(t % H)] = v;
  };
  if constexpr (F::R == 1) {
    const int ch = row0 >> sh, k = row0 & (src.L - 1);
    const double* xu = src.X + (size_t)ch * nx;
    if (k == src.L - 1) {
      double2 v[PER];
#pragma unroll
      for (int i = 0; i < PER; i++) v[i] = *(const double2*)(xu + col(i));
#pragma unroll
      for (int i = 0; i < PER; i++) put(i, v[i]);
      return;
    }
  }
  double2 g[PER];
#pragma unroll
  for (int i = 0; i < PER; i++) g[i] = ld_row<NT>((const double2*)(src.buf + (size_t)row(i) * nx + col(i)));
#pragma unroll
  for (int i0 = 0; i0 < PER; i0 += GRP) {
    double2 xu[GRP], xl[GRP], wk[GRP], zk[GRP];
#pragma unroll
    for (int u = 0; u < GRP; u++) {
      const int jm = row(i0 + u), c = col(i0 + u);
      const int ch = jm >> sh, k = jm & (src.L - 1);
      const int ks = k == src.L - 1 ? 0 : k;
      xu[u] = *(const double2*)(src.X + (size_t)ch * nx + c);
      xl[u] = *(const double2*)(src.X + (size_t)(ch == 0 ? src.P - 1 : ch - 1) * nx + c);
      wk[u] = *(const double2*)(src.w + (size_t)ks * nx + c);
      zk[u] = *(const double2*)(src.z + (size_t)ks * nx + c);
    }
#pragma unroll
    for (int u = 0; u < GRP; u++) {
      const bool last = F::R > 1 && (row(i0 + u) & (src.L - 1)) == src.L - 1;
      const double2 a = g[i0 + u];
      const double ex = (a.x + xl[u].x * wk[u].x) + xu[u].x * zk[u].x;
      const double ey = (a.y + xl[u].y * wk[u].y) + xu[u].y * zk[u].y;
      put(i0 + u, make_double2(last ? xu[u].x : ex, last ? xu[u].y : ey));
    }
  }
}

// REDFT01 of mode rows row0 .. row0 + R - 1 (of `rows`) into out: with PSI,
// psi[1] rows j = row + 1 with the Neumann x ghosts and the periodic y halo
// (compute_potential.c:365-380, src/fluid/boundary/psi.c; out must be
// 16-byte aligned); otherwise plain rows of N
template <int LOG, bool PSI, int AHEAD, bool NT>
__device__ __forceinline__ void dct3_wave(const ModeSrc& src, int rows, int row0, const double2* __restrict__ tw,
                                          const double2* __restrict__ t1, const double2* __restrict__ t2,
                                          double* __restrict__ out, double2* lds, int lane) {
  using F = WaveFft<LOG>;
  constexpr int M = F::M, N = F::N, G = F::G;
  const int rg = lane / G, lt = lane % G;
  const int nrows = min(F::R, rows - row0);
  if (nrows <= 0) return;  // a second unit past the last row (one wavefront, no barrier)
  const auto tw0 = pass_tw<LOG, 0>(lt, tw);
  // the pre-twiddle's factors of k and M - k for iteration i (k = lt + i G,
  // the last iteration only for k = M / 2): with AHEAD > 0 the first AHEAD
  // iterations' are loaded before the row and the others AHEAD iterations
  // ahead of their use, so that the pass does not wait for L2 in each
  // iteration (AHEAD 0: each at its use, for kernels short of registers)
  struct PreTw {
    double2 t1, t2a, t2b;  // t1[k], t2[k], t2[k + M]
  };
  constexpr int NPRE = M / 2 / G + 1;
  PreTw ta[NPRE], tb[NPRE];
  auto fetch = [&ta, &tb, t1, t2, lt](int i) {
    const int k = min(lt + i * G, M / 2), kb = M - max(k, 1);  // kb: unused for k = 0
    ta[i] = PreTw{t1[k], t2[k], t2[k + M]};
    tb[i] = PreTw{t1[kb], t2[kb], t2[kb + M]};
  };
#pragma unroll
  for (int i = 0; i < AHEAD; i++) fetch(i);
  if (src.L == 0)
    rows_to_lds<LOG>(lds, lane, nrows, [&src, row0](int rr, int h) { return src.solved(row0 + rr, h); });
  else
    expanded_to_lds<LOG, NT>(lds, lane, nrows, src, row0);
  F::sync();
  double2* Z = lds + rg * F::PM;
  // pre-twiddle in place: the values of k and M - k (0 <= k <= M/2) need
  // exactly x[k], x[M-k], x[M+k] and x[N-k], so A[k] takes the slots of
  // x[k] and x[M-k], and A[M-k] those of x[M+k] and x[N-k]; no lane writes
  // a slot that another lane reads
  auto pre = [](const PreTw& t, double xk, double xnk, double xmk, double xkm) {
    // xk = x[k], xnk = x[N-k] (0 for k = 0), xmk = x[M-k], xkm = x[k+M]
    const double pr = xk, pim = -xnk;
    const double2 vk = wmul(make_double2(pr, pim), make_double2(t.t2a.x, -t.t2a.y));
    const double qr = xkm, qim = -xmk;
    const double2 vm = wmul(make_double2(qr, qim), make_double2(t.t2b.x, -t.t2b.y));
    const double sr = vk.x + vm.x, si = vk.y + vm.y;
    const double2 f = wmul(make_double2(vk.x - vm.x, vk.y - vm.y), make_double2(t.t1.x, -t.t1.y));
    return make_double2(sr - f.y, -(si + f.x));
  };
#pragma unroll
  for (int i = 0; i < NPRE; i++) {
    if (i + AHEAD < NPRE) fetch(i + AHEAD);
    const int k = lt + i * G;
    if (2 * k > M) break;
    if (k == 0) {
      const double x0 = zd(Z, 0), xm = zd(Z, M);
      const double2 a = pre(ta[i], x0, 0., xm, xm);
      zd(Z, 0) = a.x;
      zd(Z, M) = a.y;
    } else if (2 * k == M) {
      const double xa = zd(Z, k), xb = zd(Z, M + k);
      const double2 a = pre(ta[i], xa, xb, xa, xb);
      zd(Z, k) = a.x;
      zd(Z, M + k) = a.y;
    } else {
      const double xk = zd(Z, k), xmk = zd(Z, M - k), xkm = zd(Z, M + k), xnk = zd(Z, N - k);
      const double2 a = pre(ta[i], xk, xnk, xmk, xkm);
      const double2 b = pre(tb[i], xmk, xkm, xk, xnk);
      zd(Z, k) = a.x;
      zd(Z, M - k) = a.y;
      zd(Z, M + k) = b.x;
      zd(Z, N - k) = b.y;
    }
  }
  F::sync();
  auto ld0 = [Z](int j) {
    if (j == 0) return make_double2(zd(Z, 0), zd(Z, M));
    if (2 * j == M) return make_double2(zd(Z, j), zd(Z, M + j));
    if (2 * j < M) return make_double2(zd(Z, j), zd(Z, M - j));
    return make_double2(zd(Z, 2 * M - j), zd(Z, M + j));
  };
  wave_fft<LOG>(Z, lt, tw, ld0, tw0);
  // y[2x] (x < M) and y[2(N-1-x)+1] (x >= M) take z[x/2].x for even x and
  // -z[x/2].y for odd x
  if (rg >= nrows) return;
  auto y = [Z](int i) {
    const int x = (i & 1) ? N - 1 - (i >> 1) : (i >> 1);
    const double2 zz = Z[fpad(x >> 1)];
    return (x & 1) ? -zz.y : zz.x;
  };
  const int r = row0 + rg;
  if constexpr (PSI) {
    // the row with both ghosts is nx + 2 = N + 2 doubles from an even
    // index: pair q holds columns 2q, 2q + 1, i.e. y[2q - 1], y[2q], with
    // the ghosts y[-1] = y[0] and y[N] = y[N - 1]
    const int j = r + 1;
    double2* row = (double2*)(out + L::cc(0, j, N));
    double2* halo = j == rows ? (double2*)(out + L::cc(0, 0, N))
                              : (j == 1 ? (double2*)(out + L::cc(0, rows + 1, N)) : nullptr);
    constexpr int Q = (N / 2 + 1 + G - 1) / G;
#pragma unroll
    for (int u = 0; u < Q; u++) {
      const int q = lt + u * G;
      if (q > N / 2) break;
      const double2 v = make_double2(y(max(2 * q - 1, 0)), y(min(2 * q, N - 1)));
      st_row<NT>(row + q, v);
      if (halo) st_row<NT>(halo + q, v);
    }
  } else {
    double* dst = out + (size_t)r * N;
#pragma unroll
    for (int u = 0; u < N / G; u++) {
      const int i = lt + u * G;
      dst[i] = y(i);
    }
  }
}

// two wavefronts' worth of VGPRs (256): the twiddles of the next pass are
// in flight beside the current one, and LDS bounds these kernels to about 2
// wavefronts per SIMD anyway; the cap keeps the unrolled loops from
// hoisting every load
#define SBFS_ROWS_ATTR(LG) __launch_bounds__(wave_fft_u(LG)) __attribute__((amdgpu_waves_per_eu(2)))

template <int LOG, bool NT>
__global__ SBFS_ROWS_ATTR(LOG) void k_dct2_rows(const double* __restrict__ in, double* __restrict__ out, int rows,
                                                const double2* __restrict__ tw, const double2* __restrict__ t1,
                                                const double2* __restrict__ t2) {
  extern __shared__ double2 lds[];
  dct2_wave<LOG, NT>(in, out, rows, (int)blockIdx.x * WaveFft<LOG>::R, tw, t1, t2, lds, (int)threadIdx.x);
}

template <int LOG>
__global__ SBFS_ROWS_ATTR(LOG) void k_dct3_rows(const double* __restrict__ in, double* __restrict__ out, int rows,
                                                const double2* __restrict__ tw, const double2* __restrict__ t1,
                                                const double2* __restrict__ t2) {
  constexpr int N = WaveFft<LOG>::N;
  extern __shared__ double2 lds[];
  const ModeSrc src{in, nullptr, nullptr, nullptr, N, 0, 0};
  dct3_wave<LOG, false, 1, false>(src, rows, (int)blockIdx.x * WaveFft<LOG>::R, tw, t1, t2, out, lds, (int)threadIdx.x);
}

template <int LOG, bool NT>
__global__ SBFS_ROWS_ATTR(LOG) void k_dct3_psi(ModeSrc src, double* __restrict__ psi, int ny,
                                               const double2* __restrict__ tw, const double2* __restrict__ t1,
                                               const double2* __restrict__ t2) {
  extern __shared__ double2 lds[];
  // Rows in XCD order.  One row per unit with a partitioned solve: XCD x
  // takes chunk half x / 4 and quarter x % 4 of the rows within a chunk, so
  // its L2 holds half the separators and a quarter of the spike tables (3 MB
  // of 8 at 4096 x 8192).  Otherwise XCD-major rows: an XCD's L2 holds the
  // separators of its own chunks.
  int unit;
  const int n = (int)gridDim.x, b = (int)blockIdx.x;
  if (WaveFft<LOG>::R == 1 && src.L >= 4 && (src.P & 1) == 0 && (n & 7) == 0) {
    const int x = b & 7, i = b >> 3, kq = src.L / 4;
    unit = ((x >> 2) * (src.P / 2) + i / kq) * src.L + (x & 3) * kq + i % kq;
  } else {
    unit = xcd_group(b, n);
  }
  dct3_wave<LOG, true, 1, NT>(src, ny, unit * WaveFft<LOG>::R, tw, t1, t2, psi, lds, (int)threadIdx.x);
}

#define SBFS_LOG_CASES(X) \
  X(0) X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11)

static int ilog2(int m) { return 31 - __builtin_clz((unsigned)m); }

// rows, threads and LDS bytes of one unit of the row transforms for m
// complex points
int wave_rows(int m) { return wave_fft_r(ilog2(m)); }
int wave_unit(int m) { return wave_fft_u(ilog2(m)); }
size_t wave_lds(int m) { return wave_fft_lds(ilog2(m)); }

// ---------------------------------------------------------------------------
// tridiagonal solves, reference order: one Thomas + Sherman-Morrison chain
// per x-mode (src/tdm.c:17-106 with constant l, c, u)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void tdm_sequential_body(double* __restrict__ q, double* __restrict__ v,
                                                    double* __restrict__ q1, int nx, int ny,
                                                    const double* __restrict__ evals, double l, double u,
                                                    int pin_mode0, int mo) {
  if (mo >= nx) return;
  const double c = -l - u + evals[mo];
  const int n = ny, N = ny - 1;
#define Q(r) q[(size_t)(r) * nx + mo]
#define V(r) v[(size_t)(r) * nx + mo]
#define Q1(r) q1[(size_t)(r) * nx + mo]
  for (int r = 0; r < N; r++) Q1(r) = (r == 0) ? -l : (r == n - 2) ? -u : 0.;
  V(0) = u / c;
  Q1(0) = Q1(0) / c;
  Q(0) = Q(0) / c;
  for (int r = 1; r < N - 1; r++) {
    const double val = 1. / (c - l * V(r - 1));
    V(r) = val * u;
    Q1(r) = val * (Q1(r) - l * Q1(r - 1));
    Q(r) = val * (Q(r) - l * Q(r - 1));
  }
  {
    const double val = c - l * V(N - 2);
    if (fabs(val) > DBL_EPSILON) {
      Q1(N - 1) = 1. / val * (Q1(N - 1) - l * Q1(N - 2));
      Q(N - 1) = 1. / val * (Q(N - 1) - l * Q(N - 2));
    } else {
      Q1(N - 1) = 0.;
      Q(N - 1) = 0.;
    }
  }
  for (int r = N - 2; r >= 0; r--) {
    Q1(r) -= V(r) * Q1(r + 1);
    Q(r) -= V(r) * Q(r + 1);
  }
  const double num = Q(n - 1) - u * Q(0) - l * Q(n - 2);
  const double den = c + u * Q1(0) + l * Q1(n - 2);
  const double xl = ((pin_mode0 && mo == 0) || fabs(den) < DBL_EPSILON) ? 0. : num / den;
  Q(n - 1) = xl;
  for (int r = 0; r < n - 1; r++) Q(r) = Q(r) + xl * Q1(r);
#undef Q
#undef V
#undef Q1
}

__global__ void k_tdm_sequential(double* __restrict__ q, double* __restrict__ v,
                                 double* __restrict__ q1, int nx, int ny,
                                 const double* __restrict__ evals, double l, double u,
                                 int pin_mode0) {
  tdm_sequential_body(q, v, q1, nx, ny, evals, l, u, pin_mode0, blockIdx.x * blockDim.x + threadIdx.x);
}

// ensembles: one map entry (case, group of rows or of 64 modes) per
// workgroup; the row transforms switch on the case's transform length,
// which is uniform over the wavefront
__global__ __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(2))) void k_dct2_rows_batch(
    const RowArgs* __restrict__ tab, const uint2* __restrict__ map) {
  extern __shared__ double2 lds[];
  const uint2 mp = map[blockIdx.x];
  const RowArgs& a = tab[mp.x];
  switch (31 - __clz(a.m)) {
#define SBFS_DCT2_CASE(LG)                                                                              \
  case LG: {                                                                                            \
    using F = WaveFft<LG>;                                                                              \
    const int units = (int)blockDim.x / F::U, u = (int)threadIdx.x / F::U;                              \
    dct2_wave<LG, false>(a.rhs, a.buf, a.ny, ((int)mp.y * units + u) * F::R, a.tw, a.t1, a.t2,                   \
                  lds + (size_t)u * F::R * F::PM, (int)threadIdx.x % F::U);                              \
  } break;
    SBFS_LOG_CASES(SBFS_DCT2_CASE)
#undef SBFS_DCT2_CASE
  }
}

__global__ void k_tdm_sequential_batch(const RowArgs* __restrict__ tab, const uint2* __restrict__ map) {
  const uint2 mp = map[blockIdx.x];
  const RowArgs& a = tab[mp.x];
  tdm_sequential_body(a.buf, a.v, a.q1, a.nx, a.ny, a.evals, a.l, a.l, a.pin, (int)mp.y * 64 + threadIdx.x);
}

__global__ __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(2))) void k_dct3_psi_batch(
    const RowArgs* __restrict__ tab, const uint2* __restrict__ map) {
  extern __shared__ double2 lds[];
  const uint2 mp = map[blockIdx.x];
  const RowArgs& a = tab[mp.x];
  const int K = a.L - 1;
  const ModeSrc src{a.buf, a.X, a.L ? a.ptab + (size_t)2 * K * a.nx : nullptr,
                    a.L ? a.ptab + (size_t)3 * K * a.nx : nullptr, a.nx, a.L, a.P};
  switch (31 - __clz(a.m)) {
#define SBFS_DCT3_CASE(LG)                                                                              \
  case LG: {                                                                                            \
    using F = WaveFft<LG>;                                                                              \
    const int units = (int)blockDim.x / F::U, u = (int)threadIdx.x / F::U;                              \
    dct3_wave<LG, true, 0, false>(src, a.ny, ((int)mp.y * units + u) * F::R, a.tw, a.t1, a.t2, a.psi,              \
                        lds + (size_t)u * F::R * F::PM, (int)threadIdx.x % F::U);                        \
  } break;
    SBFS_LOG_CASES(SBFS_DCT3_CASE)
#undef SBFS_DCT3_CASE
  }
}


// ---------------------------------------------------------------------------
// plan
// ---------------------------------------------------------------------------
static int plan_build(PoissonPlan* p, const domain_t* d) {
  const int nx = p->nx, ny = p->ny, m = p->m;
  const size_t cells = (size_t)nx * ny;
  SBFS_HIP(hipMalloc(&p->buf, sizeof(double) * cells));
  SBFS_HIP(hipMalloc(&p->v, sizeof(double) * cells));
  SBFS_HIP(hipMalloc(&p->q1, sizeof(double) * cells));
  SBFS_HIP(hipMalloc(&p->evals, sizeof(double) * nx));
  SBFS_HIP(hipMalloc(&p->tw, sizeof(double2) * (m > 0 ? m : 1)));
  SBFS_HIP(hipMalloc(&p->t1, sizeof(double2) * (m + 1)));
  SBFS_HIP(hipMalloc(&p->t2, sizeof(double2) * nx));
  // eigenvalues of the x operator, compute_potential.c:233-261
  std::vector<double> ev(nx);
  const double sl = 2. * (double)d->glsizes[0];
  const double gsz = d->lengths[0] / (double)d->glsizes[0];
  for (int i = 0; i < nx; i++) {
    const double s = sin(kPiPoisson * (double)i / sl);
    ev[i] = -4. / (gsz * gsz) * (s * s);
  }
  std::vector<double> tw(2 * (m > 0 ? m : 1)), t1(2 * (m + 1)), t2(2 * nx);
  for (int e = 0; e < m; e++) twiddle(e, m, &tw[2 * e], &tw[2 * e + 1]);
  for (int k = 0; k <= m; k++) twiddle(k, nx, &t1[2 * k], &t1[2 * k + 1]);
  for (int k = 0; k < nx; k++) twiddle(k, 4 * (size_t)nx, &t2[2 * k], &t2[2 * k + 1]);
  SBFS_HIP(hipMemcpy(p->evals, ev.data(), sizeof(double) * nx, hipMemcpyHostToDevice));
  SBFS_HIP(hipMemcpy(p->tw, tw.data(), sizeof(double) * tw.size(), hipMemcpyHostToDevice));
  SBFS_HIP(hipMemcpy(p->t1, t1.data(), sizeof(double) * t1.size(), hipMemcpyHostToDevice));
  SBFS_HIP(hipMemcpy(p->t2, t2.data(), sizeof(double) * t2.size(), hipMemcpyHostToDevice));
  return 0;
}


static PoissonPlan* plan_get(const domain_t* d, DomainCtx* c) {
  const int nx = (int)c->nx, ny = (int)c->ny;
  if (c->poisson && c->poisson->nx == nx && c->poisson->ny == ny) return c->poisson;
  if (c->poisson) {
    poisson_destroy(c->poisson);
    c->poisson = nullptr;
  }
  if (nx < 2 || (nx & (nx - 1)) != 0 || nx > 4096) {
    set_error("Poisson solver: Nx = %d must be a power of two in [2, 4096]", nx);
    return nullptr;
  }
  if (ny < 3) {
    set_error("Poisson solver: Ny = %d must be at least 3", ny);
    return nullptr;
  }
  PoissonPlan* p = new PoissonPlan();
  p->nx = nx;
  p->ny = ny;
  p->m = nx / 2;
  p->lx = d->lengths[0];
  if (plan_build(p, d)) {
    poisson_destroy(p);
    return nullptr;
  }
  c->poisson = p;
  return p;
}

static int row_threads(int m) { return m >= 1024 ? 512 : 256; }

static int poisson_tail(const domain_t* d, PoissonPlan* p, fluid_t* f);

int poisson_solve(const domain_t* d, DomainCtx* c, double dt_old, double dt_new, fluid_t* f) {
  PoissonPlan* p = plan_get(d, c);
  if (!p) return 1;
  const int nx = p->nx, ny = p->ny, m = p->m;
  // coefficients incl. the FFT normalisation, compute_potential.c:303-314
  const double norm = 2. * (double)d->glsizes[0];
  const double refden = f->refden;
  const double c0 = 1. / norm / dt_new * refden;
  const double c1 = 1. / norm * -1. / dt_new * dt_old;
  const size_t lds = sizeof(double2) * 2 * (size_t)(m > 0 ? m : 1);
  SBFS_TIMED("k_dct2_rhs",
             hipLaunchKernelGGL(k_dct2_rhs, dim3(ny), dim3(row_threads(m)), lds, stream(),
                                (const double*)f->ux.data, (const double*)f->uy.data,
                                (const double*)f->psi[0].data, (const double*)f->den[1].data,
                                c->hxxf, c->jdxf, c->jdxc, c->hy, refden, c0, c1, nx, m, p->tw,
                                p->t1, p->t2, p->buf));
  if (check_launch("poisson: rhs + DCT-II")) return 1;
  return poisson_tail(d, p, f);
}

// tridiagonal solves + DCT-III into psi[1]
static int poisson_tail(const domain_t* d, PoissonPlan* p, fluid_t* f) {
  const int nx = p->nx, ny = p->ny, m = p->m;
  // tridiagonal matrix l = u = 1/hy^2 (compute_potential.c:173-176)
  const double l = 1. / d->hy / d->hy;
  const int pin = options().gauge == SBFS_GAUGE_PINNED ? 1 : 0;
  const int L = poisson_chunk(ny);
  ModeSrc src{p->buf, nullptr, nullptr, nullptr, nx, 0, 0};
  if (options().tdm == SBFS_TDM_PARTITIONED && L != 0 && pin && ny / L <= kMaxSeparators) {
    if (p->L != L && partition_setup(p, d, L)) return 1;
    // reduce + separators; the expansion happens as k_dct3_psi loads its rows
    if (partition_solve(p, pin)) return 1;
    const int K = L - 1;
    src = ModeSrc{p->buf, p->X, p->ptab + (size_t)2 * K * nx, p->ptab + (size_t)3 * K * nx, nx, L, p->P};
  } else {
    SBFS_TIMED("k_tdm_sequential",
               hipLaunchKernelGGL(k_tdm_sequential, dim3(cdiv(nx, 64)), dim3(64), 0, stream(),
                                  p->buf, p->v, p->q1, nx, ny, p->evals, l, l, pin));
    if (check_launch("poisson: tridiagonal")) return 1;
  }
  double* psi = (double*)f->psi[1].data;
  SBFS_REQUIRE(((uintptr_t)psi & 15) == 0, "poisson: psi[1] must be 16-byte aligned (hipMalloc'ed)");
  const bool nt = stream_fields(nx, ny);
  switch (ilog2(m)) {
#define SBFS_TAIL_CASE(LG)                                                                          \
  case LG:                                                                                          \
    SBFS_TIMED("k_dct3_psi", hipLaunchKernelGGL((nt ? k_dct3_psi<LG, true> : k_dct3_psi<LG, false>),      \
                                                dim3(cdiv(ny, WaveFft<LG>::R)), dim3(WaveFft<LG>::U),      \
                                                wave_fft_lds(LG), stream(), src, psi, ny, p->tw, p->t1,    \
                                                p->t2));                                            \
    break;
    SBFS_LOG_CASES(SBFS_TAIL_CASE)
#undef SBFS_TAIL_CASE
  }
  return check_launch("poisson: DCT-III + psi");
}

// marching path: `rhs` (filled by k_march) -> DCT-II rows -> tridiagonal ->
// DCT-III into psi[1].  Called with rhs_out non-null it only returns the
// right-hand-side buffer (allocating the plan on first use).
int poisson_transform_fused(const domain_t* d, DomainCtx* c, fluid_t* f, double** rhs_out) {
  PoissonPlan* p = plan_get(d, c);
  if (!p) return 1;
  if (!p->rhs) SBFS_HIP(hipMalloc(&p->rhs, sizeof(double) * (size_t)p->nx * p->ny));
  if (rhs_out) {
    *rhs_out = p->rhs;
    return 0;
  }
  const int ny = p->ny;
  const bool nt = stream_fields(p->nx, ny);
  switch (ilog2(p->m)) {
#define SBFS_ROWS_CASE(LG)                                                                              \
  case LG:                                                                                              \
    SBFS_TIMED("k_dct2_rows", hipLaunchKernelGGL((nt ? k_dct2_rows<LG, true> : k_dct2_rows<LG, false>),    \
                                                 dim3(cdiv(ny, WaveFft<LG>::R)), dim3(WaveFft<LG>::U),     \
                                                 wave_fft_lds(LG), stream(), (const double*)p->rhs, p->buf,  \
                                                 ny, p->tw, p->t1, p->t2));                             \
    break;
    SBFS_LOG_CASES(SBFS_ROWS_CASE)
#undef SBFS_ROWS_CASE
  }
  if (check_launch("poisson: DCT-II")) return 1;
  return poisson_tail(d, p, f);
}

// the Poisson arguments of one case (psi: the stage's psi[1])
int poisson_row_args(const domain_t* d, DomainCtx* c, double* psi, RowArgs* a) {
  PoissonPlan* p = plan_get(d, c);
  if (!p) return 1;
  if (!p->rhs) SBFS_HIP(hipMalloc(&p->rhs, sizeof(double) * (size_t)p->nx * p->ny));
  a->rhs = p->rhs;
  a->buf = p->buf;
  a->v = p->v;
  a->q1 = p->q1;
  SBFS_REQUIRE(((uintptr_t)psi & 15) == 0, "poisson: psi[1] must be 16-byte aligned (hipMalloc'ed)");
  a->psi = psi;
  a->evals = p->evals;
  a->tw = p->tw;
  a->t1 = p->t1;
  a->t2 = p->t2;
  a->l = 1. / d->hy / d->hy;
  a->nx = p->nx;
  a->ny = p->ny;
  a->m = p->m;
  a->pin = options().gauge == SBFS_GAUGE_PINNED ? 1 : 0;
  // the tridiagonal algorithm poisson_tail would take for this case
  const int L = poisson_chunk(p->ny);
  a->L = 0;
  a->P = 0;
  a->ptab = a->htab = nullptr;
  a->gfl = a->X = nullptr;
  a->pa = 0.;
  if (options().tdm == SBFS_TDM_PARTITIONED && L != 0 && a->pin && p->ny / L <= kMaxSeparators) {
    if (p->L != L && partition_setup(p, d, L)) return 1;
    a->L = L;
    a->P = p->P;
    a->ptab = p->ptab;
    a->htab = p->htab;
    a->gfl = p->gfl;
    a->X = p->X;
    a->pa = p->a;
  }
  return 0;
}

}  // namespace sbfs

using namespace sbfs;

// ---------------------------------------------------------------------------
// transform entry points for tests (device in / out, howmany rows of n)
// ---------------------------------------------------------------------------
extern "C" int sbfs_dct_rows(int inverse, size_t n, size_t howmany, const double* in_host,
                             double* out_host) {
  if (init_runtime()) return 1;
  SBFS_REQUIRE(n >= 2 && (n & (n - 1)) == 0 && n <= 4096, "n must be a power of two in [2, 4096]");
  const int m = (int)(n / 2);
  std::vector<double> tw(2 * m), t1(2 * (m + 1)), t2(2 * n);
  for (int e = 0; e < m; e++) twiddle(e, m, &tw[2 * e], &tw[2 * e + 1]);
  for (int k = 0; k <= m; k++) twiddle(k, n, &t1[2 * k], &t1[2 * k + 1]);
  for (size_t k = 0; k < n; k++) twiddle(k, 4 * n, &t2[2 * k], &t2[2 * k + 1]);
  double *din = nullptr, *dout = nullptr;
  double2 *dtw = nullptr, *dt1 = nullptr, *dt2 = nullptr;
  const size_t bytes = sizeof(double) * n * howmany;
  SBFS_HIP(hipMalloc(&din, bytes));
  SBFS_HIP(hipMalloc(&dout, bytes));
  SBFS_HIP(hipMalloc(&dtw, sizeof(double) * tw.size()));
  SBFS_HIP(hipMalloc(&dt1, sizeof(double) * t1.size()));
  SBFS_HIP(hipMalloc(&dt2, sizeof(double) * t2.size()));
  SBFS_HIP(hipMemcpy(din, in_host, bytes, hipMemcpyHostToDevice));
  SBFS_HIP(hipMemcpy(dtw, tw.data(), sizeof(double) * tw.size(), hipMemcpyHostToDevice));
  SBFS_HIP(hipMemcpy(dt1, t1.data(), sizeof(double) * t1.size(), hipMemcpyHostToDevice));
  SBFS_HIP(hipMemcpy(dt2, t2.data(), sizeof(double) * t2.size(), hipMemcpyHostToDevice));
  const int rows = (int)howmany;
  switch (ilog2(m)) {
#define SBFS_TEST_CASE(LG)                                                                                 \
  case LG:                                                                                                 \
    if (inverse)                                                                                           \
      hipLaunchKernelGGL(k_dct3_rows<LG>, dim3(cdiv(rows, WaveFft<LG>::R)), dim3(WaveFft<LG>::U), wave_fft_lds(LG), stream(), \
                         din, dout, rows, dtw, dt1, dt2);                                                  \
    else                                                                                                   \
      hipLaunchKernelGGL((k_dct2_rows<LG, false>), dim3(cdiv(rows, WaveFft<LG>::R)), dim3(WaveFft<LG>::U), wave_fft_lds(LG), stream(), \
                         din, dout, rows, dtw, dt1, dt2);                                                  \
    break;
    SBFS_LOG_CASES(SBFS_TEST_CASE)
#undef SBFS_TEST_CASE
  }
  int rc = check_launch("sbfs_dct_rows");
  if (!rc) {
    SBFS_HIP(hipStreamSynchronize(stream()));
    SBFS_HIP(hipMemcpy(out_host, dout, bytes, hipMemcpyDeviceToHost));
  }
  (void)hipFree(din);
  (void)hipFree(dout);
  (void)hipFree(dtw);
  (void)hipFree(dt1);
  (void)hipFree(dt2);
  return rc;
}
