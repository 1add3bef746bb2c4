// Row transforms with one wavefront per group of rows.
//
// The complex FFT is the Stockham autosort sequence of fft_lds (poisson.hip):
// radix-4 stages with strides s = 1, 4, 16, ... and a final radix-2 stage
// when log2(M) is odd, with the same butterflies, the same twiddle products
// and the same order of operations, so every value has the same bits.  What
// changes is where the data live between stages:
//  * a thread owns the 16 points that 4 butterflies of stage t produce and
//    4 butterflies of stage t+1 consume, so two radix-4 stages run in
//    registers per pass through LDS (an M = 1024 transform makes 3 passes
//    instead of 5 barrier-separated stages);
//  * a pass reads all its points, then writes them back in place, so a row
//    needs one M-point buffer;
//  * up to M = 1024 the G <= 64 lanes of a row sit in one wavefront, whose
//    LDS accesses complete in order, so the read and write phases are
//    separated by a wavefront-level fence and no workgroup barrier; a
//    wavefront holds R = 64 / G rows, contiguous in memory, and loads and
//    stores them with all 64 lanes;
//  * M = 2048 takes G = 128 lanes (two wavefronts and a barrier), so every
//    length keeps 16 points per lane and the same register footprint;
//  * the buffer is padded by one point in 16 so that the stride-16 stores
//    of the first pass spread over the LDS banks.
// A "unit" is the U = R G threads that share one group of rows.
#pragma once

#include <hip/hip_runtime.h>

namespace sbfs {

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ int fpad(int i) { return i + (i >> 4); }

__device__ __forceinline__ double2 wmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

__host__ __device__ constexpr int wave_fft_g(int log) { return log >= 4 ? (1 << log) / 16 : 1; }
__host__ __device__ constexpr int wave_fft_r(int log) { return wave_fft_g(log) >= 64 ? 1 : 64 / wave_fft_g(log); }
__host__ __device__ constexpr int wave_fft_u(int log) { return wave_fft_r(log) * wave_fft_g(log); }

template <int LOG>
struct WaveFft {
  static constexpr int M = 1 << LOG;           // complex points
  static constexpr int N = 2 * M;              // real points
  static constexpr int G = wave_fft_g(LOG);    // lanes per row
  static constexpr int R = wave_fft_r(LOG);    // rows per unit
  static constexpr int U = wave_fft_u(LOG);    // threads per unit: 64, or 128 for M = 2048
  static constexpr int PM = M + (M >> 4);      // padded buffer, points
  static constexpr int NR4 = LOG / 2;          // radix-4 stages
  static __device__ __forceinline__ void sync() {
    if constexpr (U > 64) __syncthreads();
    else wave_sync();
  }
};

// LDS bytes of one unit's rows
__host__ __device__ constexpr size_t wave_fft_lds(int log) {
  return sizeof(double2) * (size_t)wave_fft_r(log) * (size_t)((1 << log) + ((1 << log) >> 4));
}

// radix-4 butterfly of fft_lds: outputs 1..3 take the twiddles w1..w3
__device__ __forceinline__ void bfly4(double2 a, double2 bb, double2 c, double2 d, double2 w1, double2 w2,
                                      double2 w3, double2* o) {
  const double t0r = a.x + c.x, t0i = a.y + c.y;
  const double t1r = a.x - c.x, t1i = a.y - c.y;
  const double t2r = bb.x + d.x, t2i = bb.y + d.y;
  const double t3r = bb.x - d.x, t3i = bb.y - d.y;
  o[0] = make_double2(t0r + t2r, t0i + t2i);
  o[1] = wmul(make_double2(t1r + t3i, t1i - t3r), w1);
  o[2] = wmul(make_double2(t0r - t2r, t0i - t2i), w2);
  o[3] = wmul(make_double2(t1r - t3i, t1i + t3r), w3);
}

struct ZLoad {
  const double2* z;
  __device__ __forceinline__ double2 operator()(int i) const { return z[fpad(i)]; }
};

// The twiddles of a pass, in the order the pass uses them.  They depend on
// the lane only, so each pass's set is loaded one pass ahead (the first
// before the row arrives): the loads are in flight while the previous pass
// runs, since a wavefront-level fence does not wait for them.
struct NoTw {};
template <int N>
struct TwSet {
  double2 w[N];
};

template <int LOG, int T>
__device__ __forceinline__ auto pass_tw(int lt, const double2* __restrict__ tw) {
  using F = WaveFft<LOG>;
  constexpr int M = F::M, NR4 = F::NR4;
  if constexpr (T + 1 < NR4) {  // stages T, T+1 (one group of 4 + 4 butterflies per lane)
    static_assert((M / 16) / F::G == 1, "one group per lane");
    constexpr int s = 1 << (2 * T), Q = M / (16 * s);
    TwSet<15> t;
    const int p2 = lt >> (2 * T);
#pragma unroll
    for (int k2 = 0; k2 < 4; k2++) {
      const int p = p2 + k2 * Q;
      t.w[3 * k2] = tw[p * s];
      t.w[3 * k2 + 1] = tw[2 * p * s];
      t.w[3 * k2 + 2] = tw[3 * p * s];
    }
    t.w[12] = tw[p2 * 4 * s];
    t.w[13] = tw[2 * p2 * 4 * s];
    t.w[14] = tw[3 * p2 * 4 * s];
    return t;
  } else if constexpr (T < NR4) {  // the last radix-4 stage alone
    constexpr int s = 1 << (2 * T), PER = (M / 4) / F::G;
    TwSet<3 * PER> t;
#pragma unroll
    for (int i = 0; i < PER; i++) {
      const int p = (lt + i * F::G) >> (2 * T);
      t.w[3 * i] = tw[p * s];
      t.w[3 * i + 1] = tw[2 * p * s];
      t.w[3 * i + 2] = tw[3 * p * s];
    }
    return t;
  } else if constexpr (LOG & 1) {  // the final radix-2 stage
    TwSet<1> t;
    t.w[0] = tw[0];
    return t;
  } else {
    return NoTw{};
  }
}

// stages T and T+1 (s = 4^T): group (p2, q) runs the stage-T butterflies
// p = p2 + k2 Q (k2 = 0..3) at q, whose outputs k are the inputs of the
// stage-(T+1) butterfly (p2, q + s k)
template <int LOG, int T, typename Ld>
__device__ __forceinline__ void pass_pair(double2* Z, int lt, const TwSet<15>& t, Ld ld) {
  using F = WaveFft<LOG>;
  constexpr int M = F::M, s = 1 << (2 * T), q4 = M / (4 * s), Q = q4 / 4;
  const int q = lt & (s - 1), p2 = lt >> (2 * T);
  double2 r[16];
  {
    double2 v[4][4];
#pragma unroll
    for (int k2 = 0; k2 < 4; k2++)
#pragma unroll
      for (int k = 0; k < 4; k++) v[k2][k] = ld(q + s * (p2 + k2 * Q + k * q4));
    double2 o[4][4];
#pragma unroll
    for (int k2 = 0; k2 < 4; k2++)
      bfly4(v[k2][0], v[k2][1], v[k2][2], v[k2][3], t.w[3 * k2], t.w[3 * k2 + 1], t.w[3 * k2 + 2], o[k2]);
#pragma unroll
    for (int k = 0; k < 4; k++) bfly4(o[0][k], o[1][k], o[2][k], o[3][k], t.w[12], t.w[13], t.w[14], &r[4 * k]);
  }
  F::sync();
#pragma unroll
  for (int k = 0; k < 4; k++)
#pragma unroll
    for (int kk = 0; kk < 4; kk++) Z[fpad(q + s * k + 4 * s * (4 * p2 + kk))] = r[4 * k + kk];
  F::sync();
}

// one radix-4 stage T (the last one when their count is odd)
template <int LOG, int T, typename Ld, int N>
__device__ __forceinline__ void pass_r4(double2* Z, int lt, const TwSet<N>& t, Ld ld) {
  using F = WaveFft<LOG>;
  constexpr int M = F::M, s = 1 << (2 * T), q4 = M / (4 * s);
  constexpr int PER = (M / 4) / F::G;
  static_assert(N == 3 * PER, "twiddles of the stage");
  double2 r[PER][4];
#pragma unroll
  for (int i = 0; i < PER; i++) {
    const int b = lt + i * F::G;
    const int p = b >> (2 * T), q = b & (s - 1);
    bfly4(ld(q + s * p), ld(q + s * (p + q4)), ld(q + s * (p + 2 * q4)), ld(q + s * (p + 3 * q4)), t.w[3 * i],
          t.w[3 * i + 1], t.w[3 * i + 2], r[i]);
  }
  F::sync();
#pragma unroll
  for (int i = 0; i < PER; i++) {
    const int b = lt + i * F::G;
    const int p = b >> (2 * T), q = b & (s - 1);
#pragma unroll
    for (int k = 0; k < 4; k++) Z[fpad(q + s * (4 * p + k))] = r[i][k];
  }
  F::sync();
}

// the final radix-2 stage (s = M / 2)
template <int LOG, typename Ld>
__device__ __forceinline__ void pass_r2(double2* Z, int lt, const TwSet<1>& t, Ld ld) {
  using F = WaveFft<LOG>;
  constexpr int s = F::M / 2;
  constexpr int PER = s / F::G;
  double2 r0[PER], r1[PER];
  const double2 w = t.w[0];
#pragma unroll
  for (int i = 0; i < PER; i++) {
    const int q = lt + i * F::G;
    const double2 a = ld(q), b = ld(q + s);
    r0[i] = make_double2(a.x + b.x, a.y + b.y);
    r1[i] = wmul(make_double2(a.x - b.x, a.y - b.y), w);
  }
  F::sync();
#pragma unroll
  for (int i = 0; i < PER; i++) {
    const int q = lt + i * F::G;
    Z[fpad(q)] = r0[i];
    Z[fpad(q + s)] = r1[i];
  }
  F::sync();
}

// the stages from T on, with the twiddles of the pass at T already loaded;
// the first pass reads its input through `ld`
template <int LOG, int T, typename Ld, typename Tw>
__device__ __forceinline__ void fft_from(double2* Z, int lt, const double2* __restrict__ tw, Ld ld, const Tw& cur) {
  constexpr int NR4 = WaveFft<LOG>::NR4;
  if constexpr (T + 1 < NR4) {
    const auto nxt = pass_tw<LOG, T + 2>(lt, tw);
    pass_pair<LOG, T>(Z, lt, cur, ld);
    fft_from<LOG, T + 2>(Z, lt, tw, ZLoad{Z}, nxt);
  } else if constexpr (T < NR4) {
    const auto nxt = pass_tw<LOG, T + 1>(lt, tw);
    pass_r4<LOG, T>(Z, lt, cur, ld);
    fft_from<LOG, T + 1>(Z, lt, tw, ZLoad{Z}, nxt);
  } else if constexpr (LOG & 1) {
    pass_r2<LOG>(Z, lt, cur, ld);
  }
}

// forward M-point FFT of the points ld(0..M-1) into Z (natural order);
// tw0 = pass_tw<LOG, 0>(lt, tw), loaded by the caller as early as it likes
template <int LOG, typename Ld, typename Tw>
__device__ __forceinline__ void wave_fft(double2* Z, int lt, const double2* __restrict__ tw, Ld ld, const Tw& tw0) {
  if constexpr (LOG == 0) {
    const double2 v = ld(0);
    wave_sync();
    Z[0] = v;
    wave_sync();
  } else {
    fft_from<LOG, 0>(Z, lt, tw, ld, tw0);
  }
}

// double d of the padded row buffer Z
__device__ __forceinline__ double& zd(double2* Z, int d) { return ((double*)(Z + fpad(d >> 1)))[d & 1]; }
__device__ __forceinline__ double zd(const double2* Z, int d) {
  return ((const double*)(Z + fpad(d >> 1)))[d & 1];
}

}  // namespace sbfs
