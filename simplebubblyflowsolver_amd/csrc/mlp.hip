// Learned curvature: a 26 -> 64 -> 64 -> 1 MLP (GELU) that maps a 5 x 5
// volume-fraction stencil to the grid-scaled curvature kappa * h of the
// centre cell, as a learned correction of the plain curvature: its inputs
// are the stencil and the plain curvature of the centre cell, kp (the
// reference's corner-normal divergence on the stencil), and its output y is
// added to kp, so kappa h = kp + y.  On the wide tanh profile of the
// solver's initial condition the plain value is the curvature of the level
// set through the cell centre, 1/r; the network learns the curvature of the
// interface itself, 1/R, the same on every cell of the band (a force
// kappa grad(vof) with one kappa across the band is a discrete gradient,
// which the pressure balances).  Everything runs on the device:
//   * k_mlp_samples   rasterises training stencils of the solver's own
//                     initial profile, tanh of the distance to a circle of
//                     known radius, at random sub-cell offsets of the grid,
//                     slopes and normal offsets, in the 8 symmetries; the
//                     target is the label minus kp, weighted by 1 / label^2
//                     (a relative squared error);
//   * k_mlp_grad      forward + backward of a tile of 32 samples on BF16
//                     MFMA with FP32 accumulation, eight wavefronts per tile
//                     (16 samples by a block of 16 hidden units each);
//                     writes one gradient partial per tile (a 4096-sample
//                     batch is 128 tiles);
//   * k_mlp_adam      sums the partials in a fixed order and takes an Adam
//                     step on the FP32 master weights, refreshing the BF16
//                     copies the MFMA operands read;
//   * k_curv_gate     lists the cells of the solver's vof in the interface
//                     band;
//   * k_curv_mlp      inference on those cells, 64 per wavefront; writes
//                     kappa (NaN where the plain curvature stays).
//
// MFMA: gfx950's v_mfma_f32_16x16x32_bf16, whose lane l holds eight k,
// A[l%16][8(l/16)+j] and B[8(l/16)+j][l%16] (j = 0..7), and
// D[4(l/16)+r][l%16] (r = 0..3); one weight-gradient product uses
// v_mfma_f32_16x16x16bf16_1k (four k per lane).  Activations are
// kept as (hidden x samples) matrices, the B operand, with the weights as the
// A operand (out x in), so one layer's D output is the next layer's B operand
// lane for lane.
//
// The reference computes the curvature from the normalised vof gradient
// (src/interface/curvature_tensor.c:100-135); this network replaces it in
// the interface band (cut cells and their neighbours) of diffuse interfaces,
// where the stencil lies inside the domain and the prediction falls inside
// the training range.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <cstring>
#include <vector>

#include "march.hpp"
#include "sbfs.h"
#include "sbfs_internal.hpp"

namespace sbfs {
namespace mlp {

constexpr int NI = 25, NF = 26, NIP = 32, NH = 64;  // stencil values, features used, padded
// offsets of the FP32 parameter block
constexpr int OW1 = 0;               // W1 [NH][NIP] (columns NF.. stay zero)
constexpr int OB1 = OW1 + NH * NIP;  // b1 [NH]
constexpr int OW2 = OB1 + NH;        // W2 [NH][NH] (out x in)
constexpr int OB2 = OW2 + NH * NH;   // b2 [NH]
constexpr int OW3 = OB2 + NH;        // w3 [NH]
constexpr int OB3 = OW3 + NH;        // b3
constexpr int NP = OB3 + 1;          // 6337 parameters
constexpr int NPART = NP + 1;        // + the loss sum of the tile
constexpr double kPi = 3.14159265358979323846;
// A cell gets the learned value if its 3 x 3 block holds an interface cell
// (vof in [kUseLo, kUseHi]), its 5 x 5 stencil avoids the wall ghosts and is
// not a sharp interface (near-pure cells of both phases, vof < kPure and
// > 1 - kPure: outside the training profiles), and |kp + y| lies in
// [kKhLo, kKhHi] (kappa h; the training circles span 1/64 .. 1/6).
constexpr double kUseLo = kCutLo, kUseHi = kCutHi, kPure = 1e-3;
constexpr double kKhLo = 1. / 96., kKhHi = 0.25;

typedef short s4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------
// counter-based random numbers: splitmix64 finaliser of (seed, sample, k)
// ---------------------------------------------------------------------------
__host__ __device__ inline uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__host__ __device__ inline double uni(uint64_t seed, uint64_t i, uint64_t k) {
  const uint64_t key = seed * 0x9E3779B97F4A7C15ull + i * 16ull + k;
  return (double)(mix64(key) >> 11) * 0x1.0p-53;
}

// FP32 -> BF16, round to nearest even (finite inputs): the host form
__host__ __device__ inline unsigned short f2bf(float x) {
  uint32_t u;
  memcpy(&u, &x, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
// the same on the device, by gfx950's v_cvt_pk_bf16_f32 (round to nearest
// even under the default mode; the same bits as f2bf for finite values)
__device__ inline unsigned short bf(float x) { return __builtin_bit_cast(unsigned short, (__bf16)x); }

// GELU in the tanh form of Hendrycks & Gimpel, x (1 + tanh u) / 2 with
// u = sqrt(2/pi) (x + 0.044715 x^3), written as x s with s = sigma(2u) =
// 1 / (1 + 2^(-2u log2 e)): one exponential and one reciprocal.  It stays
// within 4.8e-4 of the erf form x Phi(x) (at x = 2.7), 20 times below the
// BF16 step of the activations there; training and inference share it.
// x -> -inf: 2^(...) overflows to inf and s to 0, so gelu -> -0.
constexpr float kGeluA = 2.3022082f;    // 2 sqrt(2/pi) log2 e
constexpr float kGeluB = 0.10294324f;   // 2 sqrt(2/pi) 0.044715 log2 e
constexpr float kGeluC = 1.5957692f;    // 2 sqrt(2/pi)
constexpr float kGeluD = 0.21406445f;   // 6 sqrt(2/pi) 0.044715
__device__ inline float gelu_s(float x) {
  const float e = __builtin_amdgcn_exp2f(-x * fmaf(kGeluB, x * x, kGeluA));
  return __builtin_amdgcn_rcpf(1.f + e);
}
__device__ inline float gelu(float x) { return x * gelu_s(x); }
// d/dx x s = s + x s (1 - s) d(2u)/dx
__device__ inline float dgelu(float x) {
  const float sg = gelu_s(x);
  return fmaf(x * fmaf(-sg, sg, sg), fmaf(kGeluD, x * x, kGeluC), sg);
}

__device__ inline f4 mma(s4 a, s4 b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0); }
// gfx950's v_mfma_f32_16x16x32_bf16: lane l holds A[l%16][8(l/16)+j] and
// B[8(l/16)+j][l%16] (j = 0..7); D as for the 16x16x16 form
typedef __bf16 b8 __attribute__((ext_vector_type(8)));
__device__ inline f4 mma32(b8 a, b8 b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0); }
__device__ inline b8 ld8(const unsigned short* p) { return *(const b8*)p; }
__device__ inline s4 ld4(const unsigned short* p) { return *(const s4*)p; }
__device__ inline s4 pack4(float a, float b, float c, float d) {
  s4 r;
  r[0] = (short)bf(a);
  r[1] = (short)bf(b);
  r[2] = (short)bf(c);
  r[3] = (short)bf(d);
  return r;
}

// plain curvature kappa h of the centre cell of a 5 x 5 stencil on square
// cells (h = 1): the corner unit gradients of the four corners of the centre
// cell (src/interface/curvature_tensor.c:16-49) and their divergence
// (:100-135), in the reference's operation order
__host__ __device__ inline double plain_kh(const double* v) {
  double nx[2][2], ny[2][2];  // corner (ci, cj): ci 0 = left, cj 0 = bottom
  for (int cj = 0; cj < 2; cj++)
    for (int ci = 0; ci < 2; ci++) {
      const int ai = 2 + ci, bj = 2 + cj;  // the corner between columns ai-1, ai and rows bj-1, bj
      const double mm = v[(bj - 1) * 5 + ai - 1], pm = v[(bj - 1) * 5 + ai];
      const double mp = v[bj * 5 + ai - 1], pp = v[bj * 5 + ai];
      const double gx = 1. / 1. * (-mm + pm - mp + pp);
      const double gy = 1. / 1. * (-mm - pm + mp + pp);
      const double ninv = 1. / fmax(sqrt(gx * gx + gy * gy), DBL_EPSILON);
      nx[cj][ci] = gx * ninv;
      ny[cj][ci] = gy * ninv;
    }
  const double nxm = 0.5 * nx[0][0] + 0.5 * nx[1][0];
  const double nxp = 0.5 * nx[0][1] + 0.5 * nx[1][1];
  const double nym = 0.5 * ny[0][0] + 0.5 * ny[0][1];
  const double nyp = 0.5 * ny[1][0] + 0.5 * ny[1][1];
  return -1. * (-nxm + nxp - nym + nyp);
}

// The stencil in canonical orientation: of its 8 images under the mirrors
// and the transpose of the grid, the one whose vof gradient over the 3 x 3
// core (gx, gy) has gx >= gy >= 0.  Curvature does not change under these
// maps, so the network sees one octant of orientations only, and two
// stencils that are mirror images of each other (the top and the bottom of
// a bubble) give the same features bit for bit: the learned curvature has
// no orientation bias that could push a bubble.  The sums negate exactly
// under a mirror, so the choice is consistent.
// w(di, dj) = v(sx dj, sy di) if transposed, else v(sx di, sy dj): the x
// mirror, then the y mirror, then the transpose, each a select between two
// fixed positions, so that no register array is indexed at run time.
__host__ __device__ inline void canonical(const double* v, double* w) {
  double gx = 0., gy = 0.;
  for (int r = 1; r <= 3; r++) gx += v[r * 5 + 3] - v[r * 5 + 1];
  for (int c = 1; c <= 3; c++) gy += v[3 * 5 + c] - v[1 * 5 + c];
  const bool mx = gx < 0., my = gy < 0., tr = fabs(gy) > fabs(gx);
  double a[NI], u[NI];
#pragma unroll
  for (int r = 0; r < 5; r++)
#pragma unroll
    for (int c = 0; c < 5; c++) a[r * 5 + c] = mx ? v[r * 5 + 4 - c] : v[r * 5 + c];
#pragma unroll
  for (int r = 0; r < 5; r++)
#pragma unroll
    for (int c = 0; c < 5; c++) u[r * 5 + c] = my ? a[(4 - r) * 5 + c] : a[r * 5 + c];
#pragma unroll
  for (int r = 0; r < 5; r++)
#pragma unroll
    for (int c = 0; c < 5; c++) w[r * 5 + c] = tr ? u[c * 5 + r] : u[r * 5 + c];
}

// features of a canonical 5 x 5 stencil (row-major, centre at 12):
// differences to the centre value, the centre value minus 1/2, and the
// plain curvature kp
__host__ __device__ inline void features(const double* v, double kp, float* f) {
  for (int k = 0; k < NI; k++) f[k] = (float)(k == 12 ? v[12] - 0.5 : v[k] - v[12]);
  f[NI] = (float)kp;
}

// ---------------------------------------------------------------------------
// training data
// ---------------------------------------------------------------------------
// One sample per thread.  Random numbers k of sample i:
//   1 radius R = 6 + 58 u cells, 4 point angle t, 5, 6 sub-cell offsets of
//   the grid, 8 slope s = 0.5 * 2^u per cell, 9 symmetry (8 transforms),
//   10 normal offset of the centre cell, 11 complement (if < 1/2); streams
//   0, 2, 3, 7 and 12 are unused.
// Cells hold 1/2 (1 + tanh(s (R - d))) at the cell centre, d the distance to
// the circle's centre: the solver's initial profile
// (initial_condition/main.py:40-41), whose slope per cell is 1/2 on every
// configured grid; THINC's own profile (beta = 1) has slope 1.  The centre
// cell sits at a normal offset uniform in +-2/s cells from the interface,
// where its vof is about 0.02 .. 0.98.  The label is the curvature of the
// interface (the vof = 1/2 circle) times h, +1/R with the gas inside, the
// reference's sign convention; the complement (liquid inside,
// vof -> 1 - vof) flips its sign.  The training target is label - kp,
// weighted by 1 / label^2.
__global__ void k_mlp_samples(uint64_t seed, long n0, long n, unsigned short* __restrict__ X,
                              float* __restrict__ T, float* __restrict__ W, float* __restrict__ Lab,
                              float* __restrict__ F) {
  const long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const uint64_t i = (uint64_t)(n0 + s);
  double u[13];
  for (int k = 0; k < 13; k++) u[k] = uni(seed, i, (uint64_t)k);
  const double b = 6. + 58. * u[1];
  const double t = 2. * kPi * u[4];
  const double ox = u[5], oy = u[6];
  const double slope = 0.5 * pow(2., u[8]);
  const int sym = (int)(u[9] * 8.);
  const double off = (2. * u[10] - 1.) * 2. / slope;
  const double px = (b + off) * cos(t), py = (b + off) * sin(t);
  const bool comp = u[11] < 0.5;
  const double label = (comp ? -1. : 1.) / b;
  const double ic = floor(px - ox), jc = floor(py - oy);
  double v[NI];
  for (int dj = -2; dj <= 2; dj++) {
    for (int di = -2; di <= 2; di++) {
      const double cx = ic + di + ox + 0.5, cy = jc + dj + oy + 0.5;
      const double val = 0.5 * (1. + tanh(slope * (b - sqrt(cx * cx + cy * cy))));
      // 8-fold symmetry: bit 0 transposes, bit 1 mirrors x, bit 2 mirrors y
      int ti = di, tj = dj;
      if (sym & 1) {
        const int w = ti;
        ti = tj;
        tj = w;
      }
      if (sym & 2) ti = -ti;
      if (sym & 4) tj = -tj;
      v[(tj + 2) * 5 + (ti + 2)] = comp ? 1. - val : val;
    }
  }
  double c[NI];
  canonical(v, c);
  const double kp = plain_kh(c);
  float f[NF];
  features(c, kp, f);
  for (int k = 0; k < NIP; k++) X[(size_t)s * NIP + k] = k < NF ? bf(f[k]) : 0;
  T[s] = (float)(label - kp);
  W[s] = (float)(1. / (label * label));
  if (Lab) Lab[s] = (float)label;
  if (F)
    for (int k = 0; k < NI; k++) F[(size_t)s * NI + k] = (float)v[k];
}

// ---------------------------------------------------------------------------
// forward pass of one 16-sample subtile.  The weights stay resident as A
// operands: aw1[mt][ks] = W1[16 mt + l%16][16 ks + 4 (l/16) ..],
// aw2[mt][ks] = W2[16 mt + l%16][16 ks + 4 (l/16) ..].
// ---------------------------------------------------------------------------
struct Weights {
  const unsigned short* w1;   // [NH][NIP] bf16
  const unsigned short* w2;   // [NH][NH] bf16
  const unsigned short* w2t;  // [NH][NH] bf16, transposed
  const float* p;             // FP32 parameters (biases, w3, b3)
};

// Inference uses the 16x16x32 form: layer 1 is one product per block of 16
// hidden units (K = the 32 padded features), layer 2 two (K = 64 hidden
// units in two chunks).  Layer 2's k order within chunk q: lane group lg's
// slots j < 4 take hidden unit 32q + 4 lg + j and slots j >= 4 hidden unit
// 32q + 16 + 4 lg + (j - 4), which is where layer 1's D layout leaves them
// (blocks 2q and 2q + 1), so the activations become the B operand without a
// shuffle; the weights are loaded in the same order.
struct AOps {
  b8 w1[4];     // [block of 16 hidden units]
  b8 w2[4][2];  // [block of 16 outputs][chunk of 32 hidden units]
};

struct B8Halves {
  s4 lo, hi;
};
__device__ inline b8 cat8(s4 lo, s4 hi) { return __builtin_bit_cast(b8, B8Halves{lo, hi}); }

__device__ inline void load_aops(const Weights& w, AOps& a) {
  const int l = threadIdx.x & 63, lr = l & 15, lg = l >> 4;
#pragma unroll
  for (int mt = 0; mt < 4; mt++) {
    a.w1[mt] = ld8(w.w1 + (16 * mt + lr) * NIP + 8 * lg);
#pragma unroll
    for (int q = 0; q < 2; q++) {
      const unsigned short* r = w.w2 + (16 * mt + lr) * NH + 32 * q + 4 * lg;
      a.w2[mt][q] = cat8(ld4(r), ld4(r + 16));
    }
  }
}

// x: the B operand of the subtile (features 8 (l/16) .. +7 of sample l%16);
// h1, h2: the pre-activations (D layout: hidden 16 mt + 4 (l/16) + r,
// sample l%16); returns y of sample l%16 of the subtile (in every lane group)
__device__ inline float forward16(const Weights& w, const AOps& a, b8 x, f4 (&h1)[4], f4 (&h2)[4]) {
  const int lg = (threadIdx.x & 63) >> 4;
#pragma unroll
  for (int mt = 0; mt < 4; mt++) {
    f4 c = mma32(a.w1[mt], x, f4{0.f, 0.f, 0.f, 0.f});
#pragma unroll
    for (int r = 0; r < 4; r++) c[r] += w.p[OB1 + 16 * mt + 4 * lg + r];
    h1[mt] = c;
  }
  b8 ab[2];
#pragma unroll
  for (int q = 0; q < 2; q++) {
#pragma unroll
    for (int r = 0; r < 4; r++) {
      ab[q][r] = (__bf16)gelu(h1[2 * q][r]);
      ab[q][4 + r] = (__bf16)gelu(h1[2 * q + 1][r]);
    }
  }
#pragma unroll
  for (int mt = 0; mt < 4; mt++) {
    f4 c = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int q = 0; q < 2; q++) c = mma32(a.w2[mt][q], ab[q], c);
#pragma unroll
    for (int r = 0; r < 4; r++) c[r] += w.p[OB2 + 16 * mt + 4 * lg + r];
    h2[mt] = c;
  }
  float s = 0.f;
#pragma unroll
  for (int mt = 0; mt < 4; mt++)
#pragma unroll
    for (int r = 0; r < 4; r++) s += w.p[OW3 + 16 * mt + 4 * lg + r] * gelu(h2[mt][r]);
  s += __shfl_xor(s, 16);
  s += __shfl_xor(s, 32);
  return s + w.p[OB3];
}

// B operand of subtile nt (16 samples) from [sample][NIP] bf16 in LDS
__device__ inline b8 load_x(const unsigned short* X, int nt) {
  const int l = threadIdx.x & 63, lr = l & 15, lg = l >> 4;
  return ld8(X + (16 * nt + lr) * NIP + 8 * lg);
}

// sum over the 16 lanes l%16 of a group: a butterfly through DPP lane
// moves (quad_perm [1,0,3,2] and [2,3,0,1], then the half-row and row
// mirrors, which after the first two steps fetch the same partial sums as
// lane l^4 and l^8 would)
template <int CTRL>
__device__ __forceinline__ float dpp_mov(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, true));
}
__device__ inline float sum16(float v) {
  v += dpp_mov<0xB1>(v);
  v += dpp_mov<0x4E>(v);
  v += dpp_mov<0x141>(v);
  v += dpp_mov<0x140>(v);
  return v;
}


// ---------------------------------------------------------------------------
// training: one workgroup = one tile of S = 16 ST samples and 4 ST
// wavefronts.  Wavefront w takes the samples of subtile st = w / 4 and the
// hidden units of block mt = w % 4 (16 of each layer's 64), so the GELU
// work, the bias sums and the dependent chain of one subtile are split four
// ways; the layers exchange their activations and deltas through LDS,
// stored both as [sample][hidden] (the B operand of the next product over
// hidden units) and [hidden][sample] (the operands of the weight gradients,
// which contract over the samples).  part[tile][NPART].
// ---------------------------------------------------------------------------
constexpr int kGradST = 2;               // subtiles per tile
constexpr int kGradS = 16 * kGradST;     // samples per tile
constexpr int kGradWaves = 4 * kGradST;  // wavefronts per tile

template <int ST>
__global__ __launch_bounds__(256 * ST) void k_mlp_grad(Weights w, const unsigned short* __restrict__ X,
                                                       const float* __restrict__ T, const float* __restrict__ Wt,
                                                       long n, long first, long stride, float scale,
                                                       float* __restrict__ part) {
  constexpr int S = 16 * ST, NW = 4 * ST;
  constexpr int XP = NIP + 8, HP = NH + 8, SP = S + 8;  // padded rows (bf16), 16-byte multiples
  __shared__ __align__(16) unsigned short sx[S * XP];    // x [sample][feature]
  __shared__ __align__(16) unsigned short sxt[NIP * SP]; // x [feature][sample]
  __shared__ __align__(16) unsigned short a1s[S * HP];   // gelu(h1) [sample][hidden]
  __shared__ __align__(16) unsigned short a1h[NH * SP];  // gelu(h1) [hidden][sample]
  __shared__ __align__(16) unsigned short d2s[S * HP];   // dh2 [sample][hidden]
  __shared__ __align__(16) unsigned short d2h[NH * SP];  // dh2 [hidden][sample]
  __shared__ __align__(16) unsigned short d1h[NH * SP];  // dh1 [hidden][sample]
  __shared__ float sy[4][S];          // output partial of each hidden block
  __shared__ float stg[S], swt[S];    // targets, weights
  __shared__ float sb[ST][3][NH];     // per-subtile sums: gw3, gb2, gb1
  __shared__ float sl[ST][2];         // per-subtile sums: loss, db3
  const int t = (int)threadIdx.x, wv = t >> 6, l = t & 63, lr = l & 15, lg = l >> 4;
  const int st = wv >> 2, mt = wv & 3;
  const long tile = blockIdx.x;
  // the tile's samples (a stride permutation of the data set), 16 threads
  // per sample, 2 features each
  {
    const int sm = t >> 4, fp = t & 15;
    const long idx = (first + (tile * S + sm) * stride) % n;
    const unsigned v = *(const unsigned*)(X + (size_t)idx * NIP + 2 * fp);
    *(unsigned*)(sx + sm * XP + 2 * fp) = v;
    sxt[(2 * fp) * SP + sm] = (unsigned short)(v & 0xFFFFu);
    sxt[(2 * fp + 1) * SP + sm] = (unsigned short)(v >> 16);
    if (fp == 0) {
      stg[sm] = T[idx];
      swt[sm] = Wt[idx];
    }
  }
  const int hb = 16 * mt + 4 * lg;  // this lane's hidden units hb + r of the D layout
  const int sc = 16 * st + lr;      // and its sample
  const b8 aw1 = ld8(w.w1 + (16 * mt + lr) * NIP + 8 * lg);
  b8 aw2[2], aw2t[2];
#pragma unroll
  for (int ks = 0; ks < 2; ks++) {
    aw2[ks] = ld8(w.w2 + (16 * mt + lr) * NH + 32 * ks + 8 * lg);
    aw2t[ks] = ld8(w.w2t + (16 * mt + lr) * NH + 32 * ks + 8 * lg);
  }
  float b1[4], b2[4], w3[4];
#pragma unroll
  for (int r = 0; r < 4; r++) {
    b1[r] = w.p[OB1 + hb + r];
    b2[r] = w.p[OB2 + hb + r];
    w3[r] = w.p[OW3 + hb + r];
  }
  __syncthreads();
  // layer 1: h1 = W1 x + b1 on this block's 16 hidden units
  f4 h1 = mma32(aw1, ld8(sx + sc * XP + 8 * lg), f4{0.f, 0.f, 0.f, 0.f});
  {
    unsigned short ab[4];
#pragma unroll
    for (int r = 0; r < 4; r++) {
      h1[r] += b1[r];
      ab[r] = bf(gelu(h1[r]));
      a1h[(hb + r) * SP + sc] = ab[r];
    }
    *(uint2*)(a1s + sc * HP + hb) = make_uint2(ab[0] | ((unsigned)ab[1] << 16), ab[2] | ((unsigned)ab[3] << 16));
  }
  __syncthreads();
  // layer 2 and this block's part of the output y = w3 . gelu(h2) + b3
  f4 h2 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int ks = 0; ks < 2; ks++) h2 = mma32(aw2[ks], ld8(a1s + sc * HP + 32 * ks + 8 * lg), h2);
  float ys = 0.f;
#pragma unroll
  for (int r = 0; r < 4; r++) {
    h2[r] += b2[r];
    ys += w3[r] * gelu(h2[r]);
  }
  ys += __shfl_xor(ys, 16);
  ys += __shfl_xor(ys, 32);
  if (lg == 0) sy[mt][sc] = ys;
  __syncthreads();
  // output layer: loss w (y - t)^2, dy = scale w (y - t), scale = 2 / batch
  const float y = ((sy[0][sc] + sy[1][sc]) + (sy[2][sc] + sy[3][sc])) + w.p[OB3];
  const float e = y - stg[sc];
  const float we = swt[sc] * e;
  const float dy = scale * we;
  float g3[4], g2[4], g1[4];
  {
    unsigned short db[4];
#pragma unroll
    for (int r = 0; r < 4; r++) {
      g3[r] = gelu(h2[r]) * dy;
      g2[r] = w3[r] * dy * dgelu(h2[r]);
      db[r] = bf(g2[r]);
      d2h[(hb + r) * SP + sc] = db[r];
    }
    *(uint2*)(d2s + sc * HP + hb) = make_uint2(db[0] | ((unsigned)db[1] << 16), db[2] | ((unsigned)db[3] << 16));
  }
  __syncthreads();
  // dh1 = gelu'(h1) (W2^T dh2) on this block's hidden units
  {
    f4 c = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < 2; ks++) c = mma32(aw2t[ks], ld8(d2s + sc * HP + 32 * ks + 8 * lg), c);
#pragma unroll
    for (int r = 0; r < 4; r++) {
      g1[r] = c[r] * dgelu(h1[r]);
      d1h[(hb + r) * SP + sc] = bf(g1[r]);
    }
  }
  // bias gradients: sums over the subtile's 16 samples
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const float s3 = sum16(g3[r]), s2 = sum16(g2[r]), s1 = sum16(g1[r]);
    if (lr == 0) {
      sb[st][0][hb + r] = s3;
      sb[st][1][hb + r] = s2;
      sb[st][2][hb + r] = s1;
    }
  }
  if (mt == 0) {
    const float sls = sum16(we * e), sdy = sum16(dy);
    if (l == 0) {
      sl[st][0] = sls;
      sl[st][1] = sdy;
    }
  }
  __syncthreads();
  float* out = part + (size_t)tile * NPART;
  // weight gradients, contracting over the tile's S samples: dW2 =
  // dh2 gelu(h1)^T (16 blocks of 16 x 16), dW1 = dh1 x (8 blocks)
  for (int b = wv; b < 24; b += NW) {
    const bool two = b < 16;
    const int ra = two ? b >> 2 : (b - 16) >> 1, cb = two ? b & 3 : (b - 16) & 1;
    const unsigned short* A = (two ? d2h : d1h) + (16 * ra + lr) * SP;
    const unsigned short* B = (two ? a1h : sxt) + (16 * cb + lr) * SP;
    f4 c = {0.f, 0.f, 0.f, 0.f};
    if constexpr (S % 32 == 0) {
#pragma unroll
      for (int ks = 0; ks < S / 32; ks++) c = mma32(ld8(A + 32 * ks + 8 * lg), ld8(B + 32 * ks + 8 * lg), c);
    } else {
      c = mma(ld4(A + 4 * lg), ld4(B + 4 * lg), c);
    }
    float* o = out + (two ? OW2 + (16 * ra + 4 * lg) * NH : OW1 + (16 * ra + 4 * lg) * NIP) + 16 * cb + lr;
    const int ld = two ? NH : NIP;
#pragma unroll
    for (int r = 0; r < 4; r++) o[r * ld] = c[r];
  }
  if (t < NH) {
    float s3 = 0.f, s2 = 0.f, s1 = 0.f;
#pragma unroll
    for (int u = 0; u < ST; u++) {
      s3 += sb[u][0][t];
      s2 += sb[u][1][t];
      s1 += sb[u][2][t];
    }
    out[OW3 + t] = s3;
    out[OB2 + t] = s2;
    out[OB1 + t] = s1;
  } else if (t == NH) {
    float sls = 0.f, sdy = 0.f;
#pragma unroll
    for (int u = 0; u < ST; u++) {
      sls += sl[u][0];
      sdy += sl[u][1];
    }
    out[NP] = sls;
    out[OB3] = sdy;
  }
}

// gradient sums over the tiles (fixed order), then Adam on the FP32 master
// weights and a refresh of the BF16 copies; `adam` 0 only sums (gradient check)
struct AdamState {
  float* p;    // parameters
  float* m;
  float* v;
  unsigned short* w1b;
  unsigned short* w2b;
  unsigned short* w2tb;
};

// A workgroup of 256 threads takes 32 parameters, 8 threads each: thread
// `sl` sums the tiles [sl tiles / 8, (sl + 1) tiles / 8) in order, 16 loads
// in flight at a time, and the first thread adds the 8 slice sums in order.
constexpr int kAdamSlices = 8;
__global__ __launch_bounds__(256) void k_mlp_adam(AdamState s, const float* __restrict__ part, int tiles, int adam,
                                                  float lr, float b1t, float b2t, float* __restrict__ grad,
                                                  float* __restrict__ loss) {
  __shared__ float slice[256];
  const int sl = (int)threadIdx.x % kAdamSlices;
  const int k = (int)blockIdx.x * (256 / kAdamSlices) + (int)threadIdx.x / kAdamSlices;
  const int t1 = (int)(((long)tiles * (sl + 1)) / kAdamSlices);
  int t = (int)(((long)tiles * sl) / kAdamSlices);
  float gs = 0.f;
  if (k <= NP) {
    for (; t + 16 <= t1; t += 16) {
      float v[16];
#pragma unroll
      for (int u = 0; u < 16; u++) v[u] = part[(size_t)(t + u) * NPART + k];
#pragma unroll
      for (int u = 0; u < 16; u++) gs += v[u];
    }
    for (; t < t1; t++) gs += part[(size_t)t * NPART + k];
  }
  slice[threadIdx.x] = gs;
  __syncthreads();
  if (sl != 0 || k > NP) return;
  float g = 0.f;
#pragma unroll
  for (int u = 0; u < kAdamSlices; u++) g += slice[threadIdx.x + u];
  if (grad) grad[k] = g;
  if (k == NP) {
    if (loss) *loss = g;
    return;
  }
  if (!adam) return;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float m = b1 * s.m[k] + (1.f - b1) * g;
  const float v = b2 * s.v[k] + (1.f - b2) * g * g;
  s.m[k] = m;
  s.v[k] = v;
  const float p = s.p[k] - lr * (m / b1t) / (sqrtf(v / b2t) + eps);
  s.p[k] = p;
  if (k >= OW1 && k < OB1) s.w1b[k - OW1] = bf(p);
  if (k >= OW2 && k < OB2) {
    const int o = (k - OW2) / NH, i = (k - OW2) % NH;
    s.w2b[o * NH + i] = bf(p);
    s.w2tb[i * NH + o] = bf(p);
  }
}

__global__ void k_mlp_bf16(AdamState s) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= NP) return;
  const float p = s.p[k];
  if (k >= OW1 && k < OB1) s.w1b[k - OW1] = bf(p);
  if (k >= OW2 && k < OB2) {
    const int o = (k - OW2) / NH, i = (k - OW2) % NH;
    s.w2b[o * NH + i] = bf(p);
    s.w2tb[i * NH + o] = bf(p);
  }
}

// prediction for host-given features [n][NIP] bf16: y[n]
__global__ __launch_bounds__(64) void k_mlp_predict(Weights w, const unsigned short* __restrict__ X, long n,
                                                    float* __restrict__ y) {
  __shared__ __align__(16) unsigned short sx[64 * NIP];
  const int l = threadIdx.x, lr = l & 15, lg = l >> 4;
  const long s = (long)blockIdx.x * 64 + l;
  for (int k = 0; k < NIP; k++) sx[l * NIP + k] = s < n ? X[(size_t)s * NIP + k] : 0;
  AOps a;
  load_aops(w, a);
  __syncthreads();
#pragma unroll 1
  for (int nt = 0; nt < 4; nt++) {
    f4 h1[4], h2[4];
    const float v = forward16(w, a, load_x(sx, nt), h1, h2);
    const long o = (long)blockIdx.x * 64 + 16 * nt + lr;
    if (lg == 0 && o < n) y[o] = v;
  }
}

// ---------------------------------------------------------------------------
// inference on the solver's vof.  A gate (k_curv_gate from the vof field,
// k_curv_gate_cut from the stage kernel's cut-cell bitmask) lists the cells
// whose rows j-1..j+1, columns i-1..i+1 hold a cut cell (vof in [kUseLo,
// kUseHi]) and whose stencil stays off the wall ghosts; k_curv_mlp walks
// that list 64 cells per wavefront with a fixed grid and writes
// kappa = (kp + y) / h in the cells the rule below admits (NaN stays in the
// others).  The list's order follows the gates' atomics and differs from
// run to run; each cell's value does not depend on it.  Layout of kappa
// and of the list entries: interior cells, (j-1) Nx + (i-1).
// ---------------------------------------------------------------------------
// The kappa field and its lists of cells (list[0] the count, then the cell
// indices) are kept together: every cell outside the last launch's list
// holds NaN.  A new field is filled with NaN and its lists emptied once.
// The launches of a step cycle through three lists (stage k writes list
// k); a launch's gate, besides listing its own cells, puts the cells of the
// previous launch's list back to NaN and empties the list the next launch
// will write, which was last read two launches before, so the field is
// never written whole again and no launch of its own does either.
__global__ void k_curv_fill(float* __restrict__ kappa, long n, unsigned* l0, unsigned* l1, unsigned* l2) {
  const long stride = (long)gridDim.x * blockDim.x;
  for (long c = (long)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += stride) kappa[c] = __int_as_float(0x7FC00000);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    l0[0] = 0u;
    if (l1) l1[0] = 0u;
    if (l2) l2[0] = 0u;
  }
}
// the gates' share of the turnover above (prev, next: null for none)
__device__ inline void curv_turnover(float* __restrict__ kappa, const unsigned* __restrict__ prev,
                                     unsigned* __restrict__ next) {
  const unsigned nt = gridDim.x * gridDim.y * blockDim.x;
  const unsigned t = (blockIdx.y * gridDim.x + blockIdx.x) * blockDim.x + threadIdx.x;
  if (prev) {
    const unsigned n = prev[0];
    for (unsigned q = t; q < n; q += nt) kappa[prev[1 + q]] = __int_as_float(0x7FC00000);
  }
  if (next && t == 0) next[0] = 0u;
}

// Reserves `total` (uniform over the wavefront) list slots for each of the
// NW wavefronts of a workgroup with one atomic on the shared counter and
// returns the wavefront's first slot; every wavefront of the workgroup calls
// it.  Atomics on one address serialise in its L2 channel: with one per
// wavefront the gates took 46-70 us on the 4096 x 8192 swarm, most of it
// waiting on the counter.
template <int NW>
__device__ inline unsigned reserve_slots(unsigned* count, unsigned total) {
  __shared__ unsigned part[NW + 1];
  const int w = (int)(threadIdx.x >> 6);
  if ((threadIdx.x & 63) == 0) part[w] = total;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned sum = 0u;
    for (int k = 0; k < NW; k++) {
      const unsigned t = part[k];
      part[k] = sum;
      sum += t;
    }
    part[NW] = sum ? atomicAdd(count, sum) : 0u;
  }
  __syncthreads();
  return part[NW] + part[w];
}

// one strip of kGateCols columns and kGateRows rows per wavefront, NW per
// workgroup: lane l holds column kGateCols bx + l, lanes 0 and 63 the
// columns on either side of the strip, so the 3 x 3 block of every strip
// cell lies in the wavefront's own loads (one per lane and row, all issued
// before the first test).  16 rows (72 VGPRs) keep more wavefronts in
// flight than 32 (152 VGPRs): 56 against 76 us at 4096 x 8192.
constexpr int kGateCols = 62, kGateRows = 16, kGateWaves = 8;
template <int NW>
__global__ __launch_bounds__(64 * NW) void k_curv_gate(const double* __restrict__ vof, int nx, int ny,
                                                       unsigned* __restrict__ count, unsigned* __restrict__ list,
                                                       float* __restrict__ kappa, const unsigned* __restrict__ prev,
                                                       unsigned* __restrict__ next) {
  curv_turnover(kappa, prev, next);
  const int l = threadIdx.x & 63;
  const int i = (int)blockIdx.x * kGateCols + l;
  // wavefronts past row ny load the clamped last row and list nothing
  const int j0 = 1 + ((int)blockIdx.y * NW + (int)(threadIdx.x >> 6)) * kGateRows;
  const size_t wc = (size_t)nx + 2;
  auto in_use = [](double v) { return (v >= kUseLo) & (v <= kUseHi); };
  // columns outside 1..nx never count; rows past ny + 1 of the last strip
  // are clamped and unused
  const bool col = i >= 1 && i <= nx;
  const int ic = min(max(i, 1), nx);
  bool cc[kGateRows + 2];
#pragma unroll
  for (int k = 0; k < kGateRows + 2; k++) {
    const double* r = vof + (size_t)(min(j0 - 1 + k, ny + 1) + 1) * wc;
    cc[k] = col & in_use(r[ic]);
  }
  const bool inside = l >= 1 && l <= kGateCols && i >= 3 && i <= nx - 2;
  uint64_t adm[kGateRows];
  unsigned total = 0u;
#pragma unroll
  for (int k = 0; k < kGateRows; k++) {
    const int j = j0 + k;
    // columns with a cut cell in rows j-1..j+1, then their 3-column
    // neighbourhood
    const uint64_t m = __ballot(cc[k] | cc[k + 1] | cc[k + 2]);
    const uint64_t nb = m | (m << 1) | (m >> 1);
    adm[k] = __ballot(j <= ny && inside && ((nb >> l) & 1ull));
    total += (unsigned)__popcll(adm[k]);
  }
  unsigned base = reserve_slots<NW>(count, total);
  if (total == 0u) return;
  const uint64_t below = l == 0 ? 0ull : (~0ull >> (64 - l));
#pragma unroll
  for (int k = 0; k < kGateRows; k++) {
    if ((adm[k] >> l) & 1ull)
      list[base + (unsigned)__popcll(adm[k] & below)] = (unsigned)(j0 + k - 1) * (unsigned)nx + (unsigned)(i - 1);
    base += (unsigned)__popcll(adm[k]);
  }
}

// Three wavefronts per SIMD: the body fits 168 VGPRs without spills (the FP64
// stencil and its canonical image hold about 100 of them), and the third
// wavefront hides more of the stencil gathers (measured 95 -> 88 us per
// launch on the 4096 x 8192 swarm; four would spill about 40 registers).
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3))) void k_curv_mlp(
    Weights w, const double* __restrict__ vof, float* __restrict__ kappa, int nx, int ny, float inv_h,
    const unsigned* __restrict__ count, const unsigned* __restrict__ list) {
  __shared__ __align__(16) unsigned short sx[64 * NIP];
  const int l = threadIdx.x, lg = l >> 4;
  const unsigned n = *count;
  if ((unsigned)blockIdx.x * 64u >= n) return;  // no cell for this workgroup (all of them when none is listed)
  const size_t wc = (size_t)nx + 2;
  AOps a;
  load_aops(w, a);
  // every wavefront leaves the loop after at most ceil(n / (64 gridDim.x))
  // passes; a listed cell has 3 <= i <= nx - 2, so its 5 x 5 stencil lies in
  // columns 1..nx and, with the 2 halo rows of vof, in rows j-2..j+2
  for (unsigned q0 = (unsigned)blockIdx.x * 64u; q0 < n; q0 += (unsigned)gridDim.x * 64u) {
    const unsigned q = q0 + (unsigned)l;
    const bool has = q < n;
    const unsigned c = has ? list[q] : 2u;  // lanes past the end read cell (3, 1) and write nothing
    const int i = 1 + (int)(c % (unsigned)nx), j = 1 + (int)(c / (unsigned)nx);
    double v[NI];
#pragma unroll
    for (int r = 0; r < 5; r++)
#pragma unroll
      for (int d = -2; d <= 2; d++) v[r * 5 + d + 2] = vof[(size_t)(j + r - 2 + 1) * wc + (i + d)];
    double lo = 1., hi = 0.;
#pragma unroll
    for (int k = 0; k < NI; k++) {
      lo = fmin(lo, v[k]);
      hi = fmax(hi, v[k]);
    }
    const bool ok = has && !(lo < kPure && hi > 1. - kPure);
    if (__ballot(ok) == 0) continue;  // uniform: the NaN written by the gate stays
    double cv[NI];
    canonical(v, cv);
    const double kp = plain_kh(cv);
    __syncthreads();  // the previous pass's reads of sx are done
    {
      float f[NF];
      features(cv, kp, f);
      for (int k = 0; k < NIP; k++) sx[l * NIP + k] = k < NF ? bf(f[k]) : 0;
    }
    __syncthreads();
    // lane l = 16 lg + l%16 keeps the prediction of subtile lg
    float y = 0.f;
#pragma unroll 1
    for (int nt = 0; nt < 4; nt++) {
      f4 h1[4], h2[4];
      const float p = forward16(w, a, load_x(sx, nt), h1, h2);
      if (lg == nt) y = p;
    }
    const double kh = kp + (double)y;
    const bool use = ok && fabs(kh) >= kKhLo && fabs(kh) <= kKhHi;
    if (use) kappa[c] = (float)(kh * (double)inv_h);
  }
}

// The same list from the stage kernel's cut-cell bitmask of this vof (8
// bytes per kStrip cells of a row, march.hpp) instead of the FP64 field.  A
// wavefront takes one bitmask strip (bit t: column 1 + b kStrip + t) and 64
// rows, lane l row j0 + l: the lane dilates its row's word by one column
// each way (with the edge bits of the neighbouring strips), the rows above
// and below arrive by lane shuffles, and a prefix sum over the lanes places
// each row's cells in the list behind one atomic per workgroup of NW
// wavefronts.
constexpr int kGateCutWaves = 16;
template <int NW>
__global__ __launch_bounds__(64 * NW) void k_curv_gate_cut(const unsigned long long* __restrict__ cut, int strips,
                                                           int nx, int ny, unsigned* __restrict__ count,
                                                           unsigned* __restrict__ list, float* __restrict__ kappa,
                                                           const unsigned* __restrict__ prev,
                                                           unsigned* __restrict__ next) {
  curv_turnover(kappa, prev, next);
  const int l = threadIdx.x & 63, b = (int)blockIdx.x;
  const int j0 = 1 + ((int)blockIdx.y * NW + (int)(threadIdx.x >> 6)) * 64;
  const bool live = j0 <= ny;  // uniform: wavefronts past row ny list nothing
  const int j = j0 + l;
  // columns of the strip within one column of a cut cell of row r (rows
  // wrap periodically)
  auto near = [&](int r) -> uint64_t {
    int rw = r - 1;
    if (rw < 0) rw += ny;
    while (rw >= ny) rw -= ny;
    const unsigned long long* w = cut + (size_t)rw * strips + b;
    const uint64_t c = w[0];
    const uint64_t lw = b > 0 ? w[-1] : 0ull, rwd = b + 1 < strips ? w[1] : 0ull;
    return c | (c << 1) | (c >> 1) | ((lw >> (kStrip - 1)) & 1ull) | ((rwd & 1ull) << (kStrip - 1));
  };
  uint64_t mine = 0ull, edge = 0ull;
  if (live) {
    mine = near(j);
    edge = l == 0 ? near(j0 - 1) : (l == 63 ? near(j0 + 64) : 0ull);
  }
  // both shuffles with every lane active (a lane masked off during a lane
  // permute would hand its neighbour zero), then the edge rows selected
  const uint64_t dn_ = __shfl_up(mine, 1), up_ = __shfl_down(mine, 1);
  const uint64_t dn = l == 0 ? edge : dn_;   // row j - 1
  const uint64_t up = l == 63 ? edge : up_;  // row j + 1
  // columns whose 5 x 5 stencil stays off the wall ghosts, 3 <= i <= nx - 2
  const int i0 = 1 + b * kStrip;
  const int lo = max(3 - i0, 0), hi = min(nx - 2 - i0, kStrip - 1);
  const uint64_t inside = hi < lo ? 0ull : ((~0ull >> (63 - hi)) & (~0ull << lo));
  uint64_t adm = j <= ny ? (dn | mine | up) & inside : 0ull;
  const unsigned n = (unsigned)__popcll(adm);
  unsigned incl = n;  // inclusive prefix sum over the lanes
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const unsigned v = __shfl_up(incl, d);
    if (l >= d) incl += v;
  }
  const unsigned total = __shfl(incl, 63);
  const unsigned base = reserve_slots<NW>(count, total);
  if (total == 0u) return;
  unsigned pos = base + incl - n;
  const unsigned row = (unsigned)(j - 1) * (unsigned)nx + (unsigned)(i0 - 1);
  while (adm) {
    list[pos++] = row + (unsigned)(__ffsll((unsigned long long)adm) - 1);
    adm &= adm - 1ull;
  }
}

// ---------------------------------------------------------------------------
// host state: one network per process
// ---------------------------------------------------------------------------
struct Net {
  float* p = nullptr;  // NP parameters, then m, v
  float* m = nullptr;
  float* v = nullptr;
  unsigned short* w1b = nullptr;
  unsigned short* w2b = nullptr;
  unsigned short* w2tb = nullptr;
  bool trained = false;
  int step = 0;
};
static Net g_net;

static int net_alloc() {
  if (g_net.p) return 0;
  SBFS_HIP(hipMalloc(&g_net.p, sizeof(float) * NP * 3));
  g_net.m = g_net.p + NP;
  g_net.v = g_net.p + 2 * NP;
  SBFS_HIP(hipMalloc(&g_net.w1b, sizeof(unsigned short) * (NH * NIP + 2 * NH * NH)));
  g_net.w2b = g_net.w1b + NH * NIP;
  g_net.w2tb = g_net.w2b + NH * NH;
  // on the library stream (non-blocking): a null-stream memset is not
  // ordered with the parameter upload queued right after it
  SBFS_HIP(hipMemsetAsync(g_net.p, 0, sizeof(float) * NP * 3, stream()));
  return 0;
}

static Weights weights() {
  Weights w;
  w.w1 = g_net.w1b;
  w.w2 = g_net.w2b;
  w.w2t = g_net.w2tb;
  w.p = g_net.p;
  return w;
}

static AdamState adam_state() {
  AdamState s;
  s.p = g_net.p;
  s.m = g_net.m;
  s.v = g_net.v;
  s.w1b = g_net.w1b;
  s.w2b = g_net.w2b;
  s.w2tb = g_net.w2tb;
  return s;
}

// upload FP32 parameters, zero the Adam moments, refresh the BF16 copies
static int set_params(const float* host) {
  if (net_alloc()) return 1;
  SBFS_HIP(hipMemcpyAsync(g_net.p, host, sizeof(float) * NP, hipMemcpyHostToDevice, stream()));
  SBFS_HIP(hipMemsetAsync(g_net.m, 0, sizeof(float) * NP * 2, stream()));
  hipLaunchKernelGGL(k_mlp_bf16, dim3(cdiv(NP, 256)), dim3(256), 0, stream(), adam_state());
  if (check_launch("mlp: bf16 weights")) return 1;
  g_net.step = 0;
  return 0;
}

// Glorot-uniform weights from the counter generator, zero biases
static void init_params(uint64_t seed, std::vector<float>& p) {
  p.assign(NP, 0.f);
  const double r1 = sqrt(6. / (NF + NH)), r2 = sqrt(6. / (NH + NH)), r3 = sqrt(6. / (NH + 1));
  for (int o = 0; o < NH; o++)
    for (int k = 0; k < NF; k++) p[OW1 + o * NIP + k] = (float)(r1 * (2. * uni(seed, 1000000000ull + o * NIP + k, 0) - 1.));
  for (int k = 0; k < NH * NH; k++) p[OW2 + k] = (float)(r2 * (2. * uni(seed, 2000000000ull + k, 0) - 1.));
  for (int k = 0; k < NH; k++) p[OW3 + k] = (float)(r3 * (2. * uni(seed, 3000000000ull + k, 0) - 1.));
}

struct Samples {
  unsigned short* x = nullptr;
  float* t = nullptr;  // targets
  float* w = nullptr;  // weights
  long n = 0;
  uint64_t seed = 0;
};

static int make_samples(uint64_t seed, long n, Samples& s) {
  SBFS_REQUIRE(n > 0, "mlp: no samples");
  SBFS_HIP(hipMalloc(&s.x, sizeof(unsigned short) * NIP * (size_t)n));
  SBFS_HIP(hipMalloc(&s.t, sizeof(float) * 2 * (size_t)n));
  s.w = s.t + n;
  s.n = n;
  s.seed = seed;
  hipLaunchKernelGGL(k_mlp_samples, dim3(cdiv(n, 128)), dim3(128), 0, stream(), seed, 0L, n, s.x, s.t, s.w,
                     (float*)nullptr, (float*)nullptr);
  return check_launch("mlp: samples");
}

static void free_samples(Samples& s) {
  (void)hipFree(s.x);
  (void)hipFree(s.t);
  s = Samples();
}

// one pass of `tiles` tiles starting at permuted position `first`
static int grad_pass(const Samples& s, long first, long stride, int tiles, float scale, float* part) {
  SBFS_TIMED("k_mlp_grad", hipLaunchKernelGGL(k_mlp_grad<kGradST>, dim3(tiles), dim3(64 * kGradWaves), 0, stream(), weights(), s.x,
                                              s.t, s.w, s.n, first, stride, scale, part));
  return check_launch("mlp: gradients");
}

}  // namespace mlp

// the kappa field of `cells` cells with its list (cells + 1 entries):
// NaN everywhere, empty list
int curv_field_init(float* kappa, unsigned* const* lists, int nlists, long cells, hipStream_t st) {
  SBFS_REQUIRE(nlists >= 1 && nlists <= 3, "curv_field_init: 1 to 3 lists");
  hipLaunchKernelGGL(mlp::k_curv_fill, dim3(1024), dim3(256), 0, st, kappa, cells, lists[0],
                     nlists > 1 ? lists[1] : nullptr, nlists > 2 ? lists[2] : nullptr);
  return check_launch("k_curv_fill");
}

// the learned curvature of a stage's input vof: the gate lists the cells of
// this vof into `list` (whose count is zero), puts the cells of `prev` back
// to NaN and empties `next`; then the network fills the listed cells
int curv_mlp_launch(const double* vof, const unsigned long long* cut, float* kappa, unsigned* list,
                    const unsigned* prev, unsigned* next, int nx, int ny, double h, hipStream_t st) {
  using namespace mlp;
  SBFS_REQUIRE(g_net.trained, "learned curvature: no trained network (sbfs_mlp_train or sbfs_mlp_set_params)");
  SBFS_REQUIRE(prev != list && next != list && (!prev || prev != next), "curv_mlp_launch: lists must differ");
  if (cut)
    SBFS_TIMED("k_curv_gate_cut",
               hipLaunchKernelGGL(k_curv_gate_cut<kGateCutWaves>, dim3(march_strips(nx), cdiv(ny, kGateCutWaves * 64)),
                                  dim3(64 * kGateCutWaves), 0, st, cut, march_strips(nx), nx, ny, list, list + 1,
                                  kappa, prev, next));
  else
    SBFS_TIMED("k_curv_gate",
               hipLaunchKernelGGL(k_curv_gate<kGateWaves>, dim3(cdiv(nx, kGateCols), cdiv(ny, kGateWaves * kGateRows)),
                                  dim3(64 * kGateWaves), 0, st, vof, nx, ny, list, list + 1, kappa, prev, next));
  SBFS_TIMED("k_curv_mlp", hipLaunchKernelGGL(k_curv_mlp, dim3(4096), dim3(64), 0, st, weights(), vof, kappa,
                                              nx, ny, (float)(1. / h), (const unsigned*)list,
                                              (const unsigned*)(list + 1)));
  return check_launch("k_curv_mlp");
}

}  // namespace sbfs

using namespace sbfs;
using namespace sbfs::mlp;

extern "C" int sbfs_mlp_num_params(void) { return NP; }

extern "C" int sbfs_mlp_samples(uint64_t seed, size_t first, size_t n, float* features, float* labels,
                                float* targets, float* weights, float* stencils) {
  SBFS_REQUIRE(n > 0 && features && labels, "sbfs_mlp_samples: bad arguments");
  unsigned short* x = nullptr;
  float *t = nullptr, *f = nullptr;
  SBFS_HIP(hipMalloc(&x, sizeof(unsigned short) * NIP * n));
  SBFS_HIP(hipMalloc(&t, sizeof(float) * 3 * n));
  if (stencils) SBFS_HIP(hipMalloc(&f, sizeof(float) * NI * n));
  hipLaunchKernelGGL(k_mlp_samples, dim3(cdiv((long)n, 128)), dim3(128), 0, stream(), seed, (long)first, (long)n,
                     x, t + n, t + 2 * n, t, f);
  int rc = check_launch("sbfs_mlp_samples");
  std::vector<unsigned short> hx(NIP * n);
  if (!rc) {
    SBFS_HIP(hipMemcpyAsync(hx.data(), x, sizeof(unsigned short) * NIP * n, hipMemcpyDeviceToHost, stream()));
    SBFS_HIP(hipMemcpyAsync(labels, t, sizeof(float) * n, hipMemcpyDeviceToHost, stream()));
    if (targets) SBFS_HIP(hipMemcpyAsync(targets, t + n, sizeof(float) * n, hipMemcpyDeviceToHost, stream()));
    if (weights) SBFS_HIP(hipMemcpyAsync(weights, t + 2 * n, sizeof(float) * n, hipMemcpyDeviceToHost, stream()));
    if (stencils) SBFS_HIP(hipMemcpyAsync(stencils, f, sizeof(float) * NI * n, hipMemcpyDeviceToHost, stream()));
    SBFS_HIP(hipStreamSynchronize(stream()));
    for (size_t k = 0; k < NIP * n; k++) {
      const uint32_t u = (uint32_t)hx[k] << 16;
      memcpy(&features[k], &u, 4);
    }
  }
  (void)hipFree(x);
  (void)hipFree(t);
  (void)hipFree(f);
  return rc;
}

extern "C" int sbfs_mlp_set_params(const float* params, size_t n) {
  SBFS_REQUIRE(params && n == (size_t)NP, "sbfs_mlp_set_params: need %d parameters", NP);
  if (set_params(params)) return 1;
  SBFS_HIP(hipStreamSynchronize(stream()));
  g_net.trained = true;
  return 0;
}

extern "C" int sbfs_mlp_get_params(float* params, size_t n) {
  SBFS_REQUIRE(params && n == (size_t)NP, "sbfs_mlp_get_params: need %d parameters", NP);
  SBFS_REQUIRE(g_net.p, "sbfs_mlp_get_params: no network");
  SBFS_HIP(hipMemcpyAsync(params, g_net.p, sizeof(float) * NP, hipMemcpyDeviceToHost, stream()));
  SBFS_HIP(hipStreamSynchronize(stream()));
  return 0;
}

// gradient of the weighted mean squared error mean(w (y - t)^2) over n
// samples (n a multiple of 32) given as BF16-representable features [n][32],
// targets and weights (null: all 1), at the current parameters; grad[NP],
// *loss = the weighted mean squared error
extern "C" int sbfs_mlp_gradient(const float* features, const float* targets, const float* weights, size_t n,
                                 float* grad, double* loss) {
  SBFS_REQUIRE(features && targets && grad && loss && n > 0 && n % kGradS == 0, "sbfs_mlp_gradient: bad arguments");
  SBFS_REQUIRE(g_net.p, "sbfs_mlp_gradient: no parameters set");
  std::vector<unsigned short> hx(NIP * n);
  for (size_t k = 0; k < NIP * n; k++) hx[k] = f2bf(features[k]);
  std::vector<float> hw(n, 1.f);
  if (weights) memcpy(hw.data(), weights, sizeof(float) * n);
  Samples s;
  SBFS_HIP(hipMalloc(&s.x, sizeof(unsigned short) * NIP * n));
  SBFS_HIP(hipMalloc(&s.t, sizeof(float) * 2 * n));
  s.w = s.t + n;
  s.n = (long)n;
  const int tiles = (int)(n / kGradS);
  float *part = nullptr, *g = nullptr;
  SBFS_HIP(hipMalloc(&part, sizeof(float) * NPART * tiles));
  SBFS_HIP(hipMalloc(&g, sizeof(float) * NPART));
  SBFS_HIP(hipMemcpyAsync(s.x, hx.data(), sizeof(unsigned short) * NIP * n, hipMemcpyHostToDevice, stream()));
  SBFS_HIP(hipMemcpyAsync(s.t, targets, sizeof(float) * n, hipMemcpyHostToDevice, stream()));
  SBFS_HIP(hipMemcpyAsync(s.w, hw.data(), sizeof(float) * n, hipMemcpyHostToDevice, stream()));
  int rc = grad_pass(s, 0, 1, tiles, 2.f / (float)n, part);
  if (!rc) {
    hipLaunchKernelGGL(k_mlp_adam, dim3(cdiv(NPART * kAdamSlices, 256)), dim3(256), 0, stream(), adam_state(), (const float*)part,
                       tiles, 0, 0.f, 1.f, 1.f, g, (float*)nullptr);
    rc = check_launch("sbfs_mlp_gradient");
  }
  std::vector<float> hg(NPART);
  if (!rc) {
    SBFS_HIP(hipMemcpyAsync(hg.data(), g, sizeof(float) * NPART, hipMemcpyDeviceToHost, stream()));
    SBFS_HIP(hipStreamSynchronize(stream()));
    memcpy(grad, hg.data(), sizeof(float) * NP);
    *loss = hg[NP] / (double)n;
  }
  (void)hipFree(part);
  (void)hipFree(g);
  free_samples(s);
  return rc;
}

// the learned kappa h = kp + y for n stencils of 25 vof values (row-major,
// centre 12)
extern "C" int sbfs_mlp_predict(const double* stencils, size_t n, double* out) {
  SBFS_REQUIRE(stencils && out && n > 0, "sbfs_mlp_predict: bad arguments");
  SBFS_REQUIRE(g_net.trained, "sbfs_mlp_predict: no trained network");
  std::vector<unsigned short> hx(NIP * n, 0);
  std::vector<double> kp(n);
  for (size_t s = 0; s < n; s++) {
    float f[NF];
    double c[NI];
    canonical(stencils + s * NI, c);
    kp[s] = plain_kh(c);
    features(c, kp[s], f);
    for (int k = 0; k < NF; k++) hx[s * NIP + k] = f2bf(f[k]);
  }
  unsigned short* x = nullptr;
  float* y = nullptr;
  SBFS_HIP(hipMalloc(&x, sizeof(unsigned short) * NIP * n));
  SBFS_HIP(hipMalloc(&y, sizeof(float) * n));
  SBFS_HIP(hipMemcpyAsync(x, hx.data(), sizeof(unsigned short) * NIP * n, hipMemcpyHostToDevice, stream()));
  hipLaunchKernelGGL(k_mlp_predict, dim3(cdiv((long)n, 64)), dim3(64), 0, stream(), weights(), x, (long)n, y);
  int rc = check_launch("sbfs_mlp_predict");
  std::vector<float> hy(n);
  if (!rc) {
    SBFS_HIP(hipMemcpyAsync(hy.data(), y, sizeof(float) * n, hipMemcpyDeviceToHost, stream()));
    SBFS_HIP(hipStreamSynchronize(stream()));
    for (size_t s = 0; s < n; s++) out[s] = kp[s] + (double)hy[s];
  }
  (void)hipFree(x);
  (void)hipFree(y);
  return rc;
}

// train from Glorot-uniform weights on `nsamples` generated stencils (a
// multiple of `batch`, itself a multiple of 32) for `epochs` passes with Adam
// (learning rate `lr`, cosine decay to lr / 20); history[e] = weighted mean
// squared error of epoch e (may be null)
extern "C" int sbfs_mlp_train(uint64_t seed, size_t nsamples, int epochs, int batch, double lr,
                              double* history) {
  SBFS_REQUIRE(nsamples > 0 && epochs > 0 && batch > 0 && batch % kGradS == 0 && nsamples % (size_t)batch == 0,
               "sbfs_mlp_train: nsamples must be a multiple of batch, batch a multiple of 32");
  std::vector<float> p0;
  init_params(seed, p0);
  if (set_params(p0.data())) return 1;
  Samples s;
  if (make_samples(seed, (long)nsamples, s)) return 1;
  const int tiles = batch / kGradS;
  const long steps = (long)(nsamples / (size_t)batch);
  float *part = nullptr, *loss = nullptr;
  SBFS_HIP(hipMalloc(&part, sizeof(float) * NPART * tiles));
  SBFS_HIP(hipMalloc(&loss, sizeof(float) * steps * epochs));
  int rc = 0;
  const long n = (long)nsamples;
  for (int e = 0; e < epochs && !rc; e++) {
    // stride permutation of the data set per epoch (a stride prime to n)
    long stride = (long)(mix64(seed + 7919ull * (uint64_t)e) % (uint64_t)n) | 1;
    auto gcd = [](long x, long y) {
      while (y) {
        const long r = x % y;
        x = y;
        y = r;
      }
      return x;
    };
    while (gcd(stride, n) != 1) stride += 2;
    const long offset = (long)(mix64(seed ^ (0xABCDull + (uint64_t)e)) % (uint64_t)n);
    for (long k = 0; k < steps && !rc; k++) {
      const double frac = (double)(e * steps + k) / (double)(epochs * steps);
      const float lrk = (float)(lr * (0.05 + 0.95 * 0.5 * (1. + cos(M_PI * frac))));
      g_net.step++;
      const float b1t = 1.f - powf(0.9f, (float)g_net.step), b2t = 1.f - powf(0.999f, (float)g_net.step);
      rc = grad_pass(s, offset + k * batch * stride, stride, tiles, 2.f / (float)batch, part);
      if (!rc) {
        SBFS_TIMED("k_mlp_adam", hipLaunchKernelGGL(k_mlp_adam, dim3(cdiv(NPART * kAdamSlices, 256)), dim3(256), 0, stream(),
                                                    adam_state(), (const float*)part, tiles, 1, lrk, b1t, b2t,
                                                    (float*)nullptr, loss + e * steps + k));
        rc = check_launch("sbfs_mlp_train");
      }
    }
  }
  if (!rc && history) {
    std::vector<float> hl((size_t)steps * epochs);
    SBFS_HIP(hipMemcpyAsync(hl.data(), loss, sizeof(float) * hl.size(), hipMemcpyDeviceToHost, stream()));
    SBFS_HIP(hipStreamSynchronize(stream()));
    for (int e = 0; e < epochs; e++) {
      double acc = 0.;
      for (long k = 0; k < steps; k++) acc += hl[(size_t)e * steps + k];
      history[e] = acc / (double)nsamples;
    }
  }
  SBFS_HIP(hipStreamSynchronize(stream()));
  (void)hipFree(part);
  (void)hipFree(loss);
  free_samples(s);
  if (!rc) g_net.trained = true;
  return rc;
}

namespace sbfs {
namespace mlp {
__global__ void k_count_learned(const float* __restrict__ k, long n, unsigned long long* __restrict__ cnt) {
  unsigned long long c = 0;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    c += k[i] == k[i] ? 1ull : 0ull;
  for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
  if ((threadIdx.x & 63) == 0) atomicAdd(cnt, c);
}
}  // namespace mlp
}  // namespace sbfs

// fraction of the interior cells of the current vof that the learned
// curvature would take from the network
extern "C" int sbfs_mlp_coverage(const domain_t* domain, const interface_t* interface, double* fraction) {
  SBFS_REQUIRE(domain && interface && fraction, "sbfs_mlp_coverage: null argument");
  if (check_array(domain, &interface->vof, NA_VOF, 8, "vof")) return 1;
  const int nx = (int)domain->mysizes[0], ny = (int)domain->mysizes[1];
  const long n = (long)nx * ny;
  float* k = nullptr;
  unsigned* list = nullptr;
  unsigned long long* cnt = nullptr;
  SBFS_HIP(hipMalloc(&k, sizeof(float) * n));
  SBFS_HIP(hipMalloc(&list, sizeof(unsigned) * (n + 1)));
  SBFS_HIP(hipMalloc(&cnt, sizeof(unsigned long long)));
  SBFS_HIP(hipMemsetAsync(cnt, 0, sizeof(unsigned long long), stream()));
  int rc = curv_field_init(k, &list, 1, n, stream());
  if (!rc)
    rc = curv_mlp_launch((const double*)interface->vof.data, nullptr, k, list, nullptr, nullptr, nx, ny, domain->hy,
                         stream());
  unsigned long long h = 0;
  if (!rc) {
    hipLaunchKernelGGL(k_count_learned, dim3(1024), dim3(256), 0, stream(), (const float*)k, n, cnt);
    rc = check_launch("sbfs_mlp_coverage");
  }
  if (!rc) {
    SBFS_HIP(hipMemcpyAsync(&h, cnt, sizeof(h), hipMemcpyDeviceToHost, stream()));
    SBFS_HIP(hipStreamSynchronize(stream()));
    *fraction = (double)h / (double)n;
  }
  (void)hipFree(k);
  (void)hipFree(list);
  (void)hipFree(cnt);
  return rc;
}
