// k_march: the whole Runge-Kutta stage up to the Poisson right-hand side,
// marching down the rows.
//
// One wavefront owns a strip of kStrip output columns [i0, i0 + kStrip) and
// a band of output rows.  Lane l holds column c = i0 - 3 + l (64 lanes cover
// the strip plus 3 halo columns on each side).  The wavefront walks the band row
// by row; at step s it loads one new row of vof, u, p, psi and computes one
// new row of every intermediate of the reference's stages:
//
//   corner gradients     row s+2   (curvature_tensor.c:16-49)
//   normals + intercept  row s+1   (curvature_tensor.c:51-97, band only)
//   curvature            row s     (curvature_tensor.c:100-135)
//   THINC / mass flux x  row s     (update/fluxx.c, mass_flux.c)
//   THINC / mass flux y  face s+1  (update/fluxy.c, mass_flux.c)
//   txx row s, tyy row s, txy corner row s+1  (predict/txx.c, tyy.c, txy.c)
//   new vof              row s     (update/main.c:54-120)
//   predicted ux, uy     row s     (predict/ux.c, uy.c, force.c)
//   Poisson rhs          row s-1   (compute_potential.c:285-363)
//
// Older rows stay in register rings, x-neighbours come from DPP wavefront
// shifts, so nothing is recomputed across rows and no LDS tile or barrier is
// needed; a term shared by two neighbouring faces or rows is formed once.  Every value is formed with the reference's operation order (the
// per-column quotients come from the coefficient tables), which keeps the
// result bitwise equal to the staged kernels.
#include <cfloat>
#include <type_traits>
#include <cmath>

#include "march.hpp"
#include "sbfs_internal.hpp"
#include "stencils.hpp"

#ifndef SBFS_FP64_WAVES
#define SBFS_FP64_WAVES 2
#endif
#ifndef SBFS_MIXED_WAVES
#define SBFS_MIXED_WAVES 3
#endif

namespace sbfs {

// value of the neighbouring lane (l-1 / l+1) through DPP wave shifts; the
// lane shifted in from outside the wavefront reads 0 (bound_ctrl), so the
// destination needs no clearing move first (every shift runs with all lanes
// active, where this equals keeping an old value of 0)
__device__ __forceinline__ double from_left(double v) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  int lo = (int)(unsigned)(b & 0xffffffffull), hi = (int)(unsigned)(b >> 32);
  lo = __builtin_amdgcn_update_dpp(0, lo, 0x138, 0xf, 0xf, true);  // wave_shr:1
  hi = __builtin_amdgcn_update_dpp(0, hi, 0x138, 0xf, 0xf, true);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

__device__ __forceinline__ double from_right(double v) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  int lo = (int)(unsigned)(b & 0xffffffffull), hi = (int)(unsigned)(b >> 32);
  lo = __builtin_amdgcn_update_dpp(0, lo, 0x130, 0xf, 0xf, true);  // wave_shl:1
  hi = __builtin_amdgcn_update_dpp(0, hi, 0x130, 0xf, 0xf, true);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

// Two ways to address a wave-uniform row, so that every access is a scalar
// base plus the lane's byte offset:
//  * lane_at: the row's 64-bit address, pinned to SGPRs (readfirstlane of a
//    uniform value), plus the lane offset; all arrays share one offset VGPR.
//    The FP32 bodies use it: they run three wavefronts per SIMD at 168 VGPRs
//    and have none to spare.
//  * RowRef: the field's base and a 32-bit byte offset of lane 0's element,
//    plus the lane offset; rows of one layout share one offset for all
//    their fields, where a 64-bit row address per field and row costs
//    several scalar or vector operations.  The FP64 bodies use it (2 %
//    faster there; in the FP32 bodies its offsets push the register count
//    into scratch).  The fused path keeps every field below 4 GiB.  Lanes
//    whose column lies before the array wrap the 32-bit offset; they are
//    the masked halo lanes and access nothing.
// Both go through a global-address-space pointer, so the accesses stay
// global_load / global_store (a generic pointer made from an integer would
// compile to flat accesses, which also count against the LDS wait counter).
template <typename T>
using GlobalBytes =
    __attribute__((address_space(1))) typename std::conditional<std::is_const<T>::value, const char, char>::type;

template <typename T>
__device__ __forceinline__ T& lane_at(T* row, unsigned loff) {
  const unsigned long long u = (unsigned long long)row;
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)u);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(u >> 32));
  GlobalBytes<T>* base = (GlobalBytes<T>*)(((unsigned long long)hi << 32) | lo);
  return *(T*)(base + loff);
}

template <typename T>
struct RowRef {
  T* base;
  unsigned off;
  __device__ __forceinline__ RowRef operator+(int d) const { return {base, off + (unsigned)(d * (int)sizeof(T))}; }
};
template <typename T>
__device__ __forceinline__ T& lane_at(RowRef<T> r, unsigned loff) {
  GlobalBytes<T>* base = (GlobalBytes<T>*)r.base;
  return *(T*)(base + (r.off + loff));
}

// the lane's element of a row as a global pointer (for streaming stores)
template <typename T>
__device__ __forceinline__ __attribute__((address_space(1))) T* lane_gp(RowRef<T> r, unsigned loff) {
  GlobalBytes<T>* base = (GlobalBytes<T>*)r.base;
  return (__attribute__((address_space(1))) T*)(base + (r.off + loff));
}

// A wave-uniform value pinned to SGPRs.  FP64 arithmetic has no scalar
// form, so a constant such as dt * gamma is formed on the vector unit and
// would otherwise hold a VGPR pair for the whole kernel.
__device__ __forceinline__ double uni(double x) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_readfirstlane((int)b);
  const int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
__device__ __forceinline__ float uni(float x) {
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(x)));
}

__device__ __forceinline__ int wrap_r(int j, int ny) {
  while (j < 1) j += ny;
  while (j > ny) j -= ny;
  return j;
}

enum { T_RHF, T_RHC, T_RJF, T_RJC, T_QF, T_QC, T_EQF, T_ECY, T_CYC, T_CYF, T_DFX, T_DCX, T_HXXF, T_HXXC, T_N };

// Arithmetic of the interface and momentum stencils.  FP64 is the reference's
// arithmetic, expression for expression (stencils.hpp); FP32 is the mixed-
// precision step's, with the hardware's reciprocal, square root, exponential
// and logarithm (v_rcp_f32, v_sqrt_f32, v_exp_f32, v_log_f32, about an ulp
// each).  __fdividef, __fsqrt_rn and __logf do not map to these without
// fast-math flags: they compile to the IEEE sequences (11, 6 and 13 VALU
// instructions).
// The FP64 parts (the density-weighted updates, the Poisson right-hand
// side, the folded projection) divide by the column width and hy as the
// reference does in the FP64 body; the mixed body, whose error budget is
// FP32's, multiplies by the stored reciprocals (rhf, 1/hy) instead (per_h,
// per_hy in march_body).
template <typename R>
struct Arith;

template <>
struct Arith<double> {
  static __device__ __forceinline__ double left(double v) { return from_left(v); }
  static __device__ __forceinline__ double right(double v) { return from_right(v); }
  static __device__ __forceinline__ double div(double a, double b) { return a / b; }
  static __device__ __forceinline__ double sqr(double x) { return sqrt(x); }
  static __device__ __forceinline__ double eps() { return DBL_EPSILON; }
  // 0.5 (1 + tanh(beta sf)) (update/main.c:18-35)
  static __device__ __forceinline__ double indicator(double sf) { return 0.5 * (1. + tanh(kVofBeta * sf)); }
  // -0.5 / beta ln(1 / vof - 1) (curvature_tensor.c:92)
  static __device__ __forceinline__ double intercept(double v) { return -0.5 / kVofBeta * log(1. / v - 1.); }
};

template <>
struct Arith<float> {
  static __device__ __forceinline__ float left(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x138, 0xf, 0xf, true));
  }
  static __device__ __forceinline__ float right(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x130, 0xf, 0xf, true));
  }
  static __device__ __forceinline__ float div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
  static __device__ __forceinline__ float sqr(float x) { return __builtin_amdgcn_sqrtf(x); }
  static __device__ __forceinline__ float eps() { return (float)DBL_EPSILON; }
  // tanh from one exponential: 1 - 2 / (e^{2|x|} + 1), sign restored
  static __device__ __forceinline__ float indicator(float sf) {
    const float x = (float)kVofBeta * sf;
    const float t = 1.f - 2.f * __builtin_amdgcn_rcpf(__expf(2.f * fabsf(x)) + 1.f);
    return 0.5f * (1.f + copysignf(t, x));
  }
  // 0.5 / beta (ln vof - ln(1 - vof)) from base-2 logarithms; 1 - vof in
  // FP64 keeps the band edge
  static __device__ __forceinline__ float intercept(double v) {
    constexpr float c = (float)(0.5 / kVofBeta * 0.69314718055994530942);
    return c * (__builtin_amdgcn_logf((float)v) - __builtin_amdgcn_logf((float)(1. - v)));
  }
};

template <typename R>
struct V3R {
  R x, y, z;
};

// THINC face value from the upwind cell, averaged over the 2-point Gauss
// rule (update/fluxx.c:53-68, fluxy.c:53-68); the band test reads the FP64 vof
template <typename R, bool XFACE>
__device__ __forceinline__ R thinc_flux(R vel, double v, V3R<R> n) {
  using A = Arith<R>;
  const R h = vel < R(0) ? R(-0.5) : R(+0.5);
  if (!in_band(v)) return vel * R(v);
  const R gp = R(kGaussP), gw = R(kGaussW);
  R fl = R(0);
  if (XFACE) {
    fl += gw * A::indicator(n.x * h + n.y * -gp + n.z);
    fl += gw * A::indicator(n.x * h + n.y * +gp + n.z);
  } else {
    fl += gw * A::indicator(n.x * -gp + n.y * h + n.z);
    fl += gw * A::indicator(n.x * +gp + n.y * h + n.z);
  }
  return vel * fl;
}

template <typename R>
__device__ __forceinline__ R mixr(double ratio, double lo, double hi, R v) {
  R r = R(1) + (R(ratio) - R(1)) * v;
  r = fmax(R(lo), r);
  r = fmin(R(hi), r);
  return r;
}

// R = double: the plain FP64 step.  R = float: the stencils of the
// interface (gradients, normals, curvature, THINC and mass fluxes) and of the
// momentum sources (viscosities, stresses, advection, surface tension,
// buoyancy) in FP32; the vof update, the density-weighted velocity update,
// the pressure gradient and the Poisson right-hand side stay FP64.
// (bx, by): the strip and band of this wavefront
// NT: the FP64 bodies store their output rows streaming (see ST below)
template <typename R, bool CORR, bool NT = false>
__device__ __forceinline__ void march_body(const MarchArgs& a, int bx, int by) {
  using A = Arith<R>;
  using V3 = V3R<R>;
  __shared__ double st[T_N][64];
  const int lane = threadIdx.x;
  const int nx = a.nx, ny = a.ny;
  const int i0 = 1 + bx * kStrip;
  const int cl = i0 - 3;  // column held by lane 0
  const int c = cl + lane;
  const int jb = 1 + by * a.band;
  const int je = min(ny, jb + a.band - 1);
  const bool c_vof = c >= 0 && c <= nx + 1;   // cells incl. wall ghosts
  const bool c_in = c >= 1 && c <= nx;        // interior cells
  const bool c_fx = c >= 1 && c <= nx + 1;    // x faces incl. walls
  const bool c_out = c >= i0 && c < i0 + kStrip && c <= nx;
  const int cc = c < 0 ? 0 : (c > nx + 1 ? nx + 1 : c);
  {
    const double* tabs[T_N] = {a.cf.rhf, a.cf.rhc, a.cf.rjf, a.cf.rjc, a.cf.qf, a.cf.qc, a.cf.eqf,
                               a.cf.ecy, a.cf.cyc, a.cf.cyf, a.cf.dfx, a.cf.dcx, a.hxxf1, a.hxxc1};
#pragma unroll
    for (int t = 0; t < T_N; t++) st[t][lane] = tabs[t][cc];
  }
  // The folded projection needs more VGPRs than the plain body has left, so
  // that variant keeps its older rows of the Poisson right-hand side, the
  // potential, the pressure and the new density in LDS instead (each lane
  // reads back only what it wrote, so no barrier is needed).
  constexpr bool LR = CORR;
  enum { RG_QUXS, RG_CYCUYS, RG_D1X, RG_CGY, RG_PSM1, RG_PM1, RG_DNM1, RG_N };
  __shared__ double rings[LR ? RG_N : 1][64];
  if (LR)
    for (int k = 0; k < RG_N; k++) rings[k][lane] = 0.;
  __syncthreads();
  const int ll = lane > 0 ? lane - 1 : 0, lr = lane < 63 ? lane + 1 : 63;
  // tables in FP64 (the FP64 parts) and in the stencil type
  auto TB = [&](int t) { return st[t][lane]; };
  auto TR = [&](int t) { return st[t][lr]; };
  auto SB = [&](int t) { return R(st[t][lane]); };
  auto SL = [&](int t) { return R(st[t][ll]); };
  auto SR = [&](int t) { return R(st[t][lr]); };
  // x / HXXF of the lane's column and x / hy: divided in the FP64 body,
  // multiplied by the stored reciprocals in the mixed one
  constexpr bool kExactQ = std::is_same<R, double>::value;
  auto per_h = [&](double x) { return kExactQ ? x / TB(T_HXXF) : x * TB(T_RHF); };
  const double hy = a.hy, rhy = a.cf.rhy;
  auto per_hy = [&](double x) { return kExactQ ? x / hy : x * rhy; };
  const R hyr = uni(R(hy)), rhyr = uni(R(rhy));
  const double dt = uni(*a.dt);
  // (dt * a.ca equals a.ca * dt bit for bit)
  const double dtca = uni(dt * a.ca), dtcb = uni(dt * a.cb), dtcg = uni(dt * a.cg);
  const double dt_old = a.g_old * dt, dt_new = a.g_new * dt;
  const double c0 = uni(1. / a.norm / dt_new * a.refden);
  const double c1 = uni(1. / a.norm * -1. / dt_new * dt_old);
  const bool has_b = a.stage != 0;
  const R denr1 = uni(R(a.denr) - R(1));
  // the previous stage's projection (correct_velocity/ux.c, uy.c): the same
  // expressions as k_post, with the lane's face distance in the divisions
  const double pdtn = a.pg_new * dt, pdto = a.pg_old * dt;
  const double pcoef = -1. / pdtn * pdto;
  const double pdtc = uni(pdtn * pcoef), irefden = uni(1. / a.refden);
  const R refdeninv = uni(R(a.refdeninv)), tension = uni(R(a.tension)), grav = uni(R(a.grav));

  // Rows of each layout, shifted to lane 0's column: a pinned row pointer
  // (FP32 bodies) or a (field, byte offset) pair (FP64 bodies), see RowRef.
  const ptrdiff_t wc = nx + 2, wf = nx + 1;
  auto row = [&](auto* f, ptrdiff_t e) {
    if constexpr (std::is_same<R, float>::value) return f + e;
    else return RowRef<std::remove_reference_t<decltype(*f)>>{f, (unsigned)(e * (ptrdiff_t)sizeof(*f))};
  };
  auto vof_row = [&](auto* f, int j) { return row(f, (ptrdiff_t)(j + 1) * wc + cl); };         // L::vof
  auto cc_row = [&](auto* f, int j) { return row(f, (ptrdiff_t)j * wc + cl); };                // L::cc
  auto xf_row = [&](auto* f, int j) { return row(f, (ptrdiff_t)j * wf + cl - 1); };            // L::xf
  auto ci_row = [&](auto* f, int j) { return row(f, (ptrdiff_t)(j - 1) * nx + cl - 1); };      // L::ci
  auto fx_row = [&](auto* f, int j) { return row(f, (ptrdiff_t)(j - 1) * (nx - 1) + cl - 2); };  // L::fx
  const unsigned loff = (unsigned)lane * 8u;
  auto AT = [&](auto r) -> auto& { return lane_at(r, loff); };
  // Output rows: with NT (fields too large for the Infinity Cache to keep,
  // stream_fields) the FP64 bodies store them streaming (non-temporal), which
  // measured k_march 1198 -> 1164 us and k_march_corr 1502 -> 1446 us at
  // 4096 x 8192; the FP32 bodies slowed by 4-5 % with it and keep ordinary
  // stores.
  static_assert(!NT || std::is_same<R, double>::value, "streaming stores in the FP64 bodies only");
  auto ST = [&](auto r, double v) {
    if constexpr (NT) __builtin_nontemporal_store(v, lane_gp(r, loff));
    else AT(r) = v;
  };
  // loads of row jw, already wrapped into 1..ny
  auto ld_vof = [&](int jw) { return c_vof ? AT(vof_row(a.vof, jw)) : 0.; };
  auto ld_ux = [&](int jw) { return c_fx ? AT(xf_row(a.ux, jw)) : 0.; };
  auto ld_cc = [&](const double* f, int jw) { return c_vof ? AT(cc_row(f, jw)) : 0.; };
  auto DEN = [&](double v) { return mix_clamp(a.denr, a.dlo, a.dhi, v); };
  // viscosity at x face c from the vof on its two sides (viscosity.c:37-53)
  auto VISUX = [&](int col, R vl, R vr) {
    R v;
    if (col == 1) v = vl;
    else if (col == nx + 1) v = vr;
    else v = R(0.5) * vl + R(0.5) * vr;
    return mixr<R>(a.visr, a.vlo, a.vhi, v);
  };
  auto VISUY = [&](R vb, R vt) { return mixr<R>(a.visr, a.vlo, a.vhi, R(0.5) * vb + R(0.5) * vt); };

  // Rings (row offsets relative to the current step s).  Several terms of
  // the reference's stencils are the same expression evaluated at the next
  // column or row (the right face of cell c is the left face of cell c+1);
  // they are formed once and taken from the neighbouring lane (DPP) or from
  // the previous step, which gives the same bits.
  double v0 = 0., vp1 = 0., vp2 = 0.;                       // vof s..s+2
  R rvm1 = 0;                                               // vof s-1, stencil type
  R g0x = 0, g0y = 0, g1x = 0, g1y = 0, g2x = 0, g2y = 0;   // corners s..s+2
  R g1xr = 0, g1yr = 0;                                     // corners of c+1, row s+1
  R n0x = 0, n0y = 0, n0z = 0, n1x = 0, n1y = 0, n1z = 0;   // normals s, s+1
  R km1 = 0, k0 = 0;                                        // curvature s-1, s
  R mfxm1 = 0, mfx0 = 0;                                    // x mass fluxes s-1, s
  R fy0 = 0, fyp1 = 0, mfy0 = 0, mfyp1 = 0;                 // y faces s, s+1
  R visx0 = 0, visx1 = 0, visy0 = 0, visy1 = 0;             // viscosities rows / faces s, s+1
  R txy0 = 0, cyf_txy0 = 0, cyc_tyym1 = 0;                  // stress terms of rows s, s-1
  double d_0 = 0.;                                          // density of vof row s
  double vn0 = 0., dnm1 = 0., dn0 = 0.;                     // new vof row s, its density s-1, s
  double ux0 = 0., uxp1 = 0., uy0 = 0., uyp1 = 0.;
  double pm1 = 0., p0 = 0.;
  double psm1 = 0., ps0 = 0., psp1 = 0.;                    // psi[0] s-1..s+1 (s+1: folded projection)
  double pop1 = 0.;                                         // previous potential s+1
  R qy_ux = 0, qy_uy = 0;                                   // y momentum fluxes at face s
  double q_uxs_m1 = 0., cyc_uys_m1 = 0.;                    // rhs terms of row s-1
  double d1x_m1 = 0., d1y = 0., cg_ym = 0.;                 // rhs densities / y gradient
  const int s_first = jb - (CORR ? kWarmCorr : kWarm);
  const double pcx0 = uni(pdtn / a.refden), pcy = uni(pdtn / a.refden / hy);
  auto ring_ld = [&](int k, double reg) { return LR ? rings[k][lane] : reg; };
  auto ring_st = [&](int k, double& reg, double v) {
    if (LR) rings[k][lane] = v;
    else reg = v;
  };
  // Every input of step s+1 is loaded during step s, unconditionally (the
  // rows wrap, so the last step's loads stay in bounds and go unused), and
  // is first read by the fence below, after the bulk of the step's
  // arithmetic and before its first store.  The vector-memory counter
  // retires loads and stores in issue order, so a wait placed after the
  // step's stores would also wait for them; at the fence only the step's
  // loads and the previous step's stores are outstanding.
  // w1..w3: rows s+1..s+3 wrapped into 1..ny, advanced with s
  int w1 = wrap_r(s_first + 1, ny), w2 = wrap_r(s_first + 2, ny), w3 = wrap_r(s_first + 3, ny);
  // prefetched inputs of the next step (the folded projection needs psi one
  // row further ahead)
  double nv = ld_vof(w2), nux = ld_ux(w1), nuy = ld_cc(a.uy, w1);
  double np = ld_cc(a.p, wrap_r(s_first, ny)), nps = ld_cc(a.psi0, CORR ? w1 : wrap_r(s_first, ny));
  double npo = CORR ? ld_cc(a.psi_old, w1) : 0.;
  // learned curvature of row s, prefetched a row ahead (0 where absent)
  // (no listed cell: every value is NaN, so none is read)
  const float* kap = (a.kappa && *a.kappa_n != 0u) ? a.kappa : nullptr;
  auto ld_kap = [&](int jw) { return (kap && c_in) ? lane_at(ci_row(kap, jw), loff >> 1) : 0.f; };
  float nkap = ld_kap(wrap_r(s_first, ny));
  // Momentum sources between stages: the FP32 bodies form them in FP32 and
  // store them so (4-byte elements in the same layout, the first half of the
  // buffer), which the next stage reads back exactly; half the bytes and the
  // same bits as an FP64 copy (measured 1287 -> 1189 us for the folded mixed
  // kernel at 4096 x 8192).  The vof source is formed in FP64 and stays so.
  auto msrc = [&](auto* f, auto rowf, int j) -> auto& {
    if constexpr (std::is_same<R, float>::value) {
      using F = std::conditional_t<std::is_const<std::remove_pointer_t<decltype(f)>>::value, const float, float>;
      return lane_at(rowf((F*)f, j), loff >> 1);
    } else {
      return AT(rowf(f, j));
    }
  };
  // beta sources of row s (stages 1, 2): the vof's for every marched row,
  // uy's from the first output row on (the predicted uy of row je+1 enters
  // the Poisson rhs of row je), ux's on output rows
  // (the FP32 stage-0 body keeps the prefetched momentum sources as float
  // until they are used, after the fence: widened at the load, each waits
  // for its load and with it for every load issued before it, which cost
  // k_march_mixed 928 against 860 us at 4096 x 8192; the folded body
  // measured faster with the early widening, 1158 against 1189-1205 us)
  using MS = std::conditional_t<std::is_same<R, float>::value && !CORR, float, double>;
  auto ld_sb = [&](int s_, int jw, double& bv, MS& by, MS& bx) {
    bv = 0.;
    by = bx = MS(0);
    if (has_b && c_in) {
      bv = AT(ci_row(a.srcb_vof, jw));
      if (s_ >= jb) by = msrc(a.srcb_uy, ci_row, jw);
      if (s_ >= jb && s_ <= je && c >= 2) bx = msrc(a.srcb_ux, fx_row, jw);
    }
  };
  double nsbv;
  MS nsby, nsbx;
  ld_sb(s_first, wrap_r(s_first, ny), nsbv, nsby, nsbx);
  // the fence below, for the first step's inputs: without it the loop entry
  // would carry outstanding loads into every step's first reads
  auto fence = [&]() {
    asm volatile("" : "+v"(nv), "+v"(nux), "+v"(nuy), "+v"(np), "+v"(nps), "+v"(nkap), "+v"(nsbv), "+v"(nsby),
                 "+v"(nsbx));
    if constexpr (CORR) asm volatile("" : "+v"(npo));
  };
  // the vof ring holds rows s_first and s_first + 1 before the first step,
  // so the first step's corner gradients are already valid
  vp1 = ld_vof(wrap_r(s_first, ny));
  vp2 = ld_vof(w1);
  asm volatile("" : "+v"(vp1), "+v"(vp2));
  fence();
  unsigned weight = 0;  // marched rows with an interface cell
  for (int s = s_first; s <= je + 1; s++) {
    // -- new input rows -----------------------------------------------------
    rvm1 = R(v0); v0 = vp1; vp1 = vp2; vp2 = nv;
    ux0 = uxp1; uxp1 = nux;
    uy0 = uyp1; uyp1 = nuy;
    if (!LR) pm1 = p0;
    p0 = np;
    if (CORR) {
      ps0 = psp1; psp1 = nps;   // row s-1 of the potential is in LDS
      pop1 = npo;
    } else {
      psm1 = ps0; ps0 = nps;
    }
    const bool outrow = s >= jb && s <= je;
    if (a.weight) weight += __ballot(c_vof && in_band(v0)) != 0ull ? 1u : 0u;
    const float kap = nkap;
    const double sbv = nsbv;
    const MS sby = nsby, sbx = nsbx;
    // (the folded bodies load the beta sources first: the FP32 folded body
    // widens them at the load, and the wait that needs then covers them
    // alone rather than every prefetch of the step, 1158 -> 1139 us at
    // 4096 x 8192)
    if (CORR) ld_sb(s + 1, w1, nsbv, nsby, nsbx);
    nv = ld_vof(w3);
    nux = ld_ux(w2);
    nuy = ld_cc(a.uy, w2);
    np = ld_cc(a.p, w1);
    nps = ld_cc(a.psi0, CORR ? w2 : w1);
    if (CORR) npo = ld_cc(a.psi_old, w2);
    nkap = ld_kap(w1);
    if (!CORR) ld_sb(s + 1, w1, nsbv, nsby, nsbx);
    w1 = w2;
    w2 = w3;
    w3 = w3 == ny ? 1 : w3 + 1;
    const R rux0 = R(ux0), ruxp1 = R(uxp1), ruy0 = R(uy0), ruyp1 = R(uyp1);
    const R rv0 = R(v0), rvp1 = R(vp1);
    // -- corner gradients, row s+2 (curvature_tensor.c:16-49) ----------------
    g0x = g1x; g0y = g1y; g1x = g2x; g1y = g2y;
    // (corners of c+1, row s: shifted again rather than kept, one register
    // pair fewer)
    const R g0xr = A::right(g0x), g0yr = A::right(g0y);
    {
      const double mm = from_left(vp1), pm = vp1, mp = from_left(vp2), pp = vp2;
      R gx = 0, gy = 0;
      if (c_fx && !(mm == pm && mm == mp && mm == pp)) {
        const R rmm = R(mm), rpm = R(pm), rmp = R(mp), rpp = R(pp);
        gx = SB(T_RHF) * (-rmm + rpm - rmp + rpp);
        gy = rhyr * (-rmm - rpm + rmp + rpp);
        const R norm = A::sqr(gx * gx + gy * gy);
        const R ninv = A::div(R(1), fmax(norm, A::eps()));
        gx = gx * ninv;
        gy = gy * ninv;
      }
      g2x = gx;
      g2y = gy;
    }
    const R g2xr = A::right(g2x), g2yr = A::right(g2y);
    g1xr = A::right(g1x);
    g1yr = A::right(g1y);
    // -- normals, row s+1 (curvature_tensor.c:51-97) ---------------------------
    n0x = n1x; n0y = n1y; n0z = n1z;
    {
      R nxv = 0, nyv = 0, nz = 0;
      if (c_in && in_band(vp1)) {
        nxv = g1x + g1xr + g2x + g2xr;
        nyv = g1y + g1yr + g2y + g2yr;
        nxv = A::div(nxv, SB(T_HXXC));
        nyv = A::div(nyv, hyr);
        const R norm = A::sqr(nxv * nxv + nyv * nyv);
        const R ninv = A::div(R(1), fmax(norm, A::eps()));
        nxv *= ninv;
        nyv *= ninv;
        nz = A::intercept(vp1);
      }
      n1x = nxv;
      n1y = nyv;
      n1z = nz;
    }
    // -- curvature, row s (curvature_tensor.c:100-135) -------------------------
    km1 = k0;
    {
      R kk = 0;
      // the learned value of row s replaces the plain one where the network
      // gave one
      const float kl = kap;
      if (c_in) {
        const R nxm = R(0.5) * g0x + R(0.5) * g1x;
        const R nxp = R(0.5) * g0xr + R(0.5) * g1xr;
        const R nym = R(0.5) * g0y + R(0.5) * g0yr;
        const R nyp = R(0.5) * g1y + R(0.5) * g1yr;
        const R div = SB(T_RJC) * (-SB(T_QF) * nxm + SR(T_QF) * nxp - SB(T_CYC) * nym + SB(T_CYC) * nyp);
        kk = R(-1) * div;
        if (kl == kl && kl != 0.f) kk = R(kl);
      }
      k0 = kk;
    }
    // -- THINC and mass fluxes at x face c, row s (fluxx.c, mass_flux.c) -------
    mfxm1 = mfx0;
    R fx0 = 0;
    {
      const double vl = from_left(v0);
      const R nlx = A::left(n0x), nly = A::left(n0y), nlz = A::left(n0z);
      R mf = 0;
      if (c_fx) {
        const R vel = rux0;
        if (c != 1 && c != nx + 1) {
          const bool up = vel < R(0);
          const double v = up ? v0 : vl;
          const V3 n = up ? V3{n0x, n0y, n0z} : V3{nlx, nly, nlz};
          fx0 = thinc_flux<R, true>(vel, v, in_band(v) ? n : V3{0, 0, 0});
        }
        mf = vel + denr1 * fx0;
      }
      mfx0 = mf;
    }
    // -- THINC and mass fluxes at y face s+1 (fluxy.c, mass_flux.c) ------------
    fy0 = fyp1;
    mfy0 = mfyp1;
    {
      R f = 0, mf = 0;
      if (c_vof) {
        const R vel = ruyp1;
        if (c != 0 && c != nx + 1) {
          const bool up = vel < R(0);
          const double v = up ? vp1 : v0;
          const V3 n = up ? V3{n1x, n1y, n1z} : V3{n0x, n0y, n0z};
          f = thinc_flux<R, false>(vel, v, in_band(v) ? n : V3{0, 0, 0});
        }
        mf = vel + denr1 * f;
      }
      fyp1 = f;
      mfyp1 = mf;
    }
    // -- viscosities (viscosity.c) ------------------------------------------------
    // (every DPP shift runs in uniform control flow: an inactive source lane
    // would not deliver its value)
    const double v0l = from_left(v0);
    const R rv0l = R(v0l), rvp1l = A::left(rvp1);
    const R rux0r = A::right(rux0);
    const R ruyp1l = A::left(ruyp1);
    visx0 = visx1;
    visx1 = VISUX(c, rvp1l, rvp1);   // x face c, row s+1
    visy0 = visy1;
    visy1 = VISUY(rv0, rvp1);        // y face s+1
    // -- new vof, row s (update/main.c:54-120), FP64 ---------------------------
    double src = 0.;
    const double fx0d = double(fx0);
    const double fx0r = A::right(fx0);
    {
      double v = 0.;
      if (c_in) {
        src = TB(T_RJC) * (+TB(T_QF) * fx0d - TR(T_QF) * fx0r + TB(T_CYC) * double(fy0) - TB(T_CYC) * double(fyp1));
        v = v0;
        v += dtca * src;
        if (has_b) v += dtcb * sbv;
      }
      vn0 = v;
    }
    // -- densities (density.c) and their face averages -------------------------
    const double d_m1 = d_0;
    d_0 = DEN(v0);
    if (!LR) dnm1 = dn0;
    dn0 = DEN(vn0);
    const double d_0l = from_left(d_0), dn0l = from_left(dn0);
    const double dfx = 0.5 * d_0l + 0.5 * d_0;    // x face c, row s, old vof
    const double dfy = 0.5 * d_m1 + 0.5 * d_0;    // y face s, old vof
    const double d1x = 0.5 * dn0l + 0.5 * dn0;    // x face c, row s, new vof
    d1y = 0.5 * ring_ld(RG_DNM1, dnm1) + 0.5 * dn0;  // y face s, new vof
    // Rows before jb - 1 only fill the interface and density rings: their
    // stresses, momentum fluxes, predicted velocity and Poisson rhs reach no
    // output, since every ring that row jb's predicted velocity reads is
    // refilled by row jb - 1, and the rhs of row jb is formed at step jb + 1
    // from the rings of step jb.  So the stresses and momentum fluxes start
    // at row jb - 1, the predicted velocity and the rhs at row jb.  (The
    // flags are uniform: the DPP shifts inside stay in uniform control flow.)
    // Measured on the 2048 x 4096 swarm: the folded body gains from both
    // (k_march_corr 452 -> 414 us); the plain one gains from the later start
    // of the predicted velocity and the rhs (k_march 296 -> 288 us) but loses
    // with the stresses and fluxes skipped too (361 us, its code laid out
    // worse), so it computes those on every row.
    const bool mom = !CORR || s >= jb - 1, pred = s >= jb;
    const R txy_prev = txy0;   // corner row s
    R txx0 = 0, tyy0 = 0, txy1 = 0;
    R qx = 0, qm = 0, qxl = 0, qmr = 0, qt = 0, qtl = 0, qf_txy0 = 0, qf_txy0r = 0;
    R qy_ux_new = 0, qy_uy_new = 0, cyf_txy1 = 0, cyc_tyy0 = 0;
    if (mom) {
      // -- stress tensor (predict/txx.c, tyy.c, txy.c) --------------------------
      const R visx0r = A::right(visx0), visy1l = A::left(visy1);
      if (c_in) {
        const R vx = R(0.5) * visx0 + R(0.5) * visx0r;
        const R l = SB(T_RHC) * (-rux0 + rux0r);
        txx0 = vx * l + vx * l;
      }
      if (c_in) {
        const R vy = R(0.5) * visy0 + R(0.5) * visy1;
        const R l = rhyr * (-ruy0 + ruyp1);
        tyy0 = vy * l + vy * l;
      }
      if (c_fx) {
        const R vx = R(0.5) * visx0 + R(0.5) * visx1;
        const R vy = R(0.5) * visy1l + R(0.5) * visy1;
        const R lxy = rhyr * (-rux0 + ruxp1);
        const R lyx = SB(T_RHF) * (-ruyp1l + ruyp1);
        txy1 = vx * lxy + vy * lyx;
      }
      // -- momentum fluxes shared between neighbouring faces -------------------
      const R mfx0r = A::right(mfx0), mfyp1l = A::left(mfyp1);
      const R ruy0l = A::left(ruy0);
      // ux, x direction: cell-centre flux (mp * up) of the cell right of face c
      qx = (SB(T_EQF) * mfx0 + SR(T_EQF) * mfx0r) * (R(0.5) * rux0 + R(0.5) * rux0r);
      // uy, x direction: corner flux (mm * um) at face c
      qm = (SB(T_EQF) * mfxm1 + SB(T_EQF) * mfx0) * (R(0.5) * ruy0l + R(0.5) * ruy0);
      qxl = A::left(qx);
      qmr = A::right(qm);
      qt = SB(T_QC) * txx0;
      qtl = A::left(qt);
      qf_txy0 = SB(T_QF) * txy_prev;
      qf_txy0r = A::right(qf_txy0);
      // ux, y direction: corner flux (mp * up) at face s+1
      qy_ux_new = (SL(T_ECY) * mfyp1l + SB(T_ECY) * mfyp1) * (R(0.5) * rux0 + R(0.5) * ruxp1);
      // uy, y direction: cell-centre flux (mp * up) of row s
      qy_uy_new = (SB(T_ECY) * mfy0 + SB(T_ECY) * mfyp1) * (R(0.5) * ruy0 + R(0.5) * ruyp1);
      cyf_txy1 = SB(T_CYF) * txy1;
      cyc_tyy0 = SB(T_CYC) * tyy0;
    }
    // The fence: an empty statement that reads and redefines the next step's
    // inputs, so the compiler waits for their loads here, ahead of this
    // step's first store, rather than where the loop carries them over.
    fence();
    double uxs = 0., uys = 0.;
    if (mom) {
      if (pred) {
        const double p0l = from_left(p0);
        const R k0l = A::left(k0);
        // -- predicted ux, row s (predict/ux.c, force.c) -----------------------
        if (c >= 2 && c <= nx) {
          R s_ = 0;
          s_ -= SB(T_RJF) * (-qxl + qx);
          s_ -= SB(T_RJF) * (-qy_ux + qy_ux_new);
          s_ += SB(T_DFX) * (-qtl + qt);
          s_ += SB(T_DFX) * (-cyf_txy0 + cyf_txy1);
          double g = 0.;
          g -= TB(T_RHF) * (-p0l + p0);
          const R rdfx = R(dfx);
          {
            const R k = R(0.5) * k0l + R(0.5) * k0;
            const R delta = SB(T_RHF) * (-rv0l + rv0);
            s_ += rdfx * refdeninv * tension * k * delta;
          }
          s_ += grav * rdfx;
          const double sd = double(s_);
          if (a.srca_ux && outrow && c_out) msrc(a.srca_ux, fx_row, s) = s_;
          uxs = dfx * ux0 + dtca * sd + dtcb * double(sbx) + dtcg * g;
          uxs /= d1x;
        }
        // -- predicted uy, row s (predict/uy.c, force.c) -----------------------
        if (c_in) {
          R s_ = 0;
          s_ -= SB(T_RJC) * (-qm + qmr);
          s_ -= SB(T_RJC) * (-qy_uy + qy_uy_new);
          s_ += SB(T_DCX) * (-qf_txy0 + qf_txy0r);
          s_ += SB(T_DCX) * (-cyc_tyym1 + cyc_tyy0);
          double g = 0.;
          g -= rhy * (-ring_ld(RG_PM1, pm1) + p0);
          {
            const R k = R(0.5) * km1 + R(0.5) * k0;
            const R delta = rhyr * (-rvm1 + rv0);
            s_ += R(dfy) * refdeninv * tension * k * delta;
          }
          const double sd = double(s_);
          if (a.srca_uy && outrow && c_out) msrc(a.srca_uy, ci_row, s) = s_;
          uys = dfy * uy0 + dtca * sd + dtcb * double(sby) + dtcg * g;
          uys /= d1y;
        }
      }
      qy_ux = qy_ux_new;
      qy_uy = qy_uy_new;
      cyf_txy0 = cyf_txy1;
      cyc_tyym1 = cyc_tyy0;
      txy0 = txy1;
      // -- outputs of row s -----------------------------------------------------
      const double txy0r = double(A::right(txy_prev));
      if (a.cut) {
        const uint64_t m = __ballot(outrow && c_out && vn0 >= kCutLo && vn0 <= kCutHi);
        if (outrow && lane == 0) a.cut[(size_t)(s - 1) * march_strips(nx) + bx] = m >> 3;
      }
      if (outrow && c_out) {
        auto put = [&](double* f, int nh, auto row, double v) {
          ST(row(f, s), v);
          if (s > ny - nh) ST(row(f, s - ny), v);
          if (s <= nh) ST(row(f, s + ny), v);
        };
        auto put_at = [&](double* f, int nh, auto row, int off, double v) {
          ST(row(f, s) + off, v);
          if (s > ny - nh) ST(row(f, s - ny) + off, v);
          if (s <= nh) ST(row(f, s + ny) + off, v);
        };
        auto rv = [&](double* f, int j) { return vof_row(f, j); };
        auto rc = [&](double* f, int j) { return cc_row(f, j); };
        auto rx = [&](double* f, int j) { return xf_row(f, j); };
        put(a.vofn, 2, rv, vn0);
        if (c == 1) put_at(a.vofn, 2, rv, -1, 0.);
        if (c == nx) put_at(a.vofn, 2, rv, 1, 0.);
        if (a.srca_vof) ST(ci_row(a.srca_vof, s), src);
        put(a.uxs, 1, rx, c == 1 ? 0. : uxs);
        if (c == nx) put_at(a.uxs, 1, rx, 1, 0.);
        put(a.uys, 1, rc, uys);
        if (c == 1) put_at(a.uys, 1, rc, -1, kUyXm);
        if (c == nx) put_at(a.uys, 1, rc, 1, kUyXp);
        if (a.den1) {
          const double dg = DEN(0.);
          put(a.den1, 1, rc, dn0);
          if (c == 1) put_at(a.den1, 1, rc, -1, dg);
          if (c == nx) put_at(a.den1, 1, rc, 1, dg);
        }
        if (a.txy) {
          put(a.txy, 1, rx, double(txy_prev));
          if (c == nx) put_at(a.txy, 1, rx, 1, txy0r);
        }
        if (CORR && a.p_out) {
          // the previous stage's updated pressure, with its Neumann ghosts
          put(a.p_out, 1, rc, p0);
          if (c == 1) put_at(a.p_out, 1, rc, -1, p0);
          if (c == nx) put_at(a.p_out, 1, rc, 1, p0);
        }
      }
    }
    // -- Poisson right-hand side of row s-1 (compute_potential.c:315-360), FP64
    if (pred) {
      // u* divergence: the x-face terms of cell c+1 come from lane c+1
      const double q_uxs = ring_ld(RG_QUXS, q_uxs_m1);     // QF * ux* at face c, row s-1
      const double q_uxsr = from_right(q_uxs);
      const double cyc_uys0 = TB(T_CYC) * uys;
      // explicit split-pressure term: the gradient at face c+1 is lane c+1's
      // gradient at its left face; the top y face of row s-1 is the bottom
      // face of row s, kept for the next step
      const double psm1v = ring_ld(RG_PSM1, psm1);
      const double psl = from_left(psm1v);
      const double dxm = ring_ld(RG_D1X, d1x_m1);
      const double gxm = per_h((a.refden / dxm - 1.) * (-psl + psm1v));
      const double qgx = TB(T_QF) * gxm;
      const double qgxr = from_right(qgx);
      const double gyp = per_hy((a.refden / d1y - 1.) * (-psm1v + ps0));
      const double cg_yp = TB(T_CYC) * gyp;
      const int r = s - 1;
      if (r >= jb && r <= je && c_out) {
        const double div = TB(T_RJC) * (-q_uxs + q_uxsr - ring_ld(RG_CYCUYS, cyc_uys_m1) + cyc_uys0);
        const double add = TB(T_RJC) * (-qgx + qgxr - ring_ld(RG_CGY, cg_ym) + cg_yp);
        ST(ci_row(a.rhs, r), c0 * div + c1 * add);
      }
      ring_st(RG_CGY, cg_ym, cg_yp);
      ring_st(RG_CYCUYS, cyc_uys_m1, cyc_uys0);
    }
    ring_st(RG_QUXS, q_uxs_m1, TB(T_QF) * uxs);
    ring_st(RG_D1X, d1x_m1, d1x);
    if (CORR) ring_st(RG_PSM1, psm1, ps0);
    if (LR) {
      rings[RG_PM1][lane] = p0;
      rings[RG_DNM1][lane] = dn0;
    }
    if (CORR) {
      // The previous stage's projection of the rows prefetched for the next
      // step (unused after the last step), formed here where the stencil's
      // temporaries are dead: ux of row
      // s+2, uy of face s+2 (correct_velocity/ux.c:26-45, uy.c:26-43) and the
      // pressure of row s+1 (update_pressure.c:13-24, written with the
      // outputs of that row), with k_post's expressions and the lane's own
      // face distance.  The vof is this stage's input, i.e. the previous
      // stage's new vof.
      const double da = DEN(vp1), db = DEN(vp2);   // rows s+1, s+2
      const double dbl = from_left(db);
      const double nps_l = from_left(nps), npo_l = from_left(npo);
      {
        double u = 0.;
        if (c >= 2 && c <= nx) {
          u = nux;
          u -= per_h(pcx0) * (-nps_l + nps);
          const double d = 0.5 * dbl + 0.5 * db;
          u += per_h(pdtc * (1. / d - irefden)) * (-npo_l + npo);
        }
        nux = u;
      }
      {
        double u = 0.;
        if (c == 0) u = kUyXm;
        else if (c == nx + 1) u = kUyXp;
        else if (c_in) {
          u = nuy;
          u -= pcy * (-psp1 + nps);
          const double d = 0.5 * da + 0.5 * db;
          u += per_hy(pdtc * (1. / d - irefden)) * (-pop1 + npo);
        }
        nuy = u;
      }
      // (stage 2 of a folded step: the input pressure plus stage 0's
      // potential, which stage 1 formed and did not store, then stage 1's)
      if (c_in) np = (a.p_twice ? np + pop1 : np) + psp1;
    }
  }
  if (a.weight && lane == 0) a.weight[by * march_strips(nx) + bx] = weight;
}

__global__ __launch_bounds__(1024) void k_march_order(const unsigned* __restrict__ weight,
                                                      unsigned* __restrict__ order, int n, int max_weight) {
  constexpr int kB = 32, kPer = 8, kT = 1024;
  __shared__ unsigned cnt[kB], next[kB];
  const int t = threadIdx.x;
  const unsigned mw = max_weight > 0 ? (unsigned)max_weight : 1u;
  // the buckets of kPer units per thread, their loads issued together
  // (kB: no unit)
  auto buckets = [&](int base, int* b) {
    unsigned w[kPer];
#pragma unroll
    for (int i = 0; i < kPer; i++) {
      const int u = base + i * kT + t;
      w[i] = u < n ? weight[u] : 0u;
    }
#pragma unroll
    for (int i = 0; i < kPer; i++) {
      const int u = base + i * kT + t;
      b[i] = u < n ? (int)((unsigned long long)min(w[i], mw) * (kB - 1) / mw) : kB;
    }
  };
  // Adds the wavefront's lanes with a unit to their buckets of h and gives
  // each such lane a distinct slot: one atomic per distinct bucket in the
  // wavefront (most units share the bucket of weight 0, and lane-by-lane
  // atomics on one LDS word serialise).  Called in uniform control flow.
  const int lane = t & 63;
  auto add = [&](unsigned* h, int b) -> unsigned {
    unsigned slot = 0;
    unsigned long long left = __ballot(b < kB);
    while (left) {
      const int lead = __ffsll((long long)left) - 1;
      const int bl = __shfl(b, lead);
      const unsigned long long m = __ballot(b == bl);
      unsigned first = 0;
      if (lane == lead) first = atomicAdd(&h[bl], (unsigned)__popcll(m));
      first = __shfl(first, lead);
      if ((m >> lane) & 1ull) slot = first + (unsigned)__popcll(m & ((1ull << lane) - 1ull));
      left &= ~m;
    }
    return slot;
  };
  if (t < kB) cnt[t] = 0;
  __syncthreads();
  for (int base = 0; base < n; base += kPer * kT) {
    int b[kPer];
    buckets(base, b);
#pragma unroll
    for (int i = 0; i < kPer; i++) add(cnt, b[i]);
  }
  __syncthreads();
  if (t == 0) {
    unsigned s = 0;
    for (int b = kB - 1; b >= 0; b--) {
      next[b] = s;
      s += cnt[b];
    }
  }
  __syncthreads();
  for (int base = 0; base < n; base += kPer * kT) {
    int b[kPer];
    buckets(base, b);
#pragma unroll
    for (int i = 0; i < kPer; i++) {
      const unsigned slot = add(next, b[i]);
      if (b[i] < kB) order[slot] = (unsigned)(base + i * kT + t);
    }
  }
}

// The FP64 bodies need about 230-250 VGPRs, two wavefronts per SIMD (three
// would spill dozens of registers to scratch, measured 1.4x slower).  The
// FP32 bodies fit 168 VGPRs, three wavefronts per SIMD, which hides more of
// their latency (the folded variant 1.35x faster than at two).
// The arguments arrive through a pointer to device memory, not by value:
// held as kernel arguments, the ~50 uniform pointers and constants of
// MarchArgs outgrow the scalar registers, and the compiler spills them into
// vector lanes (v_writelane / v_readlane in the row loop, up to 185 of them
// in the FP64 folded body); from memory it reloads them with scalar loads.
// the (strip, band) of this workgroup: from the grid, or from the dispatch
// order (false: a slot past the last unit)
__device__ __forceinline__ bool march_unit(const MarchArgs& a, int& bx, int& by) {
  bx = blockIdx.x;
  by = blockIdx.y;
  if (a.order) {
    const unsigned u = a.order[blockIdx.x];
    if (u >= (unsigned)march_units(a.nx, a.ny, a.band)) return false;
    const int strips = march_strips(a.nx);
    bx = (int)u % strips;
    by = (int)u / strips;
  }
  return true;
}
#define SBFS_MARCH_KERNEL(name, R, CORR, WAVES)                                            \
  __global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(WAVES))) void name( \
      const MarchArgs* __restrict__ args) {                                               \
    const MarchArgs& a = *args;                                                           \
    int bx, by;                                                                           \
    if (march_unit(a, bx, by)) march_body<R, CORR>(a, bx, by);                            \
  }                                                                                       \
  __global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(WAVES))) void name##_batch( \
      const MarchArgs* __restrict__ tab, const uint2* __restrict__ map) {                \
    const uint2 m = map[blockIdx.x];                                                      \
    const MarchArgs& a = tab[m.x];                                                        \
    const int strips = (a.nx + kStrip - 1) / kStrip;                                      \
    march_body<R, CORR>(a, (int)m.y % strips, (int)m.y / strips);                         \
  }
SBFS_MARCH_KERNEL(k_march, double, false, SBFS_FP64_WAVES)
SBFS_MARCH_KERNEL(k_march_corr, double, true, SBFS_FP64_WAVES)
SBFS_MARCH_KERNEL(k_march_mixed, float, false, SBFS_MIXED_WAVES)
SBFS_MARCH_KERNEL(k_march_mixed_corr, float, true, SBFS_MIXED_WAVES)
#undef SBFS_MARCH_KERNEL
// the FP64 stage kernels with streaming output stores (single cases whose
// fields the Infinity Cache cannot keep, stream_fields)
#define SBFS_MARCH_NT_KERNEL(name, CORR)                                                         \
  __global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(SBFS_FP64_WAVES))) void name( \
      const MarchArgs* __restrict__ args) {                                                      \
    const MarchArgs& a = *args;                                                                  \
    int bx, by;                                                                                  \
    if (march_unit(a, bx, by)) march_body<double, CORR, true>(a, bx, by);                        \
  }
SBFS_MARCH_NT_KERNEL(k_march_nt, false)
SBFS_MARCH_NT_KERNEL(k_march_corr_nt, true)
#undef SBFS_MARCH_NT_KERNEL

}  // namespace sbfs
