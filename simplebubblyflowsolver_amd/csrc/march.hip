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
// shifts, so nothing is recomputed across rows and no LDS or barrier is
// needed.  Every value is formed with the reference's operation order (the
// per-column quotients come from the coefficient tables), which keeps the
// result bitwise equal to the staged kernels.
#include <cfloat>
#include <cmath>

#include "march.hpp"
#include "sbfs_internal.hpp"
#include "stencils.hpp"

namespace sbfs {

// value of the neighbouring lane (l-1 / l+1) through DPP wave shifts
__device__ __forceinline__ double from_left(double v) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  int lo = (int)(unsigned)(b & 0xffffffffull), hi = (int)(unsigned)(b >> 32);
  lo = __builtin_amdgcn_update_dpp(0, lo, 0x138, 0xf, 0xf, false);  // wave_shr:1
  hi = __builtin_amdgcn_update_dpp(0, hi, 0x138, 0xf, 0xf, false);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

__device__ __forceinline__ double from_right(double v) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  int lo = (int)(unsigned)(b & 0xffffffffull), hi = (int)(unsigned)(b >> 32);
  lo = __builtin_amdgcn_update_dpp(0, lo, 0x130, 0xf, 0xf, false);  // wave_shl:1
  hi = __builtin_amdgcn_update_dpp(0, hi, 0x130, 0xf, 0xf, false);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

__device__ __forceinline__ int wrap_r(int j, int ny) {
  while (j < 1) j += ny;
  while (j > ny) j -= ny;
  return j;
}

enum { T_RHF, T_RHC, T_RJF, T_RJC, T_QF, T_QC, T_EQF, T_ECY, T_CYC, T_CYF, T_DFX, T_DCX, T_HXXF, T_HXXC, T_N };

__global__ __launch_bounds__(64) void k_march(MarchArgs a) {
  __shared__ double st[T_N][64];
  const int lane = threadIdx.x;
  const int nx = a.nx, ny = a.ny;
  const int i0 = 1 + blockIdx.x * kStrip;
  const int c = i0 - 3 + lane;
  const int jb = 1 + blockIdx.y * a.band;
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
  __syncthreads();
  const int ll = lane > 0 ? lane - 1 : 0, lr = lane < 63 ? lane + 1 : 63;
  auto TB = [&](int t) { return st[t][lane]; };
  auto TL = [&](int t) { return st[t][ll]; };
  auto TR = [&](int t) { return st[t][lr]; };
  const double hy = a.hy, rhy = a.cf.rhy;
  const double dt = *a.dt;
  const double dtca = dt * a.ca, dtcb = dt * a.cb;
  const double dt_old = a.g_old * dt, dt_new = a.g_new * dt;
  const double c0 = 1. / a.norm / dt_new * a.refden;
  const double c1 = 1. / a.norm * -1. / dt_new * dt_old;
  const bool has_b = a.stage != 0;

  auto ld_vof = [&](int j) { return c_vof ? a.vof[L::vof(c, wrap_r(j, ny), nx)] : 0.; };
  auto ld_ux = [&](int j) { return c_fx ? a.ux[L::xf(c, wrap_r(j, ny), nx)] : 0.; };
  auto ld_cc = [&](const double* f, int j) { return c_vof ? f[L::cc(c, wrap_r(j, ny), nx)] : 0.; };
  auto DEN = [&](double v) { return mix_clamp(a.denr, a.dlo, a.dhi, v); };
  // viscosity at x face c from the vof on its two sides (viscosity.c:37-53)
  auto VISUX = [&](int col, double vl, double vr) {
    double v;
    if (col == 1) v = vl;
    else if (col == nx + 1) v = vr;
    else v = 0.5 * vl + 0.5 * vr;
    return mix_clamp(a.visr, a.vlo, a.vhi, v);
  };
  auto VISUY = [&](double vb, double vt) { return mix_clamp(a.visr, a.vlo, a.vhi, 0.5 * vb + 0.5 * vt); };

  // rings (row offsets relative to the current step s)
  double vm1 = 0., v0 = 0., vp1 = 0., vp2 = 0.;             // vof s-1..s+2
  double g0x = 0., g0y = 0., g1x = 0., g1y = 0., g2x = 0., g2y = 0.;  // corners s..s+2
  double g0xr = 0., g0yr = 0., g1xr = 0., g1yr = 0.;         // corners of c+1, rows s, s+1
  double n0x = 0., n0y = 0., n0z = 0., n1x = 0., n1y = 0., n1z = 0.;  // normals s, s+1
  double km1 = 0., k0 = 0.;                                 // curvature s-1, s
  double fx0 = 0., mfxm1 = 0., mfx0 = 0.;                   // x fluxes s-1, s
  double fy0 = 0., fyp1 = 0., mfym1 = 0., mfy0 = 0., mfyp1 = 0.;  // y faces s-1..s+1
  double txx0 = 0., tyym1 = 0., tyy0 = 0., txy0 = 0., txy1 = 0.;
  double vnm2 = 0., vnm1 = 0., vn0 = 0.;                    // new vof s-2..s
  double uxm1 = 0., ux0 = 0., uxp1 = 0., uym1 = 0., uy0 = 0., uyp1 = 0.;
  double pm1 = 0., p0 = 0.;
  double psm2 = 0., psm1 = 0., ps0 = 0.;                    // psi[0] s-2..s
  double uxs_m1 = 0., uys_m1 = 0., uys0 = 0., uxs0 = 0.;    // predicted u rows s-1, s
  const int s_first = jb - kWarm;
  // prefetched inputs of the next step
  double nv = ld_vof(s_first + 2), nux = ld_ux(s_first + 1), nuy = ld_cc(a.uy, s_first + 1);
  double np = ld_cc(a.p, s_first), nps = ld_cc(a.psi0, s_first);
  for (int s = s_first; s <= je + 1; s++) {
    // -- new input rows -----------------------------------------------------
    vm1 = v0; v0 = vp1; vp1 = vp2; vp2 = nv;
    uxm1 = ux0; ux0 = uxp1; uxp1 = nux;
    uym1 = uy0; uy0 = uyp1; uyp1 = nuy;
    pm1 = p0; p0 = np;
    psm2 = psm1; psm1 = ps0; ps0 = nps;
    const bool outrow = s >= jb && s <= je;
    double sbv = 0., sbx = 0., sby = 0.;
    if (has_b && c_in) {
      // the new vof of every marched row needs its beta source; the
      // predicted uy of row je+1 enters the Poisson rhs of row je
      sbv = a.srcb_vof[L::ci(c, wrap_r(s, ny), nx)];
      if (s >= jb) sby = a.srcb_uy[L::ci(c, wrap_r(s, ny), nx)];
      if (outrow && c >= 2) sbx = a.srcb_ux[L::fx(c, s, nx)];
    }
    if (s + 1 <= je + 1) {
      nv = ld_vof(s + 3);
      nux = ld_ux(s + 2);
      nuy = ld_cc(a.uy, s + 2);
      np = ld_cc(a.p, s + 1);
      nps = ld_cc(a.psi0, s + 1);
    }
    // -- corner gradients, row s+2 ---------------------------------------------
    g0x = g1x; g0y = g1y; g1x = g2x; g1y = g2y;
    g0xr = g1xr; g0yr = g1yr;
    {
      const double mm = from_left(vp1), pm = vp1, mp = from_left(vp2), pp = vp2;
      double gx = 0., gy = 0.;
      if (c_fx && !(mm == pm && mm == mp && mm == pp)) {
        gx = TB(T_RHF) * (-mm + pm - mp + pp);
        gy = rhy * (-mm - pm + mp + pp);
        const double norm = sqrt(gx * gx + gy * gy);
        const double ninv = 1. / fmax(norm, DBL_EPSILON);
        gx = gx * ninv;
        gy = gy * ninv;
      }
      g2x = gx;
      g2y = gy;
    }
    const double g2xr = from_right(g2x), g2yr = from_right(g2y);
    g1xr = from_right(g1x);
    g1yr = from_right(g1y);
    // -- normals, row s+1 -------------------------------------------------------
    n0x = n1x; n0y = n1y; n0z = n1z;
    {
      double nxv = 0., nyv = 0., nz = 0.;
      if (c_in && in_band(vp1)) {
        nxv = g1x + g1xr + g2x + g2xr;
        nyv = g1y + g1yr + g2y + g2yr;
        nxv /= TB(T_HXXC);
        nyv /= hy;
        const double norm = sqrt(nxv * nxv + nyv * nyv);
        const double ninv = 1. / fmax(norm, DBL_EPSILON);
        nxv *= ninv;
        nyv *= ninv;
        nz = -0.5 / kVofBeta * log(1. / vp1 - 1.);
      }
      n1x = nxv;
      n1y = nyv;
      n1z = nz;
    }
    // -- curvature, row s -----------------------------------------------------
    km1 = k0;
    {
      double kk = 0.;
      if (c_in) {
        const double nxm = 0.5 * g0x + 0.5 * g1x;
        const double nxp = 0.5 * g0xr + 0.5 * g1xr;
        const double nym = 0.5 * g0y + 0.5 * g0yr;
        const double nyp = 0.5 * g1y + 0.5 * g1yr;
        const double div = TB(T_RJC) * (-TB(T_QF) * nxm + TR(T_QF) * nxp - TB(T_CYC) * nym + TB(T_CYC) * nyp);
        kk = -1. * div;
      }
      k0 = kk;
    }
    // -- x fluxes at face c, row s ----------------------------------------------
    mfxm1 = mfx0;
    {
      const double vl = from_left(v0);
      const double nlx = from_left(n0x), nly = from_left(n0y), nlz = from_left(n0z);
      double f = 0., mf = 0.;
      if (c_fx) {
        const double vel = ux0;
        if (c != 1 && c != nx + 1) {
          const bool up = vel < 0.;
          const double v = up ? v0 : vl;
          const V3 n = up ? V3{n0x, n0y, n0z} : V3{nlx, nly, nlz};
          f = thinc_flux_x(vel, v, in_band(v) ? n : V3{0., 0., 0.});
        }
        mf = vel + (a.denr - 1.) * f;
      }
      fx0 = f;
      mfx0 = mf;
    }
    // -- y fluxes at face s+1 (between rows s and s+1) ----------------------------
    fy0 = fyp1;
    mfym1 = mfy0; mfy0 = mfyp1;
    {
      double f = 0., mf = 0.;
      if (c_vof) {
        const double vel = uyp1;
        if (c != 0 && c != nx + 1) {
          const bool up = vel < 0.;
          const double v = up ? vp1 : v0;
          const V3 n = up ? V3{n1x, n1y, n1z} : V3{n0x, n0y, n0z};
          f = thinc_flux_y(vel, v, in_band(v) ? n : V3{0., 0., 0.});
        }
        mf = vel + (a.denr - 1.) * f;
      }
      fyp1 = f;
      mfyp1 = mf;
    }
    // -- stress tensor ----------------------------------------------------------
    // (every DPP shift runs in uniform control flow: an inactive source lane
    // would not deliver its value)
    const double v0l = from_left(v0), v0r = from_right(v0), vp1l = from_left(vp1);
    const double ux0l = from_left(ux0), ux0r = from_right(ux0);
    const double uyp1l = from_left(uyp1);
    {
      double t = 0.;
      if (c_in) {
        const double vx = 0.5 * VISUX(c, v0l, v0) + 0.5 * VISUX(c + 1, v0, v0r);
        const double l = TB(T_RHC) * (-ux0 + ux0r);
        t = vx * l + vx * l;
      }
      txx0 = t;
    }
    tyym1 = tyy0;
    {
      double t = 0.;
      if (c_in) {
        const double vy = 0.5 * VISUY(vm1, v0) + 0.5 * VISUY(v0, vp1);
        const double l = rhy * (-uy0 + uyp1);
        t = vy * l + vy * l;
      }
      tyy0 = t;
    }
    txy0 = txy1;
    {
      double t = 0.;
      if (c_fx) {
        const double vx = 0.5 * VISUX(c, v0l, v0) + 0.5 * VISUX(c, vp1l, vp1);
        const double vy = 0.5 * VISUY(v0l, vp1l) + 0.5 * VISUY(v0, vp1);
        const double lxy = rhy * (-ux0 + uxp1);
        const double lyx = TB(T_RHF) * (-uyp1l + uyp1);
        t = vx * lxy + vy * lyx;
      }
      txy1 = t;
    }
    // -- new vof, row s ---------------------------------------------------------
    vnm2 = vnm1; vnm1 = vn0;
    double src = 0.;
    const double fx0r = from_right(fx0);
    {
      double v = 0.;
      if (c_in) {
        src = TB(T_RJC) * (+TB(T_QF) * fx0 - TR(T_QF) * fx0r + TB(T_CYC) * fy0 - TB(T_CYC) * fyp1);
        v = v0;
        v += dtca * src;
        if (has_b) v += dtcb * sbv;
      }
      vn0 = v;
    }
    // -- predicted velocities, row s ------------------------------------------
    const double mfx0l = from_left(mfx0), mfx0r = from_right(mfx0);
    const double mfy0l = from_left(mfy0), mfyp1l = from_left(mfyp1);
    const double txx0l = from_left(txx0), p0l = from_left(p0), k0l = from_left(k0);
    const double vn0l = from_left(vn0);
    const double mfxm1r = from_right(mfxm1);
    const double uy0l = from_left(uy0), uy0r = from_right(uy0), txy0r = from_right(txy0);
    double uxs = 0.;
    if (c >= 2 && c <= nx) {
      double s_ = 0.;
      {
        const double mm = TL(T_EQF) * mfx0l + TB(T_EQF) * mfx0;
        const double mp = TB(T_EQF) * mfx0 + TR(T_EQF) * mfx0r;
        const double um = 0.5 * ux0l + 0.5 * ux0;
        const double up = 0.5 * ux0 + 0.5 * ux0r;
        s_ -= TB(T_RJF) * (-mm * um + mp * up);
      }
      {
        const double mm = TL(T_ECY) * mfy0l + TB(T_ECY) * mfy0;
        const double mp = TL(T_ECY) * mfyp1l + TB(T_ECY) * mfyp1;
        const double um = 0.5 * uxm1 + 0.5 * ux0;
        const double up = 0.5 * ux0 + 0.5 * uxp1;
        s_ -= TB(T_RJF) * (-mm * um + mp * up);
      }
      s_ += TB(T_DFX) * (-TL(T_QC) * txx0l + TB(T_QC) * txx0);
      s_ += TB(T_DFX) * (-TB(T_CYF) * txy0 + TB(T_CYF) * txy1);
      double g = 0.;
      g -= TB(T_RHF) * (-p0l + p0);
      {
        const double d = 0.5 * DEN(v0l) + 0.5 * DEN(v0);
        const double k = 0.5 * k0l + 0.5 * k0;
        const double delta = TB(T_RHF) * (-v0l + v0);
        s_ += d * a.refdeninv * a.tension * k * delta;
      }
      s_ += a.grav * (0.5 * DEN(v0l) + 0.5 * DEN(v0));
      if (a.srca_ux && outrow && c_out) a.srca_ux[L::fx(c, s, nx)] = s_;
      const double d0 = 0.5 * DEN(v0l) + 0.5 * DEN(v0);
      uxs = d0 * ux0 + a.ca * dt * s_ + a.cb * dt * sbx + a.cg * dt * g;
      const double d1 = 0.5 * DEN(vn0l) + 0.5 * DEN(vn0);
      uxs /= d1;
    }
    double uys = 0.;
    if (c_in) {
      double s_ = 0.;
      {
        const double mm = TB(T_EQF) * mfxm1 + TB(T_EQF) * mfx0;
        const double mp = TR(T_EQF) * mfxm1r + TR(T_EQF) * mfx0r;
        const double um = 0.5 * uy0l + 0.5 * uy0;
        const double up = 0.5 * uy0 + 0.5 * uy0r;
        s_ -= TB(T_RJC) * (-mm * um + mp * up);
      }
      {
        const double mm = TB(T_ECY) * mfym1 + TB(T_ECY) * mfy0;
        const double mp = TB(T_ECY) * mfy0 + TB(T_ECY) * mfyp1;
        const double um = 0.5 * uym1 + 0.5 * uy0;
        const double up = 0.5 * uy0 + 0.5 * uyp1;
        s_ -= TB(T_RJC) * (-mm * um + mp * up);
      }
      s_ += TB(T_DCX) * (-TB(T_QF) * txy0 + TR(T_QF) * txy0r);
      s_ += TB(T_DCX) * (-TB(T_CYC) * tyym1 + TB(T_CYC) * tyy0);
      double g = 0.;
      g -= rhy * (-pm1 + p0);
      {
        const double d = 0.5 * DEN(vm1) + 0.5 * DEN(v0);
        const double k = 0.5 * km1 + 0.5 * k0;
        const double delta = rhy * (-vm1 + v0);
        s_ += d * a.refdeninv * a.tension * k * delta;
      }
      if (a.srca_uy && outrow && c_out) a.srca_uy[L::ci(c, s, nx)] = s_;
      const double d0 = 0.5 * DEN(vm1) + 0.5 * DEN(v0);
      uys = d0 * uy0 + a.ca * dt * s_ + a.cb * dt * sby + a.cg * dt * g;
      const double d1 = 0.5 * DEN(vnm1) + 0.5 * DEN(vn0);
      uys /= d1;
    }
    uxs_m1 = uxs0; uxs0 = uxs;
    uys_m1 = uys0; uys0 = uys;
    // -- outputs of row s -------------------------------------------------------
    if (outrow && c_out) {
      auto ivof = [nx](int p, int q) { return L::vof(p, q, nx); };
      auto icc = [nx](int p, int q) { return L::cc(p, q, nx); };
      auto ixf = [nx](int p, int q) { return L::xf(p, q, nx); };
      store_with_halo(a.vofn, c, s, ny, 2, vn0, ivof);
      if (c == 1) store_with_halo(a.vofn, 0, s, ny, 2, 0., ivof);
      if (c == nx) store_with_halo(a.vofn, nx + 1, s, ny, 2, 0., ivof);
      if (a.srca_vof) a.srca_vof[L::ci(c, s, nx)] = src;
      store_with_halo(a.uxs, c, s, ny, 1, c == 1 ? 0. : uxs, ixf);
      if (c == nx) store_with_halo(a.uxs, nx + 1, s, ny, 1, 0., ixf);
      store_with_halo(a.uys, c, s, ny, 1, uys, icc);
      if (c == 1) store_with_halo(a.uys, 0, s, ny, 1, kUyXm, icc);
      if (c == nx) store_with_halo(a.uys, nx + 1, s, ny, 1, kUyXp, icc);
      if (a.den1) {
        const double dg = DEN(0.);
        store_with_halo(a.den1, c, s, ny, 1, DEN(vn0), icc);
        if (c == 1) store_with_halo(a.den1, 0, s, ny, 1, dg, icc);
        if (c == nx) store_with_halo(a.den1, nx + 1, s, ny, 1, dg, icc);
      }
      if (a.txy) {
        store_with_halo(a.txy, c, s, ny, 1, txy0, ixf);
        if (c == nx) store_with_halo(a.txy, nx + 1, s, ny, 1, txy0r, ixf);
      }
    }
    // -- Poisson right-hand side of row s-1 (compute_potential.c:315-360) ----
    {
      const double uxsr = from_right(uxs_m1);
      const double vnl = from_left(vnm1), vnr = from_right(vnm1);
      const double psl = from_left(psm1), psr = from_right(psm1);
      const int r = s - 1;
      if (r >= jb && r <= je && c_out) {
        const double div = TB(T_RJC) * (-TB(T_QF) * uxs_m1 + TR(T_QF) * uxsr - TB(T_CYC) * uys_m1 + TB(T_CYC) * uys0);
        const double d00 = DEN(vnm1);
        const double dxm = 0.5 * DEN(vnl) + 0.5 * d00;
        const double dxp = 0.5 * d00 + 0.5 * DEN(vnr);
        const double dym = 0.5 * DEN(vnm2) + 0.5 * d00;
        const double dyp = 0.5 * d00 + 0.5 * DEN(vn0);
        const double pxm = -psl + psm1;
        const double pxp = -psm1 + psr;
        const double pym = -psm2 + psm1;
        const double pyp = -psm1 + ps0;
        const double gxm = (a.refden / dxm - 1.) * pxm / TB(T_HXXF);
        const double gxp = (a.refden / dxp - 1.) * pxp / TR(T_HXXF);
        const double gym = (a.refden / dym - 1.) * pym / hy;
        const double gyp = (a.refden / dyp - 1.) * pyp / hy;
        const double add = TB(T_RJC) * (-TB(T_QF) * gxm + TR(T_QF) * gxp - TB(T_CYC) * gym + TB(T_CYC) * gyp);
        a.rhs[(size_t)(r - 1) * nx + (c - 1)] = c0 * div + c1 * add;
      }
    }
  }
}

}  // namespace sbfs
