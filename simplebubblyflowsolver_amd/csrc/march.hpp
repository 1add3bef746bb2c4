// Arguments of the marching stage kernel (march.hip).
#pragma once

#include "sbfs_internal.hpp"
#include "poisson.hpp"

namespace sbfs {

// output columns per wavefront: 58 + 3 halo columns on each side = 64 lanes.
// The right halo needs 3: the Poisson rhs of the last output cell reads the
// predicted ux of the face to its right, whose mass flux upwinds into the
// next cell, whose normal needs the corner gradient one column further.
constexpr int kStrip = 58;
// Rows marched before the first output row jb of a band.  The vof ring
// holds rows s_first .. s_first + 2 from the first step on (two rows are
// loaded before the loop), so the first valid corner gradients are those of
// row s_first + 2 and the first valid normals those of row s_first + 2.  The
// y flux into row jb - 1 needs the normals of row jb - 2, hence s_first <=
// jb - 4.  With the folded projection, ux and uy are valid from row and face
// s_first + 2, which the first output row needs from jb - 1 on.  (Without
// the two rows loaded ahead, 4 rows of fill broke the fused path's bitwise
// match with the staged one.)
constexpr int kWarm = 4;
constexpr int kWarmCorr = 4;

// per-column quotients of the metrics (see fused.hip)
struct Coef {
  const double* rhf;
  const double* rhc;
  const double* rjf;
  const double* rjc;
  const double* qf;
  const double* qc;
  const double* eqf;
  const double* ecy;
  const double* cyc;
  const double* cyf;
  const double* dfx;
  const double* dcx;
  const double* hxxf1;  // HXXF(i), 1-based
  const double* hxxc1;  // HXXC(i), 1-based
  double rhy;
};

struct MarchArgs {
  const double* vof;
  const double* ux;
  const double* uy;
  const double* p;
  const double* psi0;
  // Runge-Kutta sources: beta (the previous stage's, read) and alpha (this
  // stage's, written).  The FP32 bodies keep the momentum sources (ux, uy)
  // as float in the buffers' first halves; the vof source is always FP64.
  const double* srcb_vof;
  const double* srcb_ux;
  const double* srcb_uy;
  double* vofn;
  double* uxs;
  double* uys;
  double* rhs;
  double* srca_vof;
  double* srca_ux;
  double* srca_uy;
  double* den1;
  double* txy;
  const double* hxxf1;
  const double* hxxc1;
  Coef cf;
  double hy;
  double denr, dlo, dhi, visr, vlo, vhi;
  double refdeninv, tension, grav, refden, norm;
  const double* dt;
  double ca, cb, cg, g_old, g_new;
  // projection of the previous stage, folded into this one (stages 1, 2):
  // ux / uy then hold the previous stage's predicted velocity, p its pressure
  // before the update, psi0 its new potential and psi_old the potential it
  // started from; the updated pressure of the rows this wavefront owns is
  // written to p_out
  // p_out is null in stage 1 of a folded step, whose updated pressure is not
  // stored: stage 2 reads the step's input pressure and, with p_twice set,
  // forms it again as p + psi_old (the same operands in the same order)
  // before adding psi0
  const double* psi_old;
  double* p_out;
  int p_twice;
  double pg_old, pg_new;  // Runge-Kutta gamma of the previous stage's dt_old / dt_new
  // learned curvature of the stage's input vof (interior cells, (j-1) Nx +
  // (i-1)); NaN where the plain curvature stays; null: plain everywhere.
  // kappa_n: the gate's count of listed cells (0: no cell learned, every
  // value is NaN, and the kernel skips the kappa reads)
  const float* kappa;
  const unsigned* kappa_n;
  // cut-cell bitmask of the new vof for the next stage's learned curvature
  // (null: not written): word (j-1) strips + bx, bit t for column
  // 1 + bx kStrip + t, set where the vof lies in [kCutLo, kCutHi]
  unsigned long long* cut;
  // dispatch order of the (strip, band) units, unit = band * strips + strip
  // (null: the 2D grid's own order), and each unit's count of marched rows
  // with interface cells, written for the next order (null: not counted)
  const unsigned* order;
  unsigned* weight;
  int stage;
  int nx, ny, band;
};

// units of a stage kernel launch, and the largest weight one can have
__host__ __device__ inline int march_strips(int nx) { return (nx + kStrip - 1) / kStrip; }
__host__ __device__ inline int march_units(int nx, int ny, int band) {
  return march_strips(nx) * ((ny + band - 1) / band);
}
__host__ __device__ inline int march_max_weight(int band) { return band + kWarmCorr + 2; }
// Orders the units by weight, heaviest first (a counting sort over 32
// weight buckets, one workgroup of 1024 threads): the dispatcher starts the
// units with the most interface rows first and fills in behind them.  Any
// weights give a permutation of 0..n-1.
__global__ void k_march_order(const unsigned* weight, unsigned* order, int n, int max_weight);

// projection of u and the pressure update after the last stage (k_post):
// a thread takes kPostRows consecutive rows of one column, a workgroup
// 64 columns and 4 kPostRows rows (4096 x 8192, us per launch: 1 row 496,
// 2 rows 487-498, 4 rows 506-510, 8 rows 523-529; one cell per thread
// without the shared quotients 558)
constexpr int kPostRows = 2;
struct PostArgs {
  const double* uxs;
  const double* uys;
  const double* psi1;
  const double* psi0;
  const double* vofn;
  double* ux;
  double* uy;
  double* p;
  const double* hxxf;
  double hy;
  double denr, dlo, dhi, refden;
  const double* dt;
  double g_new, g_old;
  int nx, ny;
};

// an ensemble case's step (fused.hip plans it, ensemble.hip launches it)
struct CaseScratch {
  double* vof;
  double* ux[2];
  double* uy[2];
  double* p;
};
struct CaseStep {
  DtArgs dt;
  MarchArgs march[3];
  RowArgs rows[3];
  PostArgs post;
};
int coef_build(const domain_t* d, double Re, double** buf, Coef* out);
int ensemble_plan_step(const domain_t* domain, DomainCtx* c, fluid_t* fluid, interface_t* interface,
                       const Coef& cf, int band, CaseScratch& scr, CaseStep& out);
__global__ void k_march_batch(const MarchArgs* tab, const uint2* map);
__global__ void k_march_corr_batch(const MarchArgs* tab, const uint2* map);
__global__ void k_march_mixed_batch(const MarchArgs* tab, const uint2* map);
__global__ void k_march_mixed_corr_batch(const MarchArgs* tab, const uint2* map);
__global__ void k_post_batch(const PostArgs* tab, const uint2* map);

}  // namespace sbfs
