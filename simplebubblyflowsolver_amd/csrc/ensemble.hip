// Ensembles of independent cases of uneven sizes.  Each kernel of the fused
// step runs once per stage for all cases together: a block map assigns
// every workgroup of a launch to (case, block within the case), and the
// kernel reads that case's arguments from a table.
//
// The pointers a step uses alternate between two states (vof and psi swap an
// odd number of times per step), so the argument tables of both states are
// planned once (ensemble_plan_step, fused.hip, on copies of each case's
// structs) and each state's launch sequence is captured into a HIP graph the
// second time it runs.  Per case the step is the single-case fused step with
// the projection folded and the tridiagonal algorithm the options select
// for that case (partitioned or sequential), so a case run in an ensemble
// gives the same bits as the case run alone.  The band height of the stage kernel
// is the only tile shape the ensemble chooses; sbfs_ensemble_autotune times
// the candidates on the ensemble itself.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "march.hpp"
#include "poisson.hpp"
#include "reduce.hpp"
#include "sbfs.h"
#include "sbfs_internal.hpp"

namespace sbfs {

struct CaseState {
  fluid_t f;
  interface_t s;
  CaseScratch scr;
};

static bool same_state(const CaseState& a, const CaseState& b) {
  const array_t* fa[] = {&a.f.ux, &a.f.uy, &a.f.p, &a.f.psi[0], &a.f.psi[1], &a.f.den[0], &a.f.den[1],
                         &a.f.txy, &a.f.srcux[0], &a.f.srcux[1], &a.f.srcux[2], &a.f.srcuy[0],
                         &a.f.srcuy[1], &a.f.srcuy[2], &a.s.vof, &a.s.src[0], &a.s.src[1]};
  const array_t* fb[] = {&b.f.ux, &b.f.uy, &b.f.p, &b.f.psi[0], &b.f.psi[1], &b.f.den[0], &b.f.den[1],
                         &b.f.txy, &b.f.srcux[0], &b.f.srcux[1], &b.f.srcux[2], &b.f.srcuy[0],
                         &b.f.srcuy[1], &b.f.srcuy[2], &b.s.vof, &b.s.src[0], &b.s.src[1]};
  for (int k = 0; k < 17; k++)
    if (fa[k]->data != fb[k]->data) return false;
  return a.scr.vof == b.scr.vof && a.scr.p == b.scr.p;
}

static void set_state(fluid_t* f, interface_t* s, CaseScratch* scr, const CaseState& st) {
  *f = st.f;
  *s = st.s;
  *scr = st.scr;
}

__global__ void k_gather_dt(const DtArgs* __restrict__ tab, int n, double* __restrict__ out) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < n) out[c] = *tab[c].dt;
}

}  // namespace sbfs

using namespace sbfs;

struct sbfs_ensemble {
  struct Case {
    domain_t d{};
    fluid_t f{};
    interface_t s{};
    DomainCtx* ctx = nullptr;
    double* coef = nullptr;
    Coef cf{};
    CaseScratch scr{};
    CaseState start[2];  // pointer state at the start of an even / odd step
  };
  std::vector<Case> cs;
  int band = 32;
  bool planned = false;
  int parity = 0;
  // device: argument tables [parity] and block maps
  DtArgs* dt[2] = {};
  MarchArgs* march[2] = {};  // [3 * n]
  RowArgs* rows[2] = {};     // [3 * n]
  PostArgs* post[2] = {};
  uint2* maps = nullptr;
  unsigned nb_dt = 0, nb_march = 0, nb_row = 0, nb_mode = 0, nb_post = 0, nb_part = 0, nb_conv = 0;
  unsigned off_march = 0, off_row = 0, off_mode = 0, off_post = 0, off_part = 0, off_conv = 0;
  size_t lds = 0, lds_conv = 0;
  int row_threads = 64;
  double* dts = nullptr;
  hipGraphExec_t graph[2] = {};
  int runs[2] = {};
  int graph_precision = -1;
};

namespace sbfs {

static void drop_graphs(sbfs_ensemble* e) {
  for (int p = 0; p < 2; p++) {
    if (e->graph[p]) (void)hipGraphExecDestroy(e->graph[p]);
    e->graph[p] = nullptr;
    e->runs[p] = 0;
  }
}

static void free_tables(sbfs_ensemble* e) {
  for (int p = 0; p < 2; p++) {
    (void)hipFree(e->dt[p]);
    (void)hipFree(e->march[p]);
    (void)hipFree(e->rows[p]);
    (void)hipFree(e->post[p]);
    e->dt[p] = nullptr;
    e->march[p] = nullptr;
    e->rows[p] = nullptr;
    e->post[p] = nullptr;
  }
  (void)hipFree(e->maps);
  e->maps = nullptr;
  e->planned = false;
}

// plan both pointer states from the cases' current pointers (which must be
// the state at the start of an even step) and upload tables and maps
static int plan(sbfs_ensemble* e) {
  (void)hipStreamSynchronize(stream());
  drop_graphs(e);
  free_tables(e);
  const int n = (int)e->cs.size();
  std::vector<CaseStep> steps[2];
  steps[0].resize(n);
  steps[1].resize(n);
  for (int c = 0; c < n; c++) {
    auto& k = e->cs[c];
    CaseState st{k.f, k.s, k.scr};
    k.start[0] = st;
    for (int p = 0; p < 2; p++) {
      if (p == 1) k.start[1] = st;
      if (ensemble_plan_step(&k.d, k.ctx, &st.f, &st.s, k.cf, e->band, st.scr, steps[p][c])) return 1;
    }
    SBFS_REQUIRE(same_state(st, k.start[0]), "ensemble: the step's pointer state does not return after two steps");
  }
  // block maps: dt partials, stage kernel, groups of Poisson rows, 64-mode groups, projection tiles
  std::vector<uint2> map;
  auto add = [&](int c, unsigned count) {
    for (unsigned b = 0; b < count; b++) map.push_back(make_uint2((unsigned)c, b));
  };
  for (int c = 0; c < n; c++) add(c, (unsigned)steps[0][c].dt.nblk);
  e->nb_dt = (unsigned)map.size();
  e->off_march = (unsigned)map.size();
  for (int c = 0; c < n; c++) {
    const auto& m = steps[0][c].march[0];
    add(c, cdiv(m.nx, kStrip) * cdiv(m.ny, m.band));
  }
  e->nb_march = (unsigned)map.size() - e->off_march;
  e->off_row = (unsigned)map.size();
  // row transforms: a workgroup of row_threads threads holds
  // row_threads / unit units of the case's transform length
  e->row_threads = 64;
  for (int c = 0; c < n; c++) e->row_threads = std::max(e->row_threads, wave_unit(steps[0][c].rows[0].m));
  size_t lds_rows = 0;
  for (int c = 0; c < n; c++) {
    const auto& r = steps[0][c].rows[0];
    const int units = e->row_threads / wave_unit(r.m);
    add(c, cdiv(r.ny, units * wave_rows(r.m)));
    lds_rows = std::max(lds_rows, units * wave_lds(r.m));
  }
  e->nb_row = (unsigned)map.size() - e->off_row;
  // tridiagonal: sequential cases by 64-mode groups; partitioned cases by
  // tiles of 64 modes x 4 chunks and by groups of the convolution's modes
  e->off_mode = (unsigned)map.size();
  for (int c = 0; c < n; c++)
    if (steps[0][c].rows[0].L == 0) add(c, cdiv(steps[0][c].rows[0].nx, 64));
  e->nb_mode = (unsigned)map.size() - e->off_mode;
  e->off_part = (unsigned)map.size();
  int pmax = 1;
  for (int c = 0; c < n; c++) {
    const auto& r = steps[0][c].rows[0];
    if (r.L != 0) {
      add(c, cdiv(r.nx, 64) * cdiv(r.P, 4));
      pmax = std::max(pmax, r.P);
    }
  }
  e->nb_part = (unsigned)map.size() - e->off_part;
  e->off_conv = (unsigned)map.size();
  for (int c = 0; c < n; c++) {
    const auto& r = steps[0][c].rows[0];
    if (r.L != 0) add(c, cdiv(r.nx, kConvModesBatch));
  }
  e->nb_conv = (unsigned)map.size() - e->off_conv;
  e->lds_conv = sizeof(double) * 3 * (size_t)pmax * kConvModesBatch;
  e->off_post = (unsigned)map.size();
  for (int c = 0; c < n; c++) {
    const auto& q = steps[0][c].post;
    add(c, cdiv(q.nx + 2, 64) * cdiv(q.ny, 4 * kPostRows));
  }
  e->nb_post = (unsigned)map.size() - e->off_post;
  e->lds = lds_rows;
  SBFS_HIP(hipMalloc(&e->maps, sizeof(uint2) * map.size()));
  SBFS_HIP(hipMemcpy(e->maps, map.data(), sizeof(uint2) * map.size(), hipMemcpyHostToDevice));
  for (int p = 0; p < 2; p++) {
    std::vector<DtArgs> dt(n);
    std::vector<MarchArgs> ma(3 * (size_t)n);
    std::vector<RowArgs> ro(3 * (size_t)n);
    std::vector<PostArgs> po(n);
    for (int c = 0; c < n; c++) {
      dt[c] = steps[p][c].dt;
      po[c] = steps[p][c].post;
      for (int k = 0; k < 3; k++) {
        ma[(size_t)k * n + c] = steps[p][c].march[k];
        ro[(size_t)k * n + c] = steps[p][c].rows[k];
      }
    }
    SBFS_HIP(hipMalloc(&e->dt[p], sizeof(DtArgs) * n));
    SBFS_HIP(hipMalloc(&e->march[p], sizeof(MarchArgs) * ma.size()));
    SBFS_HIP(hipMalloc(&e->rows[p], sizeof(RowArgs) * ro.size()));
    SBFS_HIP(hipMalloc(&e->post[p], sizeof(PostArgs) * n));
    SBFS_HIP(hipMemcpy(e->dt[p], dt.data(), sizeof(DtArgs) * n, hipMemcpyHostToDevice));
    SBFS_HIP(hipMemcpy(e->march[p], ma.data(), sizeof(MarchArgs) * ma.size(), hipMemcpyHostToDevice));
    SBFS_HIP(hipMemcpy(e->rows[p], ro.data(), sizeof(RowArgs) * ro.size(), hipMemcpyHostToDevice));
    SBFS_HIP(hipMemcpy(e->post[p], po.data(), sizeof(PostArgs) * n, hipMemcpyHostToDevice));
  }
  e->planned = true;
  e->parity = 0;
  return 0;
}

// the launches of one step of parity p
static int launch_step(sbfs_ensemble* e, int p) {
  const int n = (int)e->cs.size();
  const bool mixed = options().precision == SBFS_PRECISION_MIXED;
  const uint2* mp = e->maps;
  SBFS_TIMED("k_dt_partial", hipLaunchKernelGGL(k_dt_partial_batch, dim3(e->nb_dt), dim3(kRedThreads), 0, stream(),
                                                (const DtArgs*)e->dt[p], mp));
  SBFS_TIMED("k_dt_final",
             hipLaunchKernelGGL(k_dt_final_batch, dim3(n), dim3(kRedThreads), 0, stream(), (const DtArgs*)e->dt[p]));
  for (int k = 0; k < 3; k++) {
    const MarchArgs* ma = e->march[p] + (size_t)k * n;
    const RowArgs* ro = e->rows[p] + (size_t)k * n;
    const uint2* mm = mp + e->off_march;
    if (mixed && k > 0)
      SBFS_TIMED("k_march_mixed_corr",
                 hipLaunchKernelGGL(k_march_mixed_corr_batch, dim3(e->nb_march), dim3(64), 0, stream(), ma, mm));
    else if (mixed)
      SBFS_TIMED("k_march_mixed",
                 hipLaunchKernelGGL(k_march_mixed_batch, dim3(e->nb_march), dim3(64), 0, stream(), ma, mm));
    else if (k > 0)
      SBFS_TIMED("k_march_corr",
                 hipLaunchKernelGGL(k_march_corr_batch, dim3(e->nb_march), dim3(64), 0, stream(), ma, mm));
    else
      SBFS_TIMED("k_march", hipLaunchKernelGGL(k_march_batch, dim3(e->nb_march), dim3(64), 0, stream(), ma, mm));
    SBFS_TIMED("k_dct2_rows", hipLaunchKernelGGL(k_dct2_rows_batch, dim3(e->nb_row), dim3(e->row_threads), e->lds, stream(),
                                                 ro, mp + e->off_row));
    if (e->nb_mode)
      SBFS_TIMED("k_tdm_sequential", hipLaunchKernelGGL(k_tdm_sequential_batch, dim3(e->nb_mode), dim3(64), 0,
                                                        stream(), ro, mp + e->off_mode));
    if (e->nb_part) {
      SBFS_TIMED("k_part_reduce", hipLaunchKernelGGL(k_part_reduce_batch, dim3(e->nb_part), dim3(64, 4), 0,
                                                     stream(), ro, mp + e->off_part));
      SBFS_TIMED("k_part_conv", hipLaunchKernelGGL(k_part_conv_batch, dim3(e->nb_conv), dim3(256), e->lds_conv,
                                                   stream(), ro, mp + e->off_conv));
    }
    SBFS_TIMED("k_dct3_psi", hipLaunchKernelGGL(k_dct3_psi_batch, dim3(e->nb_row), dim3(e->row_threads), e->lds, stream(), ro,
                                                mp + e->off_row));
  }
  SBFS_TIMED("k_post", hipLaunchKernelGGL(k_post_batch, dim3(e->nb_post), dim3(64, 4), 0, stream(),
                                          (const PostArgs*)e->post[p], mp + e->off_post));
  return check_launch("ensemble step");
}

static int run_step(sbfs_ensemble* e) {
  const int p = e->parity;
  for (auto& k : e->cs) {
    const CaseState now{k.f, k.s, k.scr};
    SBFS_REQUIRE(same_state(now, k.start[p]),
                 "ensemble: a case's arrays were swapped outside the ensemble step");
  }
  const int prec = options().precision;
  if (prec != e->graph_precision) {
    drop_graphs(e);
    e->graph_precision = prec;
  }
  const bool graphs = !prof_active() && !(getenv("SBFS_GRAPH") && getenv("SBFS_GRAPH")[0] == '0');
  if (!graphs || e->runs[p] == 0) {
    if (launch_step(e, p)) return 1;
  } else {
    if (!e->graph[p]) {
      hipGraph_t g = nullptr;
      SBFS_HIP(hipStreamBeginCapture(stream(), hipStreamCaptureModeRelaxed));
      const int rc = launch_step(e, p);
      const hipError_t ce = hipStreamEndCapture(stream(), &g);
      if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return 1;
      }
      SBFS_HIP(ce);
      const hipError_t ie = hipGraphInstantiate(&e->graph[p], g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      SBFS_HIP(ie);
    }
    SBFS_HIP(hipGraphLaunch(e->graph[p], stream()));
  }
  e->runs[p]++;
  const int q = p ^ 1;
  for (auto& k : e->cs) set_state(&k.f, &k.s, &k.scr, k.start[q]);
  e->parity = q;
  return 0;
}

// device copies of every array of every case (for the autotuner to restore)
static int snapshot(sbfs_ensemble* e, std::vector<void*>& save, bool restore) {
  size_t idx = 0;
  for (auto& k : e->cs) {
    array_t* arrs[] = {&k.f.ux, &k.f.uy, &k.f.p, &k.f.psi[0], &k.f.psi[1], &k.f.den[0], &k.f.den[1],
                       &k.f.txy, &k.f.srcux[0], &k.f.srcux[1], &k.f.srcux[2], &k.f.srcuy[0],
                       &k.f.srcuy[1], &k.f.srcuy[2], &k.s.vof, &k.s.src[0], &k.s.src[1]};
    for (array_t* a : arrs) {
      if (!restore) {
        void* b = nullptr;
        SBFS_HIP(hipMalloc(&b, a->datasize));
        SBFS_HIP(hipMemcpyAsync(b, a->data, a->datasize, hipMemcpyDeviceToDevice, stream()));
        save.push_back(b);
      } else {
        SBFS_HIP(hipMemcpyAsync(a->data, save[idx++], a->datasize, hipMemcpyDeviceToDevice, stream()));
      }
    }
  }
  SBFS_HIP(hipStreamSynchronize(stream()));
  return 0;
}

}  // namespace sbfs

extern "C" int sbfs_ensemble_create(const sbfs_case_t* cases, int n, sbfs_ensemble_t** out) {
  SBFS_REQUIRE(cases && n > 0 && out, "sbfs_ensemble_create: bad arguments");
  *out = nullptr;
  auto* e = new sbfs_ensemble();
  e->cs.resize(n);
  int rc = 0;
  for (int c = 0; c < n && !rc; c++) {
    auto& k = e->cs[c];
    const sbfs_case_t& q = cases[c];
    if (!fp64_march_addressable(q.nx, q.ny)) {
      set_error("sbfs_ensemble_create: case %d (%zu x %zu): fields of 4 GiB or more", c, q.nx, q.ny);
      rc = 1;
      break;
    }
    rc = sbfs_domain_create(q.nx, q.ny, q.lx, q.ly, &k.d);
    if (!rc) rc = sbfs_fluid_create(&k.d, q.Re, q.Fr, q.denr, q.visr, &k.f);
    if (!rc) rc = sbfs_interface_create(&k.d, q.We, &k.s);
    if (!rc) {
      k.ctx = domain_ctx(&k.d);
      rc = k.ctx ? 0 : 1;
    }
    if (!rc) rc = coef_build(&k.d, q.Re, &k.coef, &k.cf);
    if (!rc) {
      double** bufs[] = {&k.scr.vof, &k.scr.ux[0], &k.scr.ux[1], &k.scr.uy[0], &k.scr.uy[1], &k.scr.p};
      const size_t sizes[] = {k.s.vof.datasize, k.f.ux.datasize, k.f.ux.datasize,
                              k.f.uy.datasize,  k.f.uy.datasize, k.f.p.datasize};
      for (int b = 0; b < 6 && !rc; b++) {
        if (hipMalloc(bufs[b], sizes[b]) != hipSuccess ||
            hipMemsetAsync(*bufs[b], 0, sizes[b], stream()) != hipSuccess) {
          set_error("sbfs_ensemble_create: out of device memory");
          rc = 1;
        }
      }
    }
  }
  if (rc) {
    sbfs_ensemble_destroy(e);
    return 1;
  }
  SBFS_HIP(hipMalloc(&e->dts, sizeof(double) * n));
  *out = e;
  return 0;
}

extern "C" int sbfs_ensemble_case(sbfs_ensemble_t* e, int k, domain_t** domain, fluid_t** fluid,
                                  interface_t** interface) {
  SBFS_REQUIRE(e && k >= 0 && k < (int)e->cs.size(), "sbfs_ensemble_case: bad case index %d", k);
  if (domain) *domain = &e->cs[k].d;
  if (fluid) *fluid = &e->cs[k].f;
  if (interface) *interface = &e->cs[k].s;
  return 0;
}

extern "C" int sbfs_ensemble_step(sbfs_ensemble_t* e, double* dts) {
  SBFS_REQUIRE(e, "sbfs_ensemble_step: null ensemble");
  if (!e->planned && plan(e)) return 1;
  if (run_step(e)) return 1;
  const int n = (int)e->cs.size();
  if (dts) {
    // the dt of the step just taken (the tables of its parity)
    hipLaunchKernelGGL(k_gather_dt, dim3(cdiv(n, 64)), dim3(64), 0, stream(), (const DtArgs*)e->dt[e->parity ^ 1], n,
                       e->dts);
    if (check_launch("sbfs_ensemble_step: dt")) return 1;
    SBFS_HIP(hipMemcpyAsync(dts, e->dts, sizeof(double) * n, hipMemcpyDeviceToHost, stream()));
  }
  SBFS_HIP(hipStreamSynchronize(stream()));
  return 0;
}

extern "C" int sbfs_ensemble_autotune(sbfs_ensemble_t* e, int* band) {
  SBFS_REQUIRE(e, "sbfs_ensemble_autotune: null ensemble");
  // start from an even step so the pointer state after two trial steps is
  // the one restored
  if (e->planned && e->parity != 0) {
    SBFS_REQUIRE(false, "sbfs_ensemble_autotune: call it after an even number of steps");
  }
  std::vector<void*> save;
  if (snapshot(e, save, false)) return 1;
  int ny_max = 0;
  for (auto& k : e->cs) ny_max = std::max(ny_max, (int)k.d.mysizes[1]);
  hipEvent_t t0, t1;
  SBFS_HIP(hipEventCreate(&t0));
  SBFS_HIP(hipEventCreate(&t1));
  int best = e->band;
  float best_ms = 1e30f;
  int rc = 0;
  for (int b : {8, 16, 32, 64, 128, 256}) {
    if (b > ny_max && b != 8) break;
    e->band = b;
    e->planned = false;
    rc = plan(e);
    // one untimed pair of steps, then a timed pair, each pair returning the
    // pointer state; the contents are restored after every candidate
    for (int rep = 0; rep < 2 && !rc; rep++) {
      if (rep == 1) rc = hipEventRecord(t0, stream()) != hipSuccess;
      for (int s = 0; s < 2 && !rc; s++) rc = launch_step(e, s);
      if (rep == 1 && !rc) rc = hipEventRecord(t1, stream()) != hipSuccess;
    }
    if (!rc) rc = hipEventSynchronize(t1) != hipSuccess;
    float ms = 0.f;
    if (!rc) rc = hipEventElapsedTime(&ms, t0, t1) != hipSuccess;
    if (!rc) rc = snapshot(e, save, true);
    if (rc) break;
    if (ms < best_ms) {
      best_ms = ms;
      best = b;
    }
  }
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  for (void* b : save) (void)hipFree(b);
  if (rc) {
    set_error("sbfs_ensemble_autotune: a trial step failed");
    return 1;
  }
  e->band = best;
  e->planned = false;
  if (plan(e)) return 1;
  if (band) *band = best;
  return 0;
}

extern "C" int sbfs_ensemble_set_band(sbfs_ensemble_t* e, int band) {
  SBFS_REQUIRE(e && band >= 1, "sbfs_ensemble_set_band: bad arguments");
  SBFS_REQUIRE(!e->planned || e->parity == 0, "sbfs_ensemble_set_band: call it after an even number of steps");
  e->band = band;
  e->planned = false;
  return 0;
}

extern "C" int sbfs_ensemble_destroy(sbfs_ensemble_t* e) {
  if (!e) return 0;
  (void)hipStreamSynchronize(stream());
  drop_graphs(e);
  free_tables(e);
  for (auto& k : e->cs) {
    double* bufs[] = {k.scr.vof, k.scr.ux[0], k.scr.ux[1], k.scr.uy[0], k.scr.uy[1], k.scr.p, k.coef};
    for (double* b : bufs) (void)hipFree(b);
    if (k.s.vof.data) sbfs_interface_destroy(&k.s);
    if (k.f.ux.data) sbfs_fluid_destroy(&k.f);
    if (k.d.xf) sbfs_domain_destroy(&k.d);
  }
  (void)hipFree(e->dts);
  delete e;
  return 0;
}
