#if !defined(SBFS_H)
#define SBFS_H

#include <stddef.h>
#include <stdint.h>
#include "domain.h"
#include "array.h"
#include "fluid.h"
#include "interface.h"

/*
 * Library-level calls that have no counterpart in the reference.  They let
 * a caller set up a case without NPY files, move fields between host and
 * device, read the device diagnostics and choose the solver options.
 * Every call returns 0 on success, non-zero on failure.
 */
SBFS_EXTERN_C_BEGIN

/* text of the most recent failure ("" if none) */
extern const char * sbfs_last_error(void);

/* wait for all queued device work; reports any asynchronous error */
extern int sbfs_synchronize(void);

/* name of the device the library runs on (e.g. "AMD Instinct MI355X") */
extern int sbfs_device_name(char * buf, size_t len);

/*
 * decide_dt (include/decide_dt.h) without any device allocation, for
 * callers that cannot allocate: code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc.  No counterpart in the reference.
 * The reduction runs through a fixed 66 KB device array of the library
 * instead of buffers allocated with the domain; the result is bitwise the
 * same as decide_dt's.  The domain must already have its device context
 * (fluid_init, interface_init, sbfs_domain_create or one decide_dt call
 * create it); otherwise, or for a domain wider than about 2 million
 * columns, the call fails.  All calls on one GPU share that array: calls
 * that overlap on one GPU (several host threads) race.
 */
extern int decide_dt_noalloc(const domain_t * domain, const fluid_t * fluid,
                             const interface_t * interface, double * dt);

/* safety factors of the three time-step limits (overrides the environment
 * variables coef_dt_adv, coef_dt_dif, coef_dt_int) */
extern int sbfs_set_dt_coefs(double adv, double dif, double intf);

/* integrate() execution path.  The fused FP64 step addresses each field by
 * a 32-bit byte offset: on grids whose fields reach 4 GiB (about 23000^2
 * cells) integrate() runs the staged path instead (the same results), and
 * the learned curvature, which runs on the fused path only, needs mixed
 * precision there. */
#define SBFS_PATH_STAGED 0  /* one kernel group per reference stage */
#define SBFS_PATH_FUSED  1  /* fused per-substep kernels (default) */
extern int sbfs_set_path(int path);
extern int sbfs_get_path(void);

/* gauge of the singular (x-mode 0) periodic tridiagonal system */
#define SBFS_GAUGE_PINNED    0 /* x_{Ny-1} = 0, the reference's singular
                                  branch (src/tdm.c:94) taken for mode 0 */
#define SBFS_GAUGE_REFERENCE 1 /* num/den closure with the DBL_EPSILON guard
                                  exactly as src/tdm.c:92-94 */
extern int sbfs_set_gauge(int gauge);

/* y-direction tridiagonal algorithm */
#define SBFS_TDM_SEQUENTIAL  0 /* one Thomas + Sherman-Morrison chain per
                                  x-mode, the reference's order (src/tdm.c) */
#define SBFS_TDM_PARTITIONED 1 /* chunked Thomas + reduced periodic system
                                  (default for Ny >= 64) */
extern int sbfs_set_tdm(int kind, int chunk);

/* arithmetic of the fused step's stage kernel (the staged path is always
 * FP64).  MIXED evaluates the interface stencils (gradients, normals,
 * curvature, THINC and mass fluxes) and the momentum sources (viscous
 * stresses, advection, surface tension, buoyancy) in FP32; the vof update,
 * the density-weighted velocity update, the pressure gradient, the Poisson
 * right-hand side and solve, the projection and every field of the state
 * stay FP64, so the gas volume is conserved to FP64 round-off.  The
 * momentum sources that one Runge-Kutta stage passes to the next
 * (fluid->srcux, srcuy) are FP32 values and are stored as such, in the
 * first half of each buffer; they carry no meaning after a step. */
#define SBFS_PRECISION_FP64  0
#define SBFS_PRECISION_MIXED 1
extern int sbfs_set_precision(int precision);

/* curvature of the fused step's stage kernel (the staged path always uses
 * the plain one).  LEARNED evaluates the trained network (below) on the
 * stage's input vof and takes its value in the interface band: in cells
 * whose 3 x 3 block holds a cell with vof in [0.02, 0.98], whose 5 x 5
 * stencil avoids the wall ghosts and is not a sharp interface (cells below
 * 1e-3 and above 1 - 1e-3 both present), and whose learned |kappa h| lies
 * in [1/96, 1/4].  The plain curvature (src/interface/curvature_tensor.c:
 * 100-135) stays everywhere else.  Needs uniform square cells and a trained
 * or loaded network. */
#define SBFS_CURVATURE_PLAIN   0
#define SBFS_CURVATURE_LEARNED 1
extern int sbfs_set_curvature(int mode);

/*
 * The curvature network: a 26 -> 64 -> 64 -> 1 MLP with GELU (the tanh
 * form, x (1 + tanh(sqrt(2/pi) (x + 0.044715 x^3))) / 2), BF16 weights
 * and FP32 accumulation, a learned correction of the plain curvature: its
 * inputs are a 5 x 5 vof stencil (the differences to the centre value and
 * the centre value minus 1/2) and the plain curvature kp of the centre cell
 * (the reference's corner-normal divergence), and kappa h = kp + y of its
 * output y.  One network per process.
 */
/* number of FP32 parameters (W1 64x32 with columns 26.. zero, b1, W2 64x64,
 * b2, w3, b3) */
extern int sbfs_mlp_num_params(void);
/* train from seeded Glorot weights with Adam on `nsamples` stencils the
 * device rasterises (the initial condition's tanh profile around circles
 * of radius 6 .. 64 cells, slopes 1/2 .. 1 per cell), minimising
 * mean(w (y - t)^2) with target t = label - kp and weight w = 1 / label^2;
 * `batch` a multiple of 32, `nsamples` a multiple of `batch`;
 * history[e] = that loss over epoch e */
extern int sbfs_mlp_train(uint64_t seed, size_t nsamples, int epochs, int batch,
                          double lr, double * history);
extern int sbfs_mlp_set_params(const float * params, size_t n);
extern int sbfs_mlp_get_params(float * params, size_t n);
/* training samples first .. first+n-1 of a seed: features [n][32] (BF16
 * values), labels [n] (kappa h), and optionally the targets [n], the
 * weights [n] and the raw stencils [n][25] */
extern int sbfs_mlp_samples(uint64_t seed, size_t first, size_t n,
                            float * features, float * labels, float * targets,
                            float * weights, float * stencils);
/* gradient of mean(w (y - t)^2) over n (a multiple of 32) samples at the
 * current parameters, by the training kernels (weights null: all 1);
 * *loss = that mean */
extern int sbfs_mlp_gradient(const float * features, const float * targets,
                             const float * weights, size_t n, float * grad,
                             double * loss);
/* the learned kappa h = kp + y for n stencils of 25 vof values (row-major,
 * centre 12) */
extern int sbfs_mlp_predict(const double * stencils, size_t n, double * out);
/* fraction of the interior cells of the current vof where LEARNED takes
 * the network's value */
extern int sbfs_mlp_coverage(const domain_t * domain,
                             const interface_t * interface, double * fraction);

/* uniform-x domain built as the reference's IC generator does
 * (initial_condition/main.py:15-24): xf = linspace(0, lx, Nx+1),
 * xc = [0, face midpoints, lx]; metrics as src/domain.c:98-146 */
extern int sbfs_domain_create(size_t nx, size_t ny, double lx, double ly,
                              domain_t * domain);
extern int sbfs_domain_destroy(domain_t * domain);

/* allocate every array on the device (zero-filled) and set the parameters;
 * refden = min(1, denr) as src/fluid/init.c:106 */
extern int sbfs_fluid_create(const domain_t * domain, double Re, double Fr,
                             double denr, double visr, fluid_t * fluid);
extern int sbfs_fluid_destroy(fluid_t * fluid);
extern int sbfs_interface_create(const domain_t * domain, double We,
                                 interface_t * interface);
extern int sbfs_interface_destroy(interface_t * interface);

/* whole-buffer copies between host memory and array_t.data; `bytes` must
 * equal array->datasize */
extern int sbfs_array_upload(array_t * array, const void * host, size_t bytes);
extern int sbfs_array_download(const array_t * array, void * host,
                               size_t bytes);

/*
 * Diagnostics reduced on the device in FP64 (fixed reduction order, so the
 * result is reproducible run to run):
 *   out[0] gas volume      V   = sum vof * JDXC
 *   out[1] centroid        x_c = sum vof * XC * JDXC / V
 *   out[2] rise velocity   u_r = sum vof * (UX(i)+UX(i+1))/2 * JDXC / V
 *   out[3] max |div u|         (src/logging/divergence.c:41-60)
 *   out[4] min vof, out[5] max vof, out[6] mean vof
 *   out[7] interface area      (src/logging/interface.c:120-137)
 */
#define SBFS_NDIAG 8
extern int sbfs_diagnostics(const domain_t * domain, const fluid_t * fluid,
                            const interface_t * interface,
                            double out[SBFS_NDIAG]);

/* every quantity the reference's logging writes (one file per quantity
 * under src/logging), reduced on the device in FP64 in a fixed order, plus
 * the bubble's volume, centroid and rise velocity (which the reference does
 * not log).  Uses
 * den[1] and txy as the last Runge-Kutta stage of integrate() leaves them. */
#define SBFS_LOG_DIVMAX     0  /* max |div u|               divergence.c:41-60 */
#define SBFS_LOG_MOMX       1  /* total x momentum          momentum.c:31-40   */
#define SBFS_LOG_MOMY       2  /* total y momentum          momentum.c:42-50   */
#define SBFS_LOG_EXTX       3  /* max |ux|                  extrema.c:25-31    */
#define SBFS_LOG_EXTY       4  /* max |uy|                  extrema.c:33-38    */
#define SBFS_LOG_ENGX       5  /* x kinetic energy          energy.c:31-40     */
#define SBFS_LOG_ENGY       6  /* y kinetic energy          energy.c:42-50     */
#define SBFS_LOG_STRESS_XM  7  /* wall shear stress, x = 0  stress.c:28-36     */
#define SBFS_LOG_STRESS_XP  8  /* wall shear stress, x = lx stress.c:28-36     */
#define SBFS_LOG_VOFMIN     9  /* min vof                   interface.c:104-117 */
#define SBFS_LOG_VOFMAX    10  /* max vof                   interface.c:104-117 */
#define SBFS_LOG_VOFMEAN   11  /* volume-weighted mean vof  interface.c:104-117 */
#define SBFS_LOG_AREA      12  /* interface area            interface.c:119-137 */
#define SBFS_LOG_VOLUME    13  /* gas volume sum vof J                          */
#define SBFS_LOG_CENTROID  14  /* gas centroid in x                             */
#define SBFS_LOG_RISE      15  /* gas-volume mean of ux (+x is up)              */
#define SBFS_NLOG 16
extern int sbfs_logging_values(const domain_t * domain, const fluid_t * fluid,
                               const interface_t * interface,
                               double out[SBFS_NLOG]);

/*
 * Ensembles of independent cases of uneven sizes (no counterpart in the
 * reference, which runs one case per process).  Every kernel of the fused
 * step runs once per Runge-Kutta stage for all cases together, and each of
 * the two pointer states a step alternates between is replayed as a HIP
 * graph.  Per case the step is integrate()'s fused step (projection folded
 * into the next stage, the tridiagonal algorithm and precision the options
 * select when the ensemble is first stepped, plain curvature): a case gives
 * the same bits as when run alone with the same options.
 * Cases are set up through their structs (sbfs_ensemble_case) with the
 * usual calls (sbfs_array_upload, the boundary routines).  Between ensemble
 * steps nothing else may swap a case's arrays (integrate() on a case does).
 */
typedef struct sbfs_ensemble sbfs_ensemble_t;
typedef struct {
  size_t nx, ny;                     /* grid (nx a power of two <= 4096) */
  double lx, ly, Re, Fr, denr, visr, We;
} sbfs_case_t;
extern int sbfs_ensemble_create(const sbfs_case_t * cases, int n,
                                sbfs_ensemble_t ** ensemble);
extern int sbfs_ensemble_case(sbfs_ensemble_t * ensemble, int k,
                              domain_t ** domain, fluid_t ** fluid,
                              interface_t ** interface);
/* one step of every case; dts[k] (may be null) receives case k's dt */
extern int sbfs_ensemble_step(sbfs_ensemble_t * ensemble, double * dts);
/* band height of the stage kernel (rows per wavefront; default 32):
 * autotune times every candidate on the ensemble's current state, restores
 * it and keeps the fastest; both need an even number of steps taken */
extern int sbfs_ensemble_autotune(sbfs_ensemble_t * ensemble, int * band);
extern int sbfs_ensemble_set_band(sbfs_ensemble_t * ensemble, int band);
extern int sbfs_ensemble_destroy(sbfs_ensemble_t * ensemble);

/* step counter and time of an NPY directory (step.npy '<u8', time.npy
 * '<f8', both 0-d), as src/main.c:63-70 reads and :28-29 writes them */
extern int sbfs_read_step_time(const char dirname[], size_t * step,
                               double * time);
extern int sbfs_write_step_time(const char dirname[], size_t step,
                                double time);

/* temporal statistics (src/statistics.c): collect() adds ux, uy and vof to
 * device-resident running sums ux1 ([1:Nx+1] x [1:Ny]), uy1 and vof1
 * ([0:Nx+1] x [1:Ny]) and counts the samples; save() writes num.npy
 * ('<u8'), ux1.npy, uy1.npy and vof1.npy into an existing directory, as
 * statistics.output does (:170-214).  One set of sums per process. */
extern int sbfs_statistics_collect(const domain_t * domain, const fluid_t * fluid,
                                   const interface_t * interface);
extern int sbfs_statistics_num(size_t * num);
extern int sbfs_statistics_save(const char dirname[], const domain_t * domain);

/* kernel timing: while on, every launch is bracketed by HIP events on the
 * library stream; stop() synchronises and aggregates per kernel name */
extern int sbfs_profile_start(void);
extern int sbfs_profile_stop(void);
extern int sbfs_profile_count(void);
extern int sbfs_profile_entry(int k, char * name, size_t len, int * count,
                              double * total_ms);

/* the Poisson solver's row transforms on their own: FFTW REDFT10
 * (inverse = 0) or REDFT01 (inverse = 1) of `howmany` rows of n doubles,
 * n a power of two <= 4096; host buffers in and out */
extern int sbfs_dct_rows(int inverse, size_t n, size_t howmany,
                         const double * in, double * out);

/* device copy bandwidth: `reps` launches of a plain copy kernel (k_copy)
 * over two buffers of `bytes` each (a multiple of 16); *gbs = bytes read +
 * written per second / 1e9.  Allocates and frees its own buffers. */
extern int sbfs_copy_bandwidth(size_t bytes, int reps, double * gbs);

SBFS_EXTERN_C_END

#endif /* SBFS_H */
