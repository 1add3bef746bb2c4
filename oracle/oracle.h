#if !defined(SBFS_ORACLE_H)
#define SBFS_ORACLE_H

/*
 * TEST INFRASTRUCTURE ONLY.  CPU FP64 restatement of the reference's
 * two-phase time step, used to check the HIP library.  Only tests,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 *
 * The oracle works on the same structs as the library (include/ headers) but
 * every array_t.data is HOST memory (calloc).
 */
#include <stddef.h>
#include "domain.h"
#include "array.h"
#include "fluid.h"
#include "interface.h"

/* Poisson options */
#define ORACLE_DCT_DIRECT 0 /* O(N^2) cosine sums of the REDFT10/REDFT01 definitions */
#define ORACLE_DCT_FFT    1 /* N/2-point complex FFT, same algorithm as the GPU */
#define ORACLE_TDM_SEQUENTIAL  0 /* src/tdm.c order */
#define ORACLE_TDM_PARTITIONED 1 /* chunked, same algorithm as the GPU */
#define ORACLE_GAUGE_PINNED    0
#define ORACLE_GAUGE_REFERENCE 1

int oracle_set_poisson(int dct_kind, int tdm_kind, int chunk, int gauge);
int oracle_set_dt_coefs(double adv, double dif, double intf);
int oracle_set_threads(int nthreads);

int oracle_domain_create(size_t nx, size_t ny, double lx, double ly,
                         domain_t * domain);
int oracle_domain_destroy(domain_t * domain);
int oracle_fluid_create(const domain_t * domain, double Re, double Fr,
                        double denr, double visr, fluid_t * fluid);
int oracle_fluid_destroy(fluid_t * fluid);
int oracle_interface_create(const domain_t * domain, double We,
                            interface_t * interface);
int oracle_interface_destroy(interface_t * interface);

/* stages (reference names with an oracle_ prefix) */
int oracle_decide_dt(const domain_t *, const fluid_t *, const interface_t *,
                     double * dt);
int oracle_fluid_compute_density(const domain_t *, fluid_t *,
                                 const interface_t *, size_t index);
int oracle_fluid_compute_viscosity(const domain_t *, fluid_t *,
                                   const interface_t *);
int oracle_interface_compute_curvature_tensor(const domain_t *, interface_t *);
int oracle_interface_compute_force(const domain_t *, const fluid_t *,
                                   interface_t *);
int oracle_interface_update_vof(const domain_t *, size_t rkstep, double dt,
                                const fluid_t *, interface_t *);
int oracle_interface_compute_mass_flux(const domain_t *, const fluid_t *,
                                       interface_t *);
int oracle_fluid_predict_field(const domain_t *, size_t rkstep, double dt,
                               fluid_t *, const interface_t *);
int oracle_fluid_compute_potential(const domain_t *, double dt_old,
                                   double dt_new, fluid_t *);
int oracle_fluid_correct_velocity(const domain_t *, double dt_old,
                                  double dt_new, fluid_t *);
int oracle_fluid_update_pressure(const domain_t *, fluid_t *);
int oracle_fluid_update_boundaries_ux(const domain_t *, array_t *);
int oracle_fluid_update_boundaries_uy(const domain_t *, array_t *);
int oracle_fluid_update_boundaries_p(const domain_t *, array_t *);
int oracle_fluid_update_boundaries_psi(const domain_t *, array_t *);
int oracle_interface_update_boundaries_vof(const domain_t *, array_t *);
int oracle_integrate(const domain_t *, fluid_t *, interface_t *, double * dt);

/* building blocks exposed for their own tests */
int oracle_dct2(int kind, size_t n, size_t howmany, const double * in,
                double * out);
int oracle_dct3(int kind, size_t n, size_t howmany, const double * in,
                double * out);
int oracle_fft(size_t m, const double * in, double * out); /* interleaved complex */
int oracle_tdm_periodic(int n, double l, double c, double u, double * q,
                        int gauge);
int oracle_tdm_partitioned(int n, int chunk, double a, double b, double * q);
int oracle_rkcoefs(double out[9]); /* {alpha, beta, gamma} x 3 stages */
int oracle_poisson_evals(const domain_t * domain, double * evals);
int oracle_diagnostics(const domain_t *, const fluid_t *, const interface_t *,
                       double out[8]);

#endif /* SBFS_ORACLE_H */
