#if !defined(FLUID_SOLVER_H)
#define FLUID_SOLVER_H

#include <stddef.h>
#include "array.h"
#include "domain.h"
#include "fluid.h"
#include "interface.h"

/*
 * Flow-solver stages.  Each entry point replaces the reference function of
 * the same name and signature (include/fluid_solver.h:9-89).  All return 0
 * on success and non-zero on failure (a message goes to stderr; the text of
 * the last failure is also available from sbfs_last_error()).
 *
 * Work is queued on the library's HIP stream in call order; results are
 * visible to the host after sbfs_array_download() or sbfs_synchronize().
 */
SBFS_EXTERN_C_BEGIN

/* allocate every fluid_t array on the device, load ux, uy, p, psi from the
 * NPY directory, apply boundary conditions, read Re, Fr, denr, visr from the
 * environment.  Replaces src/fluid/init.c:84-109. */
extern int fluid_init(
    const char dirname_ic[],
    const domain_t * domain,
    fluid_t * fluid
);

/* write ux, uy, p, psi[0] and Re, Fr as NPY files.
 * Replaces src/fluid/save.c:8-27. */
extern int fluid_save(
    const char dirname[],
    const domain_t * domain,
    const fluid_t * fluid
);

/* den[index] = clamp(1 + (denr - 1) vof) on every cell incl. ghosts.
 * Replaces src/fluid/density.c:9-31. */
extern int fluid_compute_density(
    const domain_t * SBFS_RESTRICT domain,
    fluid_t * SBFS_RESTRICT fluid,
    const interface_t * SBFS_RESTRICT interface,
    const size_t index
);

/* visux / visuy from face-averaged vof, incl. halo rows.
 * Replaces src/fluid/viscosity.c:93-108. */
extern int fluid_compute_viscosity(
    const domain_t * SBFS_RESTRICT domain,
    fluid_t * SBFS_RESTRICT fluid,
    const interface_t * SBFS_RESTRICT interface
);

/* stress tensor, explicit Runge-Kutta sources and the predicted velocity.
 * Replaces src/fluid/predict/main.c:60-80. */
extern int fluid_predict_field(
    const domain_t * domain,
    const size_t rkstep,
    const double dt,
    fluid_t * fluid,
    const interface_t * interface
);

/* solve the split-pressure Poisson equation for psi[1]
 * (DCT-II in x, periodic tridiagonal in y, DCT-III in x).
 * Replaces src/fluid/compute_potential.c:421-469. */
extern int fluid_compute_potential(
    const domain_t * domain,
    const double dt_old,
    const double dt_new,
    fluid_t * fluid
);

/* project the predicted velocity with psi[1] and psi[0].
 * Replaces src/fluid/correct_velocity/main.c:14-24. */
extern int fluid_correct_velocity(
    const domain_t * domain,
    const double dt_old,
    const double dt_new,
    fluid_t * fluid
);

/* p += psi[1], boundary conditions, then swap psi[0] and psi[1].
 * Replaces src/fluid/update_pressure.c:43-53. */
extern int fluid_update_pressure(
    const domain_t * domain,
    fluid_t * fluid
);

/* impermeable walls + periodic y halo.  Replaces src/fluid/boundary/ux.c:15-42 */
extern int fluid_update_boundaries_ux(
    const domain_t * domain,
    array_t * array
);

/* wall velocity ghosts + periodic y halo.  Replaces src/fluid/boundary/uy.c:16-43 */
extern int fluid_update_boundaries_uy(
    const domain_t * domain,
    array_t * array
);

/* Neumann x ghosts + periodic y halo.  Replaces src/fluid/boundary/p.c:15-42 */
extern int fluid_update_boundaries_p(
    const domain_t * domain,
    array_t * array
);

/* Neumann x ghosts + periodic y halo.  Replaces src/fluid/boundary/psi.c:15-42 */
extern int fluid_update_boundaries_psi(
    const domain_t * domain,
    array_t * array
);

SBFS_EXTERN_C_END

#endif /* FLUID_SOLVER_H */
