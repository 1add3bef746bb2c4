#if !defined(INTERFACE_SOLVER_H)
#define INTERFACE_SOLVER_H

#include <stddef.h>
#include "domain.h"
#include "fluid.h"
#include "interface.h"

/*
 * Interface-capturing stages.  Each entry point replaces the reference
 * function of the same name and signature (include/interface_solver.h:8-48).
 * Return 0 on success, non-zero on failure.
 */
SBFS_EXTERN_C_BEGIN

/* allocate every interface_t array on the device, load vof from the NPY
 * directory, apply boundary conditions, read We from the environment.
 * Replaces src/interface/init.c:55-69. */
extern int interface_init(
    const char dirname_ic[],
    const domain_t * domain,
    interface_t * interface
);

/* write vof and We as NPY files.  Replaces src/interface/save.c:7-22. */
extern int interface_save(
    const char dirname[],
    const domain_t * domain,
    const interface_t * interface
);

/* corner unit gradients (dvof), cell normals + THINC intercepts inside the
 * interface band, and curvature.  Replaces
 * src/interface/curvature_tensor.c:137-145. */
extern int interface_compute_curvature_tensor(
    const domain_t * domain,
    interface_t * interface
);

/* continuum surface force at x and y faces.
 * Replaces src/interface/force.c:90-98. */
extern int interface_compute_force(
    const domain_t * domain,
    const fluid_t * fluid,
    interface_t * interface
);

/* overwrite the vof fluxes by mass fluxes u + (denr - 1) F.
 * Replaces src/interface/mass_flux.c:48-58. */
extern int interface_compute_mass_flux(
    const domain_t * domain,
    const fluid_t * fluid,
    interface_t * interface
);

/* THINC face fluxes, flux divergence and the Runge-Kutta vof update.
 * Replaces src/interface/update/main.c:122-136. */
extern int interface_update_vof(
    const domain_t * domain,
    const size_t rkstep,
    const double dt,
    const fluid_t * fluid,
    interface_t * interface
);

/* vof = 0 in the wall ghosts + two-row periodic y halo.
 * Replaces src/interface/boundary.c:15-42. */
extern int interface_update_boundaries_vof(
    const domain_t * domain,
    array_t * vof
);

SBFS_EXTERN_C_END

#endif /* INTERFACE_SOLVER_H */
