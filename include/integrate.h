#if !defined(INTEGRATE_H)
#define INTEGRATE_H

#include "domain.h"
#include "fluid.h"
#include "interface.h"

SBFS_EXTERN_C_BEGIN

/*
 * Advance the flow by one full time step: decide_dt followed by three
 * low-storage Runge-Kutta stages, each running every interface and fluid
 * stage in the reference order.  *dt receives the step size.
 * Replaces src/integrate.c:11-70 (include/integrate.h:5-10).
 *
 * The fused fast path keeps every state field (vof, ux, uy, p, psi[0]) and
 * the arrays the reference's logging reads after a step (den[1], txy)
 * exactly as the staged reference order leaves them; other work arrays are
 * left unspecified.  Returns 0 on success.
 */
extern int integrate(
    const domain_t * domain,
    fluid_t * fluid,
    interface_t * interface,
    double * dt
);

SBFS_EXTERN_C_END

#endif /* INTEGRATE_H */
