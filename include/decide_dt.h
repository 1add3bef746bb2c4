#if !defined(DECIDE_DT_H)
#define DECIDE_DT_H

#include "domain.h"
#include "fluid.h"
#include "interface.h"

SBFS_EXTERN_C_BEGIN

/*
 * Time-step size: min of the advective CFL limit (a device min-reduction
 * over every interior face), the two diffusive limits and the
 * surface-tension limit.  The safety factors come from the environment
 * variables coef_dt_adv, coef_dt_dif, coef_dt_int on first use (as in the
 * reference) unless sbfs_set_dt_coefs() set them first.
 * Replaces src/decide_dt.c:140-169 (include/decide_dt.h:5-10).
 */
extern int decide_dt(
    const domain_t * domain,
    const fluid_t * fluid,
    const interface_t * interface,
    double * dt
);

SBFS_EXTERN_C_END

#endif /* DECIDE_DT_H */
