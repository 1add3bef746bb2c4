#if !defined(INTERFACE_H)
#define INTERFACE_H

#include "array.h"

/*
 * Interface (volume-of-fluid) state and work arrays.  Member order and
 * types follow the reference's interface_t (include/interface.h:7-24).
 *
 *   vof          volume fraction (1 = gas), two halo rows in y
 *   ifrcx/ifrcy  surface-tension force at the x / y faces
 *   dvof         unit gradient of vof at cell corners (vector_t)
 *   normal       cell normal and THINC intercept (normal_t)
 *   curv         curvature at cell centres
 *   fluxx/fluxy  face fluxes (vof flux, then overwritten by mass flux)
 *   src[2]       Runge-Kutta sources of vof: [0] alpha, [1] beta
 *   We           Weber number
 */
typedef double vector_t[NDIMS];
typedef double normal_t[NDIMS + 1];

typedef struct {
  array_t vof;
  array_t ifrcx;
  array_t ifrcy;
  array_t dvof;
  array_t normal;
  array_t curv;
  array_t fluxx;
  array_t fluxy;
  array_t src[2];
  double We;
} interface_t;

#endif /* INTERFACE_H */
