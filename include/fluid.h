#if !defined(FLUID_H)
#define FLUID_H

#include "array.h"
#include "domain.h"

/*
 * Flow-field state and work arrays.  Member order and types follow the
 * reference's fluid_t (include/fluid.h:27-45).
 *
 *   ux, uy      staggered velocities (x faces / y faces)
 *   p, psi[2]   pressure and scalar potentials (psi[0] is the newest after
 *               each call to fluid_update_pressure)
 *   den[2]      density at the start (0) and end (1) of a Runge-Kutta stage
 *   visux/visuy dynamic viscosity at the x / y velocity points
 *   txx/txy/tyy shear-stress tensor
 *   srcux/srcuy Runge-Kutta source terms: [0] alpha, [1] beta, [2] gamma
 *   Re, Fr      Reynolds and Froude numbers
 *   denr, visr  density and viscosity ratios (gas / liquid)
 *   refden      reference density min(1, denr)
 */
typedef struct {
  array_t ux;
  array_t uy;
  array_t p;
  array_t psi[2];
  array_t den[2];
  array_t visux;
  array_t visuy;
  array_t txx;
  array_t txy;
  array_t tyy;
  array_t srcux[3];
  array_t srcuy[3];
  double Re;
  double Fr;
  double denr;
  double visr;
  double refden;
} fluid_t;

#endif /* FLUID_H */
