// Fused integrator (placeholder until the fused kernels land): runs the
// staged composition.
#include "decide_dt.h"
#include "fluid_solver.h"
#include "integrate.h"
#include "interface_solver.h"
#include "sbfs_internal.hpp"

namespace sbfs {

int integrate_fused(const domain_t* domain, fluid_t* fluid, interface_t* interface, double* dt) {
  if (decide_dt(domain, fluid, interface, dt)) return 1;
  for (size_t k = 0; k < 3; k++) {
    if (fluid_compute_density(domain, fluid, interface, 0)) return 1;
    if (fluid_compute_viscosity(domain, fluid, interface)) return 1;
    if (interface_compute_curvature_tensor(domain, interface)) return 1;
    if (interface_compute_force(domain, fluid, interface)) return 1;
    if (interface_update_vof(domain, k, *dt, fluid, interface)) return 1;
    if (fluid_compute_density(domain, fluid, interface, 1)) return 1;
    if (interface_compute_mass_flux(domain, fluid, interface)) return 1;
    if (fluid_predict_field(domain, k, *dt, fluid, interface)) return 1;
    const double dt_old = rk((k + 2) % 3).g * (*dt);
    const double dt_new = rk(k).g * (*dt);
    if (fluid_compute_potential(domain, dt_old, dt_new, fluid)) return 1;
    if (fluid_correct_velocity(domain, dt_old, dt_new, fluid)) return 1;
    if (fluid_update_pressure(domain, fluid)) return 1;
  }
  return 0;
}

}  // namespace sbfs
