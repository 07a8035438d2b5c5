// Barnes-Hut octree solver (placeholder until the device tree lands).
#include "am_internal.hpp"

namespace am {
int BarnesHutPotentialTerms(am_store* s, double theta, double* out) {
  return Fail(AM_ERR_STATE, "Barnes-Hut potential not available in this build");
}
}  // namespace am

extern "C" int am_compute_barnes_hut_coulomb_forces(am_store* s, double theta, int precision) {
  return am::Fail(AM_ERR_STATE, "Barnes-Hut not available in this build");
}
