// Fast multipole solver (placeholder until the device FMM lands).
#include "am_internal.hpp"

namespace am {
int FastMultipolePotentialEnergy(am_store* s, int order, double theta, double* pe) {
  return Fail(AM_ERR_STATE, "FMM potential not available in this build");
}
}  // namespace am

extern "C" int am_compute_fast_multipole_coulomb_forces(am_store* s, int order, double theta) {
  return am::Fail(AM_ERR_STATE, "FMM not available in this build");
}
