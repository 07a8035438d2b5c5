// Host-side initial conditions and the dt rule, for parity with the reference at small N.
//   SetParticlePositions  (src/fields/penning_trap.cc:118-161): one std::default_random_engine
//       seeded once; per particle, uniform draws per axis with rejection into the ellipsoid;
//       species blocks in enum order (the reference's std::map order).
//   SetParticleVelocities (src/fields/penning_trap.cc:51-115): per species a fresh
//       std::mt19937(seed) and normal_distribution(0, sqrt(k_B T / m)); 3*count draws, the
//       species mean removed and the set rescaled by sqrt(count / (count - 1)).
//   dt = pi * min(1e-6, min_listed m/|q|) / (2 b0)  (src/params/physics_params.cc:60-74).
// These use libstdc++'s <random> (host code), so the draws are the ones the reference's build
// makes with the same standard library.
#include <cmath>
#include <random>
#include <vector>

#include "am_internal.hpp"

using namespace am;

extern "C" {

int am_set_particle_positions(const uint64_t counts[6], const double radii[3], int seed, double* out) {
  if (!counts || !radii || !out) return Fail(AM_ERR_INVALID, "am_set_particle_positions: NULL argument");
  if (!(radii[0] > 0.0 && radii[1] > 0.0 && radii[2] > 0.0))
    return Fail(AM_ERR_INVALID, "am_set_particle_positions: every cloud radius must be positive");  // physics_params.cc:148-151
  std::default_random_engine rng(seed);
  std::uniform_real_distribution<double> ux(-radii[0], radii[0]);
  std::uniform_real_distribution<double> uy(-radii[1], radii[1]);
  std::uniform_real_distribution<double> uz(-radii[2], radii[2]);
  int64_t k = 0;
  for (int s = 0; s < 6; ++s)
    for (uint64_t i = 0; i < counts[s]; ++i, ++k) {
      double p[3];
      bool inside = false;
      while (!inside) {
        p[0] = ux(rng);
        p[1] = uy(rng);
        p[2] = uz(rng);
        const double a = p[0] / radii[0], b = p[1] / radii[1], c = p[2] / radii[2];
        inside = (a * a) + (b * b) + (c * c) <= 1.0;
      }
      out[3 * k] = p[0];
      out[3 * k + 1] = p[1];
      out[3 * k + 2] = p[2];
    }
  return AM_OK;
}

int am_set_particle_velocities(const uint64_t counts[6], const double temperatures[6], int seed, double* out) {
  if (!counts || !temperatures || !out) return Fail(AM_ERR_INVALID, "am_set_particle_velocities: NULL argument");
  uint64_t total = 0;
  for (int s = 0; s < 6; ++s) total += counts[s];
  if (total == 0)  // penning_trap.cc:55
    return Fail(AM_ERR_INVALID, "am_set_particle_velocities: the populations hold no particles");
  int64_t k = 0;
  for (int s = 0; s < 6; ++s) {
    const uint64_t count = counts[s];
    if (count == 0) continue;
    const double T = temperatures[s];
    if (T == 0.0) {  // a cold species is at rest
      for (uint64_t i = 0; i < 3 * count; ++i) out[3 * k + i] = 0.0;
      k += (int64_t)count;
      continue;
    }
    double mass, charge;
    SpeciesProperties(s, &mass, &charge);
    const double sigma = std::sqrt(kBoltzmann * T / mass);
    std::mt19937 rng(seed);
    std::normal_distribution<double> gauss(0.0, sigma);
    std::vector<double> raw(3 * count);
    for (uint64_t i = 0; i < 3 * count; ++i) raw[i] = gauss(rng);
    if (count >= 2) {
      double mean[3] = {0.0, 0.0, 0.0};
      for (uint64_t i = 0; i < count; ++i)
        for (int d = 0; d < 3; ++d) mean[d] += raw[3 * i + d];
      for (int d = 0; d < 3; ++d) mean[d] /= (double)count;
      const double rescale = std::sqrt((double)count / (double)(count - 1));
      for (uint64_t i = 0; i < count; ++i)
        for (int d = 0; d < 3; ++d) raw[3 * i + d] = (raw[3 * i + d] - mean[d]) * rescale;
    }
    for (uint64_t i = 0; i < 3 * count; ++i) out[3 * k + i] = raw[i];
    k += (int64_t)count;
  }
  return AM_OK;
}

double am_derive_dt(const int32_t listed[6], double b0) {
  double ratio = 1e-6;
  for (int s = 0; s < 6; ++s) {
    if (!listed[s]) continue;
    double m, q;
    SpeciesProperties(s, &m, &q);
    ratio = std::fmin(ratio, m / std::fabs(q));
  }
  return (ratio * kPi) / (b0 * 2);
}

}  // extern "C"
