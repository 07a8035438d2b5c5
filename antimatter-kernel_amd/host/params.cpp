// Run configuration: the reference's three parameter groups read from one params.json.
//
// The reference reads each group with nlohmann::json's value(key, default) and get<T>()
// (physics_params.cc:42-100, interaction_params.cc:30-47, stochastic_params.cc:13-26); this loader
// uses the same library, so a key's conversions are the reference's: num_steps and the
// temperatures go through size_t (so 1e2 reads as 100 and 1000.7 K as 1000 K), seed through int.
#include <cmath>
#include <cstring>
#include <fstream>
#include <limits>
#include <sstream>
#include <stdexcept>
#include <string>

#include "antimatter_amd/am_params.h"
#include "json.hpp"

namespace am {
int Fail(int code, const std::string& msg);
}  // namespace am

namespace {

constexpr double kPi = 3.14159265358979323846;

// Species names accepted in particle_populations / particle_temperatures (physics_params.cc:17-25):
// the enum's names except antihelium.
int SpeciesOfName(const std::string& name) {
  if (name == "antiproton") return AM_ANTIPROTON;
  if (name == "positron") return AM_POSITRON;
  if (name == "antihydrogen") return AM_ANTIHYDROGEN;
  if (name == "electron") return AM_ELECTRON;
  if (name == "proton") return AM_PROTON;
  throw std::invalid_argument("params.json: unknown species '" + name + "'");
}

// nlohmann's get<size_t>() / get<int>() on a floating-point value is a static_cast, which is
// undefined when the target cannot hold the value's integer part (1e300 steps, a seed of 1e20, a
// temperature of -2 K through size_t).  Such a value is rejected here; an in-range value converts
// exactly as the reference's get<T>() does (truncation toward zero), and integer values keep
// nlohmann's integer conversion.
template <typename T>
T IntegerOf(const nlohmann::json& v, const char* key) {
  if (v.is_number_float()) {
    const double d = v.get<double>();
    const double lo = (double)std::numeric_limits<T>::min() - 1.0;
    const double hi = (double)std::numeric_limits<T>::max() + 1.0;  // 2^64 or 2^31, exact in FP64
    if (!(d > lo && d < hi))
      throw std::invalid_argument(std::string("params.json: ") + key + " = " + v.dump() +
                                  " is outside the range of its integer type");
  }
  return v.get<T>();
}

template <typename T>
T IntegerValue(const nlohmann::json& j, const char* key, T fallback) {
  if (!j.is_object()) throw std::invalid_argument("params.json: the top level is not an object");
  return j.count(key) ? IntegerOf<T>(j.at(key), key) : fallback;
}

void CopyName(char* dst, const std::string& s, const char* key) {
  if (s.size() >= AM_PARAMS_METHOD_LEN)
    throw std::invalid_argument(std::string("params.json: ") + key + " '" + s + "' is not a known method");
  std::memcpy(dst, s.c_str(), s.size() + 1);
}

void ReadPhysics(const nlohmann::json& j, am_params* p) {
  p->seed = IntegerValue<int>(j, "seed", p->seed);
  p->b0 = j.value("b0", p->b0);
  p->trap_voltage = j.value("trap_voltage", p->trap_voltage);
  p->z0 = j.value("z0", p->z0);
  p->rho0 = j.value("rho0", p->rho0);
  p->bath_temp = j.value("bath_temperature", p->bath_temp);
  p->cooling_scale = j.value("cooling_scale", p->cooling_scale);
  double ratio = 1e-6;  // the smallest m/|q| of the listed species, at most 1e-6 kg/C
  if (j.count("particle_populations")) {
    for (int s = 0; s < 6; ++s) {
      p->populations[s] = 0;
      p->population_listed[s] = 0;
    }
    const nlohmann::json& pops = j.at("particle_populations");
    for (auto it = pops.begin(); it != pops.end(); ++it) {
      const int s = SpeciesOfName(it.key());
      p->populations[s] = (uint64_t)IntegerOf<size_t>(it.value(), "particle_populations");
      p->population_listed[s] = 1;
      double m, q;
      am_species_properties(s, &m, &q);
      ratio = std::fmin(ratio, m / std::fabs(q));  // a neutral species gives +inf and changes nothing
    }
  }
  p->dt = (ratio * kPi) / (p->b0 * 2);  // physics_params.cc:74, also without particle_populations
  if (j.count("particle_temperatures")) {
    for (int s = 0; s < 6; ++s) {
      p->temperatures[s] = 0.0;
      p->temperature_listed[s] = 0;
    }
    const nlohmann::json& temps = j.at("particle_temperatures");
    for (auto it = temps.begin(); it != temps.end(); ++it) {
      const int s = SpeciesOfName(it.key());
      p->temperatures[s] = (double)IntegerOf<size_t>(it.value(), "particle_temperatures");
      p->temperature_listed[s] = 1;
    }
  }
  p->num_steps = (uint64_t)IntegerValue<size_t>(j, "num_steps", (size_t)p->num_steps);
  if (j.count("initial_cloud_radii")) {
    const nlohmann::json& r = j.at("initial_cloud_radii");
    const char* keys[3] = {"radius_x", "radius_y", "radius_z"};
    for (int d = 0; d < 3; ++d)
      if (r.count(keys[d])) p->initial_cloud_radii[d] = r.at(keys[d]).get<double>();
  }
}

void ReadInteraction(const nlohmann::json& j, am_params* p) {
  p->coulomb_enabled = j.value("coulomb_forces_enabled", p->coulomb_enabled != 0) ? 1 : 0;
  CopyName(p->coulomb_method, j.value("coulomb_computation_method", std::string(p->coulomb_method)),
           "coulomb_computation_method");
  CopyName(p->energy_method, j.value("energy_computation_method", std::string(p->energy_method)),
           "energy_computation_method");
  p->cutoff_radius = j.value("cutoff_radius", p->cutoff_radius);
  p->barnes_hut_theta = j.value("barnes_hut_theta", p->barnes_hut_theta);
  p->fast_multipole_order = j.value("fast_multipole_order", p->fast_multipole_order);
  p->fast_multipole_theta = j.value("fast_multipole_theta", p->fast_multipole_theta);
}

void ReadStochastic(const nlohmann::json& j, am_params* p) {
  p->annihilation_enabled = j.value("annihilation_enabled", p->annihilation_enabled != 0) ? 1 : 0;
  p->contact_radius = j.value("contact_radius", p->contact_radius);
  p->annihilation_probability = j.value("annihilation_probability", p->annihilation_probability);
}

int Invalid(const char* msg) { return am::Fail(AM_ERR_INVALID, msg); }

// The checks shared by the energy and Coulomb halves of InteractionParams::Validate
// (interaction_params.cc:76-94 and :96-114); both are given the Coulomb method.
int CheckMethodSettings(const am_params* p, int32_t method) {
  if (method == AM_METHOD_STANDARD && p->cutoff_radius <= 0.0)
    return Invalid("params: cutoff_radius must be positive for the standard method");
  if (method == AM_METHOD_BARNES_HUT && (p->barnes_hut_theta < 0.0 || p->barnes_hut_theta > 1.0))
    return Invalid("params: barnes_hut_theta must lie in [0, 1]");
  if (method == AM_METHOD_FAST_MULTIPOLE) {
    if (p->fast_multipole_order < 1 || p->fast_multipole_order > 8)
      return Invalid("params: fast_multipole_order must lie in [1, 8]");
    if (p->fast_multipole_theta < 0.0 || p->fast_multipole_theta >= 1.0)
      return Invalid("params: fast_multipole_theta must lie in [0, 1)");
  }
  return AM_OK;
}

}  // namespace

extern "C" {

int am_params_defaults(am_params* p) {
  if (!p) return Invalid("am_params_defaults: NULL");
  std::memset(p, 0, sizeof(*p));
  p->seed = 42;
  p->b0 = 1.0;
  p->trap_voltage = 10.0;
  p->z0 = 0.01;
  p->rho0 = 0.01;
  p->dt = 1e-9;
  p->num_steps = 5000;
  p->bath_temp = 0.5;
  p->cooling_scale = 10.0;
  p->populations[AM_POSITRON] = 10;
  p->population_listed[AM_POSITRON] = 1;
  p->temperatures[AM_ELECTRON] = 4.0;
  p->temperatures[AM_POSITRON] = 4.0;
  p->temperatures[AM_ANTIPROTON] = 1000.0;
  p->temperature_listed[AM_ELECTRON] = p->temperature_listed[AM_POSITRON] = p->temperature_listed[AM_ANTIPROTON] = 1;
  p->initial_cloud_radii[0] = p->initial_cloud_radii[1] = p->initial_cloud_radii[2] = 1e-4;
  p->coulomb_enabled = 1;
  std::strcpy(p->coulomb_method, "standard");
  std::strcpy(p->energy_method, "standard");
  p->cutoff_radius = 0.0;
  p->barnes_hut_theta = 0.5;
  p->fast_multipole_order = 4.0;
  p->fast_multipole_theta = 0.5;
  p->annihilation_enabled = 0;
  p->contact_radius = 1e-9;
  p->annihilation_probability = 0.1;
  p->loaded = 0;
  return AM_OK;
}

static int ReadAll(std::istream& in, am_params* p, const char* who) {
  try {
    nlohmann::json j;
    in >> j;
    ReadPhysics(j, p);
    ReadInteraction(j, p);
    ReadStochastic(j, p);
  } catch (const std::exception& e) {
    return am::Fail(AM_ERR_INVALID, std::string(who) + ": " + e.what());
  }
  p->loaded = 1;
  return AM_OK;
}

int am_params_load(const char* path, am_params* p) {
  if (!path || !p) return Invalid("am_params_load: NULL argument");
  am_params_defaults(p);
  std::ifstream file(path);
  if (!file.is_open()) return AM_OK;  // defaults kept, as each group's LoadParams does
  return ReadAll(file, p, "am_params_load");
}

int am_params_parse(const char* json_text, am_params* p) {
  if (!json_text || !p) return Invalid("am_params_parse: NULL argument");
  am_params_defaults(p);
  std::istringstream in(json_text);
  return ReadAll(in, p, "am_params_parse");
}

int am_params_method(const char* name, int32_t* method) {
  if (!name || !method) return Invalid("am_params_method: NULL argument");
  const std::string s(name);
  if (s == "standard") *method = AM_METHOD_STANDARD;
  else if (s == "barnes-hut") *method = AM_METHOD_BARNES_HUT;
  else if (s == "fast-multipole") *method = AM_METHOD_FAST_MULTIPOLE;
  else
    return am::Fail(AM_ERR_INVALID, "params: unknown method '" + s +
                                        "' (the accepted names are standard, barnes-hut, fast-multipole)");
  return AM_OK;
}

int am_params_validate(const am_params* p) {
  if (!p) return Invalid("am_params_validate: NULL");
  // PhysicsParams::Validate (physics_params.cc:126-160)
  if (p->b0 <= 0.0) return Invalid("params: b0 must be positive");
  if (p->z0 <= 0.0 || p->rho0 <= 0.0) return Invalid("params: z0 and rho0 must be positive");
  if (p->num_steps == 0) return Invalid("params: num_steps must be positive");
  for (int s = 0; s < 6; ++s)
    if (p->temperature_listed[s] && p->temperatures[s] < 0.0) return Invalid("params: temperatures must be >= 0");
  if (p->bath_temp < 0.0) return Invalid("params: bath_temperature must be >= 0");
  if (p->cooling_scale > 1.0) return Invalid("params: cooling_scale must be at most 1");
  int listed = 0;
  for (int s = 0; s < 6; ++s) listed += p->population_listed[s] ? 1 : 0;
  if (listed == 0) return Invalid("params: particle_populations is empty");
  const double* r = p->initial_cloud_radii;
  if (r[0] <= 0.0 || r[1] <= 0.0 || r[2] <= 0.0) return Invalid("params: every cloud radius must be positive");
  if (r[0] >= p->rho0 || r[1] >= p->rho0) return Invalid("params: the radial cloud radii must be below rho0");
  if (r[2] >= p->z0) return Invalid("params: the axial cloud radius must be below z0");
  // InteractionParams::Validate (interaction_params.cc:71-115): the method name is checked even with
  // Coulomb forces off, and both halves check the Coulomb method (:74-75).
  int32_t coulomb = 0;
  int rc = am_params_method(p->coulomb_method, &coulomb);
  if (rc) return rc;
  if ((rc = CheckMethodSettings(p, coulomb))) return rc;
  if (p->coulomb_enabled && (rc = CheckMethodSettings(p, coulomb))) return rc;
  // StochasticParams::Validate (stochastic_params.cc:35-46)
  if (p->annihilation_enabled) {
    if (p->contact_radius <= 0.0) return Invalid("params: contact_radius must be positive with annihilation on");
    if (p->annihilation_probability < 0.0 || p->annihilation_probability > 1.0)
      return Invalid("params: annihilation_probability must lie in [0, 1]");
  }
  // GetEnergyMethod (antimatter-kernel.cc:103)
  int32_t energy = 0;
  return am_params_method(p->energy_method, &energy);
}

}  // extern "C"
