/*
 * antimatter-kernel for MI355X — run configuration (params.json) and the host driver loop.
 *
 * The reference reads one params.json into three BioDynaMo parameter groups (PhysicsParams,
 * InteractionParams, StochasticParams) with nlohmann::json, validates them, and runs Simulate
 * (src/antimatter-kernel.cc:42-217).  This header is the C-ABI counterpart: one plain struct with
 * the fields of the three groups, a loader with the reference's key names, defaults and conversions,
 * the three Validate() checks, and the driver loop on the device with the reference's energy CSV.
 * Same return convention as am_api.h (AM_OK or a negative code, message in am_last_error()).
 */
#ifndef ANTIMATTER_AMD_AM_PARAMS_H_
#define ANTIMATTER_AMD_AM_PARAMS_H_

#include <stdint.h>

#include "antimatter_amd/am_api.h"

#ifdef __cplusplus
extern "C" {
#endif

#define AM_PARAMS_METHOD_LEN 32

typedef struct {
  /* PhysicsParams (src/params/physics_params.h:22-47) */
  int32_t seed;
  double b0, trap_voltage, z0, rho0;
  double dt; /* derived from the populations and b0 when a file is read (physics_params.cc:60-74) */
  uint64_t num_steps;
  double bath_temp, cooling_scale;
  uint64_t populations[6];         /* particle count per species (AM_ANTIPROTON ...) */
  int32_t population_listed[6];    /* species present in particle_populations */
  double temperatures[6];          /* kelvin; read as size_t, so fractions are truncated */
  int32_t temperature_listed[6];
  double initial_cloud_radii[3];
  /* InteractionParams (src/params/interaction_params.h:37-47) */
  int32_t coulomb_enabled;
  char coulomb_method[AM_PARAMS_METHOD_LEN]; /* "standard", "barnes-hut", "fast-multipole" */
  char energy_method[AM_PARAMS_METHOD_LEN];
  double cutoff_radius, barnes_hut_theta, fast_multipole_order, fast_multipole_theta;
  /* StochasticParams (src/params/stochastic_params.h:20-23) */
  int32_t annihilation_enabled;
  double contact_radius, annihilation_probability;
  /* 1 when a file was read; 0 when it was missing and the defaults were kept (LoadParams prints a
   * notice and returns in that case, physics_params.cc:44-49). */
  int32_t loaded;
} am_params;

/* The parameter groups' defaults (physics_params.h:22-47, interaction_params.h:37-47,
 * stochastic_params.h:20-23): one positron x 10, dt = 1e-9, cooling_scale = 10 (which Validate
 * rejects), and so on. */
int am_params_defaults(am_params* p);
/* PhysicsParams/InteractionParams/StochasticParams::LoadParams on one file: keys and conversions of
 * physics_params.cc:42-100, interaction_params.cc:30-47, stochastic_params.cc:13-26.  A missing file
 * keeps the defaults (loaded = 0) and returns AM_OK; malformed JSON, a value of the wrong type or an
 * unknown species name returns AM_ERR_INVALID.  p is filled from the defaults first. */
int am_params_load(const char* path, am_params* p);
/* am_params_load on JSON text instead of a file (loaded = 1 on success). */
int am_params_parse(const char* json_text, am_params* p);
/* PhysicsParams::Validate, InteractionParams::Validate, StochasticParams::Validate in the driver's
 * order (antimatter-kernel.cc:62-64), then the energy method's name (GetEnergyMethod,
 * antimatter-kernel.cc:103): AM_ERR_INVALID with the first failure.  The reference's quirks are
 * kept: the energy checks use the Coulomb method (interaction_params.cc:74-75), a method name is
 * checked even when Coulomb forces are off, and cooling_scale > 1 is rejected. */
int am_params_validate(const am_params* p);
/* GetCoulombMethod / GetEnergyMethod (interaction_params.cc:14-28): AM_METHOD_* of a name. */
int am_params_method(const char* name, int32_t* method);

/* Options of the device driver loop (no reference counterpart: the reference's solvers are FP64 and
 * its random streams are BioDynaMo's). */
typedef struct {
  int32_t device;
  int32_t precision;     /* AM_FP64 (reference arithmetic) or AM_FP32 for the solvers and energy */
  uint64_t rng_seed;     /* key of the counter-based draws (annihilation, cooling noise) */
  int32_t print;         /* the reference's stdout lines (step energies, drift) */
  int32_t print_final_positions; /* "final position: (x, y, z)" per particle (antimatter-kernel.cc:204-210) */
  /* ParaView export (bdm.toml:6-13 turns BioDynaMo's on every step): every export_interval steps
   * (0: off) the agents go to export_dir/ChargedAntiparticle-<step>.vtu, step 0 included, and
   * export_dir/ChargedAntiparticle.pvd lists them with their times. */
  int32_t export_interval;
  const char* export_dir;
} am_run_options;

/* One row of the energy log: the 8 CSV columns of antimatter-kernel.cc:128. */
typedef struct {
  int64_t step;
  double kinetic, trap_potential, coulomb_potential, total;
  int64_t num_active_particles;
  double fractional_drift, absolute_drift;
} am_energy_row;

/* Simulate (src/antimatter-kernel.cc:76-197) on the device: initial conditions from the reference
 * generators, then per step the Coulomb solver, the fused Lorentz push + cooling, annihilation in
 * canonical order, and the time-centred energy at log steps and the step before them.  csv_path
 * (optional) receives the energy CSV with the reference's header and number format; rows (optional,
 * capacity rows_cap) receive the logged rows and n_rows their count.  p must pass
 * am_params_validate. */
int am_simulate(const am_params* p, const am_run_options* opt, const char* csv_path, am_energy_row* rows,
                int64_t rows_cap, int64_t* n_rows);

/* The agents still in the simulation (not annihilated) as a VTK XML UnstructuredGrid of vertex cells:
 * the points are the positions (m), with the point data species_id (charged_antiparticle.h:73, the
 * data member bdm.toml:13 adds to the export) and diameter_ (1e-6 m, constants.h:38), the content of
 * BioDynaMo's ParaView export of ChargedAntiparticle.  ASCII, positions to 17 significant digits. */
int am_export_vtu(am_store* store, const char* path);
/* A .pvd collection of count exported files (names relative to the .pvd) with their times (s). */
int am_export_pvd(const char* path, const char* const* files, const double* times, int64_t count);

#ifdef __cplusplus
}
#endif

#endif /* ANTIMATTER_AMD_AM_PARAMS_H_ */
