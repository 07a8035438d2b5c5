// The host driver loop of the reference, Simulate (src/antimatter-kernel.cc:42-217), over the C ABI:
// every per-particle step runs on the device; this file only sequences the calls and writes the
// energy log.  Step order (antimatter-kernel.cc:153-197): the Coulomb solver, then the behaviours
// (Lorentz push and cyclotron cooling fused, annihilation in canonical order on the pushed
// positions), then the energy at log steps and at the step before each (the time-centred potentials
// pair consecutive calls).
#include <cmath>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "antimatter_amd/am_params.h"

namespace am {
int Fail(int code, const std::string& msg);
}  // namespace am

namespace {

constexpr double kElementaryCharge = 1.602176634e-19;

// Owns the context and store of one run.
struct Run {
  am_ctx* ctx = nullptr;
  am_store* store = nullptr;
  ~Run() {
    if (store) am_store_destroy(store);
    if (ctx) am_ctx_destroy(ctx);
  }
};

}  // namespace

extern "C" int am_simulate(const am_params* p, const am_run_options* opt_in, const char* csv_path,
                           am_energy_row* rows, int64_t rows_cap, int64_t* n_rows) {
  if (!p) return am::Fail(AM_ERR_INVALID, "am_simulate: NULL params");
  if (n_rows) *n_rows = 0;
  int rc = am_params_validate(p);
  if (rc) return rc;
  am_run_options opt{0, AM_FP64, 42, 0, 0, 0, nullptr};
  if (opt_in) opt = *opt_in;
  if (opt.precision != AM_FP32 && opt.precision != AM_FP64) return am::Fail(AM_ERR_INVALID, "am_simulate: precision");
  int32_t coulomb_method = 0, energy_method = 0;
  am_params_method(p->coulomb_method, &coulomb_method);
  am_params_method(p->energy_method, &energy_method);

  // Field and initial conditions (antimatter-kernel.cc:82-100): species blocks in enum order.
  am_trap_field field;
  if ((rc = am_trap_make_from_geometry(p->b0, p->trap_voltage, p->z0, p->rho0, &field))) return rc;
  uint64_t counts[6];
  double temps[6];
  int64_t n = 0;
  for (int s = 0; s < 6; ++s) {
    counts[s] = p->population_listed[s] ? p->populations[s] : 0;
    temps[s] = p->temperature_listed[s] ? p->temperatures[s] : 0.0;  // an unlisted species starts at rest
    n += (int64_t)counts[s];
  }
  std::vector<double> pos(3 * n), vel(3 * n);
  std::vector<uint8_t> species(n);
  if ((rc = am_set_particle_positions(counts, p->initial_cloud_radii, p->seed, pos.data()))) return rc;
  if ((rc = am_set_particle_velocities(counts, temps, p->seed, vel.data()))) return rc;
  for (int s = 0, k = 0; s < 6; ++s)
    for (uint64_t c = 0; c < counts[s]; ++c) species[k++] = (uint8_t)s;

  Run run;
  if ((rc = am_ctx_create(opt.device, &run.ctx))) return rc;
  if ((rc = am_store_create(run.ctx, n > 0 ? n : 1, &run.store))) return rc;
  if ((rc = am_store_upload(run.store, n, pos.data(), vel.data(), species.data(), nullptr, nullptr, nullptr,
                            nullptr)))
    return rc;

  // Energy analysis (antimatter-kernel.cc:103-151).
  am_energy_settings settings{energy_method, p->cutoff_radius, p->barnes_hut_theta, p->fast_multipole_theta,
                              p->fast_multipole_order};
  const uint64_t interval = p->num_steps >= 100 ? p->num_steps / 100 : 1;
  am_energy_history history{0, 0.0, 0.0};
  auto evaluate = [&](am_energy_breakdown* e) {
    const int r = am_compute_total_energy_ex(run.store, &field, &settings, &history, opt.precision, e);
    if (!p->coulomb_enabled) e->energy_potential = 0.0;
    return r;
  };
  std::ofstream csv;
  if (csv_path) {
    csv.open(csv_path);
    if (!csv.is_open()) return am::Fail(AM_ERR_INVALID, std::string("am_simulate: cannot write ") + csv_path);
    csv << "step,kinetic_energy_J,trap_potential_energy_J,potential_energy_J,total_energy_J,num_active_particles,"
           "fractional_drift,absolute_drift_J\n";
  }
  int64_t logged = 0;
  auto record = [&](int64_t step, const am_energy_breakdown& e, double total, double frac, double absd) {
    if (rows && logged < rows_cap)
      rows[logged] = am_energy_row{step, e.kinetic, e.trap_potential, e.energy_potential, total,
                                   e.num_active_particles, frac, absd};
    ++logged;
    if (csv.is_open())
      csv << step << "," << e.kinetic << "," << e.trap_potential << "," << e.energy_potential << "," << total << ","
          << e.num_active_particles << "," << frac << "," << absd << "\n";
  };
  am_energy_breakdown e0;
  if ((rc = evaluate(&e0))) return rc;
  const double total0 = e0.kinetic + e0.trap_potential + e0.energy_potential;
  // Drift scale: the trap depth e |U0| (antimatter-kernel.cc:133-138).
  double scale = kElementaryCharge * std::fabs(p->trap_voltage);
  if (scale <= 0.0) scale = std::fabs(total0);
  if (scale <= 0.0) scale = 1.0;
  if (opt.print)
    std::cout << "\nstep 0  KE=" << e0.kinetic << " PE_trap=" << e0.trap_potential
              << " PE_coulomb=" << e0.energy_potential << " TOTAL=" << total0 << " J" << std::endl;
  record(0, e0, total0, 0.0, 0.0);

  // ParaView export (bdm.toml:6-13): the agents after every export_interval-th step, and a collection.
  std::vector<std::string> vtu_names;
  std::vector<double> vtu_times;
  auto export_step = [&](uint64_t step) -> int {
    const std::string name = "ChargedAntiparticle-" + std::to_string(step) + ".vtu";
    const int r = am_export_vtu(run.store, (std::string(opt.export_dir) + "/" + name).c_str());
    if (r) return r;
    vtu_names.push_back(name);
    vtu_times.push_back((double)step * p->dt);
    return AM_OK;
  };
  const bool exporting = opt.export_interval > 0 && opt.export_dir;
  if (exporting && (rc = export_step(0))) return rc;

  const int fmm_order = (int)p->fast_multipole_order;  // real -> int (antimatter-kernel.cc:158)
  for (uint64_t step = 1; step <= p->num_steps; ++step) {
    if (p->coulomb_enabled) {
      if (coulomb_method == AM_METHOD_BARNES_HUT)
        rc = am_compute_barnes_hut_coulomb_forces(run.store, p->barnes_hut_theta, opt.precision);
      else if (coulomb_method == AM_METHOD_FAST_MULTIPOLE)
        rc = am_compute_fast_multipole_coulomb_forces_ex(run.store, fmm_order, p->fast_multipole_theta,
                                                          opt.precision);
      else
        rc = am_compute_standard_coulomb_forces(run.store, p->cutoff_radius, opt.precision);
      if (rc) return rc;
    }
    if ((rc = am_push(run.store, &field, p->dt, p->cooling_scale, p->bath_temp, opt.rng_seed, step))) return rc;
    if (p->annihilation_enabled &&
        (rc = am_annihilation(run.store, p->contact_radius, p->annihilation_probability, opt.rng_seed, step,
                              nullptr)))
      return rc;
    if (exporting && step % (uint64_t)opt.export_interval == 0 && (rc = export_step(step))) return rc;
    const bool log_step = (step % interval == 0) || (step == p->num_steps);
    const bool need = log_step || ((step + 1) % interval == 0) || (step + 1 == p->num_steps);
    if (!need) continue;
    am_energy_breakdown e;
    if ((rc = evaluate(&e))) return rc;
    if (!log_step) continue;
    const double total = e.kinetic + e.trap_potential + e.energy_potential;
    const double absd = total - total0;
    const double frac = absd / scale;
    if (opt.print)
      std::cout << "step " << step << "  KE=" << e.kinetic << " PE_trap=" << e.trap_potential
                << " PE_coulomb=" << e.energy_potential << " TOTAL=" << total << " J"
                << "  (drift: " << absd << " J, " << frac << " of trap depth)" << std::endl;
    record((int64_t)step, e, total, frac, absd);
  }
  if (exporting) {
    std::vector<const char*> names;
    for (const std::string& v : vtu_names) names.push_back(v.c_str());
    if ((rc = am_export_pvd((std::string(opt.export_dir) + "/ChargedAntiparticle.pvd").c_str(), names.data(),
                            vtu_times.data(), (int64_t)names.size())))
      return rc;
  }
  if (csv.is_open()) {
    csv.close();
    if (opt.print) std::cout << "Energy analysis written to: " << csv_path << std::endl;
  }
  if (opt.print_final_positions && n > 0) {  // the agents still in the simulation
    std::vector<uint8_t> gone(n);
    if ((rc = am_store_download(run.store, pos.data(), nullptr, nullptr, gone.data(), nullptr))) return rc;
    for (int64_t i = 0; i < n; ++i)
      if (!gone[i])
        std::cout << "final position: (" << pos[3 * i] << ", " << pos[3 * i + 1] << ", " << pos[3 * i + 2] << ")"
                  << std::endl;
  }
  if (n_rows) *n_rows = logged;
  return AM_OK;
}
