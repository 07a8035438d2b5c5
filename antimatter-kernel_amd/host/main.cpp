// am_simulate: the reference executable's main (src/antimatter-kernel.cc:42-220) on the MI355X
// library.  Reads params.json (or --params FILE), validates it, prints the configuration, runs the
// simulation on the device and writes the energy CSV into the output directory; with
// --export-interval K (the reference's bdm.toml exports every step) also the ParaView files of every
// K-th step into DIR/visualization.
//
//   am_simulate [--params FILE] [--output-dir DIR] [--fp32] [--rng-seed N] [--device D]
//               [--final-positions] [--export-interval K]
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>
#include <sys/stat.h>

#include "antimatter_amd/am_params.h"

namespace {

const char* kSpeciesNames[6] = {"antiproton", "positron", "antihydrogen", "electron", "proton", "antihelium"};

void PrintParams(const am_params& p) {
  std::cout << "[physics] \nseed = " << p.seed << "\nb0 = " << p.b0 << " T\ntrap_voltage = " << p.trap_voltage
            << " V\nz0 = " << p.z0 << " m\nrho0 = " << p.rho0 << " m\ndt = " << p.dt << " s\nnum_steps = "
            << p.num_steps << std::endl;
  std::cout << "\nPARTICLE POPULATIONS:";
  for (int s = 0; s < 6; ++s)
    if (p.population_listed[s]) std::cout << '\n' << kSpeciesNames[s] << " x " << p.populations[s];
  std::cout << "\n\nPARTICLE TEMPERATURES:";
  for (int s = 0; s < 6; ++s)
    if (p.temperature_listed[s]) std::cout << '\n' << kSpeciesNames[s] << " - " << p.temperatures[s] << " Kelvin";
  std::cout << "\n\n[interaction] \ncoulomb_enabled = " << (p.coulomb_enabled ? "true" : "false")
            << "\ncoulomb_computation_method = " << p.coulomb_method;
  if (p.coulomb_enabled) {
    const std::string m = p.coulomb_method;
    if (m == "standard") std::cout << "\ncutoff_radius = " << p.cutoff_radius << " m";
    else if (m == "barnes-hut") std::cout << "\nbarnes_hut_theta = " << p.barnes_hut_theta;
    else if (m == "fast-multipole")
      std::cout << "\nfast_multipole_order = " << p.fast_multipole_order
                << "\nfast_multipole_theta = " << p.fast_multipole_theta;
  }
  std::cout << "\nenergy_computation_method = " << p.energy_method << std::endl;
  std::cout << "\n[stochastic] \nannihilation_enabled = " << (p.annihilation_enabled ? "true" : "false")
            << " \ncontact_radius = " << p.contact_radius << " m \nannihilation_probability = "
            << p.annihilation_probability << std::endl;
}

int Usage() {
  std::cerr << "usage: am_simulate [--params FILE] [--output-dir DIR] [--fp32] [--rng-seed N] [--device D] "
               "[--final-positions] [--export-interval K]"
            << std::endl;
  return 2;
}

}  // namespace

int main(int argc, char** argv) {
  const auto start = std::chrono::high_resolution_clock::now();
  std::string params_path = "params.json", out_dir = "output";
  am_run_options opt{0, AM_FP64, 42, 1, 0, 0, nullptr};
  std::string export_dir;
  for (int a = 1; a < argc; ++a) {
    const std::string s = argv[a];
    auto next = [&]() -> const char* { return a + 1 < argc ? argv[++a] : nullptr; };
    if (s == "--params") { const char* v = next(); if (!v) return Usage(); params_path = v; }
    else if (s == "--output-dir") { const char* v = next(); if (!v) return Usage(); out_dir = v; }
    else if (s == "--fp32") opt.precision = AM_FP32;
    else if (s == "--rng-seed") { const char* v = next(); if (!v) return Usage(); opt.rng_seed = std::strtoull(v, nullptr, 10); }
    else if (s == "--device") { const char* v = next(); if (!v) return Usage(); opt.device = std::atoi(v); }
    else if (s == "--final-positions") opt.print_final_positions = 1;
    else if (s == "--export-interval") { const char* v = next(); if (!v) return Usage(); opt.export_interval = std::atoi(v); }
    else return Usage();
  }
  am_params p;
  if (am_params_load(params_path.c_str(), &p) != AM_OK) {
    std::cerr << am_last_error() << std::endl;
    return 1;
  }
  if (!p.loaded) std::cout << "'" << params_path << "' not found -- using default values." << std::endl;
  if (am_params_validate(&p) != AM_OK) {
    std::cerr << am_last_error() << std::endl;
    return 1;
  }
  std::cout << std::endl;
  PrintParams(p);
  mkdir(out_dir.c_str(), 0755);
  if (opt.export_interval > 0) {  // the ParaView files beside the energy CSV
    export_dir = out_dir + "/visualization";
    mkdir(export_dir.c_str(), 0755);
    opt.export_dir = export_dir.c_str();
  }
  const std::string csv = out_dir + "/energy_analysis_data.csv";
  int64_t rows = 0;
  if (am_simulate(&p, &opt, csv.c_str(), nullptr, 0, &rows) != AM_OK) {
    std::cerr << am_last_error() << std::endl;
    return 1;
  }
  const std::chrono::duration<double> took = std::chrono::high_resolution_clock::now() - start;
  std::cout << "Simulation completed successfully!" << std::endl;
  std::cout << "Time taken " << took.count() << "s" << std::endl;
  return 0;
}
