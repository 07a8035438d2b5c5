// antimatter-kernel for MI355X — C++ layer over the C ABI with the reference's names.
//
// The reference's free functions, behaviours and standalone solvers (src/interaction/*.h,
// src/forces/*.h, src/analysis/energy_analysis.h) take a BioDynaMo ResourceManager or run per
// agent; here they take the device particle store (am_store) and act on every particle at once.
// Errors become antimatter_amd::Error (the C ABI's status code and message).
#ifndef ANTIMATTER_AMD_AM_HPP_
#define ANTIMATTER_AMD_AM_HPP_

#include <array>
#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "antimatter_amd/am_api.h"

namespace antimatter_amd {

using Real3 = std::array<double, 3>;

class Error : public std::runtime_error {
 public:
  Error(int code, const std::string& what) : std::runtime_error(what), code_(code) {}
  int code() const { return code_; }

 private:
  int code_;
};

inline void Check(int rc) {
  if (rc != AM_OK) throw Error(rc, am_last_error());
}

// ---- owning handles ------------------------------------------------------------------------------
class Context {
 public:
  explicit Context(int device = 0) { Check(am_ctx_create(device, &h_)); }
  ~Context() { am_ctx_destroy(h_); }
  Context(const Context&) = delete;
  Context& operator=(const Context&) = delete;
  am_ctx* get() const { return h_; }

 private:
  am_ctx* h_ = nullptr;
};

class Store {
 public:
  Store(am_ctx* ctx, int64_t capacity) : capacity_(capacity) { Check(am_store_create(ctx, capacity, &h_)); }
  ~Store() { am_store_destroy(h_); }
  Store(const Store&) = delete;
  Store& operator=(const Store&) = delete;
  am_store* get() const { return h_; }
  int64_t capacity() const { return capacity_; }

 private:
  am_store* h_ = nullptr;
  int64_t capacity_ = 0;
};

// ---- the reference's free functions (one call per step from the driver) ----------------------------
// ComputeStandardCoulombForces (src/interaction/coulomb_interaction.h:10)
inline void ComputeStandardCoulombForces(am_store* s, double cutoff_radius, int precision = AM_FP64) {
  Check(am_compute_standard_coulomb_forces(s, cutoff_radius, precision));
}
// ComputeBarnesHutCoulombForces (src/interaction/barnes-hut.h:76)
inline void ComputeBarnesHutCoulombForces(am_store* s, double theta, int precision = AM_FP64) {
  Check(am_compute_barnes_hut_coulomb_forces(s, theta, precision));
}
// ComputeFastMultipoleCoulombForces (src/interaction/fast-multipole.h:73)
inline void ComputeFastMultipoleCoulombForces(am_store* s, int order, double theta, int precision = AM_FP64) {
  Check(am_compute_fast_multipole_coulomb_forces_ex(s, order, theta, precision));
}
// The exact all-pairs force: the theta = 0 limit of both tree solvers (barnes-hut.h:22-24).
inline void ComputeAllPairsCoulombForces(am_store* s, int precision = AM_FP32) {
  Check(am_compute_allpairs_coulomb_forces(s, precision));
}
// ComputeTotalEnergy (src/analysis/energy_analysis.h:35)
inline am_energy_breakdown ComputeTotalEnergy(am_store* s, const am_trap_field& field,
                                              const am_energy_settings& settings, am_energy_history* history = nullptr,
                                              int precision = AM_FP64) {
  am_energy_breakdown e;
  Check(am_compute_total_energy_ex(s, &field, &settings, history, precision, &e));
  return e;
}
// IsAntiparticlePair (src/interaction/annihilation.h:9)
inline bool IsAntiparticlePair(int a, int b) { return am_is_antiparticle_pair(a, b) != 0; }

// PenningTrapField::MakeFromGeometry (src/fields/penning_trap.cc:15-20)
inline am_trap_field MakePenningTrapField(double b0, double trap_voltage, double z0, double rho0) {
  am_trap_field f;
  Check(am_trap_make_from_geometry(b0, trap_voltage, z0, rho0, &f));
  return f;
}

// ---- behaviours: Run applies them to every alive particle of the store ------------------------------
// LorentzForce (src/forces/lorentz.h:21-34): dt fixed at construction, as the reference reads
// PhysicsParams::dt in its constructor; the first Run back-staggers each particle's velocity.
class LorentzForce {
 public:
  LorentzForce(const am_trap_field* field, double dt) : field_(field), dt_(dt) {}
  void Run(am_store* s) const { Check(am_lorentz_force(s, field_, dt_)); }

 private:
  const am_trap_field* field_;  // not owned, as in the reference
  double dt_;
};

// CyclotronCooling (src/forces/cooling.h:23-25); tau_scale <= 0 switches it off.
class CyclotronCooling {
 public:
  CyclotronCooling(const am_trap_field* field, double tau_scale, double bath_temp_K, double dt)
      : field_(field), tau_scale_(tau_scale), bath_temp_(bath_temp_K), dt_(dt) {}
  void Run(am_store* s, uint64_t seed, uint64_t step) const {
    Check(am_cyclotron_cooling(s, field_, tau_scale_, bath_temp_, dt_, seed, step));
  }

 private:
  const am_trap_field* field_;
  double tau_scale_, bath_temp_, dt_;
};

// AnnihilationInteraction (src/interaction/annihilation.h, .cc:21-61) in canonical order; returns
// the number of particles marked.
class AnnihilationInteraction {
 public:
  AnnihilationInteraction(double contact_radius, double probability)
      : contact_radius_(contact_radius), probability_(probability) {}
  int64_t Run(am_store* s, uint64_t seed, uint64_t step) const {
    int64_t marked = 0;
    Check(am_annihilation(s, contact_radius_, probability_, seed, step, &marked));
    return marked;
  }

 private:
  double contact_radius_, probability_;
};

// ---- standalone solvers over sources ----------------------------------------------------------------
struct Source {
  Real3 position{0.0, 0.0, 0.0};
  double charge = 0.0;
};

namespace detail {
// Uploads sources into a store (grown when needed); species are placeholders, charges are the sources'.
inline void UploadSources(am_ctx* ctx, const std::vector<Source>& sources, am_store** s, int64_t* capacity) {
  const int64_t n = (int64_t)sources.size();
  if (!*s || n > *capacity) {
    if (*s) am_store_destroy(*s);
    *s = nullptr;
    *capacity = n > 1 ? n : 1;
    Check(am_store_create(ctx, *capacity, s));
  }
  std::vector<double> pos(3 * n), vel(3 * n, 0.0), q(n);
  std::vector<uint8_t> species(n, AM_POSITRON);
  for (int64_t i = 0; i < n; ++i) {
    for (int d = 0; d < 3; ++d) pos[3 * i + d] = sources[i].position[d];
    q[i] = sources[i].charge;
  }
  Check(am_store_upload(*s, n, pos.data(), vel.data(), species.data(), q.data(), nullptr, nullptr, nullptr));
}
}  // namespace detail

// BarnesHutTree (src/interaction/barnes-hut.h:28-74): Build over sources, then forces and
// potentials at any points, one at a time as in the reference or batched.
class BarnesHutTree {
 public:
  static constexpr size_t kNoSelf = static_cast<size_t>(-1);

  BarnesHutTree(am_ctx* ctx, double theta, size_t leaf_capacity = 8, int precision = AM_FP64)
      : ctx_(ctx), theta_(theta), precision_(precision) {
    if (leaf_capacity != 8) throw Error(AM_ERR_INVALID, "BarnesHutTree: the device tree has leaf capacity 8");
  }
  ~BarnesHutTree() {
    if (store_) am_store_destroy(store_);
  }
  BarnesHutTree(const BarnesHutTree&) = delete;
  BarnesHutTree& operator=(const BarnesHutTree&) = delete;

  void Build(const std::vector<Source>& sources) {
    detail::UploadSources(ctx_, sources, &store_, &capacity_);
    n_ = sources.size();
  }
  Real3 ComputeForce(const Real3& position, double charge, size_t self_index = kNoSelf) const {
    std::vector<Real3> f;
    ComputeForces({position}, {charge}, {self_index}, &f);
    return f[0];
  }
  double ComputePotential(const Real3& position, size_t self_index = kNoSelf) const {
    std::vector<double> phi;
    ComputePotentials({position}, {self_index}, &phi);
    return phi[0];
  }
  // Batched forms: one tree walk per point, all points in one launch.
  void ComputeForces(const std::vector<Real3>& positions, const std::vector<double>& charges,
                     const std::vector<size_t>& self_index, std::vector<Real3>* forces) const {
    Evaluate(positions, &charges, self_index, forces, nullptr);
  }
  void ComputePotentials(const std::vector<Real3>& positions, const std::vector<size_t>& self_index,
                         std::vector<double>* phi) const {
    Evaluate(positions, nullptr, self_index, nullptr, phi);
  }
  size_t NumSources() const { return n_; }
  am_store* store() const { return store_; }

 private:
  void Evaluate(const std::vector<Real3>& positions, const std::vector<double>* charges,
                const std::vector<size_t>& self_index, std::vector<Real3>* forces, std::vector<double>* phi) const {
    const int64_t m = (int64_t)positions.size();
    std::vector<double> pos(3 * m), f(forces ? 3 * m : 0);
    std::vector<int64_t> self(m, -1);
    for (int64_t k = 0; k < m; ++k) {
      for (int d = 0; d < 3; ++d) pos[3 * k + d] = positions[k][d];
      if (k < (int64_t)self_index.size() && self_index[k] != kNoSelf) self[k] = (int64_t)self_index[k];
    }
    if (phi) phi->assign(m, 0.0);
    if (forces) forces->assign(m, Real3{0.0, 0.0, 0.0});
    if (!store_) return;
    Check(am_barnes_hut_evaluate(store_, theta_, precision_, m, pos.data(), charges ? charges->data() : nullptr,
                                 self.data(), forces ? f.data() : nullptr, phi ? phi->data() : nullptr));
    if (forces)
      for (int64_t k = 0; k < m; ++k) (*forces)[k] = Real3{f[3 * k], f[3 * k + 1], f[3 * k + 2]};
  }

  am_ctx* ctx_;
  double theta_;
  int precision_;
  am_store* store_ = nullptr;
  int64_t capacity_ = 0;
  size_t n_ = 0;
};

// FastMultipoleSolver (src/interaction/fast-multipole.h:46-71): forces on the sources themselves
// and the total potential energy U = k/2 sum_i q_i phi_i.
class FastMultipoleSolver {
 public:
  FastMultipoleSolver(am_ctx* ctx, int order, double theta, size_t leaf_capacity = 32, int precision = AM_FP64)
      : ctx_(ctx), order_(order), theta_(theta), precision_(precision) {
    if (leaf_capacity != 32) throw Error(AM_ERR_INVALID, "FastMultipoleSolver: the device tree has leaf capacity 32");
  }
  ~FastMultipoleSolver() {
    if (store_) am_store_destroy(store_);
  }
  FastMultipoleSolver(const FastMultipoleSolver&) = delete;
  FastMultipoleSolver& operator=(const FastMultipoleSolver&) = delete;

  void ComputeForces(const std::vector<Source>& sources, std::vector<Real3>* forces) {
    Compute(sources, forces, nullptr);
  }
  double ComputePotentialEnergy(const std::vector<Source>& sources) {
    double u = 0.0;
    Compute(sources, nullptr, &u);
    return u;
  }
  void Compute(const std::vector<Source>& sources, std::vector<Real3>* forces, double* potential_energy = nullptr) {
    const int64_t n = (int64_t)sources.size();
    if (forces) forces->assign(n, Real3{0.0, 0.0, 0.0});
    if (potential_energy) *potential_energy = 0.0;
    if (n == 0) return;
    detail::UploadSources(ctx_, sources, &store_, &capacity_);
    if (forces) {
      Check(am_compute_fast_multipole_coulomb_forces_ex(store_, order_, theta_, precision_));
      std::vector<double> f(3 * n);
      Check(am_store_download(store_, nullptr, nullptr, f.data(), nullptr, nullptr));
      for (int64_t i = 0; i < n; ++i) (*forces)[i] = Real3{f[3 * i], f[3 * i + 1], f[3 * i + 2]};
    }
    if (potential_energy) {
      const am_trap_field none{1.0, 0.0, 1.0};  // no trap: only the Coulomb part is read
      const am_energy_settings st{AM_METHOD_FAST_MULTIPOLE, 0.0, 0.5, theta_, (double)order_};
      am_energy_breakdown e;
      Check(am_compute_total_energy(store_, &none, &st, nullptr, &e));
      *potential_energy = e.energy_potential;
    }
  }

 private:
  am_ctx* ctx_;
  int order_;
  double theta_;
  int precision_;
  am_store* store_ = nullptr;
  int64_t capacity_ = 0;
};

}  // namespace antimatter_amd

#endif  // ANTIMATTER_AMD_AM_HPP_
