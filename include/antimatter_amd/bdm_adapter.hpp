// antimatter-kernel for MI355X — adapter from the reference's BioDynaMo agents to the device store.
//
// The reference keeps its particles as ChargedAntiparticle agents in a BioDynaMo ResourceManager
// and calls ComputeStandard/BarnesHut/FastMultipoleCoulombForces(rm, ...) once per step
// (src/antimatter-kernel.cc:155-160).  AgentBridge gathers the agents into the device store,
// runs a solver, and writes the Coulomb forces back with SetCoulombForce, so those three call sites
// can be pointed at it unchanged.  Step() also moves the per-agent behaviours (Lorentz push with
// cooling, annihilation) onto the device, and ScatterState() writes the state back when the agents
// are needed on the host.
//
// It is a template over the resource manager and particle types, so it compiles against BioDynaMo's
// real classes (ResourceManager, ChargedAntiparticle, with FastCast as the cast) and against the mock
// the tests build.  What it uses:
//   rm->ForEachAgent(callable(Agent*));  Particle: GetPosition(), GetVelocity(), GetCharge(),
//   GetMass(), GetSpecies(), IsAnnihilated(), MarkAnnihilated(), SetCoulombForce(Real3),
//   SetPosition(Real3), SetVelocity(Real3); Real3 indexable [0..2] and brace-constructible.
#ifndef ANTIMATTER_AMD_BDM_ADAPTER_HPP_
#define ANTIMATTER_AMD_BDM_ADAPTER_HPP_

#include <cstdint>
#include <functional>
#include <vector>

#include "antimatter_amd/am.hpp"

namespace antimatter_amd {

template <class ResourceManager, class Agent, class Particle, class Real3T>
class AgentBridge {
 public:
  using Cast = std::function<Particle*(Agent*)>;

  explicit AgentBridge(Cast cast, int device = 0, int precision = AM_FP64) : cast_(cast), precision_(precision) {
    Check(am_ctx_create(device, &ctx_));
  }
  ~AgentBridge() {
    if (store_) am_store_destroy(store_);
    am_ctx_destroy(ctx_);
  }
  AgentBridge(const AgentBridge&) = delete;
  AgentBridge& operator=(const AgentBridge&) = delete;

  // Agents -> device: every particle in ForEachAgent order (annihilated ones flagged).  The store is
  // recreated with room to spare when the agents outgrow it.
  void Sync(ResourceManager* rm) {
    agents_.clear();
    rm->ForEachAgent([&](Agent* a) {
      if (Particle* p = cast_(a)) agents_.push_back(p);
    });
    const int64_t n = (int64_t)agents_.size();
    if (!store_ || n > capacity_) {
      if (store_) am_store_destroy(store_);
      store_ = nullptr;
      capacity_ = n + n / 4 + 16;
      Check(am_store_create(ctx_, capacity_, &store_));
    }
    std::vector<double> x(3 * n), v(3 * n), q(n), m(n);
    std::vector<uint8_t> species(n), gone(n);
    for (int64_t i = 0; i < n; ++i) {
      const Particle* p = agents_[i];
      for (int d = 0; d < 3; ++d) {
        x[3 * i + d] = p->GetPosition()[d];
        v[3 * i + d] = p->GetVelocity()[d];
      }
      q[i] = p->GetCharge();
      m[i] = p->GetMass();
      species[i] = (uint8_t)p->GetSpecies();
      gone[i] = p->IsAnnihilated() ? 1 : 0;
    }
    Check(am_store_upload(store_, n, x.data(), v.data(), species.data(), q.data(), m.data(), gone.data(), nullptr));
  }

  // Device -> agents: SetCoulombForce on every particle still in the simulation (the solvers write
  // only those, as the reference's gathers skip annihilated agents).
  void ScatterForces() {
    const int64_t n = (int64_t)agents_.size();
    std::vector<double> f(3 * n);
    if (n > 0) Check(am_store_download(store_, nullptr, nullptr, f.data(), nullptr, nullptr));
    for (int64_t i = 0; i < n; ++i)
      if (!agents_[i]->IsAnnihilated()) agents_[i]->SetCoulombForce(Real3T{f[3 * i], f[3 * i + 1], f[3 * i + 2]});
  }

  // Device -> agents after Step(): positions, velocities, and MarkAnnihilated for particles the
  // device annihilated.
  void ScatterState() {
    const int64_t n = (int64_t)agents_.size();
    if (n == 0) return;
    std::vector<double> x(3 * n), v(3 * n);
    std::vector<uint8_t> gone(n);
    Check(am_store_download(store_, x.data(), v.data(), nullptr, gone.data(), nullptr));
    for (int64_t i = 0; i < n; ++i) {
      Particle* p = agents_[i];
      p->SetPosition(Real3T{x[3 * i], x[3 * i + 1], x[3 * i + 2]});
      p->SetVelocity(Real3T{v[3 * i], v[3 * i + 1], v[3 * i + 2]});
      if (gone[i] && !p->IsAnnihilated()) p->MarkAnnihilated();
    }
  }

  // The reference's free functions (coulomb_interaction.h:10, barnes-hut.h:76, fast-multipole.h:73).
  void ComputeStandardCoulombForces(ResourceManager* rm, double cutoff_radius) {
    Sync(rm);
    antimatter_amd::ComputeStandardCoulombForces(store_, cutoff_radius, precision_);
    ScatterForces();
  }
  void ComputeBarnesHutCoulombForces(ResourceManager* rm, double theta) {
    Sync(rm);
    antimatter_amd::ComputeBarnesHutCoulombForces(store_, theta, precision_);
    ScatterForces();
  }
  void ComputeFastMultipoleCoulombForces(ResourceManager* rm, int order, double theta) {
    Sync(rm);
    antimatter_amd::ComputeFastMultipoleCoulombForces(store_, order, theta, precision_);
    ScatterForces();
  }

  // One whole step on the device, in the reference's order: the solver (`solver` receives the
  // store), the Lorentz push with cooling, annihilation.  Call Sync() before the first step (and
  // whenever the agents change on the host); the Lorentz back-stagger state stays on the device.
  void Step(const std::function<void(am_store*)>& solver, const am_trap_field& field, double dt, double tau_scale,
            double bath_temp, bool annihilation, double contact_radius, double probability, uint64_t seed,
            uint64_t step) {
    solver(store_);
    Check(am_push(store_, &field, dt, tau_scale, bath_temp, seed, step));
    if (annihilation) AnnihilationInteraction(contact_radius, probability).Run(store_, seed, step);
  }

  am_ctx* ctx() const { return ctx_; }
  am_store* store() const { return store_; }
  int64_t capacity() const { return capacity_; }
  const std::vector<Particle*>& agents() const { return agents_; }

 private:
  Cast cast_;
  int precision_;
  am_ctx* ctx_ = nullptr;
  am_store* store_ = nullptr;
  int64_t capacity_ = 0;
  std::vector<Particle*> agents_;
};

}  // namespace antimatter_amd

#endif  // ANTIMATTER_AMD_BDM_ADAPTER_HPP_
