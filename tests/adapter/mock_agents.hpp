// A minimal stand-in for the BioDynaMo agent API that AgentBridge uses (test fixture): a
// ResourceManager owning agents, a particle agent with the ChargedAntiparticle accessors, and a
// second agent type that the bridge's cast must skip.  It is written for the adapter's tests and
// holds no BioDynaMo or reference code.
#ifndef AM_TESTS_MOCK_AGENTS_HPP_
#define AM_TESTS_MOCK_AGENTS_HPP_

#include <array>
#include <cstdint>
#include <memory>
#include <vector>

namespace mock {

using Real3 = std::array<double, 3>;

enum class Species : uint8_t { kAntiproton = 0, kPositron = 1, kAntihydrogen = 2, kElectron = 3, kProton = 4 };

class Agent {
 public:
  virtual ~Agent() = default;
};

class Particle : public Agent {
 public:
  Particle(const Real3& x, Species s, double charge, double mass, const Real3& v = {0.0, 0.0, 0.0})
      : x_(x), v_(v), q_(charge), m_(mass), s_(s) {}
  const Real3& GetPosition() const { return x_; }
  void SetPosition(const Real3& x) { x_ = x; }
  const Real3& GetVelocity() const { return v_; }
  void SetVelocity(const Real3& v) { v_ = v; }
  double GetCharge() const { return q_; }
  double GetMass() const { return m_; }
  Species GetSpecies() const { return s_; }
  bool IsAnnihilated() const { return gone_; }
  void MarkAnnihilated() { gone_ = true; }
  const Real3& GetCoulombForce() const { return f_; }
  void SetCoulombForce(const Real3& f) { f_ = f; }

 private:
  Real3 x_, v_, f_{0.0, 0.0, 0.0};
  double q_, m_;
  Species s_;
  bool gone_ = false;
};

class Wall : public Agent {};  // some other agent type in the same simulation

class ResourceManager {
 public:
  void AddAgent(Agent* a) { agents_.emplace_back(a); }
  template <class F>
  void ForEachAgent(F&& f) {
    for (auto& a : agents_) f(a.get());
  }
  size_t GetNumAgents() const { return agents_.size(); }

 private:
  std::vector<std::unique_ptr<Agent>> agents_;
};

}  // namespace mock

#endif  // AM_TESTS_MOCK_AGENTS_HPP_
