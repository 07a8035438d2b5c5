// AgentBridge (include/antimatter_amd/bdm_adapter.hpp) and the C++ layer (am.hpp) against direct
// FP64 sums computed here with the reference's pair term (barnes-hut.cc:19-27): prints
// "ADAPTER OK" and exits 0 when every check passes, else prints the failures and exits 1.
#include <cmath>
#include <cstdio>
#include <random>
#include <string>
#include <vector>

#include "antimatter_amd/am.hpp"
#include "antimatter_amd/bdm_adapter.hpp"
#include "mock_agents.hpp"

namespace {

constexpr double kE = 1.602176634e-19, kMe = 9.1093837015e-31, kK = 8.9875517923e9;
int failures = 0;

void Expect(bool ok, const std::string& what) {
  if (!ok) {
    std::printf("FAIL: %s\n", what.c_str());
    ++failures;
  }
}

using Bridge = antimatter_amd::AgentBridge<mock::ResourceManager, mock::Agent, mock::Particle, mock::Real3>;

mock::Particle* Cast(mock::Agent* a) { return dynamic_cast<mock::Particle*>(a); }

// Direct FP64 force on particle i from the others (alive, within r_c if r_c > 0).
mock::Real3 Direct(const std::vector<mock::Particle*>& ps, size_t i, double rc = 0.0) {
  mock::Real3 f{0.0, 0.0, 0.0};
  const mock::Particle* t = ps[i];
  for (size_t j = 0; j < ps.size(); ++j) {
    if (j == i || ps[j]->IsAnnihilated() || ps[j]->GetCharge() == 0.0) continue;
    double r[3], d2 = 0.0;
    for (int d = 0; d < 3; ++d) {
      r[d] = t->GetPosition()[d] - ps[j]->GetPosition()[d];
      d2 += r[d] * r[d];
    }
    if (d2 < 1e-30 || (rc > 0.0 && !(d2 < rc * rc))) continue;
    const double c = kK * t->GetCharge() * ps[j]->GetCharge() / d2 / std::sqrt(d2);
    for (int d = 0; d < 3; ++d) f[d] += r[d] * c;
  }
  return f;
}

double RelErr(const std::vector<mock::Particle*>& ps, double rc, const mock::Real3& sentinel) {
  double num = 0.0, den = 0.0;
  for (size_t i = 0; i < ps.size(); ++i) {
    if (ps[i]->IsAnnihilated()) {
      if (ps[i]->GetCoulombForce() != sentinel) return 1e300;  // annihilated agents are left alone
      continue;
    }
    const mock::Real3 want = Direct(ps, i, rc), got = ps[i]->GetCoulombForce();
    for (int d = 0; d < 3; ++d) {
      num += (got[d] - want[d]) * (got[d] - want[d]);
      den += want[d] * want[d];
    }
  }
  return std::sqrt(num / den);
}

void AddCloud(mock::ResourceManager* rm, std::vector<mock::Particle*>* ps, int n, unsigned seed, double size) {
  std::mt19937 rng(seed);
  std::uniform_real_distribution<double> u(-size, size);
  for (int i = 0; i < n; ++i) {
    const bool pos = (i & 1) == 0;
    auto* p = new mock::Particle({u(rng), u(rng), u(rng)}, pos ? mock::Species::kPositron : mock::Species::kElectron,
                                 pos ? kE : -kE, kMe, {1e3, 0.0, 0.0});
    rm->AddAgent(p);
    ps->push_back(p);
    if (i % 7 == 0) rm->AddAgent(new mock::Wall());
  }
}

}  // namespace

int main() {
  try {
    mock::ResourceManager rm;
    std::vector<mock::Particle*> ps;
    AddCloud(&rm, &ps, 2000, 1, 1e-4);
    const mock::Real3 sentinel{7.0, 7.0, 7.0};
    for (size_t i = 0; i < ps.size(); ++i) {
      if (i % 23 == 0) ps[i]->MarkAnnihilated();
      ps[i]->SetCoulombForce(sentinel);
    }
    Bridge bridge(Cast);
    bridge.ComputeBarnesHutCoulombForces(&rm, 0.0);  // theta = 0: the exact sum
    Expect(bridge.agents().size() == ps.size(), "the cast keeps particles only");
    Expect(RelErr(ps, 0.0, sentinel) < 1e-12, "BarnesHut(theta=0) through the bridge = direct sum");
    bridge.ComputeFastMultipoleCoulombForces(&rm, 4, 0.0);
    Expect(RelErr(ps, 0.0, sentinel) < 1e-10, "FastMultipole(theta=0) through the bridge = direct sum");
    bridge.ComputeStandardCoulombForces(&rm, 2e-5);
    Expect(RelErr(ps, 2e-5, sentinel) < 1e-12, "Standard(r_c) through the bridge = direct cutoff sum");
    bridge.ComputeBarnesHutCoulombForces(&rm, 0.6);
    Expect(RelErr(ps, 0.0, sentinel) < 5e-2 && RelErr(ps, 0.0, sentinel) > 1e-8, "BarnesHut(theta=0.6) approximates");

    // More agents than the store holds: the bridge grows it.
    const int64_t cap0 = bridge.capacity();
    AddCloud(&rm, &ps, 1500, 2, 1e-4);
    for (auto* p : ps)
      if (p->IsAnnihilated()) p->SetCoulombForce(sentinel);
    bridge.ComputeBarnesHutCoulombForces(&rm, 0.0);
    Expect(bridge.capacity() > cap0 && bridge.agents().size() == 3500, "the store grows with the agents");
    Expect(RelErr(ps, 0.0, sentinel) < 1e-12, "forces after growth");

    // A whole step on the device: touching e+/e- pairs annihilate, everyone moves.
    mock::ResourceManager rm2;
    std::vector<mock::Particle*> ps2;
    for (int i = 0; i < 50; ++i) {
      const double x = 1e-4 * i;
      auto* a = new mock::Particle({x, 0.0, 0.0}, mock::Species::kPositron, kE, kMe);
      auto* b = new mock::Particle({x + 1e-7, 0.0, 0.0}, mock::Species::kElectron, -kE, kMe);
      rm2.AddAgent(a);
      rm2.AddAgent(b);
      ps2.push_back(a);
      ps2.push_back(b);
    }
    Bridge b2(Cast);
    b2.Sync(&rm2);
    const am_trap_field field = antimatter_amd::MakePenningTrapField(1.0, 10.0, 0.01, 0.01);
    b2.Step([](am_store* s) { antimatter_amd::ComputeBarnesHutCoulombForces(s, 0.5); }, field, 8.9e-12, 0.0, 0.0,
            true, 5e-6, 1.0, 42, 1);
    b2.ScatterState();
    int gone = 0, moved = 0;
    for (size_t k = 0; k < ps2.size(); ++k) {
      gone += ps2[k]->IsAnnihilated() ? 1 : 0;
      const double x0 = 1e-4 * (double)(k / 2) + (k % 2 ? 1e-7 : 0.0);
      moved += (ps2[k]->GetPosition()[0] != x0) ? 1 : 0;
    }
    Expect(gone == 100, "every touching pair annihilates at p = 1 (" + std::to_string(gone) + ")");
    Expect(moved > 0, "the push moved the agents");

    // Standalone solvers over sources.
    antimatter_amd::Context ctx(0);
    std::vector<antimatter_amd::Source> src;
    for (size_t i = 0; i < 600; ++i)
      src.push_back({ps[i]->GetPosition(), ps[i]->IsAnnihilated() ? 0.0 : ps[i]->GetCharge()});
    antimatter_amd::BarnesHutTree tree(ctx.get(), 0.0);
    tree.Build(src);
    double num = 0.0, den = 0.0, unum = 0.0, uden = 0.0, u_direct = 0.0;
    for (size_t i = 0; i < 600; ++i) {
      const mock::Real3 got = tree.ComputeForce(src[i].position, src[i].charge, i);
      mock::Real3 want{0.0, 0.0, 0.0};
      double phi = 0.0;
      for (size_t j = 0; j < 600; ++j) {
        if (j == i || src[j].charge == 0.0) continue;
        double r[3], d2 = 0.0;
        for (int d = 0; d < 3; ++d) {
          r[d] = src[i].position[d] - src[j].position[d];
          d2 += r[d] * r[d];
        }
        const double c = kK * src[i].charge * src[j].charge / d2 / std::sqrt(d2);
        for (int d = 0; d < 3; ++d) want[d] += r[d] * c;
        phi += src[j].charge / std::sqrt(d2);
      }
      for (int d = 0; d < 3; ++d) {
        num += (got[d] - want[d]) * (got[d] - want[d]);
        den += want[d] * want[d];
      }
      const double gphi = tree.ComputePotential(src[i].position, i);
      unum += (gphi - phi) * (gphi - phi);
      uden += phi * phi;
      u_direct += 0.5 * kK * src[i].charge * phi;
    }
    Expect(std::sqrt(num / den) < 1e-12, "BarnesHutTree::ComputeForce at theta = 0");
    Expect(std::sqrt(unum / uden) < 1e-12, "BarnesHutTree::ComputePotential at theta = 0");
    antimatter_amd::FastMultipoleSolver fmm(ctx.get(), 4, 0.0);
    std::vector<antimatter_amd::Real3> ff;
    double u = 0.0;
    fmm.Compute(src, &ff, &u);
    Expect(std::fabs(u - u_direct) < 1e-10 * std::fabs(u_direct), "FastMultipoleSolver energy at theta = 0");
    Expect(ff.size() == 600 && std::fabs(ff[5][0] - tree.ComputeForce(src[5].position, src[5].charge, 5)[0]) <
                                   1e-10 * std::fabs(ff[5][0]) + 1e-40,
           "FastMultipoleSolver forces at theta = 0");
    bool threw = false;
    try {
      antimatter_amd::BarnesHutTree bad(ctx.get(), 0.5, 4);
    } catch (const antimatter_amd::Error& e) {
      threw = e.code() == AM_ERR_INVALID;
    }
    Expect(threw, "an unsupported leaf capacity is refused");
  } catch (const std::exception& e) {
    std::printf("FAIL: exception %s\n", e.what());
    return 1;
  }
  if (failures) return 1;
  std::printf("ADAPTER OK\n");
  return 0;
}
