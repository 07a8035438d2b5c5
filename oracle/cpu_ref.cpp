// ORACLE — TEST INFRASTRUCTURE ONLY.
//
// FP64 CPU restatement of the antimatter-kernel per-step hot path, written from the
// equations in the reference sources (cited file:line, paths relative to the reference
// repository root).  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
// may load this library, and only as the checker or as the timed CPU baseline; the
// product library (antimatter-kernel_amd/) never links or calls it.
//
// Pinning (what checks each function against something other than itself):
//   orc_set_particle_positions / orc_set_particle_velocities / orc_derive_dt
//       -> notebook rows 0-4 of the shipped params.json run (simulation_analysis.ipynb:161-172)
//   orc_lorentz_push, orc_cyclotron_cooling (deterministic part), orc_standard_forces,
//   orc_energy (standard), orc_annihilate (p=1), orc_simulate
//       -> same notebook rows, the 20-step drift windows (:438-442), final drift (:367)
//          and the 200 -> 170 active count (:534); plus closed forms in tests
//          (|v| conserved under pure B, rotation angle 2 atan(w dt/2), tau_E = 2.579 s).
//   orc_allpairs_forces / orc_allpairs_potential
//       -> brute-force two-body and three-body closed forms, Newton's third law, and
//          equality with orc_standard_forces at a cutoff larger than the cloud.
//   orc_bh_* -> theta = 0 equals all-pairs (barnes-hut.h:22-24); node charge sums equal
//          totals; leaf membership equals an independent recursive partition in tests; the far
//          field at theta > 0 (MAC from the geometric centre, width 2h, both class monopoles)
//          equals hand sums on a cloud whose accept/open decisions are fixed by construction,
//          and copies of this file with each of those three rules broken fail that test.
//   orc_fmm_* -> theta = 0 equals all-pairs (fast-multipole.h:42-44); D_alpha closed
//          forms for low orders; geometric error decay in P.
//   orc_philox4x32_10 -> Random123 known-answer vectors.
//   Cooling noise statistics and annihilation at p<1 use this build's counter-based RNG in
//   place of BioDynaMo's Random (unbuildable here): "parity unpinned" against the
//   reference stream; pinned only by statistics and GPU/CPU agreement.

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <random>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// src/utils/constants.h:11-38
constexpr double kPi = 3.14159265358979323846;
constexpr double kE = 1.602176634e-19;
constexpr double kMe = 9.1093837015e-31;
constexpr double kMp = 1.67262192369e-27;
constexpr double kMn = 1.67492749804e-27;
constexpr double kCoulomb = 8.9875517923e9;
constexpr double kBoltzmann = 1.380649e-23;
constexpr double kLight = 2.99792458e8;
constexpr double kMinSq = 1e-30;  // barnes-hut.cc:22, coulomb_interaction.cc:19, fast-multipole.cc:185

// src/agents/particle_species.h:9-46 (enum order; mass, charge)
void SpeciesProps(int s, double* m, double* q) {
  switch (s) {
    case 0: *m = kMp; *q = -kE; return;             // antiproton
    case 1: *m = kMe; *q = +kE; return;             // positron
    case 2: *m = kMp + kMe; *q = 0.0; return;       // antihydrogen
    case 3: *m = kMe; *q = -kE; return;             // electron
    case 4: *m = kMp; *q = +kE; return;             // proton
    case 5: *m = 2 * (kMp + kMn); *q = 0.0; return; // antihelium
    default: *m = 0.0; *q = 0.0; return;
  }
}

struct V3 { double x, y, z; };
inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 scl(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline V3 ld(const double* p, int64_t i) { return {p[3 * i], p[3 * i + 1], p[3 * i + 2]}; }
inline void st(double* p, int64_t i, V3 v) { p[3 * i] = v.x; p[3 * i + 1] = v.y; p[3 * i + 2] = v.z; }

// ---------------------------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel Random Numbers: As Easy as 1, 2, 3",
// SC'11).  Counter-based RNG shared bit-for-bit with the GPU kernels.
inline void Philox(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
  uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
  uint32_t k0 = key_in[0], k1 = key_in[1];
  for (int r = 0; r < 10; ++r) {
    if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
    const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    c0 = hi1 ^ c1 ^ k0; c1 = lo1; c2 = hi0 ^ c3 ^ k1; c3 = lo0;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}
inline double U24(uint32_t w) { return (double)(w >> 8) * (1.0 / 16777216.0); }
// Draw for the annihilation test of candidate j scanned by i at `step` (purpose word 1).
inline double AnnihilationUniform(uint64_t seed, uint64_t step, uint32_t i, uint32_t j) {
  const uint32_t ctr[4] = {(uint32_t)step, i, j, 1u};
  const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  uint32_t o[4];
  Philox(ctr, key, o);
  return U24(o[0]);
}
// Three standard normals for particle i at `step` (purpose word 2), Box-Muller.
inline void CoolingNormals(uint64_t seed, uint64_t step, uint32_t i, double n[3]) {
  const uint32_t ctr[4] = {(uint32_t)step, i, 0u, 2u};
  const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  uint32_t o[4];
  Philox(ctr, key, o);
  const double u1 = ((double)(o[0] >> 8) + 0.5) * (1.0 / 16777216.0);
  const double u2 = (double)(o[1] >> 8) * (1.0 / 16777216.0);
  const double u3 = ((double)(o[2] >> 8) + 0.5) * (1.0 / 16777216.0);
  const double u4 = (double)(o[3] >> 8) * (1.0 / 16777216.0);
  const double r1 = std::sqrt(-2.0 * std::log(u1));
  const double r2 = std::sqrt(-2.0 * std::log(u3));
  n[0] = r1 * std::cos(2.0 * kPi * u2);
  n[1] = r1 * std::sin(2.0 * kPi * u2);
  n[2] = r2 * std::cos(2.0 * kPi * u4);
}

// ---------------------------------------------------------------------------------------------
// Penning trap field (src/fields/penning_trap.cc:15-48).  The characteristic dimension is
// stored as d = sqrt(z0^2/2 + rho0^2/4) and re-squared on use, as the reference does.
struct Trap { double b0, u0, d; };
inline Trap MakeTrap(double b0, double u0, double z0, double rho0) {
  return {b0, u0, std::sqrt((z0 * z0) / 2.0 + (rho0 * rho0) / 4.0)};
}
inline V3 EField(const Trap& t, V3 x) {
  const double d2 = t.d * t.d;
  const double radial = t.u0 / (2.0 * d2);
  const double axial = t.u0 / d2;
  return {radial * x.x, radial * x.y, -axial * x.z};
}
inline double Potential(const Trap& t, V3 x) {
  const double rho2 = x.x * x.x + x.y * x.y;
  const double d2 = t.d * t.d;
  return (t.u0 / (4.0 * d2)) * (2.0 * x.z * x.z - rho2);
}

// Boris velocity update (src/forces/lorentz.cc:18-34), B = (0, 0, b0).
inline V3 Boris(V3 v, V3 f, V3 b, double q, double m, double dt) {
  const double h = 0.5 * dt;
  const V3 kick = scl(f, h / m);
  const V3 vm = add(v, kick);
  const V3 t = scl(b, q * h / m);
  const V3 vp = add(vm, cross(vm, t));
  const double t2 = dot(t, t);
  const V3 s = scl(t, 2.0 / (1.0 + t2));
  const V3 vplus = add(vm, cross(vp, s));
  return add(vplus, kick);
}

// src/forces/cooling.cc:16-22
inline double CoolingTime(double q, double m, double b) {
  const double q2 = q * q;
  const double denom = q2 * q2 * b * b;
  if (denom <= 0.0) return INFINITY;
  return (0.75 / kCoulomb) * m * m * m * kLight * kLight * kLight / denom;
}

// Coulomb pair term as in barnes-hut.cc:19-27 (AddPointCharge).
inline void PointCharge(V3 t, double qt, V3 s, double qs, V3* f) {
  const V3 r = sub(t, s);
  const double d2 = dot(r, r);
  if (d2 < kMinSq) return;
  const double inv = 1.0 / std::sqrt(d2);
  const double mag = kCoulomb * qt * qs / d2;
  *f = add(*f, scl(r, mag * inv));
}

// ---------------------------------------------------------------------------------------------
// Barnes-Hut octree (src/interaction/barnes-hut.cc:30-269), restated.
struct BHNode {
  V3 center{0, 0, 0};
  double half = 0;
  double qp = 0, qn = 0;  // sum of positive charges, sum of |negative charges|
  V3 cp{0, 0, 0}, cn{0, 0, 0};
  int child[8] = {-1, -1, -1, -1, -1, -1, -1, -1};
  int64_t begin = 0, end = 0;
  int depth = 0;
  bool leaf = true;
};
struct BHTree {
  double theta2 = 0.25;
  int64_t leaf_cap = 8;
  std::vector<V3> pos;
  std::vector<double> q;
  std::vector<int64_t> order, scratch;
  std::vector<BHNode> nodes;

  static int Octant(V3 p, V3 c) {  // barnes-hut.cc:30-33
    return (p.x >= c.x ? 1 : 0) | (p.y >= c.y ? 2 : 0) | (p.z >= c.z ? 4 : 0);
  }
  void Build() {  // barnes-hut.cc:38-66
    const int64_t n = (int64_t)pos.size();
    nodes.clear();
    order.resize(n);
    scratch.resize(n);
    for (int64_t i = 0; i < n; ++i) order[i] = i;
    if (n == 0) return;
    V3 lo = pos[0], hi = pos[0];
    for (const V3& p : pos) {
      lo.x = std::min(lo.x, p.x); lo.y = std::min(lo.y, p.y); lo.z = std::min(lo.z, p.z);
      hi.x = std::max(hi.x, p.x); hi.y = std::max(hi.y, p.y); hi.z = std::max(hi.z, p.z);
    }
    const V3 c{0.5 * (lo.x + hi.x), 0.5 * (lo.y + hi.y), 0.5 * (lo.z + hi.z)};
    const double ext = std::max({hi.x - lo.x, hi.y - lo.y, hi.z - lo.z});
    double half = 0.5 * ext * (1.0 + 1e-9);
    if (half < 1e-30) half = 1e-30;
    nodes.reserve(n);
    BuildNode(0, n, c, half, 0);
  }
  int BuildNode(int64_t b, int64_t e, V3 c, double half, int depth) {  // barnes-hut.cc:68-164
    const int idx = (int)nodes.size();
    nodes.emplace_back();
    nodes[idx].center = c;
    nodes[idx].half = half;
    nodes[idx].begin = b;
    nodes[idx].end = e;
    nodes[idx].depth = depth;
    double qp = 0, qn = 0;
    V3 mp{0, 0, 0}, mn{0, 0, 0};
    if (e - b <= leaf_cap || depth >= 32) {
      nodes[idx].leaf = true;
      for (int64_t k = b; k < e; ++k) {
        const int64_t j = order[k];
        if (q[j] > 0.0) { qp += q[j]; mp = add(mp, scl(pos[j], q[j])); }
        else if (q[j] < 0.0) { qn += -q[j]; mn = add(mn, scl(pos[j], -q[j])); }
      }
    } else {
      nodes[idx].leaf = false;
      int64_t cnt[8] = {0}, off[9], cur[8];
      for (int64_t k = b; k < e; ++k) ++cnt[Octant(pos[order[k]], c)];
      off[0] = b;
      for (int o = 0; o < 8; ++o) off[o + 1] = off[o] + cnt[o];
      for (int o = 0; o < 8; ++o) cur[o] = off[o];
      for (int64_t k = b; k < e; ++k) scratch[cur[Octant(pos[order[k]], c)]++] = order[k];
      std::copy(scratch.begin() + b, scratch.begin() + e, order.begin() + b);
      const double ch = 0.5 * half;
      for (int o = 0; o < 8; ++o) {
        if (cnt[o] == 0) continue;
        const V3 cc{c.x + ((o & 1) ? ch : -ch), c.y + ((o & 2) ? ch : -ch), c.z + ((o & 4) ? ch : -ch)};
        const int child = BuildNode(off[o], off[o + 1], cc, ch, depth + 1);
        nodes[idx].child[o] = child;
      }
      for (int o = 0; o < 8; ++o) {
        const int ci = nodes[idx].child[o];
        if (ci < 0) continue;
        const BHNode& k = nodes[ci];
        if (k.qp > 0.0) { qp += k.qp; mp = add(mp, scl(k.cp, k.qp)); }
        if (k.qn > 0.0) { qn += k.qn; mn = add(mn, scl(k.cn, k.qn)); }
      }
    }
    BHNode& nd = nodes[idx];
    nd.qp = qp;
    nd.qn = qn;
    nd.cp = qp > 0.0 ? scl(mp, 1.0 / qp) : c;
    nd.cn = qn > 0.0 ? scl(mn, 1.0 / qn) : c;
    return idx;
  }
  void Accumulate(int ni, V3 x, double qt, int64_t self, V3* f, int64_t* count) const {  // :183-221
    const BHNode& nd = nodes[ni];
    if (nd.qp == 0.0 && nd.qn == 0.0) return;
    if (nd.leaf) {
      for (int64_t k = nd.begin; k < nd.end; ++k) {
        const int64_t j = order[k];
        if (j == self) continue;
        if (q[j] == 0.0) continue;
        PointCharge(x, qt, pos[j], q[j], f);
        ++*count;
      }
      return;
    }
    const V3 d = sub(x, nd.center);
    const double d2 = dot(d, d);
    const double w = 2.0 * nd.half;
    if (w * w < theta2 * d2) {
      if (nd.qp > 0.0) { PointCharge(x, qt, nd.cp, nd.qp, f); ++*count; }
      if (nd.qn > 0.0) { PointCharge(x, qt, nd.cn, -nd.qn, f); ++*count; }
      return;
    }
    for (int o = 0; o < 8; ++o)
      if (nd.child[o] >= 0) Accumulate(nd.child[o], x, qt, self, f, count);
  }
  void AccumulatePotential(int ni, V3 x, int64_t self, double* phi) const {  // :223-269
    const BHNode& nd = nodes[ni];
    if (nd.qp == 0.0 && nd.qn == 0.0) return;
    if (nd.leaf) {
      for (int64_t k = nd.begin; k < nd.end; ++k) {
        const int64_t j = order[k];
        if (j == self || q[j] == 0.0) continue;
        const V3 r = sub(x, pos[j]);
        const double d2 = dot(r, r);
        if (d2 < kMinSq) continue;
        *phi += q[j] / std::sqrt(d2);
      }
      return;
    }
    const V3 d = sub(x, nd.center);
    const double d2 = dot(d, d);
    const double w = 2.0 * nd.half;
    if (w * w < theta2 * d2) {
      if (nd.qp > 0.0) {
        const V3 r = sub(x, nd.cp);
        const double r2 = dot(r, r);
        if (r2 >= kMinSq) *phi += nd.qp / std::sqrt(r2);
      }
      if (nd.qn > 0.0) {
        const V3 r = sub(x, nd.cn);
        const double r2 = dot(r, r);
        if (r2 >= kMinSq) *phi -= nd.qn / std::sqrt(r2);
      }
      return;
    }
    for (int o = 0; o < 8; ++o)
      if (nd.child[o] >= 0) AccumulatePotential(nd.child[o], x, self, phi);
  }
};

// Gather of non-annihilated particles (barnes-hut.cc:273-287): returns original indices.
std::vector<int64_t> Gather(int64_t n, const uint8_t* annihilated) {
  std::vector<int64_t> idx;
  idx.reserve(n);
  for (int64_t i = 0; i < n; ++i)
    if (!annihilated || !annihilated[i]) idx.push_back(i);
  return idx;
}

// ---------------------------------------------------------------------------------------------
// Cartesian Taylor FMM (src/interaction/fast-multipole.cc:17-682), restated.
struct FmmTab {
  int P = 0, nm = 0, nd = 0, size = 0;
  std::vector<std::array<int, 3>> multi;
  std::vector<int> lookup;
  std::vector<double> fact, inv_fact, sign;
  struct Rec { int axis = 0; int delta[3] = {0, 0, 0}; int m1[3] = {-1, -1, -1}; int m2[3] = {-1, -1, -1}; };
  std::vector<Rec> rec;
  std::vector<int> m2l;                      // [g * nm + a] -> index of alpha + gamma
  std::vector<std::array<int, 3>> m2m, l2l;  // (a, b, c)
  std::vector<std::array<int, 3>> grad;
  int Idx(int a, int b, int c) const {
    if (a < 0 || b < 0 || c < 0 || a >= size || b >= size || c >= size) return -1;
    return lookup[(a * size + b) * size + c];
  }
  explicit FmmTab(int p) : P(p), size(2 * p + 1) {
    // Multi-indices of degree <= 2P, ordered by degree, then a descending, then b descending
    // (fast-multipole.cc:78-95).
    lookup.assign((size_t)size * size * size, -1);
    for (int deg = 0; deg <= 2 * p; ++deg)
      for (int a = deg; a >= 0; --a)
        for (int b = deg - a; b >= 0; --b) {
          lookup[(a * size + b) * size + (deg - a - b)] = (int)multi.size();
          multi.push_back({a, b, deg - a - b});
        }
    nd = (int)multi.size();
    nm = (p + 1) * (p + 2) * (p + 3) / 6;
    std::vector<double> f(2 * p + 1, 1.0);
    for (int i = 1; i <= 2 * p; ++i) f[i] = f[i - 1] * i;
    fact.resize(nd); inv_fact.resize(nd); sign.resize(nd);
    for (int k = 0; k < nd; ++k) {
      const auto& m = multi[k];
      fact[k] = f[m[0]] * f[m[1]] * f[m[2]];
      inv_fact[k] = 1.0 / fact[k];
      sign[k] = ((m[0] + m[1] + m[2]) % 2 == 0) ? 1.0 : -1.0;
    }
    rec.resize(nd);  // fast-multipole.cc:113-134
    for (int k = 1; k < nd; ++k) {
      const auto& m = multi[k];
      Rec& r = rec[k];
      r.axis = m[0] > 0 ? 0 : (m[1] > 0 ? 1 : 2);
      for (int j = 0; j < 3; ++j) {
        r.delta[j] = m[j];
        std::array<int, 3> t = m;
        if (m[j] >= 1) { t[j] -= 1; r.m1[j] = Idx(t[0], t[1], t[2]); }
        if (m[j] >= 2) { t[j] -= 1; r.m2[j] = Idx(t[0], t[1], t[2]); }
      }
    }
    m2l.resize((size_t)nm * nm);
    for (int g = 0; g < nm; ++g)
      for (int a = 0; a < nm; ++a)
        m2l[(size_t)g * nm + a] = Idx(multi[a][0] + multi[g][0], multi[a][1] + multi[g][1], multi[a][2] + multi[g][2]);
    for (int a = 0; a < nm; ++a)
      for (int b = 0; b < nm; ++b) {
        if (multi[b][0] <= multi[a][0] && multi[b][1] <= multi[a][1] && multi[b][2] <= multi[a][2])
          m2m.push_back({a, b, Idx(multi[a][0] - multi[b][0], multi[a][1] - multi[b][1], multi[a][2] - multi[b][2])});
        if (multi[a][0] <= multi[b][0] && multi[a][1] <= multi[b][1] && multi[a][2] <= multi[b][2])
          l2l.push_back({a, b, Idx(multi[b][0] - multi[a][0], multi[b][1] - multi[a][1], multi[b][2] - multi[a][2])});
      }
    grad.resize(nm);
    for (int g = 0; g < nm; ++g)
      for (int k = 0; k < 3; ++k) {
        std::array<int, 3> t = multi[g];
        if (t[k] >= 1) { t[k] -= 1; grad[g][k] = Idx(t[0], t[1], t[2]); }
        else grad[g][k] = -1;
      }
  }
  void Monomials(double x, double y, double z, double* out) const {  // :190-204
    double px[9], py[9], pz[9];
    px[0] = py[0] = pz[0] = 1.0;
    for (int i = 1; i <= P; ++i) { px[i] = px[i - 1] * x; py[i] = py[i - 1] * y; pz[i] = pz[i - 1] * z; }
    for (int k = 0; k < nm; ++k) out[k] = px[multi[k][0]] * py[multi[k][1]] * pz[multi[k][2]];
  }
  void Derivatives(double x, double y, double z, double* D) const {  // :207-232
    const double inv_r2 = 1.0 / (x * x + y * y + z * z);
    D[0] = std::sqrt(inv_r2);
    const double xs[3] = {x, y, z};
    for (int k = 1; k < nd; ++k) {
      const Rec& r = rec[k];
      const int i = r.axis;
      double s = (2 * r.delta[i] - 1) * xs[i] * D[r.m1[i]];
      if (r.m2[i] >= 0) s += (r.delta[i] - 1) * (r.delta[i] - 1) * D[r.m2[i]];
      for (int j = 0; j < 3; ++j) {
        if (j == i || r.delta[j] == 0) continue;
        s += 2.0 * r.delta[j] * xs[j] * D[r.m1[j]];
        if (r.m2[j] >= 0) s += r.delta[j] * (r.delta[j] - 1) * D[r.m2[j]];
      }
      D[k] = -s * inv_r2;
    }
  }
};

struct FmmCell {
  double cx = 0, cy = 0, cz = 0, half = 0, radius = 0;
  int child[8] = {-1, -1, -1, -1, -1, -1, -1, -1};
  int64_t begin = 0, end = 0;
  int depth = 0;
  bool leaf = true;
};

struct FmmStats { int64_t cells, leaves, m2l_pairs, p2p_pairs; };

// fast-multipole.cc:342-682.  want_f / want_e select outputs; stats optional.
void FmmCompute(int order, double theta, int64_t leaf_cap, const std::vector<V3>& src, const std::vector<double>& qs,
                std::vector<V3>* forces, double* pe, FmmStats* stats) {
  const int64_t n = (int64_t)src.size();
  if (forces) forces->assign(n, V3{0, 0, 0});
  if (pe) *pe = 0.0;
  if (n == 0 || (!forces && !pe)) return;
  order = std::min(std::max(order, 1), 8);
  if (leaf_cap < 1) leaf_cap = 1;
  const FmmTab t(order);
  const int nm = t.nm;
  double lo[3], hi[3];
  lo[0] = hi[0] = src[0].x; lo[1] = hi[1] = src[0].y; lo[2] = hi[2] = src[0].z;
  for (const V3& p : src) {
    lo[0] = std::min(lo[0], p.x); hi[0] = std::max(hi[0], p.x);
    lo[1] = std::min(lo[1], p.y); hi[1] = std::max(hi[1], p.y);
    lo[2] = std::min(lo[2], p.z); hi[2] = std::max(hi[2], p.z);
  }
  const double ext = std::max({hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]});
  const double L = std::max(ext, 1e-15);
  const double mid[3] = {0.5 * (lo[0] + hi[0]), 0.5 * (lo[1] + hi[1]), 0.5 * (lo[2] + hi[2])};
  const double fscale = kCoulomb / (L * L);
  const double min_sq = kMinSq / (L * L);
  std::vector<double> x(n), y(n), z(n);
  for (int64_t i = 0; i < n; ++i) {
    x[i] = (src[i].x - mid[0]) / L;
    y[i] = (src[i].y - mid[1]) / L;
    z[i] = (src[i].z - mid[2]) / L;
  }
  // Tree (Workspace::BuildCell, :257-311).
  std::vector<int64_t> ord(n), scr(n);
  for (int64_t i = 0; i < n; ++i) ord[i] = i;
  std::vector<FmmCell> cells;
  cells.reserve(2 * n / leaf_cap + 16);
  struct Builder {
    std::vector<FmmCell>& cells; std::vector<int64_t>& ord; std::vector<int64_t>& scr;
    const std::vector<double>& x; const std::vector<double>& y; const std::vector<double>& z; int64_t cap;
    static int Oct(double px, double py, double pz, double cx, double cy, double cz) {
      return (px >= cx ? 1 : 0) | (py >= cy ? 2 : 0) | (pz >= cz ? 4 : 0);
    }
    int Build(int64_t b, int64_t e, double cx, double cy, double cz, double h, int depth) {
      const int idx = (int)cells.size();
      cells.emplace_back();
      cells[idx].cx = cx; cells[idx].cy = cy; cells[idx].cz = cz; cells[idx].half = h;
      cells[idx].begin = b; cells[idx].end = e; cells[idx].depth = depth;
      if (e - b <= cap || depth >= 32) { cells[idx].leaf = true; return idx; }
      cells[idx].leaf = false;
      int64_t cnt[8] = {0}, off[9], cur[8];
      for (int64_t k = b; k < e; ++k) { const int64_t p = ord[k]; ++cnt[Oct(x[p], y[p], z[p], cx, cy, cz)]; }
      off[0] = b;
      for (int o = 0; o < 8; ++o) off[o + 1] = off[o] + cnt[o];
      for (int o = 0; o < 8; ++o) cur[o] = off[o];
      for (int64_t k = b; k < e; ++k) { const int64_t p = ord[k]; scr[cur[Oct(x[p], y[p], z[p], cx, cy, cz)]++] = p; }
      std::copy(scr.begin() + b, scr.begin() + e, ord.begin() + b);
      const double ch = 0.5 * h;
      for (int o = 0; o < 8; ++o) {
        if (cnt[o] == 0) continue;
        const int c = Build(off[o], off[o + 1], cx + ((o & 1) ? ch : -ch), cy + ((o & 2) ? ch : -ch),
                            cz + ((o & 4) ? ch : -ch), ch, depth + 1);
        cells[idx].child[o] = c;
      }
      return idx;
    }
  } builder{cells, ord, scr, x, y, z, leaf_cap};
  builder.Build(0, n, 0.0, 0.0, 0.0, 0.5 * (1.0 + 1e-9), 0);
  const int nc = (int)cells.size();
  std::vector<double> sx(n), sy(n), sz(n), sq(n);
  for (int64_t k = 0; k < n; ++k) { const int64_t p = ord[k]; sx[k] = x[p]; sy[k] = y[p]; sz[k] = z[p]; sq[k] = qs[p]; }
  for (int c = nc - 1; c >= 0; --c) {  // radii, :412-438
    FmmCell& cl = cells[c];
    if (cl.leaf) {
      double r2 = 0.0;
      for (int64_t k = cl.begin; k < cl.end; ++k) {
        const double dx = sx[k] - cl.cx, dy = sy[k] - cl.cy, dz = sz[k] - cl.cz;
        r2 = std::max(r2, dx * dx + dy * dy + dz * dz);
      }
      cl.radius = std::sqrt(r2);
    } else {
      double r = 0.0;
      for (int o = 0; o < 8; ++o) {
        if (cl.child[o] < 0) continue;
        const FmmCell& ch = cells[cl.child[o]];
        const double dx = ch.cx - cl.cx, dy = ch.cy - cl.cy, dz = ch.cz - cl.cz;
        r = std::max(r, std::sqrt(dx * dx + dy * dy + dz * dz) + ch.radius);
      }
      cl.radius = r;
    }
  }
  std::vector<double> M((size_t)nc * nm, 0.0), Lc((size_t)nc * nm, 0.0);
#pragma omp parallel
  {  // P2M, :443-461
    std::vector<double> mono(nm);
#pragma omp for schedule(dynamic, 16)
    for (int c = 0; c < nc; ++c) {
      const FmmCell& cl = cells[c];
      if (!cl.leaf) continue;
      double* m = &M[(size_t)c * nm];
      for (int64_t k = cl.begin; k < cl.end; ++k) {
        if (sq[k] == 0.0) continue;
        t.Monomials(sx[k] - cl.cx, sy[k] - cl.cy, sz[k] - cl.cz, mono.data());
        for (int a = 0; a < nm; ++a) m[a] += sq[k] * mono[a] * t.inv_fact[a];
      }
    }
  }
  {  // M2M, :463-482
    std::vector<double> mono(nm);
    for (int c = nc - 1; c >= 0; --c) {
      const FmmCell& cl = cells[c];
      if (cl.leaf) continue;
      double* mp = &M[(size_t)c * nm];
      for (int o = 0; o < 8; ++o) {
        const int k = cl.child[o];
        if (k < 0) continue;
        const FmmCell& ch = cells[k];
        t.Monomials(ch.cx - cl.cx, ch.cy - cl.cy, ch.cz - cl.cz, mono.data());
        const double* mc = &M[(size_t)k * nm];
        for (const auto& pr : t.m2m) mp[pr[0]] += mc[pr[1]] * mono[pr[2]] * t.inv_fact[pr[2]];
      }
    }
  }
  // Dual-tree traversal, :484-526.
  std::vector<std::vector<int>> m2l(nc), p2p(nc);
  const double th2 = theta * theta;
  {
    std::vector<std::pair<int, int>> stack;
    stack.emplace_back(0, 0);
    while (!stack.empty()) {
      const auto pr = stack.back();
      stack.pop_back();
      const FmmCell& T = cells[pr.first];
      const FmmCell& S = cells[pr.second];
      const double dx = T.cx - S.cx, dy = T.cy - S.cy, dz = T.cz - S.cz;
      const double dist2 = dx * dx + dy * dy + dz * dz;
      const double rs = T.radius + S.radius;
      if (rs * rs < th2 * dist2) { m2l[pr.first].push_back(pr.second); continue; }
      if (T.leaf && S.leaf) { p2p[pr.first].push_back(pr.second); continue; }
      const bool split_t = S.leaf || (!T.leaf && T.radius >= S.radius);
      if (split_t) {
        for (int o = 0; o < 8; ++o) if (T.child[o] >= 0) stack.emplace_back(T.child[o], pr.second);
      } else {
        for (int o = 0; o < 8; ++o) if (S.child[o] >= 0) stack.emplace_back(pr.first, S.child[o]);
      }
    }
  }
  if (stats) {
    stats->cells = nc; stats->leaves = 0; stats->m2l_pairs = 0; stats->p2p_pairs = 0;
    for (int c = 0; c < nc; ++c) {
      stats->leaves += cells[c].leaf ? 1 : 0;
      stats->m2l_pairs += (int64_t)m2l[c].size();
      for (int s : p2p[c]) stats->p2p_pairs += (cells[c].end - cells[c].begin) * (cells[s].end - cells[s].begin);
    }
  }
#pragma omp parallel
  {  // M2L, :528-555
    std::vector<double> D(t.nd);
#pragma omp for schedule(dynamic, 8)
    for (int ti = 0; ti < nc; ++ti) {
      if (m2l[ti].empty()) continue;
      double* l = &Lc[(size_t)ti * nm];
      for (int si : m2l[ti]) {
        const FmmCell& T = cells[ti];
        const FmmCell& S = cells[si];
        t.Derivatives(T.cx - S.cx, T.cy - S.cy, T.cz - S.cz, D.data());
        const double* m = &M[(size_t)si * nm];
        for (int g = 0; g < nm; ++g) {
          const int* ix = &t.m2l[(size_t)g * nm];
          double acc = 0.0;
          for (int a = 0; a < nm; ++a) acc += t.sign[a] * m[a] * D[ix[a]];
          l[g] += acc * t.inv_fact[g];
        }
      }
    }
  }
  {  // L2L, :557-578
    std::vector<double> mono(nm);
    for (int c = 0; c < nc; ++c) {
      const FmmCell& cl = cells[c];
      if (cl.leaf) continue;
      const double* lp = &Lc[(size_t)c * nm];
      for (int o = 0; o < 8; ++o) {
        const int k = cl.child[o];
        if (k < 0) continue;
        const FmmCell& ch = cells[k];
        t.Monomials(ch.cx - cl.cx, ch.cy - cl.cy, ch.cz - cl.cz, mono.data());
        double* lc = &Lc[(size_t)k * nm];
        for (const auto& pr : t.l2l) lc[pr[0]] += lp[pr[1]] * t.fact[pr[1]] * t.inv_fact[pr[0]] * t.inv_fact[pr[2]] * mono[pr[2]];
      }
    }
  }
  std::vector<double> ex(n, 0.0), ey(n, 0.0), ez(n, 0.0), phi(n, 0.0);
  const bool wf = forces != nullptr, we = pe != nullptr;
#pragma omp parallel
  {  // L2P + P2P, :580-661
    std::vector<double> mono(nm);
#pragma omp for schedule(dynamic, 8)
    for (int ti = 0; ti < nc; ++ti) {
      const FmmCell& T = cells[ti];
      if (!T.leaf) continue;
      const double* l = &Lc[(size_t)ti * nm];
      for (int64_t i = T.begin; i < T.end; ++i) {
        t.Monomials(sx[i] - T.cx, sy[i] - T.cy, sz[i] - T.cz, mono.data());
        double pot = 0.0, fx = 0.0, fy = 0.0, fz = 0.0;
        if (we) for (int g = 0; g < nm; ++g) pot += l[g] * mono[g];
        if (wf) {
          double gx = 0.0, gy = 0.0, gz = 0.0;
          for (int g = 1; g < nm; ++g) {
            const auto& gi = t.grad[g];
            if (gi[0] >= 0) gx += t.multi[g][0] * l[g] * mono[gi[0]];
            if (gi[1] >= 0) gy += t.multi[g][1] * l[g] * mono[gi[1]];
            if (gi[2] >= 0) gz += t.multi[g][2] * l[g] * mono[gi[2]];
          }
          fx = -gx; fy = -gy; fz = -gz;
        }
        for (int si : p2p[ti]) {
          const FmmCell& S = cells[si];
          for (int64_t j = S.begin; j < S.end; ++j) {
            if (j == i || sq[j] == 0.0) continue;
            const double rx = sx[i] - sx[j], ry = sy[i] - sy[j], rz = sz[i] - sz[j];
            const double r2 = rx * rx + ry * ry + rz * rz;
            if (r2 < min_sq) continue;
            const double inv_r = 1.0 / std::sqrt(r2);
            if (wf) { const double f = sq[j] * inv_r / r2; fx += rx * f; fy += ry * f; fz += rz * f; }
            if (we) pot += sq[j] * inv_r;
          }
        }
        if (wf) { ex[i] = fx; ey[i] = fy; ez[i] = fz; }
        if (we) phi[i] = pot;
      }
    }
  }
  if (wf)
    for (int64_t k = 0; k < n; ++k) {
      const double s = fscale * sq[k];
      (*forces)[ord[k]] = V3{s * ex[k], s * ey[k], s * ez[k]};
    }
  if (we) {
    double sum = 0.0;
    for (int64_t k = 0; k < n; ++k) sum += sq[k] * phi[k];
    *pe = 0.5 * kCoulomb / L * sum;
  }
}

inline bool IsPair(int a, int b) {  // annihilation.cc:12-19 (antiproton-proton, positron-electron)
  return (a == 0 && b == 4) || (a == 4 && b == 0) || (a == 1 && b == 3) || (a == 3 && b == 1);
}

int g_threads = 0;
void SetThreads() {
#ifdef _OPENMP
  if (g_threads > 0) omp_set_num_threads(g_threads);
#endif
}

}  // namespace

extern "C" {

// ---- RNG ------------------------------------------------------------------------------------
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) { Philox(ctr, key, out); }
double orc_annihilation_uniform(uint64_t seed, uint64_t step, uint32_t i, uint32_t j) {
  return AnnihilationUniform(seed, step, i, j);
}
void orc_cooling_normals(uint64_t seed, uint64_t step, uint32_t i, double out[3]) { CoolingNormals(seed, step, i, out); }

void orc_set_num_threads(int n) { g_threads = n; }
int orc_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

// ---- species / config -----------------------------------------------------------------------
void orc_species_props(int species, double* mass, double* charge) { SpeciesProps(species, mass, charge); }

// dt rule (src/params/physics_params.cc:60-74): listed[s] != 0 marks species present in
// "particle_populations" (even with count 0).
double orc_derive_dt(const int32_t listed[6], double b0) {
  double ratio = 1e-6;
  for (int s = 0; s < 6; ++s) {
    if (!listed[s]) continue;
    double m, q;
    SpeciesProps(s, &m, &q);
    ratio = std::min(ratio, m / std::fabs(q));
  }
  return (ratio * kPi) / (b0 * 2);
}

// ---- initial conditions (src/fields/penning_trap.cc:51-161) ------------------------------------
// counts[s] particles of species s, in enum order (the std::map order of the reference).
int orc_set_particle_positions(const uint64_t counts[6], const double radii[3], int seed, double* out) {
  std::default_random_engine rng(seed);
  std::uniform_real_distribution<double> ux(-radii[0], radii[0]);
  std::uniform_real_distribution<double> uy(-radii[1], radii[1]);
  std::uniform_real_distribution<double> uz(-radii[2], radii[2]);
  int64_t k = 0;
  for (int s = 0; s < 6; ++s)
    for (uint64_t i = 0; i < counts[s]; ++i) {
      double p[3];
      for (;;) {
        p[0] = ux(rng);
        p[1] = uy(rng);
        p[2] = uz(rng);
        const double a = p[0] / radii[0], b = p[1] / radii[1], c = p[2] / radii[2];
        if ((a * a) + (b * b) + (c * c) <= 1.0) break;
      }
      out[3 * k] = p[0]; out[3 * k + 1] = p[1]; out[3 * k + 2] = p[2];
      ++k;
    }
  return 0;
}

int orc_set_particle_velocities(const uint64_t counts[6], const double temps[6], int seed, double* out) {
  int64_t k = 0;
  for (int s = 0; s < 6; ++s) {
    const uint64_t count = counts[s];
    if (count == 0) continue;
    const double T = temps[s];
    if (T == 0.0) {
      for (uint64_t i = 0; i < 3 * count; ++i) out[3 * k + i] = 0.0;
      k += (int64_t)count;
      continue;
    }
    double m, q;
    SpeciesProps(s, &m, &q);
    const double sigma = std::sqrt(kBoltzmann * T / m);
    std::mt19937 rng(seed);  // re-seeded per species (penning_trap.cc:77)
    std::normal_distribution<double> g(0.0, sigma);
    std::vector<double> raw(3 * count);
    for (uint64_t i = 0; i < 3 * count; ++i) raw[i] = g(rng);
    if (count >= 2) {
      double mean[3] = {0, 0, 0};
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
  return 0;
}

// ---- field ----------------------------------------------------------------------------------
void orc_trap_characteristic_dimension(double z0, double rho0, double* d) { *d = MakeTrap(1, 1, z0, rho0).d; }
void orc_electric_field(double u0, double d, int64_t n, const double* pos, double* e) {
  const Trap t{0.0, u0, d};
  for (int64_t i = 0; i < n; ++i) st(e, i, EField(t, ld(pos, i)));
}
void orc_potential(double u0, double d, int64_t n, const double* pos, double* phi) {
  const Trap t{0.0, u0, d};
  for (int64_t i = 0; i < n; ++i) phi[i] = Potential(t, ld(pos, i));
}
double orc_cooling_time(double q, double m, double b) { return CoolingTime(q, m, b); }

// ---- Lorentz push (src/forces/lorentz.cc:37-71) ---------------------------------------------
// staggered[i] is the per-behaviour velocity_staggered_ flag (lorentz.h:34).  The Coulomb force
// is reset to zero afterwards, as the reference does.
void orc_lorentz_push(int64_t n, double* pos, double* vel, double* fcoul, const double* q, const double* m,
                      const uint8_t* annihilated, uint8_t* staggered, double b0, double u0, double d, double dt) {
  const Trap t{b0, u0, d};
  const V3 B{0.0, 0.0, b0};
  for (int64_t i = 0; i < n; ++i) {
    if (annihilated && annihilated[i]) continue;
    if (m[i] <= 0.0 || dt <= 0.0) continue;
    const V3 x = ld(pos, i);
    V3 v = ld(vel, i);
    const V3 E = EField(t, x);
    const V3 F = add(scl(E, q[i]), ld(fcoul, i));
    if (!staggered[i]) {
      v = Boris(v, F, B, q[i], m[i], -0.5 * dt);
      staggered[i] = 1;
    }
    const V3 vn = Boris(v, F, B, q[i], m[i], dt);
    st(vel, i, vn);
    st(pos, i, add(x, scl(vn, dt)));
    st(fcoul, i, V3{0, 0, 0});
  }
}

// ---- cyclotron cooling (src/forces/cooling.cc:24-65) ----------------------------------------
// Noise (bath_temp > 0) uses the counter-based draws of CoolingNormals keyed on
// (seed, step, index) in place of BioDynaMo's Random::Gaus.
void orc_cyclotron_cooling(int64_t n, double* vel, const double* q, const double* m, const uint8_t* annihilated,
                           double b0, double tau_scale, double bath_temp, double dt, uint64_t seed, uint64_t step) {
  for (int64_t i = 0; i < n; ++i) {
    if (annihilated && annihilated[i]) continue;
    if (q[i] == 0.0 || m[i] <= 0.0 || dt <= 0.0 || tau_scale <= 0.0) continue;
    const V3 B{0.0, 0.0, b0};
    const double bmag = std::sqrt(dot(B, B));
    if (bmag <= 0.0) continue;
    const double tau = tau_scale * CoolingTime(std::fabs(q[i]), m[i], bmag);
    if (!std::isfinite(tau)) continue;
    const V3 bh = scl(B, 1.0 / bmag);
    const V3 v = ld(vel, i);
    const double vpar = dot(v, bh);
    const V3 vparv = scl(bh, vpar);
    V3 vperp = sub(v, vparv);
    const double decay = std::exp(-dt / (2.0 * tau));
    vperp = scl(vperp, decay);
    if (bath_temp > 0.0) {
      const double sigma = std::sqrt(kBoltzmann * bath_temp / m[i] * (1.0 - decay * decay));
      double g[3];
      CoolingNormals(seed, step, (uint32_t)i, g);
      V3 noise{sigma * g[0], sigma * g[1], sigma * g[2]};
      noise = sub(noise, scl(bh, dot(noise, bh)));
      vperp = add(vperp, noise);
    }
    st(vel, i, add(vparv, vperp));
  }
}

// ---- Coulomb solvers ------------------------------------------------------------------------
// Cutoff direct (src/interaction/coulomb_interaction.cc:23-80), brute-force O(N^2) neighbour
// set: j != i, alive, q_j != 0, d2 < rc^2 (strict, radius search), d2 >= 1e-30.
void orc_standard_forces(int64_t n, const double* pos, const double* q, const uint8_t* annihilated, double rc,
                         double* force) {
  SetThreads();
  for (int64_t i = 0; i < 3 * n; ++i) force[i] = 0.0;
  if (rc <= 0.0) return;
  const double rc2 = rc * rc;
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t i = 0; i < n; ++i) {
    if (annihilated && annihilated[i]) continue;
    if (q[i] == 0.0) continue;
    const V3 xi = ld(pos, i);
    if (!std::isfinite(xi.x) || !std::isfinite(xi.y) || !std::isfinite(xi.z)) continue;
    V3 f{0, 0, 0};
    for (int64_t j = 0; j < n; ++j) {
      if (j == i || (annihilated && annihilated[j]) || q[j] == 0.0) continue;
      const V3 r = sub(xi, ld(pos, j));
      const double d2 = dot(r, r);
      if (!(d2 < rc2)) continue;
      if (d2 < kMinSq) continue;
      const double dist = std::sqrt(d2);
      const double mag = kCoulomb * q[i] * q[j] / d2;
      f = add(f, scl(r, mag / dist));
    }
    st(force, i, f);
  }
}

// All-pairs direct without cutoff (the theta = 0 limit, barnes-hut.h:22-24), FP64, OpenMP.
// Pair term as AddPointCharge (barnes-hut.cc:19-27).  Also the CPU baseline for the
// all-pairs workload.
void orc_allpairs_forces(int64_t n, const double* pos, const double* q, const uint8_t* annihilated, double* force) {
  SetThreads();
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t i = 0; i < n; ++i) {
    V3 f{0, 0, 0};
    if (!(annihilated && annihilated[i]) && q[i] != 0.0) {
      const V3 xi = ld(pos, i);
      for (int64_t j = 0; j < n; ++j) {
        if (j == i || (annihilated && annihilated[j]) || q[j] == 0.0) continue;
        PointCharge(xi, q[i], ld(pos, j), q[j], &f);
      }
    }
    st(force, i, f);
  }
}

// All-pairs forces on a subset of targets (indices), against all alive sources.
void orc_allpairs_forces_subset(int64_t n, const double* pos, const double* q, const uint8_t* annihilated,
                                int64_t nt, const int64_t* targets, double* force) {
  SetThreads();
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t t = 0; t < nt; ++t) {
    const int64_t i = targets[t];
    V3 f{0, 0, 0};
    if (!(annihilated && annihilated[i]) && q[i] != 0.0) {
      const V3 xi = ld(pos, i);
      for (int64_t j = 0; j < n; ++j) {
        if (j == i || (annihilated && annihilated[j]) || q[j] == 0.0) continue;
        PointCharge(xi, q[i], ld(pos, j), q[j], &f);
      }
    }
    st(force, t, f);
  }
}

// Total all-pairs potential energy U = 1/2 k sum_i q_i phi_i (the theta = 0 limit of
// ComputeBarnesHutEnergy, energy_analysis.cc:45-69).
double orc_allpairs_potential_energy(int64_t n, const double* pos, const double* q, const uint8_t* annihilated) {
  SetThreads();
  double sum = 0.0;
#pragma omp parallel for reduction(+ : sum) schedule(dynamic, 16)
  for (int64_t i = 0; i < n; ++i) {
    if ((annihilated && annihilated[i]) || q[i] == 0.0) continue;
    const V3 xi = ld(pos, i);
    double phi = 0.0;
    for (int64_t j = 0; j < n; ++j) {
      if (j == i || (annihilated && annihilated[j]) || q[j] == 0.0) continue;
      const V3 r = sub(xi, ld(pos, j));
      const double d2 = dot(r, r);
      if (d2 < kMinSq) continue;
      phi += q[j] / std::sqrt(d2);
    }
    sum += q[i] * phi;
  }
  return 0.5 * kCoulomb * sum;
}

// Standard (shifted, cutoff) Coulomb energy (src/analysis/energy_analysis.cc:16-43): i<j over
// alive particles in index order, d2 <= rc^2 (inclusive), d2 >= 1e-30.
double orc_standard_energy(int64_t n, const double* pos, const double* q, const uint8_t* annihilated, double rc) {
  const std::vector<int64_t> idx = Gather(n, annihilated);
  const int64_t m = (int64_t)idx.size();
  const double rc2 = rc * rc;
  double sum = 0.0;
  for (int64_t a = 0; a < m; ++a) {
    const int64_t i = idx[a];
    if (q[i] == 0.0) continue;
    const V3 xi = ld(pos, i);
    for (int64_t b = a + 1; b < m; ++b) {
      const int64_t j = idx[b];
      if (q[j] == 0.0) continue;
      const V3 r = sub(xi, ld(pos, j));
      const double d2 = dot(r, r);
      if (d2 > rc2) continue;
      if (d2 < 1e-30) continue;
      sum += kCoulomb * q[i] * q[j] * (1.0 / std::sqrt(d2) - 1.0 / rc);
    }
  }
  return sum;
}

// ---- Barnes-Hut (ComputeBarnesHutCoulombForces, barnes-hut.cc:271-299) -----------------------
// interactions (optional): per-particle count of evaluated pair/monopole interactions.
void orc_bh_forces(int64_t n, const double* pos, const double* q, const uint8_t* annihilated, double theta,
                   int64_t leaf_cap, double* force, int64_t* interactions) {
  SetThreads();
  for (int64_t i = 0; i < 3 * n; ++i) force[i] = 0.0;
  const std::vector<int64_t> idx = Gather(n, annihilated);
  if (idx.empty()) return;
  BHTree tree;
  tree.theta2 = theta * theta;
  tree.leaf_cap = leaf_cap < 1 ? 1 : leaf_cap;
  for (int64_t i : idx) { tree.pos.push_back(ld(pos, i)); tree.q.push_back(q[i]); }
  tree.Build();
  const int64_t m = (int64_t)idx.size();
#pragma omp parallel for schedule(static)
  for (int64_t k = 0; k < m; ++k) {
    V3 f{0, 0, 0};
    int64_t cnt = 0;
    if (!tree.nodes.empty() && tree.q[k] != 0.0) tree.Accumulate(0, tree.pos[k], tree.q[k], k, &f, &cnt);
    st(force, idx[k], f);
    if (interactions) interactions[idx[k]] = cnt;
  }
}

// BH potential energy (ComputeBarnesHutEnergy, energy_analysis.cc:45-69).
double orc_bh_potential_energy(int64_t n, const double* pos, const double* q, const uint8_t* annihilated,
                               double theta) {
  SetThreads();
  const std::vector<int64_t> idx = Gather(n, annihilated);
  if (idx.empty()) return 0.0;
  BHTree tree;
  tree.theta2 = theta * theta;
  tree.leaf_cap = 8;
  for (int64_t i : idx) { tree.pos.push_back(ld(pos, i)); tree.q.push_back(q[i]); }
  tree.Build();
  const int64_t m = (int64_t)idx.size();
  double sum = 0.0;
#pragma omp parallel for reduction(+ : sum) schedule(static)
  for (int64_t k = 0; k < m; ++k) {
    if (tree.q[k] == 0.0) continue;
    double phi = 0.0;
    tree.AccumulatePotential(0, tree.pos[k], k, &phi);
    sum += tree.q[k] * phi;
  }
  return 0.5 * kCoulomb * sum;
}

// Per-particle Barnes-Hut potential phi_i = BarnesHutTree::ComputePotential(x_i, i)
// (barnes-hut.cc:175-181, 223-269) in C/m (without k_C), 0 for annihilated particles; the walk of
// orc_bh_potential_energy with the leaf capacity as an argument.
void orc_bh_potentials(int64_t n, const double* pos, const double* q, const uint8_t* annihilated, double theta,
                       int64_t leaf_cap, double* phi_out) {
  SetThreads();
  for (int64_t i = 0; i < n; ++i) phi_out[i] = 0.0;
  const std::vector<int64_t> idx = Gather(n, annihilated);
  if (idx.empty()) return;
  BHTree tree;
  tree.theta2 = theta * theta;
  tree.leaf_cap = leaf_cap < 1 ? 1 : leaf_cap;
  for (int64_t i : idx) { tree.pos.push_back(ld(pos, i)); tree.q.push_back(q[i]); }
  tree.Build();
  const int64_t m = (int64_t)idx.size();
#pragma omp parallel for schedule(static)
  for (int64_t k = 0; k < m; ++k) {
    double phi = 0.0;
    tree.AccumulatePotential(0, tree.pos[k], k, &phi);
    phi_out[idx[k]] = phi;
  }
}

// BarnesHutTree::Build over the alive particles, then ComputeForce(x, q, self) and
// ComputePotential(x, self) (barnes-hut.cc:166-181) at m given points; self[k] is a particle index
// (-1 or an annihilated particle: none).  force_out 3m (N), phi_out m (C/m); either may be null.
void orc_bh_evaluate(int64_t n, const double* pos, const double* q, const uint8_t* annihilated, double theta,
                     int64_t leaf_cap, int64_t m, const double* pts, const double* charges, const int64_t* self,
                     double* force_out, double* phi_out) {
  SetThreads();
  const std::vector<int64_t> idx = Gather(n, annihilated);
  std::vector<int64_t> gathered_of(n, -1);
  for (size_t k = 0; k < idx.size(); ++k) gathered_of[idx[k]] = (int64_t)k;
  BHTree tree;
  tree.theta2 = theta * theta;
  tree.leaf_cap = leaf_cap < 1 ? 1 : leaf_cap;
  for (int64_t i : idx) { tree.pos.push_back(ld(pos, i)); tree.q.push_back(q[i]); }
  tree.Build();
#pragma omp parallel for schedule(static)
  for (int64_t k = 0; k < m; ++k) {
    const int64_t s = (self && self[k] >= 0 && self[k] < n) ? gathered_of[self[k]] : -1;
    const V3 x = ld(pts, k);
    const double qt = charges ? charges[k] : 0.0;
    if (force_out) {
      V3 f{0, 0, 0};
      int64_t cnt = 0;
      if (!tree.nodes.empty() && qt != 0.0) tree.Accumulate(0, x, qt, s, &f, &cnt);
      st(force_out, k, f);
    }
    if (phi_out) {
      double phi = 0.0;
      if (!tree.nodes.empty()) tree.AccumulatePotential(0, x, s, &phi);
      phi_out[k] = phi;
    }
  }
}

// BH tree structure for tests: returns node count; fills (if non-null, sized >= count)
// order[n_alive] (gathered-index permutation), node ranges [begin,end), depth, leaf flag,
// geometric centre and half size, charges and charge centres.  Nodes in the reference's
// DFS preorder.
int64_t orc_bh_tree(int64_t n, const double* pos, const double* q, int64_t leaf_cap, int64_t cap_nodes,
                    int64_t* order, int64_t* begin, int64_t* end, int32_t* depth, uint8_t* leaf, double* center,
                    double* half, double* qp, double* qn, double* cp, double* cn, int32_t* child) {
  BHTree tree;
  tree.leaf_cap = leaf_cap < 1 ? 1 : leaf_cap;
  for (int64_t i = 0; i < n; ++i) { tree.pos.push_back(ld(pos, i)); tree.q.push_back(q[i]); }
  tree.Build();
  const int64_t nn = (int64_t)tree.nodes.size();
  if (order) for (int64_t i = 0; i < n; ++i) order[i] = tree.order[i];
  if (cap_nodes < nn) return nn;
  for (int64_t k = 0; k < nn; ++k) {
    const BHNode& nd = tree.nodes[k];
    if (begin) begin[k] = nd.begin;
    if (end) end[k] = nd.end;
    if (depth) depth[k] = nd.depth;
    if (leaf) leaf[k] = nd.leaf ? 1 : 0;
    if (center) st(center, k, nd.center);
    if (half) half[k] = nd.half;
    if (qp) qp[k] = nd.qp;
    if (qn) qn[k] = nd.qn;
    if (cp) st(cp, k, nd.cp);
    if (cn) st(cn, k, nd.cn);
    if (child) for (int o = 0; o < 8; ++o) child[8 * k + o] = nd.child[o];
  }
  return nn;
}

// ---- FMM (ComputeFastMultipoleCoulombForces, fast-multipole.cc:684-713) ----------------------
// q == 0 particles get zero force and are not sources.  stats (optional, 4 int64):
// cells, leaves, M2L pairs, P2P particle pairs.
void orc_fmm_forces(int64_t n, const double* pos, const double* q, const uint8_t* annihilated, int order,
                    double theta, int64_t leaf_cap, double* force, int64_t* stats) {
  SetThreads();
  for (int64_t i = 0; i < 3 * n; ++i) force[i] = 0.0;
  std::vector<int64_t> idx;
  std::vector<V3> src;
  std::vector<double> qs;
  for (int64_t i = 0; i < n; ++i) {
    if (annihilated && annihilated[i]) continue;
    if (q[i] == 0.0) continue;
    idx.push_back(i);
    src.push_back(ld(pos, i));
    qs.push_back(q[i]);
  }
  if (idx.empty()) return;
  std::vector<V3> f;
  FmmStats s{};
  FmmCompute(order, theta, leaf_cap, src, qs, &f, nullptr, &s);
  for (size_t k = 0; k < idx.size(); ++k) st(force, idx[k], f[k]);
  if (stats) { stats[0] = s.cells; stats[1] = s.leaves; stats[2] = s.m2l_pairs; stats[3] = s.p2p_pairs; }
}

// FMM potential energy (ComputeFastMultipoleEnergy, energy_analysis.cc:71-81): all alive
// particles are passed as sources (zero charges contribute nothing).
double orc_fmm_potential_energy(int64_t n, const double* pos, const double* q, const uint8_t* annihilated, int order,
                                double theta) {
  SetThreads();
  std::vector<V3> src;
  std::vector<double> qs;
  for (int64_t i = 0; i < n; ++i) {
    if (annihilated && annihilated[i]) continue;
    src.push_back(ld(pos, i));
    qs.push_back(q[i]);
  }
  double pe = 0.0;
  FmmCompute(order, theta, 32, src, qs, nullptr, &pe, nullptr);
  return pe;
}

// FMM helper exposure for closed-form tests: derivative tensor D_alpha(r) of 1/|r| for all
// multi-indices of degree <= 2P, and the multi-index table.  Returns n_d.
int orc_fmm_derivatives(int order, double x, double y, double z, double* D, int32_t* multi) {
  const FmmTab t(order);
  t.Derivatives(x, y, z, D);
  if (multi)
    for (int k = 0; k < t.nd; ++k) { multi[3 * k] = t.multi[k][0]; multi[3 * k + 1] = t.multi[k][1]; multi[3 * k + 2] = t.multi[k][2]; }
  return t.nd;
}

// ---- annihilation (src/interaction/annihilation.cc:21-61), canonical order -------------------
// After all pushes, scan i ascending over alive particles; candidates j ascending over alive
// opposite-matter particles with d2 < r_a^2 (strict); for each candidate draw
// U = AnnihilationUniform(seed, step, i, j); the first U < p marks both.  Returns the number of
// particles marked.  pairs (optional, 2 per pair) receives the (i, j) pairs in scan order.
int64_t orc_annihilate(int64_t n, const double* pos, const uint8_t* species, uint8_t* annihilated, double ra,
                       double p, uint64_t seed, uint64_t step, int64_t* pairs) {
  if (ra <= 0.0) return 0;
  const double ra2 = ra * ra;
  std::vector<uint8_t> marked(n, 0);
  int64_t count = 0, np = 0;
  for (int64_t i = 0; i < n; ++i) {
    if (annihilated[i] || marked[i]) continue;
    const V3 xi = ld(pos, i);
    for (int64_t j = 0; j < n; ++j) {
      if (j == i || annihilated[j] || marked[j]) continue;
      if (!IsPair(species[i], species[j])) continue;
      const V3 r = sub(xi, ld(pos, j));
      if (!(dot(r, r) < ra2)) continue;
      if (AnnihilationUniform(seed, step, (uint32_t)i, (uint32_t)j) < p) {
        marked[i] = marked[j] = 1;
        count += 2;
        if (pairs) { pairs[2 * np] = i; pairs[2 * np + 1] = j; }
        ++np;
        break;
      }
    }
  }
  for (int64_t i = 0; i < n; ++i)
    if (marked[i]) annihilated[i] = 1;
  return count;
}

// ---- energy (src/analysis/energy_analysis.cc:84-145, without the time-centring, which the
// caller applies).  method: 0 standard, 1 Barnes-Hut, 2 FMM.  out = {KE, PE_trap, PE_coul}.
int64_t orc_energy(int64_t n, const double* pos, const double* vel, const double* q, const double* m,
                   const uint8_t* annihilated, double u0, double d, int method, double rc, double bh_theta,
                   double fmm_theta, double fmm_order, double out[3]) {
  const Trap t{0.0, u0, d};
  double ke = 0.0, pt = 0.0;
  int64_t active = 0;
  for (int64_t i = 0; i < n; ++i) {
    if (annihilated && annihilated[i]) continue;
    ++active;
    const V3 v = ld(vel, i);
    ke += 0.5 * m[i] * dot(v, v);
    pt += q[i] * Potential(t, ld(pos, i));
  }
  double pc = 0.0;
  if (method == 2) pc = orc_fmm_potential_energy(n, pos, q, annihilated, (int)fmm_order, fmm_theta);
  else if (method == 1) pc = orc_bh_potential_energy(n, pos, q, annihilated, bh_theta);
  else pc = orc_standard_energy(n, pos, q, annihilated, rc);
  out[0] = ke; out[1] = pt; out[2] = pc;
  return active;
}

// ---- driver (src/antimatter-kernel.cc:42-217) -------------------------------------------------
struct OrcSimConfig {
  uint64_t counts[6];
  int32_t listed[6];
  double temps[6];
  double radii[3];
  int32_t seed;
  double b0, u0, z0, rho0;
  int64_t num_steps;
  double cooling_scale, bath_temp;
  int32_t coulomb_enabled;
  int32_t coulomb_method;  // 0 standard, 1 BH, 2 FMM, 3 all-pairs (BH at theta = 0)
  int32_t energy_method;   // 0 standard, 1 BH, 2 FMM
  double cutoff_radius, bh_theta, fmm_theta, fmm_order;
  int32_t annihilation_enabled;
  double contact_radius, annihilation_probability;
  uint64_t rng_seed;  // key of the counter-based draws (annihilation, cooling noise)
};

// rows: up to max_rows x 8 doubles (step, KE, PE_trap, PE_coul, total, n_active, frac drift,
// abs drift).  Final state (optional) in pos/vel/annihilated (n = sum counts).  Returns the
// number of rows written, or -1 on invalid input.
int64_t orc_simulate(const OrcSimConfig* c, double* rows, int64_t max_rows, double* pos_out, double* vel_out,
                     uint8_t* ann_out) {
  uint64_t n = 0;
  for (int s = 0; s < 6; ++s) n += c->counts[s];
  if (n == 0) return -1;
  std::vector<double> pos(3 * n), vel(3 * n), f(3 * n, 0.0), q(n), m(n);
  std::vector<uint8_t> sp(n), ann(n, 0), stag(n, 0);
  orc_set_particle_positions(c->counts, c->radii, c->seed, pos.data());
  orc_set_particle_velocities(c->counts, c->temps, c->seed, vel.data());
  {
    uint64_t k = 0;
    for (int s = 0; s < 6; ++s)
      for (uint64_t i = 0; i < c->counts[s]; ++i, ++k) { SpeciesProps(s, &m[k], &q[k]); sp[k] = (uint8_t)s; }
  }
  const double dt = orc_derive_dt(c->listed, c->b0);
  const Trap trap = MakeTrap(c->b0, c->u0, c->z0, c->rho0);
  const int64_t steps = c->num_steps;
  const int64_t interval = steps >= 100 ? steps / 100 : 1;
  bool has_prev = false;
  double prev_trap = 0.0, prev_coul = 0.0;
  auto evaluate = [&](double e[4]) {
    double o[3];
    const int64_t act = orc_energy((int64_t)n, pos.data(), vel.data(), q.data(), m.data(), ann.data(), trap.u0,
                                   trap.d, c->energy_method, c->cutoff_radius, c->bh_theta, c->fmm_theta,
                                   c->fmm_order, o);
    const double raw_t = o[1], raw_c = o[2];
    double pt = raw_t, pc = raw_c;
    if (has_prev) { pt = 0.5 * (prev_trap + raw_t); pc = 0.5 * (prev_coul + raw_c); }
    prev_trap = raw_t; prev_coul = raw_c; has_prev = true;
    if (!c->coulomb_enabled) pc = 0.0;
    e[0] = o[0]; e[1] = pt; e[2] = pc; e[3] = (double)act;
  };
  int64_t nrows = 0;
  double e0[4];
  evaluate(e0);
  const double total0 = e0[0] + e0[1] + e0[2];
  double scale = 1.602176634e-19 * std::fabs(c->u0);
  if (scale <= 0.0) scale = std::fabs(total0);
  if (scale <= 0.0) scale = 1.0;
  auto put = [&](int64_t step, const double e[4], double frac, double absd) {
    if (nrows >= max_rows) return;
    double* r = rows + 8 * nrows;
    r[0] = (double)step; r[1] = e[0]; r[2] = e[1]; r[3] = e[2]; r[4] = e[0] + e[1] + e[2]; r[5] = e[3];
    r[6] = frac; r[7] = absd;
    ++nrows;
  };
  put(0, e0, 0.0, 0.0);
  for (int64_t step = 1; step <= steps; ++step) {
    if (c->coulomb_enabled) {
      if (c->coulomb_method == 1) orc_bh_forces((int64_t)n, pos.data(), q.data(), ann.data(), c->bh_theta, 8, f.data(), nullptr);
      else if (c->coulomb_method == 2) orc_fmm_forces((int64_t)n, pos.data(), q.data(), ann.data(), (int)c->fmm_order, c->fmm_theta, 32, f.data(), nullptr);
      else if (c->coulomb_method == 3) orc_allpairs_forces((int64_t)n, pos.data(), q.data(), ann.data(), f.data());
      else orc_standard_forces((int64_t)n, pos.data(), q.data(), ann.data(), c->cutoff_radius, f.data());
    }
    orc_lorentz_push((int64_t)n, pos.data(), vel.data(), f.data(), q.data(), m.data(), ann.data(), stag.data(), trap.b0,
                     trap.u0, trap.d, dt);
    orc_cyclotron_cooling((int64_t)n, vel.data(), q.data(), m.data(), ann.data(), trap.b0, c->cooling_scale,
                          c->bath_temp, dt, c->rng_seed, (uint64_t)step);
    if (c->annihilation_enabled)
      orc_annihilate((int64_t)n, pos.data(), sp.data(), ann.data(), c->contact_radius, c->annihilation_probability,
                     c->rng_seed, (uint64_t)step, nullptr);
    const bool log = (step % interval == 0) || (step == steps);
    const bool need = log || ((step + 1) % interval == 0) || (step + 1 == steps);
    if (need) {
      double e[4];
      evaluate(e);
      if (log) {
        const double absd = (e[0] + e[1] + e[2]) - total0;
        put(step, e, absd / scale, absd);
      }
    }
  }
  if (pos_out) std::memcpy(pos_out, pos.data(), sizeof(double) * 3 * n);
  if (vel_out) std::memcpy(vel_out, vel.data(), sizeof(double) * 3 * n);
  if (ann_out) std::memcpy(ann_out, ann.data(), n);
  return nrows;
}

}  // extern "C"
