// Host code under AddressSanitizer and UndefinedBehaviorSanitizer (built by tests/test_sanitize.py
// with g++, host code only): the params.json loader and Validate, the .pvd writer, and the CPU
// oracle's solvers, on ordinary, malformed, extreme and degenerate inputs.  Any sanitizer report
// aborts the program (-fno-sanitize-recover=all); the last line is "SANITIZE OK" otherwise.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <random>
#include <sstream>
#include <string>
#include <vector>

#include "antimatter_amd/am_params.h"

extern "C" {
void orc_set_num_threads(int n);
void orc_species_props(int species, double* mass, double* charge);
double orc_derive_dt(const int32_t listed[6], double b0);
int orc_set_particle_positions(const uint64_t counts[6], const double radii[3], int seed, double* out);
int orc_set_particle_velocities(const uint64_t counts[6], const double temps[6], int seed, double* out);
void orc_lorentz_push(int64_t n, double* pos, double* vel, double* fcoul, const double* q, const double* m,
                      const uint8_t* annihilated, uint8_t* staggered, double b0, double u0, double d, double dt);
void orc_cyclotron_cooling(int64_t n, double* vel, const double* q, const double* m, const uint8_t* annihilated,
                           double b0, double tau_scale, double bath_temp, double dt, uint64_t seed, uint64_t step);
void orc_standard_forces(int64_t n, const double* pos, const double* q, const uint8_t* annihilated, double rc,
                         double* force);
void orc_allpairs_forces(int64_t n, const double* pos, const double* q, const uint8_t* annihilated, double* force);
double orc_allpairs_potential_energy(int64_t n, const double* pos, const double* q, const uint8_t* annihilated);
double orc_standard_energy(int64_t n, const double* pos, const double* q, const uint8_t* annihilated, double rc);
void orc_bh_forces(int64_t n, const double* pos, const double* q, const uint8_t* annihilated, double theta,
                   int64_t leaf_cap, double* force, int64_t* interactions);
double orc_bh_potential_energy(int64_t n, const double* pos, const double* q, const uint8_t* annihilated,
                               double theta);
void orc_bh_evaluate(int64_t n, const double* pos, const double* q, const uint8_t* annihilated, double theta,
                     int64_t leaf_cap, int64_t m, const double* pts, const double* charges, const int64_t* self,
                     double* force_out, double* phi_out);
int64_t orc_bh_tree(int64_t n, const double* pos, const double* q, int64_t leaf_cap, int64_t cap_nodes,
                    int64_t* order, int64_t* begin, int64_t* end, int32_t* depth, uint8_t* leaf, double* center,
                    double* half, double* qp, double* qn, double* cp, double* cn, int32_t* child);
void orc_fmm_forces(int64_t n, const double* pos, const double* q, const uint8_t* annihilated, int order,
                    double theta, int64_t leaf_cap, double* force, int64_t* stats);
double orc_fmm_potential_energy(int64_t n, const double* pos, const double* q, const uint8_t* annihilated, int order,
                                double theta);
int64_t orc_annihilate(int64_t n, const double* pos, const uint8_t* species, uint8_t* annihilated, double ra,
                       double p, uint64_t seed, uint64_t step, int64_t* pairs);
}

static int g_fail = 0;
#define CHECK(cond)                                                       \
  do {                                                                    \
    if (!(cond)) {                                                        \
      std::fprintf(stderr, "%s:%d: check failed: %s\n", __FILE__, __LINE__, #cond); \
      ++g_fail;                                                           \
    }                                                                     \
  } while (0)

static std::string ReadFile(const char* path) {
  std::ifstream f(path);
  std::stringstream ss;
  ss << f.rdbuf();
  return ss.str();
}

static void ParamsChecks(const char* golden_path, const char* tmpdir) {
  am_params p;
  // the shipped configuration, from the file and from its text
  CHECK(am_params_load(golden_path, &p) == AM_OK && p.loaded == 1);
  CHECK(am_params_validate(&p) == AM_OK);
  const std::string text = ReadFile(golden_path);
  CHECK(am_params_parse(text.c_str(), &p) == AM_OK && p.num_steps == 100);
  // a missing file keeps the defaults; a directory is not a readable file
  CHECK(am_params_load((std::string(tmpdir) + "/does_not_exist.json").c_str(), &p) == AM_OK && p.loaded == 0);
  CHECK(am_params_validate(&p) != AM_OK);  // the default cooling_scale = 10
  (void)am_params_load(tmpdir, &p);
  CHECK(am_params_defaults(&p) == AM_OK);
  // malformed and mistyped inputs (the file is one flat object, as the shipped params.json)
  const char* bad[] = {
      "", "{", "[]", "null", "3", "\"text\"", "{\"b0\": \"x\"}", "{\"seed\": [1]}", "{\"num_steps\": true}",
      "{\"particle_populations\": {\"muon\": 3}}",
      "{\"particle_populations\": [1, 2]}",
      "{\"particle_populations\": 5}",
      "{\"particle_temperatures\": {\"positron\": \"hot\"}}",
      "{\"coulomb_computation_method\": 7}",
      "{\"initial_cloud_radii\": {\"radius_x\": \"wide\"}}",
      "{\"coulomb_forces_enabled\": \"yes\"}",
      "\x01\x02\x03",
  };
  for (const char* s : bad) {
    CHECK(am_params_parse(s, &p) != AM_OK);
    (void)am_params_validate(&p);
  }
  // json::count() on an array is 0 in the reference's reader too: the default radii stay
  CHECK(am_params_parse("{\"initial_cloud_radii\": [1, 2]}", &p) == AM_OK && p.initial_cloud_radii[0] == 1e-4);
  // numbers whose integer part the size_t / int fields cannot hold are rejected, not cast
  const char* out_of_range[] = {
      "{\"num_steps\": 1e300}",
      "{\"num_steps\": -1.5}",
      "{\"particle_populations\": {\"positron\": 1e30}}",
      "{\"particle_temperatures\": {\"positron\": 1e308}}",
      "{\"particle_temperatures\": {\"electron\": -2.5}}",
      "{\"seed\": 1e20}",
      "{\"seed\": -2147483649.0}",
      "{\"num_steps\": 18446744073709551616.0}",
  };
  for (const char* s : out_of_range) CHECK(am_params_parse(s, &p) != AM_OK);
  // in-range values convert as the reference's get<T>() does
  CHECK(am_params_parse("{\"num_steps\": 1e2, \"seed\": -7.9}", &p) == AM_OK && p.num_steps == 100 && p.seed == -7);
  CHECK(am_params_parse("{\"particle_temperatures\": {\"positron\": 1000.7, \"electron\": -0.5}}", &p) == AM_OK &&
        p.temperatures[AM_POSITRON] == 1000.0 && p.temperatures[AM_ELECTRON] == 0.0);
  CHECK(am_params_parse("{\"num_steps\": 18446744073709551615, \"seed\": 2147483647}", &p) == AM_OK);
  // extreme floating-point values go through the dt rule and Validate without trapping
  const char* extreme[] = {
      "{\"b0\": 0, \"particle_populations\": {\"antiproton\": 4}}",
      "{\"b0\": 1e-320, \"trap_voltage\": -1e308, \"z0\": 1e308}",
      "{\"b0\": -1, \"cooling_scale\": -1e308, \"bath_temperature\": 1e308}",
      "{\"initial_cloud_radii\": {\"radius_x\": 0, \"radius_y\": -1, \"radius_z\": 1e308}}",
      "{\"particle_populations\": {}}",
  };
  for (const char* s : extreme) {
    CHECK(am_params_parse(s, &p) == AM_OK);
    (void)am_params_validate(&p);
  }
  // method names longer than the struct's buffers, and unknown ones
  std::string longname(5000, 'x');
  std::string js = "{\"coulomb_computation_method\": \"" + longname + "\"}";
  CHECK(am_params_parse(js.c_str(), &p) != AM_OK);
  js = "{\"energy_computation_method\": \"" + longname + "\"}";
  CHECK(am_params_parse(js.c_str(), &p) != AM_OK);
  CHECK(std::strlen(p.coulomb_method) < AM_PARAMS_METHOD_LEN && std::strlen(p.energy_method) < AM_PARAMS_METHOD_LEN);
  CHECK(am_params_parse("{\"coulomb_computation_method\": \"barnes_hut\"}", &p) == AM_OK);
  CHECK(am_params_validate(&p) != AM_OK);
  int32_t method = 0;
  CHECK(am_params_method("barnes-hut", &method) == AM_OK);
  CHECK(am_params_method("nope", &method) != AM_OK);
  CHECK(am_params_method(nullptr, &method) != AM_OK);
  CHECK(am_params_method("standard", nullptr) != AM_OK);
  CHECK(am_params_parse(nullptr, &p) != AM_OK);
  CHECK(am_params_load(nullptr, &p) != AM_OK);
  CHECK(am_params_validate(nullptr) != AM_OK);
}

static void ExportChecks(const char* tmpdir) {
  const char* files[3] = {"a.vtu", "b.vtu", "c.vtu"};
  const double times[3] = {0.0, 1e-12, 2e-12};
  const std::string pvd = std::string(tmpdir) + "/run.pvd";
  CHECK(am_export_pvd(pvd.c_str(), files, times, 3) == AM_OK);
  CHECK(ReadFile(pvd.c_str()).find("c.vtu") != std::string::npos);
  CHECK(am_export_pvd(pvd.c_str(), files, times, 0) == AM_OK);
  CHECK(am_export_pvd(pvd.c_str(), files, times, -1) != AM_OK);
  CHECK(am_export_pvd(nullptr, files, times, 3) != AM_OK);
  CHECK(am_export_pvd((std::string(tmpdir) + "/no/such/dir/x.pvd").c_str(), files, times, 3) != AM_OK);
}

// Solvers of the oracle on one cloud: n particles, alternating e+/e-, some annihilated.
static void OracleOnCloud(int64_t n, const std::vector<double>& pos, bool with_dead) {
  std::vector<double> q(n), m(n), f(3 * n), vel(3 * n, 1.0), fc(3 * n, 0.0);
  std::vector<uint8_t> dead(n, 0), stag(n, 0), sp(n);
  for (int64_t i = 0; i < n; ++i) {
    sp[i] = (uint8_t)(i % 2 ? 3 : 1);  // electron / positron
    orc_species_props(sp[i], &m[i], &q[i]);
    if (with_dead && i % 7 == 3) dead[i] = 1;
  }
  orc_allpairs_forces(n, pos.data(), q.data(), dead.data(), f.data());
  orc_standard_forces(n, pos.data(), q.data(), dead.data(), 1e-4, f.data());
  (void)orc_allpairs_potential_energy(n, pos.data(), q.data(), dead.data());
  (void)orc_standard_energy(n, pos.data(), q.data(), dead.data(), 1e-4);
  std::vector<int64_t> inter(n);
  for (double th : {0.0, 0.5, 0.9}) {
    orc_bh_forces(n, pos.data(), q.data(), dead.data(), th, 8, f.data(), inter.data());
    (void)orc_bh_potential_energy(n, pos.data(), q.data(), dead.data(), th);
  }
  const int64_t nn = orc_bh_tree(n, pos.data(), q.data(), 8, 0, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                                 nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
  if (nn > 0) {
    std::vector<int64_t> order(n), b(nn), e(nn);
    std::vector<int32_t> depth(nn), child(8 * nn);
    std::vector<uint8_t> leaf(nn);
    std::vector<double> c(3 * nn), h(nn), qp(nn), qn(nn), cp(3 * nn), cn(3 * nn);
    CHECK(orc_bh_tree(n, pos.data(), q.data(), 8, nn, order.data(), b.data(), e.data(), depth.data(), leaf.data(),
                      c.data(), h.data(), qp.data(), qn.data(), cp.data(), cn.data(), child.data()) == nn);
  }
  const int64_t np = 5;
  std::vector<double> pts(3 * np, 1e-5), ch(np, 1.6e-19), fo(3 * np), phi(np);
  std::vector<int64_t> self(np, -1);
  if (n > 0) self[0] = 0;
  orc_bh_evaluate(n, pos.data(), q.data(), dead.data(), 0.6, 8, np, pts.data(), ch.data(), self.data(), fo.data(),
                  phi.data());
  for (int order = 1; order <= 8; ++order) {
    int64_t stats[4];
    orc_fmm_forces(n, pos.data(), q.data(), dead.data(), order, 0.5, 32, f.data(), stats);
  }
  (void)orc_fmm_potential_energy(n, pos.data(), q.data(), dead.data(), 4, 0.5);
  std::vector<double> p2(pos);
  orc_lorentz_push(n, p2.data(), vel.data(), fc.data(), q.data(), m.data(), dead.data(), stag.data(), 1.0, 10.0,
                   0.01, 1e-12);
  orc_cyclotron_cooling(n, vel.data(), q.data(), m.data(), dead.data(), 1.0, 1.0, 4.0, 1e-12, 42, 0);
  std::vector<int64_t> pairs(n + 2);
  (void)orc_annihilate(n, pos.data(), sp.data(), dead.data(), 1e-3, 0.5, 42, 1, pairs.data());
}

static void OracleChecks() {
  orc_set_num_threads(1);
  // the reference's initial conditions (100 e+ + 100 e-)
  const uint64_t counts[6] = {0, 100, 0, 100, 0, 0};
  const double radii[3] = {1e-4, 1e-4, 1e-4}, temps[6] = {0, 1000, 0, 1000, 0, 0};
  std::vector<double> pos(600), vel(600);
  CHECK(orc_set_particle_positions(counts, radii, 42, pos.data()) == 0);
  CHECK(orc_set_particle_velocities(counts, temps, 42, vel.data()) == 0);
  const int32_t listed[6] = {0, 1, 0, 1, 0, 0};
  CHECK(orc_derive_dt(listed, 1.0) > 0.0);
  OracleOnCloud(200, pos, false);
  OracleOnCloud(200, pos, true);
  // random clouds of a few sizes, and degenerate ones
  std::mt19937_64 rng(7);
  std::normal_distribution<double> g(0.0, 1e-4);
  for (int64_t n : {0, 1, 2, 3, 9, 33, 257}) {
    std::vector<double> p(3 * n);
    for (auto& v : p) v = g(rng);
    OracleOnCloud(n, p, n > 4);
  }
  {  // coincident particles, a line, and one far outlier
    std::vector<double> p(3 * 40, 0.0);
    for (int i = 20; i < 40; ++i) p[3 * i + 2] = 1e-6 * i;
    p[3 * 39] = 1.0;
    OracleOnCloud(40, p, false);
  }
}

int main(int argc, char** argv) {
  if (argc < 3) {
    std::fprintf(stderr, "usage: host_checks <golden params.json> <tmpdir>\n");
    return 2;
  }
  ParamsChecks(argv[1], argv[2]);
  ExportChecks(argv[2]);
  OracleChecks();
  if (g_fail) {
    std::printf("%d checks failed\n", g_fail);
    return 1;
  }
  std::printf("SANITIZE OK\n");
  return 0;
}
