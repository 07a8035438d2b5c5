// Internal declarations shared by the HIP translation units of libantimatter_amd.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "antimatter_amd/am_api.h"

namespace am {

// src/utils/constants.h:11-38
constexpr double kPi = 3.14159265358979323846;
constexpr double kElementaryCharge = 1.602176634e-19;
constexpr double kElectronMass = 9.1093837015e-31;
constexpr double kProtonMass = 1.67262192369e-27;
constexpr double kNeutronMass = 1.67492749804e-27;
constexpr double kCoulomb = 8.9875517923e9;
constexpr double kBoltzmann = 1.380649e-23;
constexpr double kLight = 2.99792458e8;
// Pairs closer than this (m^2) are skipped (barnes-hut.cc:22, coulomb_interaction.cc:19).
constexpr double kMinSquaredDistance = 1e-30;

// FP32 pair kernels work in a length unit of 2^-20 m (~0.95 um): power-of-two scaling is exact,
// and it keeps 1/r^3 finite in FP32 down to the 1e-15 m pair floor.
constexpr double kUnitScale = 1048576.0;          // metres -> units
constexpr double kInvUnit2 = 1048576.0 * 1048576.0;  // 1/unit^2 in 1/m^2
// Softening added to r^2 in the FP32 kernels: the 1e-30 m^2 pair floor in units^2.  It makes a
// coincident pair (including i == j) contribute exactly zero without a branch.
constexpr float kEps2Units = (float)(1e-30 * 1048576.0 * 1048576.0);

// Per-particle flag bits.
constexpr uint8_t kFlagAnnihilated = 1;
constexpr uint8_t kFlagStaggered = 2;

// ---- errors -------------------------------------------------------------------------------------
void SetError(const std::string& msg);
int Fail(int code, const std::string& msg);
int HipFail(hipError_t e, const char* what, const char* file, int line);

#define AM_HIP(call)                                                   \
  do {                                                                 \
    hipError_t am_e_ = (call);                                         \
    if (am_e_ != hipSuccess) return ::am::HipFail(am_e_, #call, __FILE__, __LINE__); \
  } while (0)

// ---- device buffers -----------------------------------------------------------------------------
struct Buffer {
  void* ptr = nullptr;
  size_t bytes = 0;
};

// Sizes of the most recent FMM call on a store (am_fmm_stats).
struct FmmRunStats {
  int64_t cells = 0, leaves = 0, m2l_pairs = 0, p2p_cell_pairs = 0, m = 0, m2l_classes = 0;
  int64_t m2l_unclassed = 0, m2l_items = 0;  // pairs outside every class; MFMA M2L work items
  int max_depth = 0;
};

// The most recent Barnes-Hut build on a store (am_bh_tree_stats / am_bh_tree_download).  order and
// nodes point into the context's scratch: valid until the next Barnes-Hut build on that context.
struct BHBuildInfo {
  int m = 0, nn = 0, max_depth = 0;
  int64_t deep_overflow = 0;
  const uint32_t* order = nullptr;
  const void* nodes = nullptr;  // WalkNode[nn]
};

}  // namespace am

struct am_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  int num_cus = 0;
  bool timing = false;
  // Unclassed M2L pairs that the FP32 MFMA M2L gives tensors of their own; past this count the FP64
  // contraction runs instead.
  int64_t fmm_max_unclassed = 1 << 20;
  bool fmm_classes = true;  // false: no M2L displacement classes (every pair unclassed; diagnostic)
  struct Pending {
    std::string name;
    hipEvent_t start, stop;
  };
  std::vector<Pending> pending;
  std::vector<hipEvent_t> event_pool;
  std::map<std::string, std::pair<double, int64_t>> timers;
  std::map<std::string, am::Buffer> scratch;  // grow-only named scratch buffers
  std::map<std::string, int> octree_depth;     // deepest non-empty level of the last build, per tree tag
};

struct am_store {
  am_ctx* ctx = nullptr;
  int64_t capacity = 0;
  int64_t n = 0;
  // FP64 master state (SoA).
  double *x = nullptr, *y = nullptr, *z = nullptr;
  double *vx = nullptr, *vy = nullptr, *vz = nullptr;
  double *fx = nullptr, *fy = nullptr, *fz = nullptr;  // Coulomb force (N)
  double *q = nullptr, *m = nullptr;                   // charge (C), mass (kg)
  uint8_t* species = nullptr;
  uint8_t* flags = nullptr;  // kFlagAnnihilated | kFlagStaggered
  uint32_t* trap = nullptr;
  // FP32 source mirror for the pair kernels: two float4 per particle,
  //   src[2i]   = {x_hi, y_hi, z_hi, q/e}   (q/e = 0 for annihilated particles)
  //   src[2i+1] = {x_lo, y_lo, z_lo, 0}
  // with x_units = x * 2^20 = x_hi + x_lo.  Refreshed by every kernel that moves particles.
  // Allocated for PadN(capacity) particles; entries [n, PadN(n)) are kept zero.
  float4* src = nullptr;
  // Spatial order cache of the sorted all-pairs kernel: perm[k] = particle at sorted position k,
  // rebuilt every few calls (the order only affects speed; the near/far split is exact per call).
  uint32_t* perm = nullptr;
  int64_t perm_n = -1;
  int perm_age = 0;
  // Bumped by every am_store_upload: objects built over the store (ensembles) record it at creation
  // and refuse to run once the particle set they were sized for has been replaced.
  uint64_t generation = 0;
  am::FmmRunStats fmm_stats;
  am::BHBuildInfo bh_info;
};

namespace am {

// Scratch buffer owned by the context, grown on demand (contents not preserved on growth).
int Scratch(am_ctx* ctx, const char* name, size_t bytes, void** out);
// Two work-queue counters [next item, finished workers] in a named scratch buffer, zeroed when first
// allocated.  The kernels that use them leave them at zero.
int QueueCounters(am_ctx* ctx, const char* name, unsigned** out);

// Kernel timing: Begin/End bracket a launch family on the context stream when timing is on.
void TimerBegin(am_ctx* ctx, const char* name, int* slot);
void TimerEnd(am_ctx* ctx, int slot);

inline int Blocks(int64_t n, int threads) { return (int)((n + threads - 1) / threads); }
// Particle counts padded to a multiple of 256 for the FP32 source mirror (padding entries are
// zero: position 0, charge 0, so they contribute nothing).
inline int PadN(int64_t n) { return (int)((n + 255) / 256 * 256); }

// Species defaults (src/agents/particle_species.h:27-46).
inline void SpeciesProperties(int s, double* mass, double* charge) {
  switch (s) {
    case AM_ANTIPROTON: *mass = kProtonMass; *charge = -kElementaryCharge; return;
    case AM_POSITRON: *mass = kElectronMass; *charge = +kElementaryCharge; return;
    case AM_ANTIHYDROGEN: *mass = kProtonMass + kElectronMass; *charge = 0.0; return;
    case AM_ELECTRON: *mass = kElectronMass; *charge = -kElementaryCharge; return;
    case AM_PROTON: *mass = kProtonMass; *charge = +kElementaryCharge; return;
    case AM_ANTIHELIUM: *mass = 2 * (kProtonMass + kNeutronMass); *charge = 0.0; return;
    default: *mass = 0.0; *charge = 0.0; return;
  }
}

// IsAntiparticlePair (src/interaction/annihilation.cc:12-19).
__host__ __device__ inline bool IsPair(int a, int b) {
  return (a == AM_ANTIPROTON && b == AM_PROTON) || (a == AM_PROTON && b == AM_ANTIPROTON) ||
         (a == AM_POSITRON && b == AM_ELECTRON) || (a == AM_ELECTRON && b == AM_POSITRON);
}

// ---- Philox4x32-10 (Salmon et al., SC'11), shared bit-for-bit by host and device ---------------
__host__ __device__ inline void Philox(uint32_t c[4], uint32_t k0, uint32_t k1) {
  for (int r = 0; r < 10; ++r) {
    if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
    const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
    const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
    c[1] = (uint32_t)p1;
    c[3] = (uint32_t)p0;
    c[0] = n0;
    c[2] = n2;
  }
}
// Uniform in [0, 1) with 24 bits, exactly representable: identical on host and device.
__host__ __device__ inline double U24(uint32_t w) { return (double)(w >> 8) * (1.0 / 16777216.0); }
// Annihilation draw for candidate j scanned by i (purpose word 1).
__host__ __device__ inline double AnnihilationUniform(uint64_t seed, uint64_t step, uint32_t i, uint32_t j) {
  uint32_t c[4] = {(uint32_t)step, i, j, 1u};
  Philox(c, (uint32_t)seed, (uint32_t)(seed >> 32));
  return U24(c[0]);
}

// ---- FP64 helpers that must not be contracted (CPU-identical rounding) -------------------------
// The library is compiled with -ffp-contract=off; these make the intent explicit.
__device__ inline double Dot3(double ax, double ay, double az, double bx, double by, double bz) {
  return __dadd_rn(__dadd_rn(__dmul_rn(ax, bx), __dmul_rn(ay, by)), __dmul_rn(az, bz));
}

__host__ __device__ inline bool Finite3(double x, double y, double z) {
  return isfinite(x) && isfinite(y) && isfinite(z);
}

// Particle i's entry of the FP32 source mirror (see am_store::src).  An annihilated particle, or one
// with a non-finite coordinate, gets an all-zero entry (position 0, charge 0): it then acts on no
// pair sum, as the direct solver skips non-finite positions (coulomb_interaction.cc:55-56), and it
// cannot turn another particle's sum into NaN.
__device__ inline void WriteMirror(float4* __restrict__ src, int64_t i, double x, double y, double z, double q,
                                   bool dead) {
  if (dead || !Finite3(x, y, z)) {
    src[2 * i] = make_float4(0.f, 0.f, 0.f, 0.f);
    src[2 * i + 1] = make_float4(0.f, 0.f, 0.f, 0.f);
    return;
  }
  const double xs = x * kUnitScale, ys = y * kUnitScale, zs = z * kUnitScale;
  const float xh = (float)xs, yh = (float)ys, zh = (float)zs;
  src[2 * i] = make_float4(xh, yh, zh, (float)(q / kElementaryCharge));
  src[2 * i + 1] = make_float4((float)(xs - (double)xh), (float)(ys - (double)yh), (float)(zs - (double)zh), 0.f);
}

// FP32 pair interaction on hi/lo coordinates (units of 2^-20 m): accumulates
// q_j (x_i - x_j) / (r^2 + eps^2)^(3/2) into (px, py, pz).  dx = (xh_i - xh_j) + (xl_i - xl_j) keeps
// ~2^-24 relative accuracy on dx even for pairs much closer than the coordinates' magnitude.
__device__ __forceinline__ void PairF32(const float4 th, const float4 tl, const float4 sh, const float4 sl,
                                        float& px, float& py, float& pz) {
  const float dx = (th.x - sh.x) + (tl.x - sl.x);
  const float dy = (th.y - sh.y) + (tl.y - sl.y);
  const float dz = (th.z - sh.z) + (tl.z - sl.z);
  const float r2 = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, kEps2Units)));
  const float ri = __builtin_amdgcn_rsqf(r2);
  const float s = (sh.w * ri) * (ri * ri);
  px = fmaf(s, dx, px);
  py = fmaf(s, dy, py);
  pz = fmaf(s, dz, pz);
}

// Launch helpers ----------------------------------------------------------------------------------
int RefreshSource(am_store* s);  // rebuild store->src from the FP64 state
// Zeroes the Coulomb force of alive particles with a non-finite coordinate (the solvers leave them
// out of their trees and pair sums, as coulomb_interaction.cc:55-56 skips them).
int ZeroNonFiniteForces(am_store* s);
int64_t CountActive(am_store* s, int* status);

// Radix sort (stable, LSD) of 64-bit keys with 32-bit values on the keys' low key_bits bits (higher
// bits are carried, not compared; from 2^30 keys on they must be zero up to the next multiple of 8).
// keys/vals are sorted in place (ping-pong through scratch).
int RadixSortPairs(am_ctx* ctx, uint64_t* keys, uint32_t* vals, int64_t n, int key_bits);
// Stable radix sort of 64-bit keys alone on bits [bit0, bit0 + key_bits), in place; n < 2^30.
int RadixSortKeys(am_ctx* ctx, uint64_t* keys, int64_t n, int bit0, int key_bits);
// Deterministic FP64 sum of n device doubles into *host_out (synchronises the stream).
int ReduceSum(am_ctx* ctx, const double* in, int64_t n, double* host_out);
// The radix sort's look-back error word on the device (set by a wait that gave up); callers read it
// with a synchronisation they make anyway and pass it to CheckSortErrorAfterSync, which fails the
// call and clears the word, so one give-up is reported once.  CheckSortError reads it with a
// synchronisation of its own.
uint32_t* SortErrorWord(am_ctx* ctx);
int CheckSortErrorAfterSync(am_ctx* ctx, uint32_t word);
int CheckSortError(am_ctx* ctx);
// Pre-sizes the scratch of RadixSortPairs / ExclusiveScan for n elements (no allocation later).
int ReserveSortScratch(am_ctx* ctx, int64_t n);
// Exclusive prefix sum of n uint32 counts into out (n+1 entries; out[n] = total).
int ExclusiveScan(am_ctx* ctx, const uint32_t* in, uint32_t* out, int64_t n);

}  // namespace am
