// Batched ensemble of independent Penning traps in one store.
//
// Every trap is an independent instance of the reference's Simulate loop
// (src/antimatter-kernel.cc:153-197) with its own field, dt, cooling and annihilation settings:
// per step, exact all-pairs Coulomb forces inside the trap (FP32 pair math on hi/lo coordinates,
// FP64 tile accumulation), LorentzForce + CyclotronCooling, then AnnihilationInteraction in the
// canonical order (scan i ascending within the trap, lowest-index passing partner).
//
// Random draws are keyed per trap: trap t uses seed_t = seed + t * 0x9E3779B97F4A7C15 and its
// particles' trap-local indices, so each trap draws exactly what a single-trap run with seed_t
// draws.
//
// GPU structure:
//   * force: a work list of (trap, 256-target block, source chunk) items, sorted by cost, drained
//     by a persistent kernel (one atomic ticket per workgroup); every item writes its own partial
//     sums, reduced per particle in chunk order, so the result does not depend on scheduling;
//     the source-chunk length is chosen by an autotuner that times the candidates on the device;
//   * push: the fused Boris + cooling kernel with per-trap parameters from a device trap table;
//   * contact: cell keys (trap | cell) sorted with the stable radix sort, per-particle candidate
//     counts, a scan, candidate lists in (i, j) order into a fixed-capacity buffer, then one
//     workgroup per trap resolves its list serially with the particle marks in LDS;
//   * no host synchronisation inside a step: the step counter lives in device memory, so one step
//     is captured once in a hipGraph and replayed.
#include <algorithm>
#include <cmath>
#include <vector>

#include "celltable.hpp"
#include "pairtile.hpp"
#include "physics.hpp"

namespace am {

struct TrapRec {
  int64_t begin, end;
  double b0, u0, d, dt, tau_scale, bath_temp, contact_radius, probability;
  double inv_h;  // 1 / contact cell edge (edge >= 2 x contact radius)
  int annihilation;
  int chunks;         // source chunks of the force work items
  int64_t part_base;  // the trap's partial sums: [chunk][3][ntile * kTile]
  int tile0, ntile;   // the trap's tiles in the sorted layout (ntile even: whole 256-target blocks)
  int64_t sorted0;    // first rank of the trap's particles in the sorted keys
};

// Force work item: a 256-target block of a trap against one chunk of the trap's source tiles.
struct EnsItem {
  int trap;
  int blk;    // target block (tiles 2 blk, 2 blk + 1 of the trap)
  int chunk;  // source tiles [chunk * tiles_per_chunk, ...)
  int cost;
};

constexpr int kEnsThreads = 256;
constexpr int kEnsTile = kTile;
constexpr int kMaxTraps = 1024;  // trap id in 10 key bits

__host__ __device__ inline uint64_t TrapSeed(uint64_t seed, uint32_t t) { return seed + (uint64_t)t * 0x9E3779B97F4A7C15ull; }

__global__ void EnsResetKernel(unsigned* __restrict__ counter) { *counter = 0u; }

__global__ void EnsStepKernel(uint64_t* __restrict__ step) { *step += 1ull; }

// ---- forces: per-trap Morton-sorted tiles (the all-pairs kernel's tile machinery) -----------------
// Every step: each trap's bounding box, 30-bit Morton keys with the trap id above them, one radix
// sort, the per-trap sorted layout (trap t's particles at tiles tile0_t.., padded to whole
// 256-target blocks), tile preparation, then the persistent force kernel over (trap, block,
// source chunk) items and the fixed-order chunk reduction.
constexpr int kEnsKeyBits = 41;  // 11 trap bits (kMaxTraps = particles outside every trap) + 30

__global__ __launch_bounds__(256) void EnsBoxKernel(const TrapRec* __restrict__ traps, const double* __restrict__ x,
                                                    const double* __restrict__ y, const double* __restrict__ z,
                                                    double* __restrict__ box) {
  __shared__ double lo[3][256], hi[3][256];
  const TrapRec& tr = traps[blockIdx.x];
  const int t = threadIdx.x;
  double l[3] = {INFINITY, INFINITY, INFINITY}, h[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int64_t i = tr.begin + t; i < tr.end; i += 256) {
    const double p[3] = {x[i], y[i], z[i]};
    for (int d = 0; d < 3; ++d) {
      l[d] = fmin(l[d], p[d]);
      h[d] = fmax(h[d], p[d]);
    }
  }
  for (int d = 0; d < 3; ++d) {
    lo[d][t] = l[d];
    hi[d][t] = h[d];
  }
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w)
      for (int d = 0; d < 3; ++d) {
        lo[d][t] = fmin(lo[d][t], lo[d][t + w]);
        hi[d][t] = fmax(hi[d][t], hi[d][t + w]);
      }
    __syncthreads();
  }
  if (t < 3) {
    const bool any = lo[t][0] <= hi[t][0];
    box[6 * blockIdx.x + t] = any ? lo[t][0] : 0.0;
    box[6 * blockIdx.x + 3 + t] = any ? hi[t][0] : 0.0;
  }
}

__global__ void EnsSortKeyKernel(int64_t n, int n_traps, const uint32_t* __restrict__ trap,
                                 const double* __restrict__ x, const double* __restrict__ y,
                                 const double* __restrict__ z, const double* __restrict__ box,
                                 uint64_t* __restrict__ keys, uint32_t* __restrict__ vals) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  vals[i] = (uint32_t)i;
  const uint32_t t = trap[i];
  if (t >= (uint32_t)n_traps) {
    keys[i] = (uint64_t)kMaxTraps << 30;
    return;
  }
  const double* b = box + 6 * t;
  const double ext = fmax(fmax(b[3] - b[0], b[4] - b[1]), fmax(b[5] - b[2], 1e-300));
  const double sc = 1023.999 / ext;
  const double p[3] = {x[i], y[i], z[i]};
  uint32_t g[3];
  for (int d = 0; d < 3; ++d) {
    double c = (p[d] - b[d]) * sc;
    c = (c == c) ? fmin(fmax(c, 0.0), 1023.0) : 0.0;
    g[d] = (uint32_t)c;
  }
  keys[i] = ((uint64_t)t << 30) | (uint64_t)(Spread10(g[0]) | (Spread10(g[1]) << 1) | (Spread10(g[2]) << 2));
}

// Rank r of the sorted keys -> the trap's sorted layout.
__global__ void EnsScatterPermKernel(int64_t n_in, const uint64_t* __restrict__ keys, const uint32_t* __restrict__ vals,
                                     const TrapRec* __restrict__ traps, uint32_t* __restrict__ perm) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n_in) return;
  const TrapRec& tr = traps[keys[r] >> 30];
  perm[(int64_t)tr.tile0 * kEnsTile + (r - tr.sorted0)] = vals[r];
}

__global__ __launch_bounds__(kEnsTile) void EnsTilePrepKernel(const int* __restrict__ tile_trap,
                                                              const TrapRec* __restrict__ traps,
                                                              uint32_t* __restrict__ perm, const double* __restrict__ x,
                                                              const double* __restrict__ y,
                                                              const double* __restrict__ z,
                                                              const float4* __restrict__ src, float4* __restrict__ shl,
                                                              float4* __restrict__ sloc, TileRec* __restrict__ tiles) {
  const int T = blockIdx.x;
  const TrapRec& tr = traps[tile_trap[T]];
  const int nvalid = (int)((tr.end - tr.begin) - (int64_t)(T - tr.tile0) * kEnsTile);
  TilePrepBody(T * kEnsTile, nvalid, T, perm, x, y, z, src, shl, sloc, tiles);
}

__global__ __launch_bounds__(kEnsThreads) __attribute__((amdgpu_waves_per_eu(8, 8))) void EnsembleForceKernel(const EnsItem* __restrict__ items, int n_items,
                                                                   unsigned* __restrict__ counter,
                                                                   const TrapRec* __restrict__ traps,
                                                                   int tiles_per_chunk,
                                                                   const uint32_t* __restrict__ perm,
                                                                   const double* __restrict__ x,
                                                                   const double* __restrict__ y,
                                                                   const double* __restrict__ z,
                                                                   const float4* __restrict__ shl,
                                                                   const float4* __restrict__ sloc,
                                                                   const TileRec* __restrict__ tiles,
                                                                   double* __restrict__ part) {
  __shared__ int s_item;
  __shared__ float4 s_stage[kTile];  // the expanded far-tile sums' source records
  for (;;) {
    __syncthreads();
    if (threadIdx.x == 0) s_item = (int)atomicAdd(counter, 1u);
    __syncthreads();
    const int it = __builtin_amdgcn_readfirstlane(s_item);
    if (it >= n_items) return;
    const EnsItem w = items[it];
    const TrapRec& tr = traps[w.trap];
    const int N = (int)(tr.end - tr.begin);
    const int local = w.blk * kEnsThreads + (int)threadIdx.x;
    const int64_t k = (int64_t)tr.tile0 * kEnsTile + local;
    const int tb = tr.tile0 + 2 * w.blk;
    const BlockBox bb = MakeBlockBox(tiles[tb], tiles[tb + 1], w.blk * kEnsThreads + kEnsTile >= N);
    const float4 th = shl[2 * k], tl = shl[2 * k + 1];
    float lx = 0.f, ly = 0.f, lz = 0.f;
    if (local < N) {
      const uint32_t i = perm[k];
      lx = (float)(x[i] * kUnitScale - bb.bc[0]);
      ly = (float)(y[i] * kUnitScale - bb.bc[1]);
      lz = (float)(z[i] * kUnitScale - bb.bc[2]);
    }
    const int t0 = tr.tile0 + w.chunk * tiles_per_chunk;
    const int t1 = min(tr.tile0 + tr.ntile, t0 + tiles_per_chunk);
    double ax = 0.0, ay = 0.0, az = 0.0;
    for (int T = t0; T < t1; ++T) {
      float px = 0.f, py = 0.f, pz = 0.f;
      TileSum(tiles[T], T * kEnsTile, bb, th, tl, lx, ly, lz, shl, sloc, s_stage, px, py, pz);
      ax += (double)px;
      ay += (double)py;
      az += (double)pz;
    }
    const int64_t npad = (int64_t)tr.ntile * kEnsTile;
    double* out = part + tr.part_base + (int64_t)w.chunk * 3 * npad;
    out[local] = ax;
    out[npad + local] = ay;
    out[2 * npad + local] = az;
  }
}

// One thread per position of the sorted layout: the fixed-order sum over chunks, scaled into N.
__global__ void EnsembleReduceKernel(int64_t n_pos, const int* __restrict__ tile_trap,
                                     const TrapRec* __restrict__ traps, const uint32_t* __restrict__ perm,
                                     const double* __restrict__ part, const double* __restrict__ x,
                                     const double* __restrict__ y, const double* __restrict__ z,
                                     const double* __restrict__ q, const uint8_t* __restrict__ flags,
                                     double* __restrict__ fx,
                                     double* __restrict__ fy, double* __restrict__ fz) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n_pos) return;
  const TrapRec& tr = traps[tile_trap[k / kEnsTile]];
  const int64_t local = k - (int64_t)tr.tile0 * kEnsTile;
  if (local >= tr.end - tr.begin) return;
  const int64_t npad = (int64_t)tr.ntile * kEnsTile;
  double sx = 0.0, sy = 0.0, sz = 0.0;
  for (int c = 0; c < tr.chunks; ++c) {
    const double* p = part + tr.part_base + (int64_t)c * 3 * npad;
    sx += p[local];
    sy += p[npad + local];
    sz += p[2 * npad + local];
  }
  const uint32_t i = perm[k];
  const bool live = !(flags[i] & kFlagAnnihilated) && q[i] != 0.0 && Finite3(x[i], y[i], z[i]);
  const double scale = live ? kCoulomb * q[i] * kElementaryCharge * kInvUnit2 : 0.0;
  fx[i] = scale * sx;
  fy[i] = scale * sy;
  fz[i] = scale * sz;
}

__global__ __launch_bounds__(256) void EnsemblePushKernel(int64_t n, double* __restrict__ x, double* __restrict__ y,
                                                          double* __restrict__ z, double* __restrict__ vx,
                                                          double* __restrict__ vy, double* __restrict__ vz,
                                                          double* __restrict__ fx, double* __restrict__ fy,
                                                          double* __restrict__ fz, const double* __restrict__ q,
                                                          const double* __restrict__ m, uint8_t* __restrict__ flags,
                                                          float4* __restrict__ src, const uint32_t* __restrict__ trap,
                                                          const TrapRec* __restrict__ traps, uint64_t seed,
                                                          const uint64_t* __restrict__ step) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t t = trap[i];
  if (t >= (uint32_t)kMaxTraps) return;
  const TrapRec& tr = traps[t];
  const PushParams p{tr.b0, tr.u0, tr.d, tr.dt, 1, tr.tau_scale, tr.bath_temp, TrapSeed(seed, t), *step};
  PushParticle(i, x, y, z, vx, vy, vz, fx, fy, fz, q, m, flags, src, p, (uint32_t)(i - tr.begin));
}

// ---- contact detection within each trap ---------------------------------------------------------
constexpr int kECellBits = 18;
constexpr uint64_t kEDeadKey = kCellEmpty;
constexpr double kEnsContactReach = 1.0 + 1e-6;  // search reach / contact radius; cell edge = 2 reach

__device__ inline int64_t ECoord(double v, double inv_h) {
  const double c = floor(v * inv_h);
  const double half = (double)(1 << (kECellBits - 1));
  if (!(c == c)) return 1 << (kECellBits - 1);
  return (int64_t)fmin(fmax(c, -half), half - 1.0) + (1 << (kECellBits - 1));
}

__device__ inline uint64_t EKey(uint64_t t, int64_t cx, int64_t cy, int64_t cz) {
  return (t << (3 * kECellBits)) | ((uint64_t)cx << (2 * kECellBits)) | ((uint64_t)cy << kECellBits) | (uint64_t)cz;
}

__global__ void EnsCellKeyKernel(int64_t n, const double* __restrict__ x, const double* __restrict__ y,
                                 const double* __restrict__ z, const uint8_t* __restrict__ flags,
                                 const uint32_t* __restrict__ trap, const TrapRec* __restrict__ traps,
                                 uint64_t* __restrict__ keys, uint32_t* __restrict__ idx) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  idx[i] = (uint32_t)i;
  const uint32_t t = trap[i];
  if (t >= (uint32_t)kMaxTraps || (flags[i] & kFlagAnnihilated) || !traps[t].annihilation) {
    keys[i] = kEDeadKey;
    return;
  }
  const double ih = traps[t].inv_h;
  keys[i] = EKey(t, ECoord(x[i], ih), ECoord(y[i], ih), ECoord(z[i], ih));
}

template <typename F>
__device__ inline void EnsForEachCandidate(int64_t i, int64_t n, const double* __restrict__ x,
                                           const double* __restrict__ y, const double* __restrict__ z,
                                           const uint8_t* __restrict__ species, const uint8_t* __restrict__ flags,
                                           const uint32_t* __restrict__ trap, const TrapRec* __restrict__ traps,
                                           const CellTable& tab, const uint32_t* __restrict__ idx, F&& f) {
  const uint32_t t = trap[i];
  const TrapRec& tr = traps[t];
  const double px = x[i], py = y[i], pz = z[i];
  const double ra2 = tr.contact_radius * tr.contact_radius;
  const int si = species[i];
  // cell edge >= 2 reach: along each axis at most the two cells from ECoord(p - reach) to
  // ECoord(p + reach) can hold a partner (ECoord is monotone and 1-Lipschitz in cells)
  const double reach = tr.contact_radius * kEnsContactReach;
  const int64_t x0 = ECoord(px - reach, tr.inv_h), x1 = ECoord(px + reach, tr.inv_h);
  const int64_t y0 = ECoord(py - reach, tr.inv_h), y1 = ECoord(py + reach, tr.inv_h);
  const int64_t z0 = ECoord(pz - reach, tr.inv_h), z1 = ECoord(pz + reach, tr.inv_h);
  for (int64_t az = z0; az <= z1; ++az)
    for (int64_t ay = y0; ay <= y1; ++ay)
      for (int64_t ax = x0; ax <= x1; ++ax) {
        const uint2 r = CellLookup(tab, EKey(t, ax, ay, az));
        for (uint32_t k = r.x; k < r.x + r.y; ++k) {
          const int64_t j = idx[k];
          if (j == i || (flags[j] & kFlagAnnihilated)) continue;
          if (!IsPair(si, species[j])) continue;
          const double rx = px - x[j], ry = py - y[j], rz = pz - z[j];
          if (!(Dot3(rx, ry, rz, rx, ry, rz) < ra2)) continue;
          f(j);
        }
      }
}

__global__ void EnsContactCountKernel(int64_t n, const double* __restrict__ x, const double* __restrict__ y,
                                      const double* __restrict__ z, const uint8_t* __restrict__ species,
                                      const uint8_t* __restrict__ flags, const uint32_t* __restrict__ trap,
                                      const TrapRec* __restrict__ traps, CellTable tab,
                                      const uint32_t* __restrict__ idx, uint32_t* __restrict__ count) {
  // Particle order: the particle's own trap, flags, species and position are read contiguously.
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t c = 0;
  const uint32_t t = trap[i];
  if (t < (uint32_t)kMaxTraps && !(flags[i] & kFlagAnnihilated) && traps[t].annihilation)
    EnsForEachCandidate(i, n, x, y, z, species, flags, trap, traps, tab, idx, [&](int64_t) { ++c; });
  count[i] = c;
}

__global__ void EnsContactWriteKernel(int64_t n, const double* __restrict__ x, const double* __restrict__ y,
                                      const double* __restrict__ z, const uint8_t* __restrict__ species,
                                      const uint8_t* __restrict__ flags, const uint32_t* __restrict__ trap,
                                      const TrapRec* __restrict__ traps, CellTable tab,
                                      const uint32_t* __restrict__ idx, const uint32_t* __restrict__ off,
                                      int64_t cap, uint2* __restrict__ pairs, uint32_t* __restrict__ overflow) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;  // particle order, as the count
  if (i >= n) return;
  const uint32_t b = off[i], e = off[i + 1];
  if (b == e) return;
  if ((int64_t)e > cap) {
    atomicOr(overflow, 1u);
    return;
  }
  uint32_t w = b;
  EnsForEachCandidate(i, n, x, y, z, species, flags, trap, traps, tab, idx, [&](int64_t j) {
    uint32_t k = w++;
    while (k > b && pairs[k - 1].y > (uint32_t)j) {
      pairs[k] = pairs[k - 1];
      --k;
    }
    pairs[k] = make_uint2((uint32_t)i, (uint32_t)j);
  });
}

constexpr int kResolveBits = 65536;  // particles per trap with marks in LDS

// One workgroup per trap: the canonical greedy scan over the trap's (i, j)-sorted candidates.
__global__ __launch_bounds__(256) void EnsResolveKernel(const TrapRec* __restrict__ traps,
                                                        const uint32_t* __restrict__ off,
                                                        const uint2* __restrict__ pairs, int64_t cap, uint64_t seed,
                                                        const uint64_t* __restrict__ step, uint8_t* __restrict__ flags,
                                                        float4* __restrict__ src, uint8_t* __restrict__ pass,
                                                        uint8_t* __restrict__ gmark,
                                                        unsigned long long* __restrict__ n_marked) {
  __shared__ uint32_t bits[kResolveBits / 32];
  const TrapRec tr = traps[blockIdx.x];
  if (!tr.annihilation) return;
  const int64_t b = std::min<int64_t>(off[tr.begin], cap), e = std::min<int64_t>(off[tr.end], cap);
  if (b >= e) return;
  const int64_t N = tr.end - tr.begin;
  const bool lds = N <= kResolveBits;
  const uint64_t st = *step;
  const uint64_t tseed = TrapSeed(seed, blockIdx.x);
  if (lds) {
    for (int k = threadIdx.x; k < kResolveBits / 32; k += blockDim.x) bits[k] = 0u;
  } else {
    for (int64_t k = threadIdx.x; k < N; k += blockDim.x) gmark[tr.begin + k] = 0;
  }
  for (int64_t q = b + threadIdx.x; q < e; q += blockDim.x) {
    const uint2 pr = pairs[q];
    pass[q] = AnnihilationUniform(tseed, st, (uint32_t)(pr.x - tr.begin), (uint32_t)(pr.y - tr.begin)) < tr.probability
                  ? 1
                  : 0;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    auto marked = [&](uint32_t p) -> bool {
      const int64_t l = (int64_t)p - tr.begin;
      return lds ? ((bits[l >> 5] >> (l & 31)) & 1u) : gmark[p] != 0;
    };
    auto mark = [&](uint32_t p) {
      const int64_t l = (int64_t)p - tr.begin;
      if (lds) bits[l >> 5] |= 1u << (l & 31);
      else gmark[p] = 1;
    };
    unsigned long long count = 0;
    int64_t q = b;
    while (q < e) {
      const uint32_t i = pairs[q].x;
      int64_t r = q;
      while (r < e && pairs[r].x == i) ++r;
      if (!marked(i)) {
        for (int64_t k = q; k < r; ++k) {
          const uint32_t j = pairs[k].y;
          if (marked(j)) continue;
          if (pass[k]) {
            mark(i);
            mark(j);
            count += 2;
            break;
          }
        }
      }
      q = r;
    }
    if (count) atomicAdd(n_marked, count);
  }
  __syncthreads();
  for (int64_t l = threadIdx.x; l < N; l += blockDim.x) {
    const bool mk = lds ? ((bits[l >> 5] >> (l & 31)) & 1u) : gmark[tr.begin + l] != 0;
    if (!mk) continue;
    flags[tr.begin + l] |= kFlagAnnihilated;
    src[2 * (tr.begin + l)].w = 0.f;
  }
}

}  // namespace am

using namespace am;

struct am_ensemble {
  am_store* store = nullptr;
  int n_traps = 0;
  std::vector<TrapRec> traps;
  TrapRec* d_traps = nullptr;
  int chunk = 4096;
  EnsItem* d_items = nullptr;
  int n_items = 0;
  double* part = nullptr;
  unsigned* counter = nullptr;
  uint64_t* step = nullptr;
  unsigned long long* n_marked = nullptr;
  uint32_t* overflow = nullptr;
  int64_t steps_done = 0;
  bool any_annihilation = false;
  int64_t contact_cap = 0;
  // contact buffers
  uint64_t* ckeys = nullptr;
  uint32_t* cidx = nullptr;
  CellSlot* hslots = nullptr;  // cell table of the sorted contact keys
  uint64_t hsize = 0;
  // force buffers: sort keys, sorted layout, tiles
  uint64_t* fkeys = nullptr;
  uint32_t* fvals = nullptr;
  double* tbox = nullptr;
  uint32_t* perm = nullptr;
  float4* shl = nullptr;
  float4* sloc = nullptr;
  TileRec* tiles = nullptr;
  int* tile_trap = nullptr;
  int n_tiles = 0;
  int64_t n_in = 0;  // particles inside traps
  int tiles_per_chunk = 32;
  uint32_t* ccount = nullptr;
  uint32_t* coff = nullptr;
  uint2* cpairs = nullptr;
  uint8_t* cpass = nullptr;
  uint8_t* gmark = nullptr;
  // graph
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  uint64_t graph_seed = 0;
  bool graph_valid = false;
  std::vector<void*> owned;
  // The store's particle set at creation: every buffer above is sized for n_at_create, and the
  // trap ids written then must still be in place.
  uint64_t store_generation = 0;
  int64_t n_at_create = 0;
};

namespace am {

// Fails (before any launch) when the store was re-uploaded after am_ensemble_create.
static int EnsCheckStore(const am_ensemble* e, const char* who) {
  if (e->store->generation != e->store_generation || e->store->n != e->n_at_create)
    return Fail(AM_ERR_INVALID, std::string(who) + ": the store was re-uploaded after am_ensemble_create "
                                                   "(buffers and trap ids are stale); create a new ensemble");
  return AM_OK;
}

static int EnsAlloc(am_ensemble* e, size_t bytes, void** p) {
  AM_HIP(hipMalloc(p, std::max<size_t>(bytes, 16)));
  e->owned.push_back(*p);
  return AM_OK;
}

// Work items and partial-buffer layout for a source-chunk length (rounded to whole tiles).
static int EnsBuildItems(am_ensemble* e, int chunk) {
  am_ctx* ctx = e->store->ctx;
  const int tpc = std::max(1, (chunk + kEnsTile - 1) / kEnsTile);
  std::vector<EnsItem> items;
  int64_t base = 0;
  for (int t = 0; t < e->n_traps; ++t) {
    TrapRec& tr = e->traps[t];
    const int64_t N = tr.end - tr.begin;
    tr.chunks = N > 0 ? (tr.ntile + tpc - 1) / tpc : 0;
    tr.part_base = base;
    base += 3 * (int64_t)tr.ntile * kEnsTile * tr.chunks;
    for (int b = 0; 2 * b < tr.ntile; ++b)
      for (int c = 0; c < tr.chunks; ++c) {
        EnsItem w;
        w.trap = t;
        w.blk = b;
        w.chunk = c;
        const int nt = (int)std::min<int64_t>(kEnsThreads, N - (int64_t)b * kEnsThreads);
        const int ns = std::min(tpc, tr.ntile - c * tpc);
        w.cost = nt * ns;
        items.push_back(w);
      }
  }
  std::stable_sort(items.begin(), items.end(), [](const EnsItem& a, const EnsItem& b) { return a.cost > b.cost; });
  void* p;
  int rc;
  if ((rc = Scratch(ctx, "ens_items", std::max<size_t>(items.size(), 1) * sizeof(EnsItem), &p))) return rc;
  e->d_items = (EnsItem*)p;
  if ((rc = Scratch(ctx, "ens_part", (size_t)std::max<int64_t>(base, 1) * 8, &p))) return rc;
  e->part = (double*)p;
  e->n_items = (int)items.size();
  e->chunk = tpc * kEnsTile;
  e->tiles_per_chunk = tpc;
  if (!items.empty())
    AM_HIP(hipMemcpyAsync(e->d_items, items.data(), items.size() * sizeof(EnsItem), hipMemcpyHostToDevice,
                          ctx->stream));
  AM_HIP(hipMemcpyAsync(e->d_traps, e->traps.data(), e->traps.size() * sizeof(TrapRec), hipMemcpyHostToDevice,
                        ctx->stream));
  AM_HIP(hipStreamSynchronize(ctx->stream));
  e->graph_valid = false;
  return AM_OK;
}

static int EnsForce(am_ensemble* e) {
  am_store* s = e->store;
  am_ctx* ctx = s->ctx;
  hipStream_t st = ctx->stream;
  const int64_t n = s->n;
  int rc;
  hipLaunchKernelGGL(EnsBoxKernel, dim3(e->n_traps), dim3(256), 0, st, (const TrapRec*)e->d_traps, s->x, s->y, s->z,
                     e->tbox);
  hipLaunchKernelGGL(EnsSortKeyKernel, dim3(Blocks(n, 256)), dim3(256), 0, st, n, e->n_traps,
                     (const uint32_t*)s->trap, s->x, s->y, s->z, (const double*)e->tbox, e->fkeys, e->fvals);
  AM_HIP(hipGetLastError());
  if ((rc = RadixSortPairs(ctx, e->fkeys, e->fvals, n, kEnsKeyBits))) return rc;
  if (e->n_in > 0)
    hipLaunchKernelGGL(EnsScatterPermKernel, dim3(Blocks(e->n_in, 256)), dim3(256), 0, st, e->n_in,
                       (const uint64_t*)e->fkeys, (const uint32_t*)e->fvals, (const TrapRec*)e->d_traps, e->perm);
  if (e->n_tiles > 0)
    hipLaunchKernelGGL(EnsTilePrepKernel, dim3(e->n_tiles), dim3(kEnsTile), 0, st, (const int*)e->tile_trap,
                       (const TrapRec*)e->d_traps, e->perm, s->x, s->y, s->z, (const float4*)s->src, e->shl, e->sloc,
                       e->tiles);
  const int grid = std::max(1, std::min(e->n_items, (ctx->num_cus > 0 ? ctx->num_cus : 256) * 8));
  hipLaunchKernelGGL(EnsResetKernel, dim3(1), dim3(1), 0, st, e->counter);
  hipLaunchKernelGGL(EnsembleForceKernel, dim3(grid), dim3(kEnsThreads), 0, st, (const EnsItem*)e->d_items,
                     e->n_items, e->counter, (const TrapRec*)e->d_traps, e->tiles_per_chunk, (const uint32_t*)e->perm,
                     s->x, s->y, s->z, (const float4*)e->shl, (const float4*)e->sloc, (const TileRec*)e->tiles,
                     e->part);
  const int64_t n_pos = (int64_t)e->n_tiles * kEnsTile;
  if (n_pos > 0)
    hipLaunchKernelGGL(EnsembleReduceKernel, dim3(Blocks(n_pos, 256)), dim3(256), 0, st, n_pos,
                       (const int*)e->tile_trap, (const TrapRec*)e->d_traps, (const uint32_t*)e->perm,
                       (const double*)e->part, s->x, s->y, s->z, s->q, s->flags, s->fx, s->fy, s->fz);
  AM_HIP(hipGetLastError());
  return AM_OK;
}

static int EnsEnqueueStep(am_ensemble* e, uint64_t seed) {
  am_store* s = e->store;
  am_ctx* ctx = s->ctx;
  hipStream_t st = ctx->stream;
  const int64_t n = s->n;
  int rc;
  if ((rc = EnsForce(e))) return rc;
  hipLaunchKernelGGL(EnsemblePushKernel, dim3(Blocks(n, 256)), dim3(256), 0, st, n, s->x, s->y, s->z, s->vx, s->vy,
                     s->vz, s->fx, s->fy, s->fz, s->q, s->m, s->flags, s->src, (const uint32_t*)s->trap,
                     (const TrapRec*)e->d_traps, seed, (const uint64_t*)e->step);
  AM_HIP(hipGetLastError());
  if (e->any_annihilation) {
    hipLaunchKernelGGL(EnsCellKeyKernel, dim3(Blocks(n, 256)), dim3(256), 0, st, n, s->x, s->y, s->z, s->flags,
                       (const uint32_t*)s->trap, (const TrapRec*)e->d_traps, e->ckeys, e->cidx);
    AM_HIP(hipGetLastError());
    if ((rc = RadixSortPairs(ctx, e->ckeys, e->cidx, n, 64))) return rc;
    if ((rc = LaunchCellTable(st, e->ckeys, n, e->hslots, e->hsize))) return rc;
    CellTable tab;
    tab.slots = e->hslots;
    tab.mask = e->hsize - 1;
    hipLaunchKernelGGL(EnsContactCountKernel, dim3(Blocks(n, 256)), dim3(256), 0, st, n, s->x, s->y, s->z,
                       s->species, s->flags, (const uint32_t*)s->trap, (const TrapRec*)e->d_traps, tab,
                       (const uint32_t*)e->cidx, e->ccount);
    AM_HIP(hipGetLastError());
    if ((rc = ExclusiveScan(ctx, e->ccount, e->coff, n))) return rc;
    hipLaunchKernelGGL(EnsContactWriteKernel, dim3(Blocks(n, 256)), dim3(256), 0, st, n, s->x, s->y, s->z,
                       s->species, s->flags, (const uint32_t*)s->trap, (const TrapRec*)e->d_traps, tab,
                       (const uint32_t*)e->cidx, (const uint32_t*)e->coff,
                       e->contact_cap, e->cpairs, e->overflow);
    hipLaunchKernelGGL(EnsResolveKernel, dim3(e->n_traps), dim3(256), 0, st, (const TrapRec*)e->d_traps,
                       (const uint32_t*)e->coff, (const uint2*)e->cpairs, e->contact_cap, seed,
                       (const uint64_t*)e->step, s->flags, s->src, e->cpass, e->gmark, e->n_marked);
    AM_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(EnsStepKernel, dim3(1), dim3(1), 0, st, e->step);
  AM_HIP(hipGetLastError());
  return AM_OK;
}

}  // namespace am

extern "C" {

int am_ensemble_create(am_store* s, int32_t n_traps, const am_trap_desc* desc, int64_t contact_capacity,
                       am_ensemble** out) {
  if (!s || !desc || !out || n_traps <= 0) return Fail(AM_ERR_INVALID, "am_ensemble_create: bad arguments");
  if (n_traps > kMaxTraps) return Fail(AM_ERR_CAPACITY, "am_ensemble_create: at most 1024 traps");
  *out = nullptr;
  AM_HIP(hipSetDevice(s->ctx->device));
  int64_t prev_end = 0;
  for (int t = 0; t < n_traps; ++t) {
    const am_trap_desc& d = desc[t];
    if (d.begin < prev_end || d.end < d.begin || d.end > s->n)
      return Fail(AM_ERR_INVALID, "am_ensemble_create: trap ranges must be ascending, disjoint and inside the store");
    if (d.end - d.begin > (int64_t)1 << 30) return Fail(AM_ERR_CAPACITY, "trap too large");
    if (!(d.d > 0.0)) return Fail(AM_ERR_INVALID, "characteristic dimension must be > 0");
    if (d.annihilation && !(d.contact_radius > 0.0))
      return Fail(AM_ERR_INVALID, "contact_radius must be > 0 when annihilation is enabled");
    prev_end = d.end;
  }
  am_ensemble* e = new am_ensemble();
  e->store = s;
  e->n_traps = n_traps;
  std::vector<uint32_t> tid(s->n, 0xffffffffu);
  for (int t = 0; t < n_traps; ++t) {
    const am_trap_desc& d = desc[t];
    TrapRec r{};
    r.begin = d.begin;
    r.end = d.end;
    r.b0 = d.b0;
    r.u0 = d.trap_voltage;
    r.d = d.d;
    r.dt = d.dt;
    r.tau_scale = d.tau_scale;
    r.bath_temp = d.bath_temp;
    r.contact_radius = d.contact_radius;
    r.probability = d.probability;
    r.annihilation = d.annihilation ? 1 : 0;
    r.inv_h = d.annihilation ? 1.0 / (2.0 * d.contact_radius * kEnsContactReach * (1.0 + 1e-9)) : 0.0;
    const int64_t N = d.end - d.begin;
    r.tile0 = e->n_tiles;
    r.ntile = (int)(2 * ((N + 2 * kEnsTile - 1) / (2 * kEnsTile)));
    r.sorted0 = e->n_in;
    e->n_tiles += r.ntile;
    e->n_in += N;
    e->traps.push_back(r);
    e->any_annihilation |= r.annihilation != 0;
    for (int64_t i = d.begin; i < d.end; ++i) tid[i] = (uint32_t)t;
  }
  int rc = AM_OK;
  void* p;
  hipStream_t st = s->ctx->stream;
  const int64_t n = s->n;
  auto fail = [&](int code) {
    am_ensemble_destroy(e);
    return code;
  };
  if ((rc = EnsAlloc(e, sizeof(TrapRec) * n_traps, &p))) return fail(rc);
  e->d_traps = (TrapRec*)p;
  {
    const int64_t n_pos = (int64_t)e->n_tiles * kEnsTile;
    if (n_pos >= ((int64_t)1 << 31)) return fail(Fail(AM_ERR_CAPACITY, "am_ensemble_create: too many particles"));
    if ((rc = EnsAlloc(e, (size_t)n * 8, &p))) return fail(rc);
    e->fkeys = (uint64_t*)p;
    if ((rc = EnsAlloc(e, (size_t)n * 4, &p))) return fail(rc);
    e->fvals = (uint32_t*)p;
    if ((rc = EnsAlloc(e, (size_t)n_traps * 6 * 8, &p))) return fail(rc);
    e->tbox = (double*)p;
    if ((rc = EnsAlloc(e, (size_t)n_pos * 4, &p))) return fail(rc);
    e->perm = (uint32_t*)p;
    if ((rc = EnsAlloc(e, (size_t)n_pos * 2 * sizeof(float4), &p))) return fail(rc);
    e->shl = (float4*)p;
    if ((rc = EnsAlloc(e, (size_t)n_pos * sizeof(float4), &p))) return fail(rc);
    e->sloc = (float4*)p;
    if ((rc = EnsAlloc(e, (size_t)e->n_tiles * sizeof(TileRec), &p))) return fail(rc);
    e->tiles = (TileRec*)p;
    if ((rc = EnsAlloc(e, (size_t)e->n_tiles * 4, &p))) return fail(rc);
    e->tile_trap = (int*)p;
    std::vector<int> tt((size_t)e->n_tiles);
    for (int t = 0; t < n_traps; ++t)
      for (int k = 0; k < e->traps[t].ntile; ++k) tt[e->traps[t].tile0 + k] = t;
    if (!tt.empty() && hipMemcpy(e->tile_trap, tt.data(), tt.size() * 4, hipMemcpyHostToDevice) != hipSuccess)
      return fail(AM_ERR_HIP);
  }
  if ((rc = EnsAlloc(e, 64, &p))) return fail(rc);
  e->counter = (unsigned*)p;
  e->step = (uint64_t*)((char*)p + 8);
  e->n_marked = (unsigned long long*)((char*)p + 16);
  e->overflow = (uint32_t*)((char*)p + 24);
  {
    const uint64_t init[4] = {0, 1, 0, 0};  // counter, step (the reference counts steps from 1), marks, overflow
    if (hipMemcpyAsync(p, init, sizeof(init), hipMemcpyHostToDevice, st) != hipSuccess) return fail(AM_ERR_HIP);
  }
  if (n > 0 && hipMemcpyAsync(s->trap, tid.data(), n * 4, hipMemcpyHostToDevice, st) != hipSuccess)
    return fail(AM_ERR_HIP);
  if (e->any_annihilation) {
    e->contact_cap = contact_capacity > 0 ? contact_capacity : std::max<int64_t>(16 * n, 1024);
    if ((rc = EnsAlloc(e, (size_t)n * 8, &p))) return fail(rc);
    e->ckeys = (uint64_t*)p;
    if ((rc = EnsAlloc(e, (size_t)n * 4, &p))) return fail(rc);
    e->cidx = (uint32_t*)p;
    e->hsize = CellTableSize(n);
    if ((rc = EnsAlloc(e, (size_t)e->hsize * sizeof(CellSlot), &p))) return fail(rc);
    e->hslots = (CellSlot*)p;
    if ((rc = EnsAlloc(e, (size_t)n * 4, &p))) return fail(rc);
    e->ccount = (uint32_t*)p;
    if ((rc = EnsAlloc(e, (size_t)(n + 1) * 4, &p))) return fail(rc);
    e->coff = (uint32_t*)p;
    if ((rc = EnsAlloc(e, (size_t)e->contact_cap * 8, &p))) return fail(rc);
    e->cpairs = (uint2*)p;
    if ((rc = EnsAlloc(e, (size_t)e->contact_cap, &p))) return fail(rc);
    e->cpass = (uint8_t*)p;
    if ((rc = EnsAlloc(e, (size_t)n, &p))) return fail(rc);
    e->gmark = (uint8_t*)p;
  }
  if ((rc = EnsBuildItems(e, 4096))) return fail(rc);
  e->store_generation = s->generation;
  e->n_at_create = s->n;
  *out = e;
  return AM_OK;
}

int am_ensemble_destroy(am_ensemble* e) {
  if (!e) return AM_OK;
  (void)hipStreamSynchronize(e->store->ctx->stream);
  if (e->exec) (void)hipGraphExecDestroy(e->exec);
  if (e->graph) (void)hipGraphDestroy(e->graph);
  for (void* p : e->owned) (void)hipFree(p);
  delete e;
  return AM_OK;
}

int am_ensemble_set_chunk(am_ensemble* e, int32_t chunk) {
  if (!e || chunk < kEnsTile) return Fail(AM_ERR_INVALID, "am_ensemble_set_chunk: chunk must be >= 128");
  if (int rc = EnsCheckStore(e, "am_ensemble_set_chunk")) return rc;
  return EnsBuildItems(e, chunk);
}

// Autotuner: times the force kernel for each candidate source-chunk length on the device and keeps
// the fastest.  times_ms (optional, n_candidates) receives the per-candidate averages.
int am_ensemble_autotune(am_ensemble* e, const int32_t* candidates, int32_t n_candidates, int32_t* chosen,
                         double* times_ms) {
  if (!e) return Fail(AM_ERR_INVALID, "am_ensemble_autotune: NULL ensemble");
  if (int rc = EnsCheckStore(e, "am_ensemble_autotune")) return rc;
  static const int32_t kDefault[] = {1024, 2048, 4096, 8192, 16384, 32768};
  if (!candidates || n_candidates <= 0) {
    candidates = kDefault;
    n_candidates = 6;
  }
  am_ctx* ctx = e->store->ctx;
  hipEvent_t a, b;
  AM_HIP(hipEventCreate(&a));
  AM_HIP(hipEventCreate(&b));
  double best = 1e30;
  int32_t best_c = e->chunk;
  for (int k = 0; k < n_candidates; ++k) {
    const int32_t c = candidates[k];
    if (c < kEnsTile) continue;
    int rc = EnsBuildItems(e, c);
    if (rc) return rc;
    if ((rc = EnsForce(e))) return rc;  // warm-up
    AM_HIP(hipEventRecord(a, ctx->stream));
    for (int r = 0; r < 3; ++r)
      if ((rc = EnsForce(e))) return rc;
    AM_HIP(hipEventRecord(b, ctx->stream));
    AM_HIP(hipEventSynchronize(b));
    float ms = 0.f;
    AM_HIP(hipEventElapsedTime(&ms, a, b));
    if (times_ms) times_ms[k] = ms / 3.0;
    if (ms < best) {
      best = ms;
      best_c = c;
    }
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  if (int rc = CheckSortError(ctx)) return rc;
  if (chosen) *chosen = best_c;
  return EnsBuildItems(e, best_c);
}

// Advance every trap by `steps` steps.  use_graph: replay one captured step (built on first use,
// rebuilt when the seed or the work layout changes).
int am_ensemble_run(am_ensemble* e, int32_t steps, uint64_t seed, int32_t use_graph) {
  if (!e || steps < 0) return Fail(AM_ERR_INVALID, "am_ensemble_run: bad arguments");
  if (int rc = EnsCheckStore(e, "am_ensemble_run")) return rc;
  am_ctx* ctx = e->store->ctx;
  hipStream_t st = ctx->stream;
  AM_HIP(hipSetDevice(ctx->device));
  int rc;
  int slot;
  TimerBegin(ctx, "ensemble_steps", &slot);
  if (!use_graph) {
    for (int k = 0; k < steps; ++k)
      if ((rc = EnsEnqueueStep(e, seed))) return rc;
  } else {
    if (!e->graph_valid || e->graph_seed != seed) {
      if ((rc = ReserveSortScratch(ctx, e->store->n))) return rc;  // nothing may allocate while capturing
      AM_HIP(hipStreamSynchronize(st));
      if (e->exec) { (void)hipGraphExecDestroy(e->exec); e->exec = nullptr; }
      if (e->graph) { (void)hipGraphDestroy(e->graph); e->graph = nullptr; }
      AM_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      rc = EnsEnqueueStep(e, seed);
      hipGraph_t g = nullptr;
      const hipError_t ce = hipStreamEndCapture(st, &g);
      if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
      AM_HIP(ce);
      e->graph = g;
      AM_HIP(hipGraphInstantiate(&e->exec, e->graph, nullptr, nullptr, 0));
      e->graph_seed = seed;
      e->graph_valid = true;
    }
    for (int k = 0; k < steps; ++k) AM_HIP(hipGraphLaunch(e->exec, st));
  }
  TimerEnd(ctx, slot);
  uint32_t ov = 0, sort_err = 0;
  AM_HIP(hipMemcpyAsync(&ov, e->overflow, 4, hipMemcpyDeviceToHost, st));
  if (const uint32_t* w = SortErrorWord(ctx)) AM_HIP(hipMemcpyAsync(&sort_err, w, 4, hipMemcpyDeviceToHost, st));
  AM_HIP(hipStreamSynchronize(st));
  e->steps_done += steps;
  if ((rc = CheckSortErrorAfterSync(ctx, sort_err))) return rc;
  if (ov) return Fail(AM_ERR_CAPACITY, "am_ensemble_run: contact candidate buffer overflowed; raise contact_capacity");
  return AM_OK;
}

int am_ensemble_stats(am_ensemble* e, int64_t* steps_done, int64_t* annihilated, int32_t* chunk, int64_t* n_items,
                      double* pair_interactions_per_step) {
  if (!e) return Fail(AM_ERR_INVALID, "am_ensemble_stats: NULL ensemble");
  if (steps_done) *steps_done = e->steps_done;
  if (annihilated) {
    unsigned long long m = 0;
    AM_HIP(hipMemcpy(&m, e->n_marked, 8, hipMemcpyDeviceToHost));
    *annihilated = (int64_t)m;
  }
  if (chunk) *chunk = e->chunk;
  if (n_items) *n_items = e->n_items;
  if (pair_interactions_per_step) {
    double s = 0.0;
    for (const auto& t : e->traps) {
      const double N = (double)(t.end - t.begin);
      s += N * (N - 1.0);
    }
    *pair_interactions_per_step = s;
  }
  return AM_OK;
}

}  // extern "C"
