// Uniform cell lists and the kernels built on them:
//   * ComputeStandardCoulombForces (src/interaction/coulomb_interaction.cc:23-80): pairs with
//     r^2 < r_c^2 (the radius search of the reference's neighbour query), in FP64 (the reference
//     arithmetic) or with FP32 hi/lo pair math and the same pair set.
//   * the shifted-cutoff standard energy (src/analysis/energy_analysis.cc:16-43), r^2 <= r_c^2.
//   * AnnihilationInteraction (src/interaction/annihilation.cc:21-61), canonical order.
//
// A cell list sorts particles by the 63-bit key (cx, cy, cz) of their cell of edge h (slightly
// larger than the interaction radius, so every pair within the radius lies in the 27-cell
// stencil) with the stable radix sort; a query looks each stencil cell's run up in a cell table
// (open addressing, built once per call).  The contact search, whose results do not depend on the
// order within a cell, groups particles through the table instead of sorting them.
// Cell coordinates are clamped to +-2^20 cells (clamping is 1-Lipschitz, so pairs within the
// radius stay within the stencil; only far-flung particles share boundary cells).
#include <algorithm>
#include <cmath>

#include "am_internal.hpp"
#include "celltable.hpp"

namespace am {

constexpr uint64_t kDeadKey = kCellEmpty;  // annihilated particles sort last and match no cell

__device__ inline int64_t CellCoord(double v, double inv_h) {
  const double c = floor(v * inv_h);
  if (!(c == c)) return 1 << 20;  // NaN
  const double lo = -(double)(1 << 20), hi = (double)(1 << 20) - 1.0;
  return (int64_t)fmin(fmax(c, lo), hi) + (1 << 20);
}

__device__ inline uint64_t CellKey(int64_t cx, int64_t cy, int64_t cz) {
  return ((uint64_t)cx << 42) | ((uint64_t)cy << 21) | (uint64_t)cz;
}

__global__ void CellKeyKernel(int64_t n, const double* __restrict__ x, const double* __restrict__ y,
                              const double* __restrict__ z, const uint8_t* __restrict__ flags, double inv_h,
                              uint64_t* __restrict__ keys, uint32_t* __restrict__ idx) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  idx[i] = (uint32_t)i;
  if (flags[i] & kFlagAnnihilated) {
    keys[i] = kDeadKey;
    return;
  }
  keys[i] = CellKey(CellCoord(x[i], inv_h), CellCoord(y[i], inv_h), CellCoord(z[i], inv_h));
}

__global__ void CellTableClearKernel(uint64_t size, CellSlot* __restrict__ slots) {
  const uint64_t h = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (h < size) slots[h] = CellSlot{kCellEmpty, 0u, 0u};
}

// One thread per sorted position; the first position of each run inserts the run.
__global__ void CellTableInsertKernel(int64_t n, const uint64_t* __restrict__ skeys, CellSlot* __restrict__ slots,
                                      uint64_t mask) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const uint64_t key = skeys[k];
  if (key == kCellEmpty || (k > 0 && skeys[k - 1] == key)) return;
  int64_t e = k + 1;
  while (e < n && skeys[e] == key) ++e;
  uint64_t h = CellHashMix(key) & mask;
  for (;;) {  // a free slot exists: at most n of the >= 2n slots are ever taken
    const unsigned long long prev =
        atomicCAS(reinterpret_cast<unsigned long long*>(&slots[h].key), (unsigned long long)kCellEmpty,
                  (unsigned long long)key);
    if (prev == kCellEmpty) {
      slots[h].start = (uint32_t)k;
      slots[h].count = (uint32_t)(e - k);
      return;
    }
    h = (h + 1) & mask;
  }
}

int LaunchCellTable(hipStream_t st, const uint64_t* skeys, int64_t n, CellSlot* slots, uint64_t size) {
  hipLaunchKernelGGL(CellTableClearKernel, dim3(Blocks((int64_t)size, 256)), dim3(256), 0, st, size, slots);
  if (n > 0)
    hipLaunchKernelGGL(CellTableInsertKernel, dim3(Blocks(n, 256)), dim3(256), 0, st, n, skeys, slots, size - 1);
  AM_HIP(hipGetLastError());
  return AM_OK;
}

// Calls f(j) for every particle j in the 27 cells around particle i's cell (including i), cells
// in (dz, dy, dx) order and each cell's particles in sorted order.
template <typename F>
__device__ inline void ForEachStencil(const CellTable& tab, const uint32_t* __restrict__ idx, double px, double py,
                                      double pz, double inv_h, F&& f) {
  const int64_t cx = CellCoord(px, inv_h), cy = CellCoord(py, inv_h), cz = CellCoord(pz, inv_h);
  const int64_t lim = (int64_t)1 << 21;
  for (int dz = -1; dz <= 1; ++dz) {
    const int64_t z = cz + dz;
    if (z < 0 || z >= lim) continue;
    for (int dy = -1; dy <= 1; ++dy) {
      const int64_t y = cy + dy;
      if (y < 0 || y >= lim) continue;
      for (int dx = -1; dx <= 1; ++dx) {
        const int64_t x = cx + dx;
        if (x < 0 || x >= lim) continue;
        const uint2 r = CellLookup(tab, CellKey(x, y, z));
        for (uint32_t k = r.x; k < r.x + r.y; ++k) f((int64_t)idx[k]);
      }
    }
  }
}

struct CellList {
  uint64_t* keys = nullptr;
  uint32_t* idx = nullptr;
  double inv_h = 0.0;
  CellTable tab;
};

static int BuildCellList(am_store* s, double radius, const char* tag, CellList* cl) {
  am_ctx* ctx = s->ctx;
  const int64_t n = s->n;
  std::string kname = std::string(tag) + "_keys", iname = std::string(tag) + "_idx";
  void *k, *v;
  int rc;
  if ((rc = Scratch(ctx, kname.c_str(), (size_t)n * 8, &k))) return rc;
  if ((rc = Scratch(ctx, iname.c_str(), (size_t)n * 4, &v))) return rc;
  cl->keys = (uint64_t*)k;
  cl->idx = (uint32_t*)v;
  cl->inv_h = 1.0 / (radius * (1.0 + 1e-9));
  hipLaunchKernelGGL(CellKeyKernel, dim3(Blocks(n, 256)), dim3(256), 0, ctx->stream, n, s->x, s->y, s->z, s->flags,
                     cl->inv_h, cl->keys, cl->idx);
  AM_HIP(hipGetLastError());
  if ((rc = RadixSortPairs(ctx, cl->keys, cl->idx, n, 64))) return rc;
  const uint64_t size = CellTableSize(n);
  void* hs;
  const std::string hsn = std::string(tag) + "_hslots";
  if ((rc = Scratch(ctx, hsn.c_str(), (size_t)size * sizeof(CellSlot), &hs))) return rc;
  if ((rc = LaunchCellTable(ctx->stream, cl->keys, n, (CellSlot*)hs, size))) return rc;
  cl->tab.slots = (const CellSlot*)hs;
  cl->tab.mask = size - 1;
  return AM_OK;
}

// ---- contact cells: grouped through the table, not sorted ---------------------------------------
// The contact search needs each cell's particles contiguous, in no particular order: a particle's
// candidates are sorted by index before they are resolved, and the counts do not depend on the
// order.  So the contact cell list is built from the table itself: every live particle finds or
// inserts its cell key and takes a rank in the cell with an atomic add on the slot's count; a scan
// of the slot counts gives each cell's start; each particle is placed at start + rank.  (The
// cutoff force and energy sums keep the sorted cell list: their summation order is the order of
// the cells' particles.)
constexpr uint32_t kNoSlot = ~0u;

__global__ void ContactCellInsertKernel(int64_t n, const double* __restrict__ x, const double* __restrict__ y,
                                        const double* __restrict__ z, const uint8_t* __restrict__ flags,
                                        double inv_h, CellSlot* __restrict__ slots, uint64_t mask,
                                        uint32_t* __restrict__ slot_of, uint32_t* __restrict__ rank) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (flags[i] & kFlagAnnihilated) {
    slot_of[i] = kNoSlot;
    return;
  }
  const uint64_t key = CellKey(CellCoord(x[i], inv_h), CellCoord(y[i], inv_h), CellCoord(z[i], inv_h));
  uint64_t h = CellHashMix(key) & mask;
  for (;;) {  // a free slot exists: at most n of the >= 2n slots are ever taken
    unsigned long long* kp = reinterpret_cast<unsigned long long*>(&slots[h].key);
    unsigned long long k = __hip_atomic_load(kp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (k == (unsigned long long)kCellEmpty) k = atomicCAS(kp, (unsigned long long)kCellEmpty, (unsigned long long)key);
    if (k == (unsigned long long)kCellEmpty || k == (unsigned long long)key) break;
    h = (h + 1) & mask;
  }
  slot_of[i] = (uint32_t)h;
  rank[i] = atomicAdd(&slots[h].count, 1u);
}

__global__ void SlotCountKernel(uint64_t size, const CellSlot* __restrict__ slots, uint32_t* __restrict__ cnt) {
  const uint64_t h = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (h < size) cnt[h] = slots[h].count;
}

__global__ void SlotStartKernel(uint64_t size, const uint32_t* __restrict__ off, CellSlot* __restrict__ slots) {
  const uint64_t h = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (h < size) slots[h].start = off[h];
}

__global__ void ContactCellPlaceKernel(int64_t n, const uint32_t* __restrict__ slot_of,
                                       const uint32_t* __restrict__ rank, const uint32_t* __restrict__ off,
                                       uint32_t* __restrict__ idx) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || slot_of[i] == kNoSlot) return;
  idx[off[slot_of[i]] + rank[i]] = (uint32_t)i;
}

static int BuildContactCells(am_store* s, double radius, CellList* cl) {
  am_ctx* ctx = s->ctx;
  hipStream_t st = ctx->stream;
  const int64_t n = s->n;
  const uint64_t size = CellTableSize(n);
  void *hs, *idx, *slot_of, *rank, *cnt, *off;
  int rc;
  if ((rc = Scratch(ctx, "cl_contact_hslots", (size_t)size * sizeof(CellSlot), &hs))) return rc;
  if ((rc = Scratch(ctx, "cl_contact_idx", (size_t)n * 4, &idx))) return rc;
  if ((rc = Scratch(ctx, "cl_contact_slot", (size_t)n * 4, &slot_of))) return rc;
  if ((rc = Scratch(ctx, "cl_contact_rank", (size_t)n * 4, &rank))) return rc;
  if ((rc = Scratch(ctx, "cl_contact_cnt", (size_t)size * 4, &cnt))) return rc;
  if ((rc = Scratch(ctx, "cl_contact_off", (size_t)(size + 1) * 4, &off))) return rc;
  cl->keys = nullptr;
  cl->idx = (uint32_t*)idx;
  cl->inv_h = 1.0 / (radius * (1.0 + 1e-9));
  CellSlot* slots = (CellSlot*)hs;
  hipLaunchKernelGGL(CellTableClearKernel, dim3(Blocks((int64_t)size, 256)), dim3(256), 0, st, size, slots);
  hipLaunchKernelGGL(ContactCellInsertKernel, dim3(Blocks(n, 256)), dim3(256), 0, st, n, s->x, s->y, s->z, s->flags,
                     cl->inv_h, slots, size - 1, (uint32_t*)slot_of, (uint32_t*)rank);
  hipLaunchKernelGGL(SlotCountKernel, dim3(Blocks((int64_t)size, 256)), dim3(256), 0, st, size,
                     (const CellSlot*)slots, (uint32_t*)cnt);
  AM_HIP(hipGetLastError());
  if ((rc = ExclusiveScan(ctx, (const uint32_t*)cnt, (uint32_t*)off, (int64_t)size))) return rc;
  hipLaunchKernelGGL(SlotStartKernel, dim3(Blocks((int64_t)size, 256)), dim3(256), 0, st, size, (const uint32_t*)off,
                     slots);
  hipLaunchKernelGGL(ContactCellPlaceKernel, dim3(Blocks(n, 256)), dim3(256), 0, st, n, (const uint32_t*)slot_of,
                     (const uint32_t*)rank, (const uint32_t*)off, (uint32_t*)idx);
  AM_HIP(hipGetLastError());
  cl->tab.slots = (const CellSlot*)slots;
  cl->tab.mask = size - 1;
  return AM_OK;
}

// coulomb_interaction.cc:60-73 (pair term r_ij * (k q_i q_j / d2 / distance)).
__global__ __launch_bounds__(256) void StandardForceKernel(int64_t n, const double* __restrict__ x,
                                                           const double* __restrict__ y, const double* __restrict__ z,
                                                           const double* __restrict__ q,
                                                           const uint8_t* __restrict__ flags, CellTable tab,
                                                           const uint32_t* __restrict__ idx, double inv_h, double rc2,
                                                           double* __restrict__ fx, double* __restrict__ fy,
                                                           double* __restrict__ fz) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double ax = 0.0, ay = 0.0, az = 0.0;
  const double qi = q[i];
  const double px = x[i], py = y[i], pz = z[i];
  const bool live = !(flags[i] & kFlagAnnihilated) && qi != 0.0 && isfinite(px) && isfinite(py) && isfinite(pz);
  if (live) {
    ForEachStencil(tab, idx, px, py, pz, inv_h, [&](int64_t j) {
      if (j == i) return;
      const double qj = q[j];
      if (qj == 0.0) return;
      const double rx = px - x[j], ry = py - y[j], rz = pz - z[j];
      const double d2 = Dot3(rx, ry, rz, rx, ry, rz);
      if (!(d2 < rc2)) return;
      if (d2 < kMinSquaredDistance) return;
      const double dist = sqrt(d2);
      const double mag = kCoulomb * qi * qj / d2;
      const double c = mag / dist;
      ax = ax + rx * c;
      ay = ay + ry * c;
      az = az + rz * c;
    });
  }
  if (!(flags[i] & kFlagAnnihilated)) {
    fx[i] = ax;
    fy[i] = ay;
    fz[i] = az;
  }
}

// The FP32 hi/lo mirror in cell-sorted order: a stencil cell's sources are then contiguous.
__global__ void CellSortedMirrorKernel(int64_t n, const uint32_t* __restrict__ idx, const float4* __restrict__ src,
                                       float4* __restrict__ ssrc) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const uint32_t i = idx[k];
  ssrc[2 * k] = src[2 * i];
  ssrc[2 * k + 1] = src[2 * i + 1];
}

// Relative half-width of the band around r_c^2 inside which the FP32 r^2 (relative error below
// ~3e-7) cannot decide the cutoff test: such pairs are decided in FP64, as the reference does.
constexpr float kCutoffBand = 1e-5f;

// ComputeStandardCoulombForces with FP32 pair math (coulomb_interaction.cc:60-73): one thread per
// cell-sorted position, so the lanes of a wave mostly share their stencil cells; sources come
// contiguously from the cell-sorted hi/lo mirror.  The pair set is the FP64 path's: r^2 < r_c^2
// from the FP32 r^2 outside a narrow band around r_c^2, from the FP64 positions inside it.  Each
// stencil cell is summed in FP32 and added to FP64 accumulators.  The 1e-30 m^2 softening makes
// the self pair and coincident pairs contribute zero; annihilated particles are in no cell.
__global__ __launch_bounds__(256) void StandardForceF32Kernel(
    int64_t n, const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ z,
    const double* __restrict__ q, const uint8_t* __restrict__ flags, CellTable tab, const uint32_t* __restrict__ idx,
    const float4* __restrict__ ssrc, double inv_h, double rc2, float rc2u, double* __restrict__ fx,
    double* __restrict__ fy, double* __restrict__ fz) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const uint32_t i = idx[k];
  if (flags[i] & kFlagAnnihilated) return;
  const double qi = q[i];
  const double px = x[i], py = y[i], pz = z[i];
  double ax = 0.0, ay = 0.0, az = 0.0;
  if (qi != 0.0 && isfinite(px) && isfinite(py) && isfinite(pz)) {
    const float4 th = ssrc[2 * k], tl = ssrc[2 * k + 1];
    const float band = kCutoffBand * rc2u;
    const int64_t cx = CellCoord(px, inv_h), cy = CellCoord(py, inv_h), cz = CellCoord(pz, inv_h);
    const int64_t lim = (int64_t)1 << 21;
    for (int dz = -1; dz <= 1; ++dz) {
      if (cz + dz < 0 || cz + dz >= lim) continue;
      for (int dy = -1; dy <= 1; ++dy) {
        if (cy + dy < 0 || cy + dy >= lim) continue;
        for (int dx_ = -1; dx_ <= 1; ++dx_) {
          if (cx + dx_ < 0 || cx + dx_ >= lim) continue;
          const uint2 r = CellLookup(tab, CellKey(cx + dx_, cy + dy, cz + dz));
          float sx = 0.f, sy = 0.f, sz = 0.f;
          for (uint32_t kk = r.x; kk < r.x + r.y; ++kk) {
            const float4 sh = ssrc[2 * kk], sl = ssrc[2 * kk + 1];
            const float dx = (th.x - sh.x) + (tl.x - sl.x);
            const float dy2 = (th.y - sh.y) + (tl.y - sl.y);
            const float dz2 = (th.z - sh.z) + (tl.z - sl.z);
            const float r2 = fmaf(dx, dx, fmaf(dy2, dy2, dz2 * dz2));
            bool in = r2 < rc2u;
            if (fabsf(r2 - rc2u) <= band) {
              const uint32_t j = idx[kk];
              const double rx = px - x[j], ry = py - y[j], rz = pz - z[j];
              in = Dot3(rx, ry, rz, rx, ry, rz) < rc2;
            }
            if (!in) continue;
            const float ri = __builtin_amdgcn_rsqf(r2 + kEps2Units);
            const float s = (sh.w * ri) * (ri * ri);
            sx = fmaf(s, dx, sx);
            sy = fmaf(s, dy2, sy);
            sz = fmaf(s, dz2, sz);
          }
          ax += (double)sx;
          ay += (double)sy;
          az += (double)sz;
        }
      }
    }
  }
  const double scale = kCoulomb * qi * kElementaryCharge * kInvUnit2;
  fx[i] = scale * ax;
  fy[i] = scale * ay;
  fz[i] = scale * az;
}

// energy_analysis.cc:22-41 over ordered pairs (the caller halves the sum): e_i = sum over
// neighbours j != i with d2 <= rc^2 and d2 >= 1e-30 of k q_i q_j (1/r - 1/rc).
__global__ __launch_bounds__(256) void StandardEnergyKernel(int64_t n, const double* __restrict__ x,
                                                            const double* __restrict__ y,
                                                            const double* __restrict__ z,
                                                            const double* __restrict__ q,
                                                            const uint8_t* __restrict__ flags, CellTable tab,
                                                            const uint32_t* __restrict__ idx, double inv_h, double rc,
                                                            double* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double e = 0.0;
  const double qi = q[i];
  if (!(flags[i] & kFlagAnnihilated) && qi != 0.0) {
    const double px = x[i], py = y[i], pz = z[i];
    const double rc2 = rc * rc;
    ForEachStencil(tab, idx, px, py, pz, inv_h, [&](int64_t j) {
      if (j == i) return;
      const double qj = q[j];
      if (qj == 0.0) return;
      const double rx = px - x[j], ry = py - y[j], rz = pz - z[j];
      const double d2 = Dot3(rx, ry, rz, rx, ry, rz);
      if (d2 > rc2) return;
      if (d2 < 1e-30) return;
      e = e + kCoulomb * qi * qj * (1.0 / sqrt(d2) - 1.0 / rc);
    });
  }
  out[i] = e;
}

__global__ void ZeroAliveForceKernel(int64_t n, const uint8_t* __restrict__ flags, double* __restrict__ fx,
                                     double* __restrict__ fy, double* __restrict__ fz) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || (flags[i] & kFlagAnnihilated)) return;
  fx[i] = 0.0;
  fy[i] = 0.0;
  fz[i] = 0.0;
}

int StandardEnergyTerms(am_store* s, double rc, double* out) {
  CellList cl;
  int rc_ = BuildCellList(s, rc, "cl_energy", &cl);
  if (rc_) return rc_;
  hipLaunchKernelGGL(StandardEnergyKernel, dim3(Blocks(s->n, 256)), dim3(256), 0, s->ctx->stream, s->n, s->x, s->y,
                     s->z, s->q, s->flags, cl.tab, cl.idx, cl.inv_h, rc, out);
  AM_HIP(hipGetLastError());
  return AM_OK;
}

// ---- annihilation ------------------------------------------------------------------------------
// The contact cell list has edge h = 4 reach, reach = r_a (1 + 1e-6): along each axis the points
// within reach of p lie in the cells from CellCoord(p - reach) to CellCoord(p + reach), at most
// two since h >= 2 reach (CellCoord is monotone and 1-Lipschitz in cells), and one when p is at
// least reach from both faces, which is half the time at h = 4 reach: 3.4 table lookups per
// particle on average instead of 8 at h = 2 reach.  Contacts are rare at the clouds' densities, so
// the lookups, not the distance tests of the larger cells, set the cost.  The margin r_a 1e-6 is far
// above the rounding of p -+ reach.
constexpr double kContactReach = 1.0 + 1e-6;
constexpr double kContactCellEdge = 4.0;  // in units of reach

template <typename F>
__device__ inline void ForEachNearCell(const CellTable& tab, const uint32_t* __restrict__ idx, double px, double py,
                                       double pz, double inv_h, double reach, F&& f) {
  const int64_t x0 = CellCoord(px - reach, inv_h), x1 = CellCoord(px + reach, inv_h);
  const int64_t y0 = CellCoord(py - reach, inv_h), y1 = CellCoord(py + reach, inv_h);
  const int64_t z0 = CellCoord(pz - reach, inv_h), z1 = CellCoord(pz + reach, inv_h);
  for (int64_t cz = z0; cz <= z1; ++cz)
    for (int64_t cy = y0; cy <= y1; ++cy)
      for (int64_t cx = x0; cx <= x1; ++cx) {
        const uint2 r = CellLookup(tab, CellKey(cx, cy, cz));
        for (uint32_t k = r.x; k < r.x + r.y; ++k) f((int64_t)idx[k]);
      }
}

// Candidates of i: alive j != i, IsAntiparticlePair(species_i, species_j), d2 < r_a^2.
template <typename F>
__device__ inline void ForEachCandidate(int64_t i, int64_t n, const double* __restrict__ x,
                                        const double* __restrict__ y, const double* __restrict__ z,
                                        const uint8_t* __restrict__ species, const uint8_t* __restrict__ flags,
                                        const CellTable& tab, const uint32_t* __restrict__ idx, double inv_h,
                                        double ra2, F&& f) {
  const double px = x[i], py = y[i], pz = z[i];
  const int si = species[i];
  ForEachNearCell(tab, idx, px, py, pz, inv_h, sqrt(ra2) * kContactReach, [&](int64_t j) {
    if (j == i || (flags[j] & kFlagAnnihilated)) return;
    if (!IsPair(si, species[j])) return;
    const double rx = px - x[j], ry = py - y[j], rz = pz - z[j];
    if (!(Dot3(rx, ry, rz, rx, ry, rz) < ra2)) return;
    f(j);
  });
}

__global__ void ContactCountKernel(int64_t n, const double* __restrict__ x, const double* __restrict__ y,
                                   const double* __restrict__ z, const uint8_t* __restrict__ species,
                                   const uint8_t* __restrict__ flags, CellTable tab,
                                   const uint32_t* __restrict__ idx, double inv_h, double ra2,
                                   uint32_t* __restrict__ count) {
  // Particle order: the particle's own position, species and flags are read contiguously.  (In
  // cell-sorted order they are gathered, about five scattered lines per particle, which costs more
  // than the table-line sharing gains when contacts are rare.)
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t c = 0;
  if (!(flags[i] & kFlagAnnihilated))
    ForEachCandidate(i, n, x, y, z, species, flags, tab, idx, inv_h, ra2, [&](int64_t) { ++c; });
  count[i] = c;
}

// Writes i's candidates into [off[i], off[i+1]) sorted by j ascending (insertion sort of the
// short segment), as (i, j) pairs.
__global__ void ContactWriteKernel(int64_t n, const double* __restrict__ x, const double* __restrict__ y,
                                   const double* __restrict__ z, const uint8_t* __restrict__ species,
                                   const uint8_t* __restrict__ flags, CellTable tab,
                                   const uint32_t* __restrict__ idx, double inv_h, double ra2,
                                   const uint32_t* __restrict__ off, uint2* __restrict__ pairs) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;  // particle order, as the count
  if (i >= n) return;
  const uint32_t b = off[i], e = off[i + 1];
  if (b == e) return;
  uint32_t w = b;
  ForEachCandidate(i, n, x, y, z, species, flags, tab, idx, inv_h, ra2, [&](int64_t j) {
    uint32_t k = w++;
    while (k > b && pairs[k - 1].y > (uint32_t)j) {
      pairs[k] = pairs[k - 1];
      --k;
    }
    pairs[k] = make_uint2((uint32_t)i, (uint32_t)j);
  });
}

constexpr int kResolveLds = 8192;      // candidate entries resolved in LDS by ContactResolveRoundsKernel
constexpr int kResolveMaxRounds = 64;  // parallel rounds before the remaining runs are walked serially

// The greedy resolution of ContactResolveKernel for lists of at most kResolveLds entries, held in
// LDS and decided in parallel rounds with the same result.  A particle's run (its entries, sorted by
// j) is decided when every particle with a smaller index among its candidates and its candidates'
// candidates is decided: only those can set the marks the run reads (its own and its candidates')
// before it in scan order, and two runs decided in one round read and write disjoint marks (a shared
// candidate, or one being the other's candidate, makes the larger wait for the smaller).  The
// smallest undecided run is always ready, so every round makes progress; after kResolveMaxRounds
// rounds thread 0 decides whatever is left in scan order.
__global__ __launch_bounds__(256) void ContactResolveRoundsKernel(const uint2* __restrict__ pairs,
                                                                  const uint32_t* __restrict__ total_ptr, double p,
                                                                  uint64_t seed, uint64_t step,
                                                                  uint8_t* __restrict__ flags,
                                                                  float4* __restrict__ src,
                                                                  unsigned int* __restrict__ n_marked) {
  __shared__ uint32_t li[kResolveLds];     // i of each entry
  __shared__ uint16_t lslot[kResolveLds];  // first entry of j's run
  __shared__ uint16_t lend[kResolveLds];   // one past the last entry of i's run (kept at run starts)
  __shared__ uint8_t lpass[kResolveLds];   // U(i, j) < p
  __shared__ uint8_t lmark[kResolveLds];   // particle marked (at run starts)
  __shared__ uint8_t lstate[kResolveLds];  // run: 0 undecided, 2 ready this round, 1 decided
  __shared__ unsigned int s_count;
  const uint32_t m = *total_ptr;
  if (m > (uint32_t)kResolveLds) return;  // the host launches the serial kernel instead
  auto lower = [&](uint32_t v) {           // first entry whose i >= v
    uint32_t lo = 0, hi = m;
    while (lo < hi) {
      const uint32_t mid = (lo + hi) >> 1;
      if (li[mid] < v) lo = mid + 1;
      else hi = mid;
    }
    return lo;
  };
  for (uint32_t e = threadIdx.x; e < m; e += blockDim.x) li[e] = pairs[e].x;
  if (threadIdx.x == 0) s_count = 0;
  __syncthreads();
  for (uint32_t e = threadIdx.x; e < m; e += blockDim.x) {
    const uint2 pr = pairs[e];
    lslot[e] = (uint16_t)lower(pr.y);
    lpass[e] = AnnihilationUniform(seed, step, pr.x, pr.y) < p ? 1 : 0;
    lmark[e] = 0;
    lstate[e] = 0;
    if (e == 0 || li[e - 1] != pr.x) {
      uint32_t end = e + 1;
      while (end < m && li[end] == pr.x) ++end;
      lend[e] = (uint16_t)end;
    }
  }
  __syncthreads();
  // Run r (a start entry) of particle i waits for any run with a smaller particle index among
  // its candidates and their candidates.
  auto ready = [&](uint32_t r) {
    const uint32_t i = li[r];
    for (uint32_t k = r; k < lend[r]; ++k) {
      const uint32_t s = lslot[k];
      if (li[s] < i && lstate[s] != 1) return false;
      for (uint32_t kk = s; kk < lend[s]; ++kk) {
        const uint32_t t = lslot[kk];
        if (li[t] < i && lstate[t] != 1) return false;
      }
    }
    return true;
  };
  auto decide = [&](uint32_t r) {
    if (lmark[r]) return;
    for (uint32_t k = r; k < lend[r]; ++k) {
      const uint32_t s = lslot[k];
      if (lmark[s]) continue;
      if (lpass[k]) {
        lmark[r] = 1;
        lmark[s] = 1;
        return;
      }
    }
  };
  bool left = true;
  for (int round = 0; round < kResolveMaxRounds && left; ++round) {
    bool mine = false;
    for (uint32_t e = threadIdx.x; e < m; e += blockDim.x) {
      const bool start = e == 0 || li[e - 1] != li[e];
      if (start && lstate[e] == 0) {
        if (ready(e)) lstate[e] = 2;
        else mine = true;
      }
    }
    __syncthreads();
    for (uint32_t e = threadIdx.x; e < m; e += blockDim.x) {
      if (lstate[e] == 2) {
        decide(e);
        lstate[e] = 1;
      }
    }
    left = __syncthreads_or(mine);
  }
  if (left && threadIdx.x == 0) {
    for (uint32_t e = 0; e < m; e = lend[e]) {
      if (lstate[e] == 0) {
        decide(e);
        lstate[e] = 1;
      }
    }
  }
  __syncthreads();
  unsigned int cnt = 0;
  for (uint32_t e = threadIdx.x; e < m; e += blockDim.x) {
    if (!lmark[e]) continue;  // marks sit at run starts
    const uint32_t i = li[e];
    flags[i] |= kFlagAnnihilated;
    src[2 * i].w = 0.f;
    ++cnt;
  }
  if (cnt) atomicAdd(&s_count, cnt);
  __syncthreads();
  if (threadIdx.x == 0) *n_marked = s_count;
}

// Greedy resolution in scan order for lists longer than ContactResolveRoundsKernel holds.
// entries: the (i, j) list sorted by (i, j); candidates are symmetric, so the distinct i values are
// exactly the involved particles.  All threads first compute for every entry whether its draw
// passes (U < p) and the slot of j (the position of j's first entry); thread 0 then walks the list
// once.  mark[] is indexed by slot.
__global__ __launch_bounds__(256) void ContactResolveKernel(const uint2* __restrict__ pairs,
                                                            const uint32_t* __restrict__ total_ptr, double p,
                                                            uint64_t seed, uint64_t step,
                                                            uint8_t* __restrict__ flags, float4* __restrict__ src,
                                                            unsigned int* __restrict__ n_marked,
                                                            uint32_t* __restrict__ slot,
                                                            uint8_t* __restrict__ pass,
                                                            uint8_t* __restrict__ mark) {
  const uint32_t m = *total_ptr;
  for (uint32_t e = threadIdx.x; e < m; e += blockDim.x) {
    const uint2 pr = pairs[e];
    uint32_t lo = 0, hi = m;
    while (lo < hi) {
      const uint32_t mid = (lo + hi) >> 1;
      if (pairs[mid].x < pr.y) lo = mid + 1;
      else hi = mid;
    }
    slot[e] = lo;
    pass[e] = AnnihilationUniform(seed, step, pr.x, pr.y) < p ? 1 : 0;
    mark[e] = 0;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned int count = 0;
    uint32_t e = 0;
    while (e < m) {
      const uint32_t i = pairs[e].x;
      const uint32_t first = e;
      uint32_t end = e;
      while (end < m && pairs[end].x == i) ++end;
      if (!mark[first]) {
        for (uint32_t k = first; k < end; ++k) {
          const uint32_t sj = slot[k];
          if (mark[sj]) continue;
          if (pass[k]) {
            mark[first] = 1;
            mark[sj] = 1;
            count += 2;
            break;
          }
        }
      }
      e = end;
    }
    *n_marked = count;
  }
  __syncthreads();
  // Apply: the first entry of each run carries the particle's mark.
  for (uint32_t e = threadIdx.x; e < m; e += blockDim.x) {
    if (!mark[e]) continue;
    const uint32_t i = pairs[e].x;
    flags[i] |= kFlagAnnihilated;
    src[2 * i].w = 0.f;
  }
}

}  // namespace am

using namespace am;

extern "C" {

int am_compute_standard_coulomb_forces(am_store* s, double cutoff_radius, int precision) {
  if (!s) return Fail(AM_ERR_INVALID, "am_compute_standard_coulomb_forces: NULL store");
  if (precision != AM_FP32 && precision != AM_FP64) return Fail(AM_ERR_INVALID, "unknown precision");
  if (s->n == 0) return AM_OK;
  AM_HIP(hipSetDevice(s->ctx->device));
  am_ctx* ctx = s->ctx;
  if (cutoff_radius <= 0.0) {  // coulomb_interaction.cc:33-37
    hipLaunchKernelGGL(ZeroAliveForceKernel, dim3(Blocks(s->n, 256)), dim3(256), 0, ctx->stream, s->n, s->flags,
                       s->fx, s->fy, s->fz);
    AM_HIP(hipGetLastError());
    return AM_OK;
  }
  CellList cl;
  int rc = BuildCellList(s, cutoff_radius, "cl_force", &cl);
  if (rc) return rc;
  int slot;
  if (precision == AM_FP32) {
    const double rcu = cutoff_radius * kUnitScale;
    if (!(rcu * rcu < 1e30)) return Fail(AM_ERR_INVALID, "am_compute_standard_coulomb_forces: cutoff too large for FP32");
    void* ssrc;
    if ((rc = Scratch(ctx, "cl_force_ssrc", (size_t)s->n * 2 * sizeof(float4), &ssrc))) return rc;
    hipLaunchKernelGGL(CellSortedMirrorKernel, dim3(Blocks(s->n, 256)), dim3(256), 0, ctx->stream, s->n, cl.idx,
                       (const float4*)s->src, (float4*)ssrc);
    TimerBegin(ctx, "standard_f32", &slot);
    hipLaunchKernelGGL(StandardForceF32Kernel, dim3(Blocks(s->n, 256)), dim3(256), 0, ctx->stream, s->n, s->x, s->y,
                       s->z, s->q, s->flags, cl.tab, cl.idx, (const float4*)ssrc, cl.inv_h,
                       cutoff_radius * cutoff_radius, (float)(rcu * rcu), s->fx, s->fy, s->fz);
    TimerEnd(ctx, slot);
    AM_HIP(hipGetLastError());
    return AM_OK;
  }
  TimerBegin(ctx, "standard", &slot);
  hipLaunchKernelGGL(StandardForceKernel, dim3(Blocks(s->n, 256)), dim3(256), 0, ctx->stream, s->n, s->x, s->y, s->z,
                     s->q, s->flags, cl.tab, cl.idx, cl.inv_h, cutoff_radius * cutoff_radius, s->fx, s->fy, s->fz);
  TimerEnd(ctx, slot);
  AM_HIP(hipGetLastError());
  return AM_OK;
}

int am_annihilation(am_store* s, double contact_radius, double probability, uint64_t seed, uint64_t step,
                    int64_t* n_annihilated) {
  if (!s) return Fail(AM_ERR_INVALID, "am_annihilation: NULL store");
  if (n_annihilated) *n_annihilated = 0;
  if (s->n == 0 || contact_radius <= 0.0) return AM_OK;  // annihilation.cc:38
  AM_HIP(hipSetDevice(s->ctx->device));
  am_ctx* ctx = s->ctx;
  const int64_t n = s->n;
  CellList cl;
  int rc = BuildContactCells(s, kContactCellEdge * contact_radius * kContactReach, &cl);
  if (rc) return rc;
  void *cnt, *off, *marked;
  if ((rc = Scratch(ctx, "contact_count", (size_t)n * 4, &cnt))) return rc;
  if ((rc = Scratch(ctx, "contact_off", (size_t)(n + 1) * 4, &off))) return rc;
  if ((rc = Scratch(ctx, "contact_marked", 8, &marked))) return rc;
  const double ra2 = contact_radius * contact_radius;
  int slot;
  TimerBegin(ctx, "annihilation", &slot);
  hipLaunchKernelGGL(ContactCountKernel, dim3(Blocks(n, 256)), dim3(256), 0, ctx->stream, n, s->x, s->y, s->z,
                     s->species, s->flags, cl.tab, cl.idx, cl.inv_h, ra2, (uint32_t*)cnt);
  AM_HIP(hipGetLastError());
  if ((rc = ExclusiveScan(ctx, (const uint32_t*)cnt, (uint32_t*)off, n))) return rc;
  uint32_t total = 0, sort_err = 0;  // (the sticky sort error word: any sort on this context so far)
  AM_HIP(hipMemcpyAsync(&total, (uint32_t*)off + n, 4, hipMemcpyDeviceToHost, ctx->stream));
  if (const uint32_t* w = SortErrorWord(ctx))
    AM_HIP(hipMemcpyAsync(&sort_err, w, 4, hipMemcpyDeviceToHost, ctx->stream));
  AM_HIP(hipStreamSynchronize(ctx->stream));
  if ((rc = CheckSortErrorAfterSync(ctx, sort_err))) return rc;
  if (total > 0) {
    void* pairs;
    if ((rc = Scratch(ctx, "contact_pairs", (size_t)total * 8, &pairs))) return rc;
    hipLaunchKernelGGL(ContactWriteKernel, dim3(Blocks(n, 256)), dim3(256), 0, ctx->stream, n, s->x, s->y, s->z,
                       s->species, s->flags, cl.tab, cl.idx, cl.inv_h, ra2, (const uint32_t*)off, (uint2*)pairs);
    AM_HIP(hipGetLastError());
    if (total <= (uint32_t)kResolveLds) {
      hipLaunchKernelGGL(ContactResolveRoundsKernel, dim3(1), dim3(256), 0, ctx->stream, (const uint2*)pairs,
                         (const uint32_t*)off + n, probability, seed, step, s->flags, s->src, (unsigned int*)marked);
    } else {
      void *gslot, *gpass, *gmark;
      if ((rc = Scratch(ctx, "contact_gslot", (size_t)total * 4, &gslot))) return rc;
      if ((rc = Scratch(ctx, "contact_gpass", (size_t)total, &gpass))) return rc;
      if ((rc = Scratch(ctx, "contact_gmark", (size_t)total, &gmark))) return rc;
      hipLaunchKernelGGL(ContactResolveKernel, dim3(1), dim3(256), 0, ctx->stream, (const uint2*)pairs,
                         (const uint32_t*)off + n, probability, seed, step, s->flags, s->src, (unsigned int*)marked,
                         (uint32_t*)gslot, (uint8_t*)gpass, (uint8_t*)gmark);
    }
    AM_HIP(hipGetLastError());
    unsigned int hm = 0;
    AM_HIP(hipMemcpyAsync(&hm, marked, 4, hipMemcpyDeviceToHost, ctx->stream));
    AM_HIP(hipStreamSynchronize(ctx->stream));
    if (n_annihilated) *n_annihilated = hm;
  }
  TimerEnd(ctx, slot);
  return AM_OK;
}

}  // extern "C"
