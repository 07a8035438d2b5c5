// Exact all-pairs Coulomb force and potential (no cutoff): the theta = 0 limit of the reference's
// Barnes-Hut (src/interaction/barnes-hut.h:22-24) with the AddPointCharge pair term
// (barnes-hut.cc:19-27), and the all-pairs workload's hot kernel.
//
// FP32 path (AllPairsF32Kernel): every lane owns one target; the wave walks the source array in
// lockstep, so source loads are wave-uniform scalar loads (s_load) and the pair math runs on the
// FP32 VALU with SGPR operands.  Coordinates are hi/lo-split floats of x * 2^20 (units of
// ~0.95 um): dx = (xh_i - xh_j) + (xl_i - xl_j) keeps ~2^-24 relative accuracy on dx even for
// pairs much closer than the cloud size.  Each tile of kTile sources is summed in FP32 and the
// tile total is added to an FP64 accumulator (blocked compensated summation).  A softening of
// 1e-30 m^2 (the reference's pair floor) on r^2 makes coincident pairs, including i == j,
// contribute exactly zero without a branch; annihilated particles carry charge 0 in the
// mirror and drop out the same way.  Grid = (target blocks) x (source chunks); chunk partials
// are reduced in a fixed order, so results are deterministic.
//
// FP64 path (AllPairsF64Kernel): one thread per target, sources j = 0..n-1 in ascending order
// with the reference pair term, no FMA contraction: the same operations in the same order as
// the CPU restatement, so results are expected to match it bit for bit.
#include <algorithm>
#include <mutex>

#include "am_internal.hpp"
#include "pairtile.hpp"

namespace am {

constexpr int kAllPairsThreads = 256;
constexpr int kSortedMinN = 4096;  // the spatially sorted kernel from this size up

// part layout: [chunk][component][n_pad] doubles (acc in e/unit^2 before scaling).
__global__ __launch_bounds__(kAllPairsThreads) void AllPairsF32Kernel(const float4* __restrict__ src, int n_pad,
                                                                       int chunk, double* __restrict__ part) {
  const int i = blockIdx.x * kAllPairsThreads + threadIdx.x;  // < n_pad by construction
  const float4 th = src[2 * i];
  const float4 tl = src[2 * i + 1];
  const int j0 = blockIdx.y * chunk;
  const int j1 = min(n_pad, j0 + chunk);
  double ax = 0.0, ay = 0.0, az = 0.0;
  for (int jt = j0; jt < j1; jt += kTile) {
    float px = 0.f, py = 0.f, pz = 0.f;
    const float4* s = src + 2 * jt;
#pragma unroll 8
    for (int k = 0; k < kTile; ++k) PairF32(th, tl, s[2 * k], s[2 * k + 1], px, py, pz);
    ax += (double)px;
    ay += (double)py;
    az += (double)pz;
  }
  double* out = part + (size_t)blockIdx.y * 3 * n_pad;
  out[i] = ax;
  out[n_pad + i] = ay;
  out[2 * (size_t)n_pad + i] = az;
}

// Sum chunk partials in order and scale to newtons: F = k_C q_i e * 2^40 * acc.
__global__ void AllPairsReduceKernel(int n, int n_pad, int chunks, const double* __restrict__ part,
                                     const double* __restrict__ x, const double* __restrict__ y,
                                     const double* __restrict__ z, const double* __restrict__ q,
                                     const uint8_t* __restrict__ flags,
                                     double* __restrict__ fx, double* __restrict__ fy, double* __restrict__ fz) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double sx = 0.0, sy = 0.0, sz = 0.0;
  for (int c = 0; c < chunks; ++c) {
    const double* p = part + (size_t)c * 3 * n_pad;
    sx += p[i];
    sy += p[n_pad + i];
    sz += p[2 * (size_t)n_pad + i];
  }
  const bool live = !(flags[i] & kFlagAnnihilated) && q[i] != 0.0 && Finite3(x[i], y[i], z[i]);
  const double scale = live ? kCoulomb * q[i] * kElementaryCharge * kInvUnit2 : 0.0;
  fx[i] = scale * sx;
  fy[i] = scale * sy;
  fz[i] = scale * sz;
}

// FP64 source record: {x, y, z, q} with q = 0 for annihilated particles.
__global__ void PackF64SourceKernel(int n, const double* __restrict__ x, const double* __restrict__ y,
                                    const double* __restrict__ z, const double* __restrict__ q,
                                    const uint8_t* __restrict__ flags, double4* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const bool skip = (flags[i] & kFlagAnnihilated) || !Finite3(x[i], y[i], z[i]);
  out[i] = make_double4(x[i], y[i], z[i], skip ? 0.0 : q[i]);
}

// AddPointCharge (barnes-hut.cc:19-27) in FP64, sources in ascending index order.
__global__ __launch_bounds__(256) void AllPairsF64Kernel(int n, const double4* __restrict__ src,
                                                          double* __restrict__ fx, double* __restrict__ fy,
                                                          double* __restrict__ fz) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int it = min(i, n - 1);
  const double4 t = src[it];
  double ax = 0.0, ay = 0.0, az = 0.0;
  if (t.w != 0.0) {
    for (int j = 0; j < n; ++j) {
      const double4 s = src[j];
      if (j == it || s.w == 0.0) continue;
      const double rx = __dsub_rn(t.x, s.x), ry = __dsub_rn(t.y, s.y), rz = __dsub_rn(t.z, s.z);
      const double d2 = Dot3(rx, ry, rz, rx, ry, rz);
      if (d2 < kMinSquaredDistance) continue;
      const double inv = 1.0 / sqrt(d2);
      const double mag = __ddiv_rn(__dmul_rn(__dmul_rn(kCoulomb, t.w), s.w), d2);
      const double c = __dmul_rn(mag, inv);
      ax = __dadd_rn(ax, __dmul_rn(rx, c));
      ay = __dadd_rn(ay, __dmul_rn(ry, c));
      az = __dadd_rn(az, __dmul_rn(rz, c));
    }
  }
  if (i < n) {
    fx[i] = ax;
    fy[i] = ay;
    fz[i] = az;
  }
}

// Potential per particle (AccumulatePotential leaf term, barnes-hut.cc:223-237): phi_i =
// sum_j q_j / r_ij over alive j != i with r^2 >= 1e-30; out[i] = q_i * phi_i (0 if dead).
__global__ __launch_bounds__(256) void AllPairsPotentialF64Kernel(int n, const double4* __restrict__ src,
                                                                   double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int it = min(i, n - 1);
  const double4 t = src[it];
  double phi = 0.0;
  if (t.w != 0.0) {
    for (int j = 0; j < n; ++j) {
      const double4 s = src[j];
      if (j == it || s.w == 0.0) continue;
      const double rx = __dsub_rn(t.x, s.x), ry = __dsub_rn(t.y, s.y), rz = __dsub_rn(t.z, s.z);
      const double d2 = Dot3(rx, ry, rz, rx, ry, rz);
      if (d2 < kMinSquaredDistance) continue;
      phi = __dadd_rn(phi, __ddiv_rn(s.w, sqrt(d2)));
    }
  }
  if (i < n) out[i] = __dmul_rn(t.w, phi);
}

// FP32 potential: part[chunk][i] = sum_{j != i} q_j / sqrt(r^2 + eps^2) in e/unit.
__global__ __launch_bounds__(kAllPairsThreads) void AllPairsPotentialF32Kernel(const float4* __restrict__ src,
                                                                                int n_pad, int chunk,
                                                                                double* __restrict__ part) {
  const int i = blockIdx.x * kAllPairsThreads + threadIdx.x;
  const float4 th = src[2 * i];
  const float4 tl = src[2 * i + 1];
  const int j0 = blockIdx.y * chunk;
  const int j1 = min(n_pad, j0 + chunk);
  double acc = 0.0;
  for (int jt = j0; jt < j1; jt += kTile) {
    float p = 0.f;
    const float4* s = src + 2 * jt;
#pragma unroll 8
    for (int k = 0; k < kTile; ++k) {
      const float4 sh = s[2 * k], sl = s[2 * k + 1];
      const float dx = (th.x - sh.x) + (tl.x - sl.x);
      const float dy = (th.y - sh.y) + (tl.y - sl.y);
      const float dz = (th.z - sh.z) + (tl.z - sl.z);
      const float r2 = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, kEps2Units)));
      // pairs under the 1e-15 m floor, the target itself among them, are skipped (q_i / eps would
      // swamp the FP32 sum)
      const float w = r2 < 2.f * kEps2Units ? 0.f : sh.w;
      p = fmaf(w, __builtin_amdgcn_rsqf(r2), p);
    }
    acc += (double)p;
  }
  part[(size_t)blockIdx.y * n_pad + i] = acc;
}

__global__ void AllPairsPotentialReduceKernel(int n, int n_pad, int chunks, const double* __restrict__ part,
                                              const float4* __restrict__ src, double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (int c = 0; c < chunks; ++c) s += part[(size_t)c * n_pad + i];
  const float qe = src[2 * i].w;
  out[i] = (double)qe * s;  // q_i phi_i in e^2/unit
}

// ---- spatially sorted near/far variant --------------------------------------------------------
// Particles are kept in a Morton order (rebuilt every kResortInterval calls); the tile machinery
// (near/far tiles, unit-charge sign folding) is shared with the ensemble (pairtile.hpp).
constexpr int kResortInterval = 16;

__global__ void IotaKernel(int64_t n, uint32_t* __restrict__ v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = (uint32_t)i;
}

// 30-bit Morton key on a 1024^3 grid over the bounding box.
__global__ void SpatialKeyKernel(int64_t n, const double* __restrict__ x, const double* __restrict__ y,
                                 const double* __restrict__ z, const double* __restrict__ box,
                                 uint64_t* __restrict__ keys, uint32_t* __restrict__ vals) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double ext = fmax(fmax(box[3] - box[0], box[4] - box[1]), fmax(box[5] - box[2], 1e-300));
  const double sc = 1023.999 / ext;
  const double p[3] = {x[i], y[i], z[i]};
  uint32_t g[3];
  for (int d = 0; d < 3; ++d) {
    double c = (p[d] - box[d]) * sc;
    c = (c == c) ? fmin(fmax(c, 0.0), 1023.0) : 0.0;
    g[d] = (uint32_t)c;
  }
  keys[i] = (uint64_t)(Spread10(g[0]) | (Spread10(g[1]) << 1) | (Spread10(g[2]) << 2));
  vals[i] = (uint32_t)i;
}

// One workgroup per 128-source tile of the store's sorted order (n particles).
__global__ __launch_bounds__(kTile) void TilePrepKernel(int n, uint32_t* __restrict__ perm,
                                                        const double* __restrict__ x, const double* __restrict__ y,
                                                        const double* __restrict__ z, const float4* __restrict__ src,
                                                        float4* __restrict__ shl, float4* __restrict__ sloc,
                                                        TileRec* __restrict__ tiles) {
  TilePrepBody(blockIdx.x * kTile, n - (int)blockIdx.x * kTile, blockIdx.x, perm, x, y, z, src, shl, sloc, tiles);
}

// Work items of the sorted kernel: (target block, source chunk), taken from a queue in chunk-major
// order by a persistent grid of kSortedWaves waves per SIMD.  A SIMD issues from its oldest waves
// first, so with one fixed (block, chunk) per workgroup the early workgroups finished long before the
// late ones and the last part of the launch ran with one or two waves per SIMD (at N = 65,536 the
// per-workgroup durations grew from 0.49 to 1.4-1.6 ms with the chunk index).  With a queue the
// waves that run ahead take more items, and items of kTilesPerItem tiles keep the final drain short.
// The counter pair [next item, finished workgroups] is reset by the last workgroup to finish.
constexpr int kSortedWaves = 8;
constexpr int kTilesPerItem = 8;
constexpr int kSortedMaxItems = 16384;  // bounds the partial sums at chunks x 24 bytes x n

__global__ __launch_bounds__(kAllPairsThreads) __attribute__((amdgpu_waves_per_eu(kSortedWaves, kSortedWaves)))
void AllPairsSortedKernel(int n, int n_pad, int tiles_per_chunk, int n_blocks, int n_items,
                          unsigned* __restrict__ queue, const uint32_t* __restrict__ perm,
                          const double* __restrict__ x, const double* __restrict__ y,
                          const double* __restrict__ z, const float4* __restrict__ shl,
                          const float4* __restrict__ sloc, const TileRec* __restrict__ tiles,
                          double* __restrict__ part) {
  __shared__ int s_item;
  __shared__ float4 s_stage[kTile];  // the expanded far-tile sums' source records
  const int n_tiles = n_pad / kTile;
  for (;;) {
    if (threadIdx.x == 0) s_item = (int)atomicAdd(&queue[0], 1u);
    __syncthreads();
    const int item = s_item;
    __syncthreads();  // every lane has read s_item before thread 0 overwrites it
    if (item >= n_items) break;
    const int blk = item % n_blocks, chunk = item / n_blocks;
    const int k = blk * kAllPairsThreads + threadIdx.x;
    const BlockBox bb = MakeBlockBox(tiles[2 * blk], tiles[2 * blk + 1], 2 * blk * kTile + kTile >= n);
    const float4 th = shl[2 * k], tl = shl[2 * k + 1];
    float lx = 0.f, ly = 0.f, lz = 0.f;
    if (k < n) {
      const uint32_t i = perm[k];
      lx = (float)(x[i] * kUnitScale - bb.bc[0]);
      ly = (float)(y[i] * kUnitScale - bb.bc[1]);
      lz = (float)(z[i] * kUnitScale - bb.bc[2]);
    }
    const int t0 = chunk * tiles_per_chunk;
    const int t1 = min(n_tiles, t0 + tiles_per_chunk);
    double ax = 0.0, ay = 0.0, az = 0.0;
    for (int T = t0; T < t1; ++T) {
      float px = 0.f, py = 0.f, pz = 0.f;
      TileSum(tiles[T], T * kTile, bb, th, tl, lx, ly, lz, shl, sloc, s_stage, px, py, pz);
      ax += (double)px;
      ay += (double)py;
      az += (double)pz;
    }
    double* out = part + (size_t)chunk * 3 * n_pad;
    out[k] = ax;
    out[n_pad + k] = ay;
    out[2 * (size_t)n_pad + k] = az;
  }
  if (threadIdx.x == 0) {
    __threadfence();
    if (atomicAdd(&queue[1], 1u) == gridDim.x - 1) {  // every other workgroup has taken its last item
      queue[0] = 0u;
      queue[1] = 0u;
      __threadfence();
    }
  }
}

__global__ void AllPairsSortedReduceKernel(int n, int n_pad, int chunks, const double* __restrict__ part,
                                           const uint32_t* __restrict__ perm, const double* __restrict__ x,
                                           const double* __restrict__ y, const double* __restrict__ z,
                                           const double* __restrict__ q,
                                           const uint8_t* __restrict__ flags, double* __restrict__ fx,
                                           double* __restrict__ fy, double* __restrict__ fz) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  double sx = 0.0, sy = 0.0, sz = 0.0;
  for (int c = 0; c < chunks; ++c) {
    const double* p = part + (size_t)c * 3 * n_pad;
    sx += p[k];
    sy += p[n_pad + k];
    sz += p[2 * (size_t)n_pad + k];
  }
  const uint32_t i = perm[k];
  const bool live = !(flags[i] & kFlagAnnihilated) && q[i] != 0.0 && Finite3(x[i], y[i], z[i]);
  const double scale = live ? kCoulomb * q[i] * kElementaryCharge * kInvUnit2 : 0.0;
  fx[i] = scale * sx;
  fy[i] = scale * sy;
  fz[i] = scale * sz;
}

// ---- sorted FP32 potential (the energy diagnostic) ----------------------------------------------
// phi of source tile R at one target, sum_j q_j / r in e/unit.  Pairs closer than the reference's
// 1e-15 m floor (energy_analysis.cc via barnes-hut.cc:244-247), the target itself among them, are
// skipped; they can only sit in near tiles, where the test costs a compare and a select per pair.
// Far unit-charge tiles across a gap of at least a half extent take the expanded form on staged
// records (4 VALU + rsq + 1 add per pair), the other far tiles the local form.
__device__ __forceinline__ float TilePotential(const TileRec& R, int tile_first, const BlockBox& bb, const float4 th,
                                               const float4 tl, float lx, float ly, float lz,
                                               const float4* __restrict__ shl, const float4* __restrict__ sloc,
                                               float4* __restrict__ stage) {
  const double gx = fmax(0.0, fabs(R.cx - bb.bc[0]) - R.hx - bb.bh[0]);
  const double gy = fmax(0.0, fabs(R.cy - bb.bc[1]) - R.hy - bb.bh[1]);
  const double gz = fmax(0.0, fabs(R.cz - bb.bc[2]) - R.hz - bb.bh[2]);
  const double size = fmax(fmax(fmax(R.hx, R.hy), R.hz), bb.bhmax);
  const double lim = 2.0 * kNearGap * size;
  const double gap2 = gx * gx + gy * gy + gz * gz;
  const int near = __builtin_amdgcn_readfirstlane(gap2 < lim * lim ? 1 : 0);
  // the force sums' expanded-form rule: a positive gap of at least kExpandGap x the larger half extent
  const int expand =
      __builtin_amdgcn_readfirstlane(gap2 > 0.0 && gap2 >= (kExpandGap * size) * (kExpandGap * size) ? 1 : 0);
  const int unit = __builtin_amdgcn_readfirstlane(R.unit);
  const int np = __builtin_amdgcn_readfirstlane(R.npos);
  const int nq = __builtin_amdgcn_readfirstlane(R.npos + R.nneg);  // charged sources come first
  const int ge = (nq + 7) & ~7;
  float p = 0.f;
  if (near) {
    // hi/lo sources in groups of 4 (32 SGPRs): groups of 8 spill SGPRs to VGPR lanes inside the loop
    const float4* sp = shl + 2 * tile_first;
    const int ge4 = (nq + 3) & ~3;
    for (int j0 = 0; j0 < ge4; j0 += 4) {
      const Src8 g = LoadSrc8(sp, 2 * j0);
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const float4 sh = g[2 * u], sl = g[2 * u + 1];
        const float dx = (th.x - sh.x) + (tl.x - sl.x);
        const float dy = (th.y - sh.y) + (tl.y - sl.y);
        const float dz = (th.z - sh.z) + (tl.z - sl.z);
        const float r2 = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, kEps2Units)));
        const float w = r2 < 2.f * kEps2Units ? 0.f : sh.w;
        p = fmaf(w, __builtin_amdgcn_rsqf(r2), p);
      }
    }
  } else if (unit && expand) {
    // +e sources [0, np), then -e sources [np, nq), on the staged expanded-form records (the whole
    // workgroup takes this branch for this tile)
    const float tx = lx + (float)(bb.bc[0] - R.cx);
    const float ty = ly + (float)(bb.bc[1] - R.cy);
    const float tz = lz + (float)(bb.bc[2] - R.cz);
    StageExpanded(sloc + tile_first, stage);
    const float tt = fmaf(tx, tx, fmaf(ty, ty, tz * tz));
    float pp = 0.f, pn = 0.f;
#pragma unroll 8
    for (int j = 0; j < np; ++j) pp += InvRExpanded(tx, ty, tz, tt, stage[j]);
#pragma unroll 8
    for (int j = np; j < nq; ++j) pn += InvRExpanded(tx, ty, tz, tt, stage[j]);
    p = pp - pn;
  } else {
    const float tx = lx + (float)(bb.bc[0] - R.cx);
    const float ty = ly + (float)(bb.bc[1] - R.cy);
    const float tz = lz + (float)(bb.bc[2] - R.cz);
    const float4* sp = sloc + tile_first;
    for (int j0 = 0; j0 < ge; j0 += 8) {
      const Src8 g = LoadSrc8(sp, j0);
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const float4 sv = g[u];
        const float dx = tx - sv.x, dy = ty - sv.y, dz = tz - sv.z;
        const float r2 = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, kEps2Units)));
        p = fmaf(sv.w, __builtin_amdgcn_rsqf(r2), p);
      }
    }
  }
  return p;
}

// part[chunk][k] = phi at sorted position k over the chunk's tiles (e/unit), FP32 per tile, FP64
// across tiles.
// The same work queue as AllPairsSortedKernel, one partial potential per (chunk, target).
__global__ __launch_bounds__(kAllPairsThreads) __attribute__((amdgpu_waves_per_eu(kSortedWaves, kSortedWaves)))
void AllPairsPotentialSortedKernel(int n, int n_pad, int tiles_per_chunk, int n_blocks, int n_items,
                                   unsigned* __restrict__ queue, const uint32_t* __restrict__ perm,
                                   const double* __restrict__ x, const double* __restrict__ y,
                                   const double* __restrict__ z, const float4* __restrict__ shl,
                                   const float4* __restrict__ sloc, const TileRec* __restrict__ tiles,
                                   double* __restrict__ part) {
  __shared__ int s_item;
  __shared__ float4 s_stage[kTile];  // the expanded far-tile sums' source records
  const int n_tiles = n_pad / kTile;
  for (;;) {
    if (threadIdx.x == 0) s_item = (int)atomicAdd(&queue[0], 1u);
    __syncthreads();
    const int item = s_item;
    __syncthreads();
    if (item >= n_items) break;
    const int blk = item % n_blocks, chunk = item / n_blocks;
    const int k = blk * kAllPairsThreads + threadIdx.x;
    const BlockBox bb = MakeBlockBox(tiles[2 * blk], tiles[2 * blk + 1], 2 * blk * kTile + kTile >= n);
    const float4 th = shl[2 * k], tl = shl[2 * k + 1];
    float lx = 0.f, ly = 0.f, lz = 0.f;
    if (k < n) {
      const uint32_t i = perm[k];
      lx = (float)(x[i] * kUnitScale - bb.bc[0]);
      ly = (float)(y[i] * kUnitScale - bb.bc[1]);
      lz = (float)(z[i] * kUnitScale - bb.bc[2]);
    }
    const int t0 = chunk * tiles_per_chunk;
    const int t1 = min(n_tiles, t0 + tiles_per_chunk);
    double acc = 0.0;
    for (int T = t0; T < t1; ++T)
      acc += (double)TilePotential(tiles[T], T * kTile, bb, th, tl, lx, ly, lz, shl, sloc, s_stage);
    part[(size_t)chunk * n_pad + k] = acc;
  }
  if (threadIdx.x == 0) {
    __threadfence();
    if (atomicAdd(&queue[1], 1u) == gridDim.x - 1) {
      queue[0] = 0u;
      queue[1] = 0u;
      __threadfence();
    }
  }
}

// out[i] = (q_i / e) phi_i in e^2/unit, zero for annihilated, neutral or non-finite particles.
__global__ void AllPairsPotentialSortedReduceKernel(int n, int n_pad, int chunks, const double* __restrict__ part,
                                                    const uint32_t* __restrict__ perm,
                                                    const float4* __restrict__ shl, double* __restrict__ out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  double sum = 0.0;
  for (int c = 0; c < chunks; ++c) sum += part[(size_t)c * n_pad + k];
  out[perm[k]] = (double)shl[2 * k].w * sum;  // the mirror's charge is 0 for every skipped particle
}

int GatheredBBox(am_ctx* ctx, const uint32_t* idx, int m, const double* x, const double* y, const double* z,
                 double* box_dev, const char* tag);

static int ChunkCount(int num_cus, int target_blocks, int n_pad);

// The store's Morton-sorted tiles, prepared for the current positions.
struct SortedTiles {
  const float4* shl = nullptr;
  const float4* sloc = nullptr;
  const TileRec* tiles = nullptr;
  int n_tiles = 0;
};

// Source chunks of the sorted kernels: kTilesPerItem tiles each, or more where that would give more
// than kSortedMaxItems (block, chunk) items.
static void SortedChunks(int n_tiles, int tblocks, int* tpc, int* nch) {
  const int max_chunks = std::max(1, kSortedMaxItems / tblocks);
  *tpc = std::max(kTilesPerItem, (n_tiles + max_chunks - 1) / max_chunks);
  *nch = (n_tiles + *tpc - 1) / *tpc;
}

static int PrepareSortedTiles(am_store* s, SortedTiles* out) {
  am_ctx* ctx = s->ctx;
  hipStream_t st = ctx->stream;
  const int n = (int)s->n;
  const int n_pad = PadN(n);
  int rc;
  if (!s->perm) AM_HIP(hipMalloc((void**)&s->perm, (size_t)PadN(s->capacity) * 4));
  if (s->perm_n != n || ++s->perm_age >= kResortInterval) {
    void *keys, *box;
    if ((rc = Scratch(ctx, "ap_keys", (size_t)n * 8, &keys))) return rc;
    if ((rc = Scratch(ctx, "ap_box", 6 * 8, &box))) return rc;
    hipLaunchKernelGGL(IotaKernel, dim3(Blocks(n, 256)), dim3(256), 0, st, (int64_t)n, s->perm);
    if ((rc = GatheredBBox(ctx, s->perm, n, s->x, s->y, s->z, (double*)box, "ap"))) return rc;
    hipLaunchKernelGGL(SpatialKeyKernel, dim3(Blocks(n, 256)), dim3(256), 0, st, (int64_t)n, s->x, s->y, s->z,
                       (const double*)box, (uint64_t*)keys, s->perm);
    AM_HIP(hipGetLastError());
    s->perm_n = -1;  // until the sort is known to have finished
    if ((rc = RadixSortPairs(ctx, (uint64_t*)keys, s->perm, n, 30))) return rc;
    // The permutation is kept for kResortInterval calls, so a sort whose look-back gave up must not
    // be kept: check it now (one synchronisation per resort).
    if ((rc = CheckSortError(ctx))) return rc;
    s->perm_n = n;
    s->perm_age = 0;
  }
  void *shl, *sloc, *tiles;
  const int n_tiles = n_pad / kTile;
  if ((rc = Scratch(ctx, "ap_shl", (size_t)n_pad * 2 * sizeof(float4), &shl))) return rc;
  if ((rc = Scratch(ctx, "ap_sloc", (size_t)n_pad * sizeof(float4), &sloc))) return rc;
  if ((rc = Scratch(ctx, "ap_tiles", (size_t)n_tiles * sizeof(TileRec), &tiles))) return rc;
  hipLaunchKernelGGL(TilePrepKernel, dim3(n_tiles), dim3(kTile), 0, st, n, s->perm, s->x, s->y,
                     s->z, (const float4*)s->src, (float4*)shl, (float4*)sloc, (TileRec*)tiles);
  AM_HIP(hipGetLastError());
  out->shl = (const float4*)shl;
  out->sloc = (const float4*)sloc;
  out->tiles = (const TileRec*)tiles;
  out->n_tiles = n_tiles;
  return AM_OK;
}

int AllPairsForcesF32Sorted(am_store* s) {
  am_ctx* ctx = s->ctx;
  hipStream_t st = ctx->stream;
  const int n = (int)s->n;
  const int n_pad = PadN(n);
  int rc;
  SortedTiles T;
  if ((rc = PrepareSortedTiles(s, &T))) return rc;
  const int n_tiles = T.n_tiles;
  void* part;
  const int tblocks = n_pad / kAllPairsThreads;
  int tpc, nch;
  SortedChunks(n_tiles, tblocks, &tpc, &nch);
  if ((rc = Scratch(ctx, "allpairs_part", (size_t)nch * 3 * n_pad * sizeof(double), &part))) return rc;
  unsigned* queue;
  if ((rc = QueueCounters(ctx, "ap_queue", &queue))) return rc;
  const int cus = ctx->num_cus > 0 ? ctx->num_cus : 256;
  int slot;
  TimerBegin(ctx, "allpairs_f32", &slot);
  hipLaunchKernelGGL(AllPairsSortedKernel, dim3(cus * kSortedWaves), dim3(kAllPairsThreads), 0, st, n, n_pad, tpc,
                     tblocks, tblocks * nch, queue, (const uint32_t*)s->perm, s->x, s->y, s->z, T.shl, T.sloc,
                     T.tiles, (double*)part);
  TimerEnd(ctx, slot);
  AM_HIP(hipGetLastError());
  hipLaunchKernelGGL(AllPairsSortedReduceKernel, dim3(Blocks(n, 256)), dim3(256), 0, st, n, n_pad, nch,
                     (const double*)part, (const uint32_t*)s->perm, s->x, s->y, s->z, s->q, s->flags, s->fx, s->fy,
                     s->fz);
  AM_HIP(hipGetLastError());
  return AM_OK;
}

static int ChunkCount(int num_cus, int target_blocks, int n_pad) {
  const int cus = num_cus > 0 ? num_cus : 256;
  const int want = 8 * cus;  // >= 8 workgroups per CU
  int chunks = (want + target_blocks - 1) / target_blocks;
  chunks = std::max(1, std::min(chunks, n_pad / kTile));
  return chunks;
}


int AllPairsForcesF32(am_store* s) {
  am_ctx* ctx = s->ctx;
  const int n = (int)s->n;
  const int n_pad = PadN(n);
  const int tblocks = n_pad / kAllPairsThreads;
  const int chunks = ChunkCount(ctx->num_cus, tblocks, n_pad);
  int chunk = (n_pad / chunks + kTile - 1) / kTile * kTile;
  const int nch = (n_pad + chunk - 1) / chunk;
  void* part = nullptr;
  int rc = Scratch(ctx, "allpairs_part", (size_t)nch * 3 * n_pad * sizeof(double), &part);
  if (rc) return rc;
  int slot;
  TimerBegin(ctx, "allpairs_f32", &slot);
  hipLaunchKernelGGL(AllPairsF32Kernel, dim3(tblocks, nch), dim3(kAllPairsThreads), 0, ctx->stream, s->src, n_pad,
                     chunk, (double*)part);
  TimerEnd(ctx, slot);
  AM_HIP(hipGetLastError());
  hipLaunchKernelGGL(AllPairsReduceKernel, dim3(Blocks(n, 256)), dim3(256), 0, ctx->stream, n, n_pad, nch,
                     (const double*)part, s->x, s->y, s->z, s->q, s->flags, s->fx, s->fy, s->fz);
  AM_HIP(hipGetLastError());
  return AM_OK;
}

int AllPairsForcesF64(am_store* s) {
  am_ctx* ctx = s->ctx;
  const int n = (int)s->n;
  void* src = nullptr;
  int rc = Scratch(ctx, "src_f64", (size_t)n * sizeof(double4), &src);
  if (rc) return rc;
  hipLaunchKernelGGL(PackF64SourceKernel, dim3(Blocks(n, 256)), dim3(256), 0, ctx->stream, n, s->x, s->y, s->z, s->q,
                     s->flags, (double4*)src);
  AM_HIP(hipGetLastError());
  int slot;
  TimerBegin(ctx, "allpairs_f64", &slot);
  hipLaunchKernelGGL(AllPairsF64Kernel, dim3(Blocks(n, 256)), dim3(256), 0, ctx->stream, n, (const double4*)src,
                     s->fx, s->fy, s->fz);
  TimerEnd(ctx, slot);
  AM_HIP(hipGetLastError());
  return AM_OK;
}

// Per-particle q_i phi_i (SI: J when multiplied by k_C/2 and summed) into out[n] (device).
int AllPairsPotentialTerms(am_store* s, int precision, double* out, double* si_scale) {
  am_ctx* ctx = s->ctx;
  const int n = (int)s->n;
  if (precision == AM_FP64) {
    void* src = nullptr;
    int rc = Scratch(ctx, "src_f64", (size_t)n * sizeof(double4), &src);
    if (rc) return rc;
    hipLaunchKernelGGL(PackF64SourceKernel, dim3(Blocks(n, 256)), dim3(256), 0, ctx->stream, n, s->x, s->y, s->z,
                       s->q, s->flags, (double4*)src);
    AM_HIP(hipGetLastError());
    hipLaunchKernelGGL(AllPairsPotentialF64Kernel, dim3(Blocks(n, 256)), dim3(256), 0, ctx->stream, n,
                       (const double4*)src, out);
    AM_HIP(hipGetLastError());
    *si_scale = 1.0;
    return AM_OK;
  }
  const int n_pad = PadN(n);
  const int tblocks = n_pad / kAllPairsThreads;
  const int chunks = ChunkCount(ctx->num_cus, tblocks, n_pad);
  *si_scale = kElementaryCharge * kElementaryCharge * kUnitScale;  // e^2/unit -> C^2/m
  if (n >= kSortedMinN) {
    SortedTiles T;
    int rc = PrepareSortedTiles(s, &T);
    if (rc) return rc;
    int tpc, nch;
    SortedChunks(T.n_tiles, tblocks, &tpc, &nch);
    void* part = nullptr;
    if ((rc = Scratch(ctx, "allpairs_part", (size_t)nch * 3 * n_pad * sizeof(double), &part))) return rc;
    unsigned* queue;
    if ((rc = QueueCounters(ctx, "ap_queue", &queue))) return rc;
    const int cus = ctx->num_cus > 0 ? ctx->num_cus : 256;
    int slot;
    TimerBegin(ctx, "allpairs_potential_f32", &slot);
    hipLaunchKernelGGL(AllPairsPotentialSortedKernel, dim3(cus * kSortedWaves), dim3(kAllPairsThreads), 0,
                       ctx->stream, n, n_pad, tpc, tblocks, tblocks * nch, queue, (const uint32_t*)s->perm, s->x,
                       s->y, s->z, T.shl, T.sloc, T.tiles, (double*)part);
    TimerEnd(ctx, slot);
    AM_HIP(hipGetLastError());
    hipLaunchKernelGGL(AllPairsPotentialSortedReduceKernel, dim3(Blocks(n, 256)), dim3(256), 0, ctx->stream, n,
                       n_pad, nch, (const double*)part, (const uint32_t*)s->perm, T.shl, out);
    AM_HIP(hipGetLastError());
    return AM_OK;
  }
  int chunk = (n_pad / chunks + kTile - 1) / kTile * kTile;
  const int nch = (n_pad + chunk - 1) / chunk;
  void* part = nullptr;
  int rc = Scratch(ctx, "allpairs_part", (size_t)nch * 3 * n_pad * sizeof(double), &part);
  if (rc) return rc;
  hipLaunchKernelGGL(AllPairsPotentialF32Kernel, dim3(tblocks, nch), dim3(kAllPairsThreads), 0, ctx->stream, s->src,
                     n_pad, chunk, (double*)part);
  AM_HIP(hipGetLastError());
  hipLaunchKernelGGL(AllPairsPotentialReduceKernel, dim3(Blocks(n, 256)), dim3(256), 0, ctx->stream, n, n_pad, nch,
                     (const double*)part, s->src, out);
  AM_HIP(hipGetLastError());
  return AM_OK;
}

}  // namespace am

using namespace am;

extern "C" int am_compute_allpairs_coulomb_forces(am_store* s, int precision) {
  if (!s) return Fail(AM_ERR_INVALID, "am_compute_allpairs_coulomb_forces: NULL store");
  if (precision != AM_FP32 && precision != AM_FP64) return Fail(AM_ERR_INVALID, "unknown precision");
  if (s->n == 0) return AM_OK;
  if (s->n > (int64_t)1 << 28) return Fail(AM_ERR_CAPACITY, "all-pairs: n too large");
  AM_HIP(hipSetDevice(s->ctx->device));
  if (precision == AM_FP64) return AllPairsForcesF64(s);
  return s->n >= kSortedMinN ? AllPairsForcesF32Sorted(s) : AllPairsForcesF32(s);
}

// ---- allocation-free variant (am_allpairs_forces_noalloc) ----------------------------------------
// The kernels of the unsorted FP32 path on caller-owned device buffers, with the FP32 source mirror
// (32 B per padded particle) and the chunk partials (24 B per padded particle and source chunk) in a
// file-scope device array.  Each GPU holds its own copy of the array for the life of the process;
// its address on a device is taken once, with that device current, by hipGetSymbolAddress, and only
// that address is ever handed to kernels.  Every kernel writes all the scratch it reads (padding
// entries of the mirror included), so nothing needs clearing between calls.
namespace am {

constexpr size_t kNoAllocScratchBytes = (size_t)16 << 20;  // 16 MiB: the mirror and partials at n = 65,536
constexpr int kNoAllocMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

struct NoAllocDevice {
  void* ptr = nullptr;  // device address of `scratch` on this device
  int num_cus = 0;
  bool ready = false;
};
static NoAllocDevice g_noalloc[kNoAllocMaxDevices];
static std::mutex g_noalloc_mu;

static int NoAllocSetup(int dev, NoAllocDevice** out) {
  std::lock_guard<std::mutex> lock(g_noalloc_mu);
  NoAllocDevice& d = g_noalloc[dev];
  if (!d.ready) {
    void* p = nullptr;
    AM_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return Fail(AM_ERR_HIP, "am_allpairs_forces_noalloc: no device address for the scratch array");
    int cus = 0;
    AM_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    d.ptr = p;
    d.num_cus = cus;
    d.ready = true;
  }
  *out = &d;
  return AM_OK;
}

// RefreshSourceKernel's mirror from interleaved positions (pos[3i..3i+2]) and caller charges (a zero
// entry for annihilated or non-finite particles).
__global__ void PackSourceAosKernel(int n, int n_pad, const double* __restrict__ pos, const double* __restrict__ q,
                                    const uint8_t* __restrict__ annihilated, float4* __restrict__ src) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_pad) return;
  if (i >= n) {
    src[2 * i] = make_float4(0.f, 0.f, 0.f, 0.f);
    src[2 * i + 1] = make_float4(0.f, 0.f, 0.f, 0.f);
    return;
  }
  WriteMirror(src, i, pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], q[i], annihilated && annihilated[i]);
}

// AllPairsReduceKernel with interleaved output force[3i..3i+2]; zero for annihilated, neutral or
// non-finite i.
__global__ void AllPairsReduceAosKernel(int n, int n_pad, int chunks, const double* __restrict__ part,
                                        const double* __restrict__ pos, const double* __restrict__ q,
                                        const uint8_t* __restrict__ annihilated,
                                        double* __restrict__ force) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double sx = 0.0, sy = 0.0, sz = 0.0;
  for (int c = 0; c < chunks; ++c) {
    const double* p = part + (size_t)c * 3 * n_pad;
    sx += p[i];
    sy += p[n_pad + i];
    sz += p[2 * (size_t)n_pad + i];
  }
  const bool live =
      !(annihilated && annihilated[i]) && q[i] != 0.0 && Finite3(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
  const double scale = live ? kCoulomb * q[i] * kElementaryCharge * kInvUnit2 : 0.0;
  force[3 * i] = scale * sx;
  force[3 * i + 1] = scale * sy;
  force[3 * i + 2] = scale * sz;
}

}  // namespace am

extern "C" int am_allpairs_forces_noalloc(int64_t n, const double* pos, const double* charge,
                                          const uint8_t* annihilated, double* force, void* stream) {
  if (n < 0 || n > AM_NOALLOC_MAX_N)
    return Fail(AM_ERR_CAPACITY, "am_allpairs_forces_noalloc: n must be in [0, AM_NOALLOC_MAX_N = 65536]");
  if (n == 0) return AM_OK;
  if (!pos || !charge || !force) return Fail(AM_ERR_INVALID, "am_allpairs_forces_noalloc: NULL buffer");
  int dev = -1;
  AM_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kNoAllocMaxDevices) return Fail(AM_ERR_INVALID, "am_allpairs_forces_noalloc: device id");
  NoAllocDevice* d = nullptr;
  int rc = NoAllocSetup(dev, &d);
  if (rc) return rc;
  const int ni = (int)n;
  const int n_pad = PadN(n);
  const int tblocks = n_pad / kAllPairsThreads;
  const int chunks = ChunkCount(d->num_cus, tblocks, n_pad);
  const int chunk = (n_pad / chunks + kTile - 1) / kTile * kTile;
  const int nch = (n_pad + chunk - 1) / chunk;
  const size_t mirror_bytes = (size_t)2 * n_pad * sizeof(float4);
  const size_t part_bytes = (size_t)nch * 3 * n_pad * sizeof(double);
  if (mirror_bytes + part_bytes > kNoAllocScratchBytes)
    return Fail(AM_ERR_CAPACITY, "am_allpairs_forces_noalloc: scratch of this size exceeds the 16 MiB array");
  float4* src = static_cast<float4*>(d->ptr);
  double* part = reinterpret_cast<double*>(static_cast<unsigned char*>(d->ptr) + mirror_bytes);
  hipStream_t st = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(PackSourceAosKernel, dim3(Blocks(n_pad, 256)), dim3(256), 0, st, ni, n_pad, pos, charge,
                     annihilated, src);
  hipLaunchKernelGGL(AllPairsF32Kernel, dim3(tblocks, nch), dim3(kAllPairsThreads), 0, st, (const float4*)src, n_pad,
                     chunk, part);
  hipLaunchKernelGGL(AllPairsReduceAosKernel, dim3(Blocks(n, 256)), dim3(256), 0, st, ni, n_pad, nch,
                     (const double*)part, pos, charge, annihilated, force);
  AM_HIP(hipGetLastError());
  return AM_OK;
}
