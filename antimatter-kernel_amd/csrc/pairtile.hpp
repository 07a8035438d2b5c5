// Morton-sorted source tiles of the FP32 pair kernels (the all-pairs kernel and the ensemble).
//
// Sources form tiles of kTile particles in a spatial (Morton) order, each with an FP64 bounding
// box.  For a (256-target block, source tile) pair whose boxes are close (gap < kNearGap x 2 x the
// larger half extent, which includes the block's own tiles) the pair math uses hi/lo coordinates.
// For the other tile pairs both sides use coordinates local to their box centre: the target
// coordinate is shifted once per tile by the FP32-rounded centre offset, then dx = tx'' - sx'
// costs one instruction instead of three.  Its rounding error is ~2^-24 (|dx| + box size), i.e.
// <~1e-6 relative for a far pair, below the FP32 accumulation error of the tile sums.
//
// Unit-charge tiles at least kExpandGap x the larger half extent away (and at a positive gap) use the
// expanded form r^2 = |t|^2 + (|s|^2 + eps^2) - 2 t.s: the workgroup stages the tile's records
// (-2s, |s|^2 + eps^2) in LDS, each wave reads them back as broadcasts (VGPR operands), and sums c and
// c(-2s) per charge sign; t sum(c) - sum(c s) is formed once per tile.  That is 10 VALU instructions
// and one rsq per pair, none reading an SGPR (a VALU instruction with an SGPR operand issues at half
// rate on gfx950).  The cancellation error of r^2 is ~2^-24 (|t| + |s|)^2 / r^2 with t and s local to
// the source tile, at most ~20 x 2^-24 at that separation.
#pragma once

#include "am_internal.hpp"

namespace am {

constexpr int kTile = 128;        // sources per tile (and per FP32 partial sum)
constexpr double kNearGap = 0.1;  // near if box gap < kNearGap * (larger half extent) * 2
constexpr double kExpandGap = 1.0;  // expanded form if box gap >= kExpandGap * (larger half extent)

struct TileRec {  // sorted source tile, units of 2^-20 m
  double cx, cy, cz;
  double hx, hy, hz;
  int npos, nneg;  // the tile holds npos charges > 0, then nneg < 0, then the zero charges
  int unit;        // every non-zero charge is exactly +-e
  int pad;
};

__device__ inline uint32_t Spread10(uint32_t v) {
  v &= 1023u;
  v = (v | (v << 16)) & 0x030000FFu;
  v = (v | (v << 8)) & 0x0300F00Fu;
  v = (v | (v << 4)) & 0x030C30C3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}

// One workgroup of kTile threads per tile: entries [first, first + kTile) of a sorted layout, the
// first nvalid of which are particles perm[first + t].  Writes the sorted hi/lo mirror (shl), the
// coordinates local to the tile's box centre (sloc) and the box (tiles[tile]).  The tile's slice of
// the permutation is reordered stably by charge sign (q > 0, q < 0, q = 0), so the pair loops can
// fold a unit charge's sign into the direction of the coordinate difference.  Padding entries
// (t >= nvalid) have q = 0 and keep their positions.  The block box of a tile without a finite
// particle is the zero box at the origin.
__device__ inline void TilePrepBody(int first, int nvalid, int tile, uint32_t* __restrict__ perm,
                                    const double* __restrict__ x, const double* __restrict__ y,
                                    const double* __restrict__ z, const float4* __restrict__ src,
                                    float4* __restrict__ shl, float4* __restrict__ sloc,
                                    TileRec* __restrict__ tiles) {
  __shared__ double lo[3][kTile], hi[3][kTile];
  __shared__ int cls_cnt[kTile / 64][3];
  __shared__ int non_unit;
  const int t = threadIdx.x;
  const int k = first + t;
  const bool valid = t < nvalid;
  double X[3] = {0.0, 0.0, 0.0};
  float4 h = make_float4(0.f, 0.f, 0.f, 0.f), l = h;
  uint32_t i = 0;
  if (valid) {
    i = perm[k];
    X[0] = x[i] * kUnitScale;
    X[1] = y[i] * kUnitScale;
    X[2] = z[i] * kUnitScale;
    h = src[2 * i];
    l = src[2 * i + 1];
  }
  // A particle with a non-finite coordinate has a zero mirror entry (charge 0): it stays out of the
  // box and gets a zero local coordinate, so it adds exactly nothing to any pair sum.
  const bool fin = valid && Finite3(X[0], X[1], X[2]);
  for (int d = 0; d < 3; ++d) {
    lo[d][t] = fin ? X[d] : INFINITY;
    hi[d][t] = fin ? X[d] : -INFINITY;
  }
  // stable partition by charge class (0: q > 0, 1: q < 0, 2: q = 0)
  const int cls = h.w > 0.f ? 0 : (h.w < 0.f ? 1 : 2);
  const int lane = t & 63, wv = t >> 6;
  const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
  const uint64_t bal[3] = {__ballot(cls == 0), __ballot(cls == 1), __ballot(cls == 2)};
  if (t == 0) non_unit = 0;
  if (lane == 0)
    for (int c = 0; c < 3; ++c) cls_cnt[wv][c] = __popcll(bal[c]);
  __syncthreads();
  if (h.w != 0.f && fabsf(h.w) != 1.f) non_unit = 1;
  int start = 0;  // first slot of this lane's class within the tile
  for (int c = 0; c < cls; ++c)
    for (int w = 0; w < kTile / 64; ++w) start += cls_cnt[w][c];
  for (int w = 0; w < wv; ++w) start += cls_cnt[w][cls];
  const int slot = first + start + __popcll(bal[cls] & below);
  for (int w = kTile / 2; w > 0; w >>= 1) {
    if (t < w)
      for (int d = 0; d < 3; ++d) {
        lo[d][t] = fmin(lo[d][t], lo[d][t + w]);
        hi[d][t] = fmax(hi[d][t], hi[d][t + w]);
      }
    __syncthreads();
  }
  const bool any = lo[0][0] <= hi[0][0];
  const double c[3] = {any ? 0.5 * (lo[0][0] + hi[0][0]) : 0.0, any ? 0.5 * (lo[1][0] + hi[1][0]) : 0.0,
                       any ? 0.5 * (lo[2][0] + hi[2][0]) : 0.0};
  if (valid) perm[slot] = i;  // every lane read its entry before the first barrier
  shl[2 * slot] = h;
  shl[2 * slot + 1] = l;
  sloc[slot] = fin ? make_float4((float)(X[0] - c[0]), (float)(X[1] - c[1]), (float)(X[2] - c[2]), h.w)
                     : make_float4(0.f, 0.f, 0.f, 0.f);
  if (t == 0) {
    TileRec r;
    r.cx = c[0];
    r.cy = c[1];
    r.cz = c[2];
    r.hx = any ? 0.5 * (hi[0][0] - lo[0][0]) : 0.0;
    r.hy = any ? 0.5 * (hi[1][0] - lo[1][0]) : 0.0;
    r.hz = any ? 0.5 * (hi[2][0] - lo[2][0]) : 0.0;
    r.npos = cls_cnt[0][0] + cls_cnt[1][0];
    r.nneg = cls_cnt[0][1] + cls_cnt[1][1];
    r.unit = non_unit ? 0 : 1;
    r.pad = 0;
    tiles[tile] = r;
  }
}

// Unit-charge pairs with the sign folded into the difference: POS gives q_j = +e, else -e.
template <bool POS>
__device__ __forceinline__ void PairF32Unit(const float4 th, const float4 tl, const float4 sh, const float4 sl,
                                            float& px, float& py, float& pz) {
  const float dx = POS ? (th.x - sh.x) + (tl.x - sl.x) : (sh.x - th.x) + (sl.x - tl.x);
  const float dy = POS ? (th.y - sh.y) + (tl.y - sl.y) : (sh.y - th.y) + (sl.y - tl.y);
  const float dz = POS ? (th.z - sh.z) + (tl.z - sl.z) : (sh.z - th.z) + (sl.z - tl.z);
  const float r2 = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, kEps2Units)));
  const float ri = __builtin_amdgcn_rsqf(r2);
  const float c = ri * (ri * ri);
  px = fmaf(c, dx, px);
  py = fmaf(c, dy, py);
  pz = fmaf(c, dz, pz);
}
template <bool POS>
__device__ __forceinline__ void PairLocalUnit(const float tx, const float ty, const float tz, const float4 s,
                                              float& px, float& py, float& pz) {
  const float dx = POS ? tx - s.x : s.x - tx, dy = POS ? ty - s.y : s.y - ty, dz = POS ? tz - s.z : s.z - tz;
  const float r2 = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, kEps2Units)));
  const float ri = __builtin_amdgcn_rsqf(r2);
  const float c = ri * (ri * ri);
  px = fmaf(c, dx, px);
  py = fmaf(c, dy, py);
  pz = fmaf(c, dz, pz);
}

// A group of 8 (or 16) sources at a multiple of 8 entries from a 128-entry tile start: 128-byte
// aligned, which lets the compiler fetch it with wide scalar loads.
__device__ __forceinline__ const float4* Group8(const float4* tile, int j0) {
  return static_cast<const float4*>(__builtin_assume_aligned(tile + j0, 128));
}

// Eight sources as two 16-dword vectors: one wide scalar load each, even where the pair loop
// leaves the charge component unused (which would otherwise split the fetch per source).
typedef float f32x16v __attribute__((ext_vector_type(16)));
struct Src8 {
  f32x16v a, b;
  __device__ __forceinline__ float4 operator[](int u) const {
    const f32x16v& v = u < 4 ? a : b;
    const int o = 4 * (u & 3);
    return make_float4(v[o], v[o + 1], v[o + 2], v[o + 3]);
  }
};
__device__ __forceinline__ Src8 LoadSrc8(const float4* tile, int j0) {
  const f32x16v* p = reinterpret_cast<const f32x16v*>(Group8(tile, j0));
  return Src8{p[0], p[1]};
}

// Far-tile pair: local coordinates (one subtraction per axis), same softening as PairF32.
__device__ __forceinline__ void PairLocal(const float tx, const float ty, const float tz, const float4 s, float& px,
                                          float& py, float& pz) {
  const float dx = tx - s.x, dy = ty - s.y, dz = tz - s.z;
  const float r2 = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, kEps2Units)));
  const float ri = __builtin_amdgcn_rsqf(r2);
  const float c = (s.w * ri) * (ri * ri);
  px = fmaf(c, dx, px);
  py = fmaf(c, dy, py);
  pz = fmaf(c, dz, pz);
}

// Expanded-form pair of a unit source of one sign: s = (-2 s_x, -2 s_y, -2 s_z, |s|^2 + eps^2) local
// to the tile centre (a staged record, read from LDS), t local to it too with tt = |t|^2; adds
// c = 1/r^3 to sc and c (-2s) to (sx, sy, sz).
__device__ __forceinline__ void PairExpanded(const float tx, const float ty, const float tz, const float tt,
                                             const float4 s, float& sc, float& sx, float& sy, float& sz) {
  const float r2 = fmaf(s.x, tx, fmaf(s.y, ty, fmaf(s.z, tz, tt + s.w)));
  const float ri = __builtin_amdgcn_rsqf(r2);
  const float c = ri * (ri * ri);
  sc += c;
  sx = fmaf(c, s.x, sx);
  sy = fmaf(c, s.y, sy);
  sz = fmaf(c, s.z, sz);
}

// Box of a 256-target block: the union of its two tiles (an empty second tile is left out).
struct BlockBox {
  double bc[3], bh[3], bhmax;
};

// Stages the kTile local records of one source tile as expanded-form records (-2s, |s|^2 + eps^2).
// Called by the whole workgroup; the first barrier keeps every wave's reads of the previously staged
// tile ahead of the overwrite.
__device__ __forceinline__ void StageExpanded(const float4* __restrict__ sp, float4* __restrict__ stage) {
  __syncthreads();
  if (threadIdx.x < kTile) {
    const float4 q = sp[threadIdx.x];
    stage[threadIdx.x] =
        make_float4(-2.f * q.x, -2.f * q.y, -2.f * q.z, fmaf(q.x, q.x, fmaf(q.y, q.y, fmaf(q.z, q.z, kEps2Units))));
  }
  __syncthreads();
}

// 1/r of one expanded-form record at target t (tt = |t|^2), both local to the source tile centre.
__device__ __forceinline__ float InvRExpanded(float tx, float ty, float tz, float tt, const float4 s) {
  return __builtin_amdgcn_rsqf(fmaf(s.x, tx, fmaf(s.y, ty, fmaf(s.z, tz, tt + s.w))));
}

__device__ inline BlockBox MakeBlockBox(const TileRec& A, const TileRec& B, bool b_empty) {
  double blo[3] = {A.cx - A.hx, A.cy - A.hy, A.cz - A.hz}, bhi[3] = {A.cx + A.hx, A.cy + A.hy, A.cz + A.hz};
  if (!b_empty) {
    blo[0] = fmin(blo[0], B.cx - B.hx); bhi[0] = fmax(bhi[0], B.cx + B.hx);
    blo[1] = fmin(blo[1], B.cy - B.hy); bhi[1] = fmax(bhi[1], B.cy + B.hy);
    blo[2] = fmin(blo[2], B.cz - B.hz); bhi[2] = fmax(bhi[2], B.cz + B.hz);
  }
  BlockBox bb;
  for (int d = 0; d < 3; ++d) {
    bb.bc[d] = 0.5 * (blo[d] + bhi[d]);
    bb.bh[d] = 0.5 * (bhi[d] - blo[d]);
  }
  bb.bhmax = fmax(fmax(bb.bh[0], bb.bh[1]), bb.bh[2]);
  return bb;
}

// Adds the sum over source tile R (entries from tile_first in shl / sloc) for one target of the
// block: hi/lo coordinates th, tl and the coordinates local to the block centre (lx, ly, lz).
__device__ __forceinline__ void TileSum(const TileRec& R, int tile_first, const BlockBox& bb, const float4 th,
                                        const float4 tl, float lx, float ly, float lz, const float4* __restrict__ shl,
                                        const float4* __restrict__ sloc, float4* __restrict__ stage, float& px,
                                        float& py, float& pz) {
  const double gx = fmax(0.0, fabs(R.cx - bb.bc[0]) - R.hx - bb.bh[0]);
  const double gy = fmax(0.0, fabs(R.cy - bb.bc[1]) - R.hy - bb.bh[1]);
  const double gz = fmax(0.0, fabs(R.cz - bb.bc[2]) - R.hz - bb.bh[2]);
  const double size = fmax(fmax(fmax(R.hx, R.hy), R.hz), bb.bhmax);
  const double lim = 2.0 * kNearGap * size;
  const double gap2 = gx * gx + gy * gy + gz * gz;
  const int near = __builtin_amdgcn_readfirstlane(gap2 < lim * lim ? 1 : 0);
  // Expanded form only across a positive gap: then r >= gap > 0 for every pair (no self or coincident
  // pair, whose r^2 would cancel to rounding noise); degenerate (zero-size) boxes at zero gap keep the
  // local form.
  const int expand =
      __builtin_amdgcn_readfirstlane(gap2 > 0.0 && gap2 >= (kExpandGap * size) * (kExpandGap * size) ? 1 : 0);
  // Unit tiles: groups of 8 sources; whole groups of one sign fold the sign into the difference,
  // the (at most two) groups straddling a sign boundary use the charge-multiplied form.
  const int unit = __builtin_amdgcn_readfirstlane(R.unit);
  const int np = __builtin_amdgcn_readfirstlane(R.npos);
  const int nq = __builtin_amdgcn_readfirstlane(R.npos + R.nneg);
  const int ga = np & ~7, gb = (np + 7) & ~7;                  // [0,ga) +e, [ga,gb) mixed
  const int gc = gb + (nq > gb ? ((nq - gb) & ~7) : 0);          // [gb,gc) -e
  const int gd = nq > gc ? (nq + 7) & ~7 : gc;                   // [gc,gd) mixed, then q = 0
  if (near) {
    // Hi/lo sources in groups of 4 (32 SGPRs per group): groups of 8 need 64 SGPRs, which at
    // 8 waves per SIMD spill to VGPR lanes.
    const float4* sp = shl + 2 * tile_first;
    if (unit) {
      for (int j0 = 0; j0 < ga; j0 += 4) {
        const Src8 g = LoadSrc8(sp, 2 * j0);
#pragma unroll
        for (int u = 0; u < 4; ++u) PairF32Unit<true>(th, tl, g[2 * u], g[2 * u + 1], px, py, pz);
      }
      for (int j0 = ga; j0 < gb; j0 += 4) {
        const Src8 g = LoadSrc8(sp, 2 * j0);
#pragma unroll
        for (int u = 0; u < 4; ++u) PairF32(th, tl, g[2 * u], g[2 * u + 1], px, py, pz);
      }
      for (int j0 = gb; j0 < gc; j0 += 4) {
        const Src8 g = LoadSrc8(sp, 2 * j0);
#pragma unroll
        for (int u = 0; u < 4; ++u) PairF32Unit<false>(th, tl, g[2 * u], g[2 * u + 1], px, py, pz);
      }
      for (int j0 = gc; j0 < gd; j0 += 4) {
        const Src8 g = LoadSrc8(sp, 2 * j0);
#pragma unroll
        for (int u = 0; u < 4; ++u) PairF32(th, tl, g[2 * u], g[2 * u + 1], px, py, pz);
      }
    } else {
#pragma unroll 8
      for (int j = 0; j < kTile; ++j) PairF32(th, tl, sp[2 * j], sp[2 * j + 1], px, py, pz);
    }
  } else {
    const float tx = lx + (float)(bb.bc[0] - R.cx);
    const float ty = ly + (float)(bb.bc[1] - R.cy);
    const float tz = lz + (float)(bb.bc[2] - R.cz);
    const float4* sp = sloc + tile_first;
    if (unit && expand) {
      // Expanded form over the staged records, by sign: the positive sources [0, np), then the
      // negative ones [np, nq), each folded into F as +-(t sum c + 0.5 sum c (-2s)).  The whole
      // workgroup takes this branch for this tile (the classification depends only on the block and
      // the tile), so it stages the tile's kTile records together.
      StageExpanded(sp, stage);
      const float tt = fmaf(tx, tx, fmaf(ty, ty, tz * tz));
      float sc = 0.f, sx = 0.f, sy = 0.f, sz = 0.f;
// unrolled 16 deep: 1.155 against 1.167 ms per launch unrolled 8 (4: 1.180; 32 is slower too)
#pragma unroll 16
      for (int j = 0; j < np; ++j) PairExpanded(tx, ty, tz, tt, stage[j], sc, sx, sy, sz);
      px = fmaf(tx, sc, fmaf(0.5f, sx, px));
      py = fmaf(ty, sc, fmaf(0.5f, sy, py));
      pz = fmaf(tz, sc, fmaf(0.5f, sz, pz));
      sc = sx = sy = sz = 0.f;
#pragma unroll 16
      for (int j = np; j < nq; ++j) PairExpanded(tx, ty, tz, tt, stage[j], sc, sx, sy, sz);
      px = fmaf(-tx, sc, fmaf(-0.5f, sx, px));
      py = fmaf(-ty, sc, fmaf(-0.5f, sy, py));
      pz = fmaf(-tz, sc, fmaf(-0.5f, sz, pz));
    } else if (unit) {
      for (int j0 = 0; j0 < ga; j0 += 8) {
        const Src8 g = LoadSrc8(sp, j0);
#pragma unroll
        for (int u = 0; u < 8; ++u) PairLocalUnit<true>(tx, ty, tz, g[u], px, py, pz);
      }
      for (int j0 = ga; j0 < gb; j0 += 8) {
        const Src8 g = LoadSrc8(sp, j0);
#pragma unroll
        for (int u = 0; u < 8; ++u) PairLocal(tx, ty, tz, g[u], px, py, pz);
      }
      for (int j0 = gb; j0 < gc; j0 += 8) {
        const Src8 g = LoadSrc8(sp, j0);
#pragma unroll
        for (int u = 0; u < 8; ++u) PairLocalUnit<false>(tx, ty, tz, g[u], px, py, pz);
      }
      for (int j0 = gc; j0 < gd; j0 += 8) {
        const Src8 g = LoadSrc8(sp, j0);
#pragma unroll
        for (int u = 0; u < 8; ++u) PairLocal(tx, ty, tz, g[u], px, py, pz);
      }
    } else {
#pragma unroll 8
      for (int j = 0; j < kTile; ++j) PairLocal(tx, ty, tz, sp[j], px, py, pz);
    }
  }
}

}  // namespace am
