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

#include "am_internal.hpp"

namespace am {

constexpr int kAllPairsThreads = 256;
constexpr int kTile = 128;  // sources per FP32 partial sum
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
                                     const double* __restrict__ q, const uint8_t* __restrict__ flags,
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
  const bool live = !(flags[i] & kFlagAnnihilated) && q[i] != 0.0;
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
  out[i] = make_double4(x[i], y[i], z[i], (flags[i] & kFlagAnnihilated) ? 0.0 : q[i]);
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
      const float w = (jt + k == i) ? 0.f : sh.w;  // the self term q_i / eps would swamp FP32
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
// Particles are kept in a Morton order (rebuilt every kResortInterval calls); sources form tiles
// of 128 with an FP64 bounding box.  For a (256-target block, source tile) pair whose boxes are
// close (gap < 0.1 x the larger box, which includes the block's own tiles) the pair math is the
// hi/lo form above.  For the other ~90% of tile pairs both sides use coordinates local to their
// box centre: the target coordinate is shifted once per tile by the FP32-rounded centre offset,
// then dx = tx'' - sx' costs one instruction instead of three.  Its rounding error is
// ~2^-24 (|dx| + box size), i.e. <~1e-6 relative for a far pair, below the FP32 accumulation
// error of the tile sums (measured on the all-pairs workload's cloud: same global error as
// hi/lo everywhere, 2.4e-7).
constexpr int kResortInterval = 16;
constexpr double kNearGap = 0.1;  // near if box gap < kNearGap * (larger half extent) * 2

struct TileRec {  // sorted source tile, units of 2^-20 m
  double cx, cy, cz;
  double hx, hy, hz;
  int npos, nneg;  // the tile holds npos charges > 0, then nneg < 0, then the zero charges
  int unit;        // every non-zero charge is exactly +-e
  int pad;
};

__global__ void IotaKernel(int64_t n, uint32_t* __restrict__ v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = (uint32_t)i;
}

__device__ inline uint32_t Spread10(uint32_t v) {
  v &= 1023u;
  v = (v | (v << 16)) & 0x030000FFu;
  v = (v | (v << 8)) & 0x0300F00Fu;
  v = (v | (v << 4)) & 0x030C30C3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
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

// One workgroup per 128-source tile: sorted hi/lo mirror, tile-local coordinates and the box.
// The tile's slice of the permutation is reordered stably by charge sign (q > 0, q < 0, q = 0), so
// the pair loops can fold a unit charge's sign into the direction of the coordinate difference.
// Padding entries (k >= n) have q = 0 and keep their positions.
__global__ __launch_bounds__(kTile) void TilePrepKernel(int n, uint32_t* __restrict__ perm,
                                                        const double* __restrict__ x, const double* __restrict__ y,
                                                        const double* __restrict__ z, const float4* __restrict__ src,
                                                        float4* __restrict__ shl, float4* __restrict__ sloc,
                                                        TileRec* __restrict__ tiles) {
  __shared__ double lo[3][kTile], hi[3][kTile];
  __shared__ int cls_cnt[kTile / 64][3];
  __shared__ int non_unit;
  const int t = threadIdx.x;
  const int k = blockIdx.x * kTile + t;
  const bool valid = k < n;
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
  for (int d = 0; d < 3; ++d) {
    lo[d][t] = valid ? X[d] : INFINITY;
    hi[d][t] = valid ? X[d] : -INFINITY;
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
  const int slot = blockIdx.x * kTile + start + __popcll(bal[cls] & below);
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
  sloc[slot] = valid ? make_float4((float)(X[0] - c[0]), (float)(X[1] - c[1]), (float)(X[2] - c[2]), h.w)
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
    tiles[blockIdx.x] = r;
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
// Eight hi/lo source pairs (16 float4) as four 16-dword vectors.
struct SrcHL8 {
  Src8 h0, h1;  // entries 0..7 and 8..15 of the interleaved hi/lo array
  __device__ __forceinline__ float4 hi(int u) const { return u < 4 ? h0[2 * u] : h1[2 * (u - 4)]; }
  __device__ __forceinline__ float4 lo(int u) const { return u < 4 ? h0[2 * u + 1] : h1[2 * (u - 4) + 1]; }
};
__device__ __forceinline__ SrcHL8 LoadSrcHL8(const float4* tile, int j0) {
  return SrcHL8{LoadSrc8(tile, 2 * j0), LoadSrc8(tile, 2 * j0 + 8)};
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

__global__ __launch_bounds__(kAllPairsThreads) void AllPairsSortedKernel(
    int n, int n_pad, int tiles_per_chunk, const uint32_t* __restrict__ perm, const double* __restrict__ x,
    const double* __restrict__ y, const double* __restrict__ z, const float4* __restrict__ shl,
    const float4* __restrict__ sloc, const TileRec* __restrict__ tiles, double* __restrict__ part) {
  const int k = blockIdx.x * kAllPairsThreads + threadIdx.x;
  const TileRec A = tiles[2 * blockIdx.x], B = tiles[2 * blockIdx.x + 1];
  // block box = union of its two tiles (an empty tile has zero extent at the origin)
  const bool b_empty = 2 * blockIdx.x * kTile + kTile >= n;
  double blo[3] = {A.cx - A.hx, A.cy - A.hy, A.cz - A.hz}, bhi[3] = {A.cx + A.hx, A.cy + A.hy, A.cz + A.hz};
  if (!b_empty) {
    blo[0] = fmin(blo[0], B.cx - B.hx); bhi[0] = fmax(bhi[0], B.cx + B.hx);
    blo[1] = fmin(blo[1], B.cy - B.hy); bhi[1] = fmax(bhi[1], B.cy + B.hy);
    blo[2] = fmin(blo[2], B.cz - B.hz); bhi[2] = fmax(bhi[2], B.cz + B.hz);
  }
  const double bc[3] = {0.5 * (blo[0] + bhi[0]), 0.5 * (blo[1] + bhi[1]), 0.5 * (blo[2] + bhi[2])};
  const double bh[3] = {0.5 * (bhi[0] - blo[0]), 0.5 * (bhi[1] - blo[1]), 0.5 * (bhi[2] - blo[2])};
  const double bhmax = fmax(fmax(bh[0], bh[1]), bh[2]);
  const float4 th = shl[2 * k], tl = shl[2 * k + 1];
  float lx = 0.f, ly = 0.f, lz = 0.f;
  if (k < n) {
    const uint32_t i = perm[k];
    lx = (float)(x[i] * kUnitScale - bc[0]);
    ly = (float)(y[i] * kUnitScale - bc[1]);
    lz = (float)(z[i] * kUnitScale - bc[2]);
  }
  const int n_tiles = n_pad / kTile;
  const int t0 = blockIdx.y * tiles_per_chunk;
  const int t1 = min(n_tiles, t0 + tiles_per_chunk);
  double ax = 0.0, ay = 0.0, az = 0.0;
  for (int T = t0; T < t1; ++T) {
    const TileRec R = tiles[T];
    const double gx = fmax(0.0, fabs(R.cx - bc[0]) - R.hx - bh[0]);
    const double gy = fmax(0.0, fabs(R.cy - bc[1]) - R.hy - bh[1]);
    const double gz = fmax(0.0, fabs(R.cz - bc[2]) - R.hz - bh[2]);
    const double size = fmax(fmax(fmax(R.hx, R.hy), R.hz), bhmax);
    const double lim = 2.0 * kNearGap * size;
    const int near = __builtin_amdgcn_readfirstlane((gx * gx + gy * gy + gz * gz) < lim * lim ? 1 : 0);
    float px = 0.f, py = 0.f, pz = 0.f;
    // Unit tiles: groups of 8 sources; whole groups of one sign fold the sign into the difference,
    // the (at most two) groups straddling a sign boundary use the charge-multiplied form.
    const int unit = __builtin_amdgcn_readfirstlane(R.unit);
    const int np = __builtin_amdgcn_readfirstlane(R.npos);
    const int nq = __builtin_amdgcn_readfirstlane(R.npos + R.nneg);
    const int ga = np & ~7, gb = (np + 7) & ~7;                  // [0,ga) +e, [ga,gb) mixed
    const int gc = gb + (nq > gb ? ((nq - gb) & ~7) : 0);          // [gb,gc) -e
    const int gd = nq > gc ? (nq + 7) & ~7 : gc;                   // [gc,gd) mixed, then q = 0
    if (near) {
      const float4* sp = shl + 2 * T * kTile;
      if (unit) {
        for (int j0 = 0; j0 < ga; j0 += 8) {
          const SrcHL8 g = LoadSrcHL8(sp, j0);
#pragma unroll
          for (int u = 0; u < 8; ++u) PairF32Unit<true>(th, tl, g.hi(u), g.lo(u), px, py, pz);
        }
        for (int j0 = ga; j0 < gb; j0 += 8) {
          const SrcHL8 g = LoadSrcHL8(sp, j0);
#pragma unroll
          for (int u = 0; u < 8; ++u) PairF32(th, tl, g.hi(u), g.lo(u), px, py, pz);
        }
        for (int j0 = gb; j0 < gc; j0 += 8) {
          const SrcHL8 g = LoadSrcHL8(sp, j0);
#pragma unroll
          for (int u = 0; u < 8; ++u) PairF32Unit<false>(th, tl, g.hi(u), g.lo(u), px, py, pz);
        }
        for (int j0 = gc; j0 < gd; j0 += 8) {
          const SrcHL8 g = LoadSrcHL8(sp, j0);
#pragma unroll
          for (int u = 0; u < 8; ++u) PairF32(th, tl, g.hi(u), g.lo(u), px, py, pz);
        }
      } else {
#pragma unroll 8
        for (int j = 0; j < kTile; ++j) PairF32(th, tl, sp[2 * j], sp[2 * j + 1], px, py, pz);
      }
    } else {
      const float tx = lx + (float)(bc[0] - R.cx);
      const float ty = ly + (float)(bc[1] - R.cy);
      const float tz = lz + (float)(bc[2] - R.cz);
      const float4* sp = sloc + T * kTile;
      if (unit) {
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
    ax += (double)px;
    ay += (double)py;
    az += (double)pz;
  }
  double* out = part + (size_t)blockIdx.y * 3 * n_pad;
  out[k] = ax;
  out[n_pad + k] = ay;
  out[2 * (size_t)n_pad + k] = az;
}

__global__ void AllPairsSortedReduceKernel(int n, int n_pad, int chunks, const double* __restrict__ part,
                                           const uint32_t* __restrict__ perm, const double* __restrict__ q,
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
  const bool live = !(flags[i] & kFlagAnnihilated) && q[i] != 0.0;
  const double scale = live ? kCoulomb * q[i] * kElementaryCharge * kInvUnit2 : 0.0;
  fx[i] = scale * sx;
  fy[i] = scale * sy;
  fz[i] = scale * sz;
}

int GatheredBBox(am_ctx* ctx, const uint32_t* idx, int m, const double* x, const double* y, const double* z,
                 double* box_dev, const char* tag);

static int ChunkCount(am_ctx* ctx, int target_blocks, int n_pad);

int AllPairsForcesF32Sorted(am_store* s) {
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
    if ((rc = RadixSortPairs(ctx, (uint64_t*)keys, s->perm, n, 30))) return rc;
    s->perm_n = n;
    s->perm_age = 0;
  }
  void *shl, *sloc, *tiles, *part;
  const int n_tiles = n_pad / kTile;
  if ((rc = Scratch(ctx, "ap_shl", (size_t)n_pad * 2 * sizeof(float4), &shl))) return rc;
  if ((rc = Scratch(ctx, "ap_sloc", (size_t)n_pad * sizeof(float4), &sloc))) return rc;
  if ((rc = Scratch(ctx, "ap_tiles", (size_t)n_tiles * sizeof(TileRec), &tiles))) return rc;
  hipLaunchKernelGGL(TilePrepKernel, dim3(n_tiles), dim3(kTile), 0, st, n, s->perm, s->x, s->y,
                     s->z, (const float4*)s->src, (float4*)shl, (float4*)sloc, (TileRec*)tiles);
  AM_HIP(hipGetLastError());
  const int tblocks = n_pad / kAllPairsThreads;
  const int chunks = ChunkCount(ctx, tblocks, n_pad);
  const int tpc = (n_tiles + chunks - 1) / chunks;
  const int nch = (n_tiles + tpc - 1) / tpc;
  if ((rc = Scratch(ctx, "allpairs_part", (size_t)nch * 3 * n_pad * sizeof(double), &part))) return rc;
  int slot;
  TimerBegin(ctx, "allpairs_f32", &slot);
  hipLaunchKernelGGL(AllPairsSortedKernel, dim3(tblocks, nch), dim3(kAllPairsThreads), 0, st, n, n_pad, tpc,
                     (const uint32_t*)s->perm, s->x, s->y, s->z, (const float4*)shl, (const float4*)sloc,
                     (const TileRec*)tiles, (double*)part);
  TimerEnd(ctx, slot);
  AM_HIP(hipGetLastError());
  hipLaunchKernelGGL(AllPairsSortedReduceKernel, dim3(Blocks(n, 256)), dim3(256), 0, st, n, n_pad, nch,
                     (const double*)part, (const uint32_t*)s->perm, s->q, s->flags, s->fx, s->fy, s->fz);
  AM_HIP(hipGetLastError());
  return AM_OK;
}

static int ChunkCount(am_ctx* ctx, int target_blocks, int n_pad) {
  const int cus = ctx->num_cus > 0 ? ctx->num_cus : 256;
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
  const int chunks = ChunkCount(ctx, tblocks, n_pad);
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
                     (const double*)part, s->q, s->flags, s->fx, s->fy, s->fz);
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
  const int chunks = ChunkCount(ctx, tblocks, n_pad);
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
  *si_scale = kElementaryCharge * kElementaryCharge * kUnitScale;  // e^2/unit -> C^2/m
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
