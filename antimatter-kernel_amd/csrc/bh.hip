// Barnes-Hut Coulomb solver (src/interaction/barnes-hut.cc:30-299) on the device.
//
// Tree build (the reference builds it recursively and serially every call, barnes-hut.cc:38-164):
//   1. gather the alive particles (ascending index, as ComputeBarnesHutCoulombForces does);
//   2. bounding cube exactly as Build (:38-66): centre = bbox mid, half = 0.5 ext (1 + 1e-9);
//   3. 63-bit Morton key per particle by descending 21 levels with the reference's own FP64 centre
//      arithmetic (child centre = parent centre +- half/2, octant = (x >= c) | (y >= c) << 1 |
//      (z >= c) << 2), so every octant decision is the reference's comparison, bit for bit;
//   4. stable LSD radix sort of (key, index): the order of the reference's recursive stable
//      octant partition;
//   5. level-synchronous count-capped split (split iff count > leaf capacity, depth < 21), child
//      ranges by binary search on the sorted keys, children in octant order 0..7;
//   6. leaves re-sorted by particle index (the reference keeps index order inside a leaf), so the
//      permutation equals the reference's order_;
//   7. dual monopoles bottom-up with the reference's accumulation order (:138-163);
//   8. DFS preorder = nodes sorted by (first particle, depth); skip pointer = first node whose
//      first particle is at or past the end of the subtree's range.
// Walk: one lane per target, targets in Morton order, the wave visits nodes in lockstep (the
// node is wave-uniform, so node and leaf-particle loads are scalar loads); each lane applies its
// own MAC (2h)^2 < theta^2 |x - c|^2 in FP64 without contraction, so the set of interactions of
// every particle is the reference's.  Pair math is FP32 on hi/lo coordinates (AM_FP32) or the
// reference's FP64 AddPointCharge (AM_FP64, visit order = the reference recursion order).
// Limitation: the tree depth is capped at 21 levels (63-bit keys); the reference splits to 32.
// A depth-21 node with more than leaf-capacity particles (particles within ~cube/2^21 of each
// other) becomes a leaf here; am_bh_tree_stats reports how many such nodes a build produced.
#include <algorithm>
#include <cmath>
#include <vector>

#include "bh.hpp"
#include "pairtile.hpp"

namespace am {

// Root cube of BarnesHutTree::Build (barnes-hut.cc:47-62) from the gathered bounding box.
__global__ void BHRootKernel(const double* __restrict__ box, RootBox* __restrict__ root) {
  const double lo[3] = {box[0], box[1], box[2]}, hi[3] = {box[3], box[4], box[5]};
  RootBox r;
  r.cx = 0.5 * (lo[0] + hi[0]);
  r.cy = 0.5 * (lo[1] + hi[1]);
  r.cz = 0.5 * (lo[2] + hi[2]);
  const double ext = fmax(fmax(hi[0] - lo[0], hi[1] - lo[1]), hi[2] - lo[2]);
  double half = 0.5 * ext * (1.0 + 1e-9);
  if (half < 1e-30) half = 1e-30;
  r.half = half;
  *root = r;
}

struct Moments {  // BFS order
  double* qp;
  double* qn;
  double* cpx;
  double* cpy;
  double* cpz;
  double* cnx;
  double* cny;
  double* cnz;
};

// Tree-ordered FP64 copy of (x, y, z, q), one record per sorted position, so that the leaf sums of
// the moment passes read their particles contiguously instead of gathering them level by level.
// xq holds the same records by particle index (written by the Morton pass).
__global__ void GatherTreeF64Kernel(int m, const uint32_t* __restrict__ order, const double4* __restrict__ xq,
                                    double4* __restrict__ sxq) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= m) return;
  sxq[k] = xq[order[k]];
}

// Dual monopoles (barnes-hut.cc:138-163): leaves sum q x over their particles in order; internal
// nodes combine their children's (charge, centre) in octant order; centre = moment * (1/charge).
// sxq is the tree-ordered (x, y, z, q) of GatherTreeF64Kernel.
__global__ void MomentKernel(int n0, int n1, NodeArrays nd, Moments mo, const double4* __restrict__ sxq) {
  const int node = n0 + blockIdx.x * blockDim.x + threadIdx.x;
  if (node >= n1) return;
  double qp = 0.0, qn = 0.0;
  double mpx = 0.0, mpy = 0.0, mpz = 0.0, mnx = 0.0, mny = 0.0, mnz = 0.0;
  if (nd.leaf[node]) {
    for (int k = nd.b[node]; k < nd.e[node]; ++k) {
      const double4 pj = sxq[k];
      const double qj = pj.w;
      if (qj > 0.0) {
        qp = qp + qj;
        mpx = mpx + pj.x * qj;
        mpy = mpy + pj.y * qj;
        mpz = mpz + pj.z * qj;
      } else if (qj < 0.0) {
        const double a = -qj;
        qn = qn + a;
        mnx = mnx + pj.x * a;
        mny = mny + pj.y * a;
        mnz = mnz + pj.z * a;
      }
    }
  } else {
    const int c0 = nd.first_child[node], c1 = c0 + nd.nchild[node];
    for (int c = c0; c < c1; ++c) {
      const double a = mo.qp[c];
      if (a > 0.0) {
        qp = qp + a;
        mpx = mpx + mo.cpx[c] * a;
        mpy = mpy + mo.cpy[c] * a;
        mpz = mpz + mo.cpz[c] * a;
      }
      const double bq = mo.qn[c];
      if (bq > 0.0) {
        qn = qn + bq;
        mnx = mnx + mo.cnx[c] * bq;
        mny = mny + mo.cny[c] * bq;
        mnz = mnz + mo.cnz[c] * bq;
      }
    }
  }
  mo.qp[node] = qp;
  mo.qn[node] = qn;
  if (qp > 0.0) {
    const double r = 1.0 / qp;
    mo.cpx[node] = mpx * r;
    mo.cpy[node] = mpy * r;
    mo.cpz[node] = mpz * r;
  } else {
    mo.cpx[node] = nd.cx[node];
    mo.cpy[node] = nd.cy[node];
    mo.cpz[node] = nd.cz[node];
  }
  if (qn > 0.0) {
    const double r = 1.0 / qn;
    mo.cnx[node] = mnx * r;
    mo.cny[node] = mny * r;
    mo.cnz[node] = mnz * r;
  } else {
    mo.cnx[node] = nd.cx[node];
    mo.cny[node] = nd.cy[node];
    mo.cnz[node] = nd.cz[node];
  }
}

// Walk-ready node record in DFS preorder.
// Second moments of each sign's charges about its charge centre, S = sum |q| (x - c)(x - c)^T
// (xx, yy, zz, xy, xz, yz), bottom-up with the parallel-axis shift; used by the learned
// far-field correction only.
struct SecondMoments {  // BFS order, 6 per node per sign
  double* sp;
  double* sn;
};

__global__ void SecondMomentKernel(int n0, int n1, NodeArrays nd, Moments mo, SecondMoments sm,
                                   const double4* __restrict__ sxq) {
  const int node = n0 + blockIdx.x * blockDim.x + threadIdx.x;
  if (node >= n1) return;
  double ap[6] = {0, 0, 0, 0, 0, 0}, an[6] = {0, 0, 0, 0, 0, 0};
  const double cp[3] = {mo.cpx[node], mo.cpy[node], mo.cpz[node]};
  const double cn[3] = {mo.cnx[node], mo.cny[node], mo.cnz[node]};
  auto add = [](double* a, double w, double dx, double dy, double dz) {
    a[0] += w * dx * dx;
    a[1] += w * dy * dy;
    a[2] += w * dz * dz;
    a[3] += w * dx * dy;
    a[4] += w * dx * dz;
    a[5] += w * dy * dz;
  };
  if (nd.leaf[node]) {
    for (int k = nd.b[node]; k < nd.e[node]; ++k) {
      const double4 pj = sxq[k];
      const double qj = pj.w;
      if (qj > 0.0) add(ap, qj, pj.x - cp[0], pj.y - cp[1], pj.z - cp[2]);
      else if (qj < 0.0) add(an, -qj, pj.x - cn[0], pj.y - cn[1], pj.z - cn[2]);
    }
  } else {
    const int c0 = nd.first_child[node], c1 = c0 + nd.nchild[node];
    for (int c = c0; c < c1; ++c) {
      if (mo.qp[c] > 0.0) {
        for (int k = 0; k < 6; ++k) ap[k] += sm.sp[6 * c + k];
        add(ap, mo.qp[c], mo.cpx[c] - cp[0], mo.cpy[c] - cp[1], mo.cpz[c] - cp[2]);
      }
      if (mo.qn[c] > 0.0) {
        for (int k = 0; k < 6; ++k) an[k] += sm.sn[6 * c + k];
        add(an, mo.qn[c], mo.cnx[c] - cn[0], mo.cny[c] - cn[1], mo.cnz[c] - cn[2]);
      }
    }
  }
  for (int k = 0; k < 6; ++k) {
    sm.sp[6 * node + k] = ap[k];
    sm.sn[6 * node + k] = an[k];
  }
}

__global__ void GatherPreorderKernel(int nn, NodeArrays nd, Moments mo, const uint64_t* __restrict__ skeys,
                                     const uint32_t* __restrict__ pre, WalkNode* __restrict__ out,
                                     WalkNodeAux* __restrict__ out32, SecondMoments sm,
                                     FeatNode* __restrict__ feat) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nn) return;
  const int node = (int)pre[p];
  if (feat) {
    FeatNode f;
    const double h = nd.half[node];
    const double cg[3] = {nd.cx[node], nd.cy[node], nd.cz[node]};
    const double cs[2][3] = {{mo.cpx[node], mo.cpy[node], mo.cpz[node]}, {mo.cnx[node], mo.cny[node], mo.cnz[node]}};
    const double Q[2] = {mo.qp[node], mo.qn[node]};
    const double* S[2] = {sm.sp + 6 * node, sm.sn + 6 * node};
    float* dst[2] = {f.fp, f.fn};
    for (int sgn = 0; sgn < 2; ++sgn) {
      for (int d = 0; d < 3; ++d) dst[sgn][d] = Q[sgn] > 0.0 ? (float)((cs[sgn][d] - cg[d]) / h) : 0.f;
      const double norm = Q[sgn] > 0.0 ? 1.0 / (Q[sgn] * h * h) : 0.0;
      for (int k = 0; k < 6; ++k) dst[sgn][3 + k] = (float)(S[sgn][k] * norm);
      dst[sgn][9] = (float)(Q[sgn] / kElementaryCharge);
    }
    f.h = (float)(h * kUnitScale);
    f.pad = 0.f;
    feat[p] = f;
  }
  WalkNode w;
  WalkNodeAux f;
  w.cx = nd.cx[node];
  w.cy = nd.cy[node];
  w.cz = nd.cz[node];
  w.half = nd.half[node];
  {  // FP32 form of the acceptance test's operands (units of 2^-20 m, exact power-of-two scaling)
    const double c[3] = {w.cx * kUnitScale, w.cy * kUnitScale, w.cz * kUnitScale};
    float hi[3], lo[3];
    for (int d = 0; d < 3; ++d) {
      hi[d] = (float)c[d];
      lo[d] = (float)(c[d] - (double)hi[d]);
    }
    w.chx = hi[0];
    w.chy = hi[1];
    w.chz = hi[2];
    w.clx = lo[0];
    w.cly = lo[1];
    w.clz = lo[2];
    const double wu = 2.0 * w.half * kUnitScale, W = wu * wu;
    w.w2hi = (float)(W * (1.0 + 0x1p-15));
    w.w2lo = (float)(W * (1.0 - 0x1p-15));
  }
  f.pqx = mo.cpx[node];
  f.pqy = mo.cpy[node];
  f.pqz = mo.cpz[node];
  w.qp = mo.qp[node];
  f.nqx = mo.cnx[node];
  f.nqy = mo.cny[node];
  f.nqz = mo.cnz[node];
  w.qn = mo.qn[node];
  w.b = nd.b[node];
  w.e = nd.e[node];
  // A twig is an internal node whose children are all leaves: a lane that opens it sums the
  // children's particles directly, in the children's order, which is what visiting them would do.
  const bool leaf = nd.leaf[node] != 0;
  uint32_t flags = (w.qp == 0.0 && w.qn == 0.0) ? kNodeEmpty : 0u;
  if (leaf) flags |= kNodeLeaf;
  if (w.qp > 0.0) flags |= kNodeHasPos;
  if (w.qn > 0.0) flags |= kNodeHasNeg;
  if (!leaf) {
    const int c0 = nd.first_child[node], c1 = c0 + nd.nchild[node];
    int all = c1 > c0 ? 1 : 0;
    for (int ch = c0; ch < c1; ++ch) all &= nd.leaf[ch] ? 1 : 0;
    if (all) flags |= kNodeTwig;
  }
  w.flags = flags;
  // skip = first preorder position whose key >= (e << 5)
  const uint64_t target = (uint64_t)w.e << 5;
  int lo = p + 1, hi = nn;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (skeys[mid] < target) lo = mid + 1;
    else hi = mid;
  }
  w.skip = lo;
  // FP32 monopoles; an empty charge class sits at the cell centre with charge 0, so the walk can
  // sum both monopoles of an accepted node without testing (a charge-0 pair adds exactly 0).
  const double qs = 1.0 / kElementaryCharge;
  {
    const bool has = w.qp > 0.0;
    const double xs = (has ? f.pqx : w.cx) * kUnitScale, ys = (has ? f.pqy : w.cy) * kUnitScale,
                 zs = (has ? f.pqz : w.cz) * kUnitScale;
    const float xh = (float)xs, yh = (float)ys, zh = (float)zs;
    f.p_hi = make_float4(xh, yh, zh, (float)(w.qp * qs));
    f.p_lo = make_float4((float)(xs - xh), (float)(ys - yh), (float)(zs - zh), 0.f);
    w.p_rel = make_float4((float)(xs - w.cx * kUnitScale), (float)(ys - w.cy * kUnitScale),
                          (float)(zs - w.cz * kUnitScale), (float)(w.qp * qs));
  }
  {
    const bool has = w.qn > 0.0;
    const double xs = (has ? f.nqx : w.cx) * kUnitScale, ys = (has ? f.nqy : w.cy) * kUnitScale,
                 zs = (has ? f.nqz : w.cz) * kUnitScale;
    const float xh = (float)xs, yh = (float)ys, zh = (float)zs;
    f.n_hi = make_float4(xh, yh, zh, (float)(-w.qn * qs));
    f.n_lo = make_float4((float)(xs - xh), (float)(ys - yh), (float)(zs - zh), 0.f);
    w.n_rel = make_float4((float)(xs - w.cx * kUnitScale), (float)(ys - w.cy * kUnitScale),
                          (float)(zs - w.cz * kUnitScale), (float)(-w.qn * qs));
  }
  out[p] = w;
  out32[p] = f;
}

// ---- walk -------------------------------------------------------------------------------------

// Minimum over the wave (every lane active), returned wave-uniform.  DPP inclusive min-scan:
// row_shr 1, 2, 4, 8 within each row of 16 lanes, then row_bcast 15 and 31 across rows, so lane 63
// ends with the minimum; lanes whose source is outside the row keep the identity.  Six dependent
// VALU steps instead of a chain of LDS permutes.
__device__ inline int WaveMin(int v) {
  constexpr int kId = 0x7fffffff;
  v = min(v, __builtin_amdgcn_update_dpp(kId, v, 0x111, 0xf, 0xf, false));  // row_shr:1
  v = min(v, __builtin_amdgcn_update_dpp(kId, v, 0x112, 0xf, 0xf, false));  // row_shr:2
  v = min(v, __builtin_amdgcn_update_dpp(kId, v, 0x114, 0xf, 0xf, false));  // row_shr:4
  v = min(v, __builtin_amdgcn_update_dpp(kId, v, 0x118, 0xf, 0xf, false));  // row_shr:8
  v = min(v, __builtin_amdgcn_update_dpp(kId, v, 0x142, 0xa, 0xf, false));  // row_bcast:15, rows 1 and 3
  v = min(v, __builtin_amdgcn_update_dpp(kId, v, 0x143, 0xc, 0xf, false));  // row_bcast:31, rows 2 and 3
  return __builtin_amdgcn_readlane(v, 63);
}

// AddPointCharge (barnes-hut.cc:19-27) without contraction.
__device__ __forceinline__ void PairF64Acc(double tx, double ty, double tz, double tq, double sx, double sy,
                                           double sz, double sq, double& ax, double& ay, double& az) {
  const double rx = __dsub_rn(tx, sx), ry = __dsub_rn(ty, sy), rz = __dsub_rn(tz, sz);
  const double d2 = Dot3(rx, ry, rz, rx, ry, rz);
  if (d2 < kMinSquaredDistance) return;
  const double inv = 1.0 / sqrt(d2);
  const double mag = __ddiv_rn(__dmul_rn(__dmul_rn(kCoulomb, tq), sq), d2);
  const double c = __dmul_rn(mag, inv);
  ax = __dadd_rn(ax, __dmul_rn(rx, c));
  ay = __dadd_rn(ay, __dmul_rn(ry, c));
  az = __dadd_rn(az, __dmul_rn(rz, c));
}

// FP32 potential of one source on hi/lo coordinates (e/unit); pairs under the reference's 1e-15 m
// floor (the target itself among them) are skipped, as barnes-hut.cc:244-247 skips them.
__device__ __forceinline__ float PotF32(const float4 th, const float4 tl, const float4 sh, const float4 sl) {
  const float dx = (th.x - sh.x) + (tl.x - sl.x);
  const float dy = (th.y - sh.y) + (tl.y - sl.y);
  const float dz = (th.z - sh.z) + (tl.z - sl.z);
  const float r2 = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, kEps2Units)));
  const float w = r2 < 2.f * kEps2Units ? 0.f : sh.w;
  return w * __builtin_amdgcn_rsqf(r2);
}

// FP32 potential of a source given by its offset s (units) from a cube centre, for a target at
// offset (dx, dy, dz) from the same centre; the same softening and self floor as PotF32.
__device__ __forceinline__ float PotLocal(float dx, float dy, float dz, const float4 s) {
  const float rx = dx - s.x, ry = dy - s.y, rz = dz - s.z;
  const float r2 = fmaf(rx, rx, fmaf(ry, ry, fmaf(rz, rz, kEps2Units)));
  const float w = r2 < 2.f * kEps2Units ? 0.f : s.w;
  return w * __builtin_amdgcn_rsqf(r2);
}

// A walk record's hot part (two wide scalar loads: the node index is wave-uniform) and its FP64 part.
typedef uint32_t u32x16v __attribute__((ext_vector_type(16)));
typedef uint32_t u32x8v __attribute__((ext_vector_type(8)));
typedef uint32_t u32x4v __attribute__((ext_vector_type(4)));
__device__ __forceinline__ WalkHot LoadHot(const WalkNode* __restrict__ nodes, int c) {
  const char* base = reinterpret_cast<const char*>(nodes + c);
  const u32x16v a = *reinterpret_cast<const u32x16v*>(__builtin_assume_aligned(base, 64));
  const u32x4v b = *reinterpret_cast<const u32x4v*>(__builtin_assume_aligned(base + 64, 16));
  WalkHot h;
  __builtin_memcpy(&h, &a, 64);
  __builtin_memcpy(reinterpret_cast<char*>(&h) + 64, &b, 16);
  return h;
}
__device__ __forceinline__ WalkCold LoadCold(const WalkNode* __restrict__ nodes, int c) {
  const char* base = reinterpret_cast<const char*>(nodes + c) + sizeof(WalkHot);
  const u32x8v a = *reinterpret_cast<const u32x8v*>(__builtin_assume_aligned(base, 16));
  const u32x4v b = *reinterpret_cast<const u32x4v*>(__builtin_assume_aligned(base + 32, 16));
  WalkCold w;
  __builtin_memcpy(&w, &a, 32);
  __builtin_memcpy(reinterpret_cast<char*>(&w) + 32, &b, 16);
  return w;
}

// kDirectGroup consecutive sources' hi/lo pairs from the tree-ordered FP32 mirror, starting at
// sorted position k0 (32-byte aligned): one wide scalar load per two sources.  The mirror has
// kSsrcPad zero entries past its last position, so a group never reads outside it.
constexpr int kDirectGroup = 2;  // sources per group of a walk's direct sum
constexpr int kSsrcPad = 8;
static_assert(kDirectGroup % 2 == 0 && kDirectGroup <= kSsrcPad, "direct-sum groups");
__device__ __forceinline__ void LoadSourceGroup(const float4* __restrict__ ssrc, int k0, float4* sh, float4* sl) {
#pragma unroll
  for (int g = 0; g < kDirectGroup / 2; ++g) {
    const u32x16v v = *reinterpret_cast<const u32x16v*>(__builtin_assume_aligned(ssrc + 2 * (k0 + 2 * g), 32));
    const char* w = reinterpret_cast<const char*>(&v);
    __builtin_memcpy(&sh[2 * g], w, 16);
    __builtin_memcpy(&sl[2 * g], w + 16, 16);
    __builtin_memcpy(&sh[2 * g + 1], w + 32, 16);
    __builtin_memcpy(&sl[2 * g + 1], w + 48, 16);
  }
}


// PairLocal into fresh accumulators (products instead of fused adds to zero).
__device__ __forceinline__ void PairLocalFirst(const float tx, const float ty, const float tz, const float4 s,
                                               float& px, float& py, float& pz) {
  const float dx = tx - s.x, dy = ty - s.y, dz = tz - s.z;
  const float r2 = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, kEps2Units)));
  const float ri = __builtin_amdgcn_rsqf(r2);
  const float c = (s.w * ri) * (ri * ri);
  px = c * dx;
  py = c * dy;
  pz = c * dz;
}

// The acceptance test (2h)^2 < theta^2 |x - c|^2 (barnes-hut.cc:201-204) in FP32.  With the target
// and the cube centre as hi/lo pairs in units of 2^-20 m, each FP32 difference is within
// 2^-23 |d| + delta of the exact one, delta = 2^-44 R (R bounds |coordinate| in units), and
// theta^2 |d|^2 then lies within 2^-19 relative of the FP64 side wherever |d| >= w_min / max(1, theta)
// (w_min = the smallest internal cube width, 2^-19 root half widths at depth <= 20) and
// 3.5 delta max(1, theta) / w_min <= 2^-20, i.e. R max(1, theta) <= 8 root half widths.  The node
// brackets W = (2h)^2 by w2hi >= W (1 + 2^-16) and w2lo <= W (1 - 2^-16), so theta^2 |d|^2 above
// w2hi accepts and below w2lo opens exactly as the FP64 test does; in between the walk evaluates the
// FP64 test.  FastMacRange checks that bound (and keeps the units^2 values in FP32's normal range).
__device__ inline bool FastMacRange(const RootBox* __restrict__ root, double theta2) {
  const RootBox r = *root;
  const double h = r.half * kUnitScale;
  const double R = (fmax(fmax(fabs(r.cx), fabs(r.cy)), fabs(r.cz)) + r.half) * kUnitScale;
  return h >= 0x1p-30 && h <= 0x1p60 && R * R * fmax(1.0, theta2) <= 64.0 * h * h;
}

// MODE 0: FP32 force; MODE 1: FP64 force; MODE 2: FP64 potential (q_i phi_i into out_x); MODE 3:
// FP32 potential ((q_i / e) phi_i in e^2/unit into out_x, FP32 per leaf, FP64 across nodes).  COUNT:
// interactions are counted into counts (a compile-time switch, so the plain walk carries no counting
// arithmetic in its loops).
// External evaluation points of am_barnes_hut_evaluate: position pos[3k..3k+2], charge q[k], and the
// source (particle index) left out as the point's "self", or 0xffffffff for none
// (BarnesHutTree::ComputeForce / ComputePotential, barnes-hut.h:28-45).
struct ExtTargets {
  const double* pos = nullptr;
  const double* q = nullptr;
  const uint32_t* self = nullptr;
};

template <int MODE, bool COUNT, bool EXT = false>
__global__ __launch_bounds__(256) void BHWalkKernel(int m, int nn, const WalkNode* __restrict__ nodes,
                                                     const WalkNodeAux* __restrict__ nodes32,
                                                     const RootBox* __restrict__ root,
                                                     const uint32_t* __restrict__ order,
                                                     const double* __restrict__ x, const double* __restrict__ y,
                                                     const double* __restrict__ z, const double* __restrict__ q,
                                                     const float4* __restrict__ src,
                                                     const float4* __restrict__ ssrc, double theta2,
                                                     double* __restrict__ out_x, double* __restrict__ out_y,
                                                     double* __restrict__ out_z, int64_t* __restrict__ counts,
                                                     const uint32_t* __restrict__ tsel, uint32_t* __restrict__ rec,
                                                     int rec_cap, uint32_t* __restrict__ rec_cnt,
                                                     ExtTargets ext = ExtTargets()) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = k < m;
  const int kc = valid ? k : m - 1;
  uint32_t i;
  int ks = -1;  // the target's position in the tree order (store targets)
  double px, py, pz, qi;
  float4 th, tl;
  bool want;
  if (EXT) {  // hi/lo coordinates of the point as the source mirror builds them
    i = ext.self[kc];
    px = ext.pos[3 * kc];
    py = ext.pos[3 * kc + 1];
    pz = ext.pos[3 * kc + 2];
    qi = ext.q[kc];
    float4 pair[2];
    WriteMirror(pair, 0, px, py, pz, qi, false);
    th = pair[0];
    tl = pair[1];
    // ComputeForce returns zero for a neutral point (barnes-hut.cc:168); ComputePotential does not
    // need a charge.
    want = valid && Finite3(px, py, pz) && (MODE >= 2 || qi != 0.0);
  } else {
    ks = tsel ? (int)tsel[kc] : kc;
    i = order[ks];
    px = x[i];
    py = y[i];
    pz = z[i];
    qi = q[i];
    th = src[2 * i];
    tl = src[2 * i + 1];
    // ComputeForce returns zero for a neutral target (barnes-hut.cc:168); the energy sum skips
    // neutral particles (energy_analysis.cc:64).
    want = valid && qi != 0.0;
  }
  uint32_t nrec = 0;
  int next = want ? 0 : nn;
  double ax = 0.0, ay = 0.0, az = 0.0;
  int64_t cnt = 0;
  // Store targets decide the acceptance test in FP32 where the root cube allows it (FastMacRange);
  // external points may lie anywhere and always take the FP64 test.
  const bool fast = !EXT && FastMacRange(root, theta2);
  const float theta2f = (float)theta2;
  // Direct sum over sorted positions [b, e): a leaf's particles, or a twig's (its leaf children's,
  // in their order).
  auto direct = [&](int b, int e) {
    // FP32 modes: sources in groups of kDirectGroup whose loads are issued together (one round trip
    // per group).  A group's entries past e are loaded clamped with their charge set to 0, so every
    // entry runs the same arithmetic (a charge-0 pair adds exactly 0) and no branch separates the
    // loads from their uses.
    if (MODE == 3) {
      float ph = 0.f;
      for (int k0 = b; k0 < e; k0 += kDirectGroup) {
        float4 sh[kDirectGroup], sl[kDirectGroup];
        LoadSourceGroup(ssrc, k0, sh, sl);
#pragma unroll
        for (int u = 0; u < kDirectGroup; ++u) {
          const bool self = EXT && order[min(k0 + u, e - 1)] == i;
          if (k0 + u >= e || self) sh[u].w = 0.f;
        }
#pragma unroll
        for (int u = 0; u < kDirectGroup; ++u) {
          ph += PotF32(th, tl, sh[u], sl[u]);
          if (COUNT) cnt += (sh[u].w != 0.f && k0 + u != ks) ? 1 : 0;
        }
      }
      ax += ph;
    } else if (MODE == 0) {
      // The range's sources are contiguous in the tree-ordered copy of the FP32 mirror.  The target
      // itself and charge-0 sources add exactly 0 (softened pair, q = 0), so they are not tested; the
      // interaction count leaves them out.
      float fx = 0.f, fy = 0.f, fz = 0.f;
      for (int k0 = b; k0 < e; k0 += kDirectGroup) {
        float4 sh[kDirectGroup], sl[kDirectGroup];
        LoadSourceGroup(ssrc, k0, sh, sl);
#pragma unroll
        for (int u = 0; u < kDirectGroup; ++u) {
          const bool self = EXT && order[min(k0 + u, e - 1)] == i;  // the point's self source, anywhere
          if (k0 + u >= e || self) sh[u].w = 0.f;
        }
#pragma unroll
        for (int u = 0; u < kDirectGroup; ++u) {
          PairF32(th, tl, sh[u], sl[u], fx, fy, fz);
          if (COUNT) cnt += (sh[u].w != 0.f && k0 + u != ks) ? 1 : 0;
        }
      }
      ax += fx;
      ay += fy;
      az += fz;
    } else {
      for (int kk = b; kk < e; ++kk) {
        const uint32_t j = order[kk];
        if (j == i) continue;
        const double qj = q[j];
        if (qj == 0.0) continue;
        if (MODE == 1) {
          PairF64Acc(px, py, pz, qi, x[j], y[j], z[j], qj, ax, ay, az);
        } else {
          const double rx = __dsub_rn(px, x[j]), ry = __dsub_rn(py, y[j]), rz = __dsub_rn(pz, z[j]);
          const double d2 = Dot3(rx, ry, rz, rx, ry, rz);
          if (d2 >= kMinSquaredDistance) ax = __dadd_rn(ax, __ddiv_rn(qj, sqrt(d2)));
        }
        if (COUNT) ++cnt;
      }
    }
  };
  // One visit of node c (record nd) by the wave; returns whether any lane opened c, so that the
  // wave descends to c + 1 (else it moves to skip(c)).  A lane reaches a node inside c's subtree
  // only by opening c, and the wave visits c once (its node only increases), so every lane that was
  // not at c already points at skip(c) or beyond, and every lane at c now points at c + 1 (opened)
  // or skip(c): the wave's next node needs no reduction over the lanes.  A lane that opens a twig
  // sums its children at once and moves to skip(c).
  auto visit = [&](const int c, const WalkHot& nd) -> bool {
    bool opened = false;
    if (next == c) {
      if (nd.flags & kNodeEmpty) {
        next = nd.skip;
      } else if (nd.flags & kNodeLeaf) {
        direct(nd.b, nd.e);
        next = nd.skip;
      } else {
        // The target's offset from the cube centre in FP32 (hi/lo): the fast acceptance test's
        // distance and the accepted monopoles' differences.
        const float dxc = (th.x - nd.chx) + (tl.x - nd.clx);
        const float dyc = (th.y - nd.chy) + (tl.y - nd.cly);
        const float dzc = (th.z - nd.chz) + (tl.z - nd.clz);
        int verdict = -1;
        if (fast) {
          const float r = theta2f * fmaf(dxc, dxc, fmaf(dyc, dyc, dzc * dzc));
          verdict = r > nd.w2hi ? 1 : (r < nd.w2lo ? 0 : -1);
        }
        if (verdict < 0) {  // the reference's test (barnes-hut.cc:201-204) in FP64 without contraction
          const WalkCold cb = LoadCold(nodes, c);
          const double dx = __dsub_rn(px, cb.cx), dy = __dsub_rn(py, cb.cy), dz = __dsub_rn(pz, cb.cz);
          const double d2 = Dot3(dx, dy, dz, dx, dy, dz);
          const double w = __dmul_rn(2.0, cb.half);
          verdict = __dmul_rn(w, w) < __dmul_rn(theta2, d2) ? 1 : 0;
        }
        if (verdict) {  // accepted (barnes-hut.cc:201-215); an empty class has charge 0 and adds 0
          if (MODE == 3) {
            ax += PotLocal(dxc, dyc, dzc, nd.p_rel) + PotLocal(dxc, dyc, dzc, nd.n_rel);
            if (COUNT) cnt += ((nd.flags & kNodeHasPos) ? 1 : 0) + ((nd.flags & kNodeHasNeg) ? 1 : 0);
          } else if (MODE == 0) {
            float fx, fy, fz;
            PairLocalFirst(dxc, dyc, dzc, nd.p_rel, fx, fy, fz);
            PairLocal(dxc, dyc, dzc, nd.n_rel, fx, fy, fz);
            if (COUNT) cnt += ((nd.flags & kNodeHasPos) ? 1 : 0) + ((nd.flags & kNodeHasNeg) ? 1 : 0);
            ax += fx;
            ay += fy;
            az += fz;
            if (rec) {
              if (nrec < (uint32_t)rec_cap) rec[(size_t)k * rec_cap + nrec] = (uint32_t)c;
              ++nrec;
            }
          } else {
            const WalkNodeAux au = nodes32[c];
            const WalkCold cq = LoadCold(nodes, c);
            if (MODE == 1) {
              if (cq.qp > 0.0) { PairF64Acc(px, py, pz, qi, au.pqx, au.pqy, au.pqz, cq.qp, ax, ay, az); if (COUNT) ++cnt; }
              if (cq.qn > 0.0) { PairF64Acc(px, py, pz, qi, au.nqx, au.nqy, au.nqz, -cq.qn, ax, ay, az); if (COUNT) ++cnt; }
            } else {
              if (cq.qp > 0.0) {
                const double rx = __dsub_rn(px, au.pqx), ry = __dsub_rn(py, au.pqy), rz = __dsub_rn(pz, au.pqz);
                const double d2p = Dot3(rx, ry, rz, rx, ry, rz);
                if (d2p >= kMinSquaredDistance) ax = __dadd_rn(ax, __ddiv_rn(cq.qp, sqrt(d2p)));
                if (COUNT) ++cnt;
              }
              if (cq.qn > 0.0) {
                const double rx = __dsub_rn(px, au.nqx), ry = __dsub_rn(py, au.nqy), rz = __dsub_rn(pz, au.nqz);
                const double d2n = Dot3(rx, ry, rz, rx, ry, rz);
                if (d2n >= kMinSquaredDistance) ax = __dsub_rn(ax, __ddiv_rn(cq.qn, sqrt(d2n)));
                if (COUNT) ++cnt;
              }
            }
          }
          next = nd.skip;
        } else if (nd.flags & kNodeTwig) {
          direct(nd.b, nd.e);  // opened; its children are leaves, summed here in their order
          next = nd.skip;
        } else {
          next = c + 1;
          opened = true;
        }
      }
    }
    return __ballot(opened) != 0;
  };
  // Node records come in by scalar loads (the node is wave-uniform).  The loop is unrolled by two
  // with the records in alternating registers: while node c is visited, the record of skip(c), the
  // wave's next node unless some lane opens c, is already being fetched.  An opened node is internal
  // and not a twig, so c + 1 < nn then.
  int c = __builtin_amdgcn_readfirstlane(WaveMin(next));
  WalkHot n0 = LoadHot(nodes, min(c, nn - 1)), n1;
  while (c < nn) {
    n1 = LoadHot(nodes, min(n0.skip, nn - 1));
    if (visit(c, n0)) {
      ++c;
      n1 = LoadHot(nodes, c);
    } else {
      c = n0.skip;
    }
    if (c >= nn) break;
    n0 = LoadHot(nodes, min(n1.skip, nn - 1));
    if (visit(c, n1)) {
      ++c;
      n0 = LoadHot(nodes, c);
    } else {
      c = n1.skip;
    }
  }
  if (!valid) return;
  if (EXT) {  // per point: the force on charge q, or phi (C/m for MODE 2, e/unit for MODE 3)
    if (MODE == 0) {
      const double scale = want ? kCoulomb * qi * kElementaryCharge * kInvUnit2 : 0.0;
      out_x[k] = scale * ax;
      out_y[k] = scale * ay;
      out_z[k] = scale * az;
    } else if (MODE == 1) {
      out_x[k] = ax;
      out_y[k] = ay;
      out_z[k] = az;
    } else {
      out_x[k] = ax;
    }
    return;
  }
  if (rec_cnt) rec_cnt[k] = nrec;
  if (MODE == 0) {
    const double scale = qi != 0.0 ? kCoulomb * qi * kElementaryCharge * kInvUnit2 : 0.0;
    out_x[i] = scale * ax;
    out_y[i] = scale * ay;
    out_z[i] = scale * az;
  } else if (MODE == 1) {
    out_x[i] = ax;
    out_y[i] = ay;
    out_z[i] = az;
  } else if (MODE == 2) {
    out_x[i] = __dmul_rn(qi, ax);
  } else {
    out_x[i] = (double)th.w * ax;
  }
  if (COUNT) counts[i] = cnt;
}

// ---- orchestration ------------------------------------------------------------------------------
// Tree-ordered copy of the FP32 source mirror (hi/lo per sorted position), followed by
// kSsrcPad zero entries: the walk's direct sums fetch sources in groups that may run past the last
// position (those entries get charge 0).
__global__ void GatherTreeSrcKernel(int m, const uint32_t* __restrict__ order, const float4* __restrict__ src,
                                    float4* __restrict__ ssrc) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= m + kSsrcPad) return;
  if (k >= m) {
    ssrc[2 * k] = make_float4(0.f, 0.f, 0.f, 0.f);
    ssrc[2 * k + 1] = make_float4(0.f, 0.f, 0.f, 0.f);
    return;
  }
  const uint32_t i = order[k];
  ssrc[2 * k] = src[2 * i];
  ssrc[2 * k + 1] = src[2 * i + 1];
}

int BuildBHTree(am_store* s, int leaf_cap, BHTree* t, bool with_features) {
  am_ctx* ctx = s->ctx;
  hipStream_t st = ctx->stream;
  const int64_t n = s->n;
  int rc;
  void *gathered, *root;
  if ((rc = Scratch(ctx, "bh_gathered", (size_t)n * 4, &gathered))) return rc;
  if ((rc = Scratch(ctx, "bh_root", sizeof(RootBox) + 6 * 8, &root))) return rc;
  double* box = (double*)((char*)root + sizeof(RootBox));
  int m = 0;
  if ((rc = GatherParticles(ctx, n, s->flags, s->q, s->x, s->y, s->z, 0, (uint32_t*)gathered, &m, "bh"))) return rc;
  t->m = m;
  t->nn = 0;
  if (m == 0) return AM_OK;
  if ((rc = GatheredBBox(ctx, (const uint32_t*)gathered, m, s->x, s->y, s->z, box, "bh"))) return rc;
  hipLaunchKernelGGL(BHRootKernel, dim3(1), dim3(1), 0, st, (const double*)box, (RootBox*)root);
  AM_HIP(hipGetLastError());
  Octree& o = t->tree;
  void* xq;
  if ((rc = Scratch(ctx, "bh_xq", (size_t)n * sizeof(double4), &xq))) return rc;
  if ((rc = BuildOctree(ctx, (const uint32_t*)gathered, m, s->x, s->y, s->z, s->q, (const RootBox*)root, leaf_cap,
                        "bh", &o, (double4*)xq)))
    return rc;
  const int nn = o.nn;
  const int nlevels = (int)o.level_start.size() - 1;
  Moments mo;
  {
    const char* names[] = {"bh_qp", "bh_qn", "bh_cpx", "bh_cpy", "bh_cpz", "bh_cnx", "bh_cny", "bh_cnz"};
    double** dst[] = {&mo.qp, &mo.qn, &mo.cpx, &mo.cpy, &mo.cpz, &mo.cnx, &mo.cny, &mo.cnz};
    for (int a = 0; a < 8; ++a) {
      void* p;
      if ((rc = Scratch(ctx, names[a], (size_t)nn * 8, &p))) return rc;
      *dst[a] = (double*)p;
    }
  }
  SecondMoments sm{nullptr, nullptr};
  void* featp = nullptr;
  if (with_features) {
    void *a, *b;
    if ((rc = Scratch(ctx, "bh_s2p", (size_t)nn * 6 * 8, &a))) return rc;
    if ((rc = Scratch(ctx, "bh_s2n", (size_t)nn * 6 * 8, &b))) return rc;
    if ((rc = Scratch(ctx, "bh_feat", (size_t)nn * sizeof(FeatNode), &featp))) return rc;
    sm.sp = (double*)a;
    sm.sn = (double*)b;
  }
  void* sxq;
  if ((rc = Scratch(ctx, "bh_sxq", (size_t)m * sizeof(double4), &sxq))) return rc;
  hipLaunchKernelGGL(GatherTreeF64Kernel, dim3(Blocks(m, 256)), dim3(256), 0, st, m, (const uint32_t*)o.order,
                     (const double4*)xq, (double4*)sxq);
  for (int d = nlevels - 1; d >= 0; --d) {
    const int n0 = o.level_start[d], n1 = o.level_start[d + 1];
    if (n1 <= n0) continue;
    hipLaunchKernelGGL(MomentKernel, dim3(Blocks(n1 - n0, 256)), dim3(256), 0, st, n0, n1, o.nd, mo,
                       (const double4*)sxq);
    if (with_features)
      hipLaunchKernelGGL(SecondMomentKernel, dim3(Blocks(n1 - n0, 256)), dim3(256), 0, st, n0, n1, o.nd, mo, sm,
                         (const double4*)sxq);
    AM_HIP(hipGetLastError());
  }
  void *wn, *wn32;
  if ((rc = Scratch(ctx, "bh_walknodes", (size_t)nn * sizeof(WalkNode), &wn))) return rc;
  if ((rc = Scratch(ctx, "bh_walknodes32", (size_t)nn * sizeof(WalkNodeAux), &wn32))) return rc;
  hipLaunchKernelGGL(GatherPreorderKernel, dim3(Blocks(nn, 256)), dim3(256), 0, st, nn, o.nd, mo,
                     (const uint64_t*)o.pkeys, (const uint32_t*)o.pre, (WalkNode*)wn, (WalkNodeAux*)wn32, sm,
                     (FeatNode*)featp);
  AM_HIP(hipGetLastError());
  void* ssrc;
  if ((rc = Scratch(ctx, "bh_ssrc", (size_t)(m + kSsrcPad) * 2 * sizeof(float4), &ssrc))) return rc;
  hipLaunchKernelGGL(GatherTreeSrcKernel, dim3(Blocks(m + kSsrcPad, 256)), dim3(256), 0, st, m, (const uint32_t*)o.order,
                     (const float4*)s->src, (float4*)ssrc);
  AM_HIP(hipGetLastError());
  t->ssrc = (float4*)ssrc;
  t->feat = (FeatNode*)featp;
  t->nn = nn;
  t->order = o.order;
  t->nodes = (WalkNode*)wn;
  t->nodes32 = (WalkNodeAux*)wn32;
  t->root = (const RootBox*)root;
  t->max_depth = nlevels - 1;
  t->deep_overflow = o.deep_overflow;
  return AM_OK;
}

int BarnesHutWalk(am_store* s, double theta, int mode, double* ox, double* oy, double* oz, int64_t* counts) {
  BHTree t;
  int rc = BuildBHTree(s, 8, &t);
  if (rc) return rc;
  BHBuildInfo& info = s->bh_info;
  info.m = t.m;
  info.nn = t.nn;
  info.max_depth = t.max_depth;
  info.deep_overflow = t.deep_overflow;
  info.order = t.order;
  info.nodes = t.nodes;
  if (mode < 2 && (rc = ZeroNonFiniteForces(s))) return rc;  // they are in no tree
  if (t.m == 0) return AM_OK;
  am_ctx* ctx = s->ctx;
  const double theta2 = theta * theta;
  int slot;
  TimerBegin(ctx, mode == 0 ? "bh_walk_f32" : (mode == 1 ? "bh_walk_f64" : (mode == 2 ? "bh_potential" : "bh_potential_f32")),
             &slot);
  const dim3 grid(Blocks(t.m, 256)), block(256);
  auto launch = [&](auto kernel) {
    hipLaunchKernelGGL(kernel, grid, block, 0, ctx->stream, t.m, t.nn, t.nodes, t.nodes32, t.root, t.order, s->x, s->y, s->z,
                       s->q, s->src, (const float4*)t.ssrc, theta2, ox, oy, oz, counts, nullptr, nullptr, 0, nullptr,
                       ExtTargets());
  };
  if (mode == 0) counts ? launch(BHWalkKernel<0, true>) : launch(BHWalkKernel<0, false>);
  else if (mode == 1) counts ? launch(BHWalkKernel<1, true>) : launch(BHWalkKernel<1, false>);
  else if (mode == 2) counts ? launch(BHWalkKernel<2, true>) : launch(BHWalkKernel<2, false>);
  else counts ? launch(BHWalkKernel<3, true>) : launch(BHWalkKernel<3, false>);
  TimerEnd(ctx, slot);
  AM_HIP(hipGetLastError());
  return AM_OK;
}

int LaunchBHWalkF32(am_store* s, const BHTree& t, double theta, const uint32_t* tsel, int n_targets, double* ox,
                    double* oy, double* oz, uint32_t* rec, int rec_cap, uint32_t* rec_cnt) {
  if (n_targets <= 0) return AM_OK;
  am_ctx* ctx = s->ctx;
  int slot;
  TimerBegin(ctx, "bh_walk_f32", &slot);
  hipLaunchKernelGGL((BHWalkKernel<0, false>), dim3(Blocks(n_targets, 256)), dim3(256), 0, ctx->stream, n_targets, t.nn,
                     t.nodes, t.nodes32, t.root, t.order, s->x, s->y, s->z, s->q, s->src, (const float4*)t.ssrc, theta * theta,
                     ox, oy, oz,
                     (int64_t*)nullptr, tsel, rec, rec_cap, rec_cnt, ExtTargets());
  TimerEnd(ctx, slot);
  AM_HIP(hipGetLastError());
  return AM_OK;
}

__global__ void ZeroF64Kernel(int64_t n, double* __restrict__ a) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0.0;
}

// Per-particle potential terms of the Barnes-Hut energy into out[n]; *si_scale turns their sum into
// C^2/m (1 for the FP64 walk, e^2 2^20 for the FP32 walk's e^2/unit).
int BarnesHutPotentialTerms(am_store* s, double theta, int precision, double* out, double* si_scale) {
  hipLaunchKernelGGL(ZeroF64Kernel, dim3(Blocks(s->n, 256)), dim3(256), 0, s->ctx->stream, s->n, out);
  AM_HIP(hipGetLastError());
  *si_scale = precision == AM_FP32 ? kElementaryCharge * kElementaryCharge * kUnitScale : 1.0;
  return BarnesHutWalk(s, theta, precision == AM_FP32 ? 3 : 2, out, nullptr, nullptr, nullptr);
}

}  // namespace am

using namespace am;

extern "C" {

int am_compute_barnes_hut_coulomb_forces(am_store* s, double theta, int precision) {
  if (!s) return Fail(AM_ERR_INVALID, "am_compute_barnes_hut_coulomb_forces: NULL store");
  if (precision != AM_FP32 && precision != AM_FP64) return Fail(AM_ERR_INVALID, "unknown precision");
  if (!(theta >= 0.0)) return Fail(AM_ERR_INVALID, "barnes_hut_theta must be >= 0");
  if (s->n == 0) return AM_OK;
  if (s->n > (int64_t)1 << 30) return Fail(AM_ERR_CAPACITY, "Barnes-Hut: n too large");
  AM_HIP(hipSetDevice(s->ctx->device));
  return BarnesHutWalk(s, theta, precision == AM_FP32 ? 0 : 1, s->fx, s->fy, s->fz, nullptr);
}

// BarnesHutTree::Build over the store's alive particles, then ComputeForce / ComputePotential at
// n_points given points (barnes-hut.h:28-45).  Host arrays: pos 3n, charge n (NULL: no force),
// self_index n (NULL or -1: no source left out); force 3n in N, potential n in C/m (either NULL).
int am_barnes_hut_evaluate(am_store* s, double theta, int precision, int64_t n_points, const double* pos,
                           const double* charge, const int64_t* self_index, double* force, double* potential) {
  if (!s || (n_points > 0 && !pos)) return Fail(AM_ERR_INVALID, "am_barnes_hut_evaluate: NULL argument");
  if (precision != AM_FP32 && precision != AM_FP64) return Fail(AM_ERR_INVALID, "unknown precision");
  if (!(theta >= 0.0)) return Fail(AM_ERR_INVALID, "barnes_hut_theta must be >= 0");
  if (n_points < 0 || n_points > (int64_t)1 << 30) return Fail(AM_ERR_CAPACITY, "am_barnes_hut_evaluate: n_points");
  if (force)
    for (int64_t k = 0; k < 3 * n_points; ++k) force[k] = 0.0;
  if (potential)
    for (int64_t k = 0; k < n_points; ++k) potential[k] = 0.0;
  if (n_points == 0 || s->n == 0 || (!force && !potential)) return AM_OK;
  std::vector<uint32_t> self(n_points, 0xffffffffu);
  if (self_index)
    for (int64_t k = 0; k < n_points; ++k) {
      if (self_index[k] >= s->n) return Fail(AM_ERR_INVALID, "am_barnes_hut_evaluate: self_index beyond the store");
      if (self_index[k] >= 0) self[k] = (uint32_t)self_index[k];
    }
  std::vector<double> q(n_points, 0.0);
  if (charge) std::copy(charge, charge + n_points, q.begin());
  am_ctx* ctx = s->ctx;
  hipStream_t st = ctx->stream;
  AM_HIP(hipSetDevice(ctx->device));
  BHTree t;
  int rc = BuildBHTree(s, 8, &t);
  if (rc) return rc;
  if (t.m == 0) return AM_OK;
  const int m = (int)n_points;
  void* buf;
  // points (3m) + charges (m) + outputs (3m) + self (m uint32)
  if ((rc = Scratch(ctx, "bh_eval", (size_t)m * (7 * 8 + 4), &buf))) return rc;
  double* dpos = (double*)buf;
  double* dq = dpos + 3 * (size_t)m;
  double* out = dq + m;
  uint32_t* dself = (uint32_t*)(out + 3 * (size_t)m);
  AM_HIP(hipMemcpyAsync(dpos, pos, (size_t)m * 24, hipMemcpyHostToDevice, st));
  AM_HIP(hipMemcpyAsync(dq, q.data(), (size_t)m * 8, hipMemcpyHostToDevice, st));
  AM_HIP(hipMemcpyAsync(dself, self.data(), (size_t)m * 4, hipMemcpyHostToDevice, st));
  ExtTargets ext;
  ext.pos = dpos;
  ext.q = dq;
  ext.self = dself;
  const double theta2 = theta * theta;
  auto launch = [&](auto kernel) {
    hipLaunchKernelGGL(kernel, dim3(Blocks(m, 256)), dim3(256), 0, st, m, t.nn, t.nodes, t.nodes32, t.root, t.order, s->x,
                       s->y, s->z, s->q, s->src, (const float4*)t.ssrc, theta2, out, out + m, out + 2 * (size_t)m,
                       (int64_t*)nullptr, (const uint32_t*)nullptr, (uint32_t*)nullptr, 0, (uint32_t*)nullptr, ext);
  };
  std::vector<double> h(3 * (size_t)m);
  if (force) {
    precision == AM_FP32 ? launch(BHWalkKernel<0, false, true>) : launch(BHWalkKernel<1, false, true>);
    AM_HIP(hipGetLastError());
    AM_HIP(hipMemcpyAsync(h.data(), out, (size_t)m * 24, hipMemcpyDeviceToHost, st));
    AM_HIP(hipStreamSynchronize(st));
    for (int k = 0; k < m; ++k) {
      force[3 * (size_t)k] = h[k];
      force[3 * (size_t)k + 1] = h[m + k];
      force[3 * (size_t)k + 2] = h[2 * (size_t)m + k];
    }
  }
  if (potential) {
    precision == AM_FP32 ? launch(BHWalkKernel<3, false, true>) : launch(BHWalkKernel<2, false, true>);
    AM_HIP(hipGetLastError());
    AM_HIP(hipMemcpyAsync(h.data(), out, (size_t)m * 8, hipMemcpyDeviceToHost, st));
    AM_HIP(hipStreamSynchronize(st));
    const double unit = precision == AM_FP32 ? kElementaryCharge * kUnitScale : 1.0;  // e/unit -> C/m
    for (int k = 0; k < m; ++k) potential[k] = unit * h[k];
  }
  return AM_OK;
}

// Barnes-Hut with per-particle interaction counts (leaf pairs + accepted monopoles), n int64.
int am_barnes_hut_forces_counted(am_store* s, double theta, int precision, int64_t* counts_host) {
  if (!s || !counts_host) return Fail(AM_ERR_INVALID, "am_barnes_hut_forces_counted: NULL argument");
  if (s->n == 0) return AM_OK;
  void* cnt;
  int rc = Scratch(s->ctx, "bh_counts", (size_t)s->n * 8, &cnt);
  if (rc) return rc;
  AM_HIP(hipMemsetAsync(cnt, 0, (size_t)s->n * 8, s->ctx->stream));
  rc = BarnesHutWalk(s, theta, precision == AM_FP32 ? 0 : 1, s->fx, s->fy, s->fz, (int64_t*)cnt);
  if (rc) return rc;
  AM_HIP(hipMemcpyAsync(counts_host, cnt, (size_t)s->n * 8, hipMemcpyDeviceToHost, s->ctx->stream));
  AM_HIP(hipStreamSynchronize(s->ctx->stream));
  return AM_OK;
}

// Tree of the most recent Barnes-Hut build: alive count, nodes, deepest level, and how many
// depth-21 leaves hold more than 8 particles (the reference would split those further).
int am_bh_tree_stats(am_store* s, int64_t* n_alive, int64_t* n_nodes, int32_t* max_depth, int64_t* deep_overflow) {
  if (!s) return Fail(AM_ERR_INVALID, "am_bh_tree_stats: NULL store");
  if (n_alive) *n_alive = s->bh_info.m;
  if (n_nodes) *n_nodes = s->bh_info.nn;
  if (max_depth) *max_depth = s->bh_info.max_depth;
  if (deep_overflow) *deep_overflow = s->bh_info.deep_overflow;
  return AM_OK;
}

// Sorted particle order and preorder node ranges of the most recent build (host copies):
// order[n_alive] particle indices; begin/end/leaf[n_nodes] (any may be NULL).
int am_bh_tree_download(am_store* s, uint32_t* order, int32_t* begin, int32_t* end, uint8_t* leaf, double* qp,
                        double* qn) {
  if (!s) return Fail(AM_ERR_INVALID, "am_bh_tree_download: NULL store");
  am_ctx* ctx = s->ctx;
  const int m = s->bh_info.m, nn = s->bh_info.nn;
  if (order && m > 0)
    AM_HIP(hipMemcpyAsync(order, s->bh_info.order, (size_t)m * 4, hipMemcpyDeviceToHost, ctx->stream));
  if (nn > 0 && (begin || end || leaf || qp || qn)) {
    std::vector<WalkNode> w(nn);
    AM_HIP(hipMemcpyAsync(w.data(), s->bh_info.nodes, (size_t)nn * sizeof(WalkNode), hipMemcpyDeviceToHost,
                          ctx->stream));
    AM_HIP(hipStreamSynchronize(ctx->stream));
    for (int k = 0; k < nn; ++k) {
      if (begin) begin[k] = w[k].b;
      if (end) end[k] = w[k].e;
      if (leaf) leaf[k] = (w[k].flags & kNodeLeaf) ? 1 : 0;
      if (qp) qp[k] = w[k].qp;
      if (qn) qn[k] = w[k].qn;
    }
  }
  AM_HIP(hipStreamSynchronize(ctx->stream));
  return AM_OK;
}

}  // extern "C"
