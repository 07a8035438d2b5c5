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

#include "am_internal.hpp"

namespace am {

constexpr int kMaxLevel = 21;

struct RootBox {
  double cx, cy, cz, half;
};

// ---- gather / bbox / keys ---------------------------------------------------------------------
__global__ void AliveFlagKernel(int64_t n, const uint8_t* __restrict__ flags, uint32_t* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (flags[i] & kFlagAnnihilated) ? 0u : 1u;
}

__global__ void CompactKernel(int64_t n, const uint8_t* __restrict__ flags, const uint32_t* __restrict__ off,
                              uint32_t* __restrict__ gathered) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && !(flags[i] & kFlagAnnihilated)) gathered[off[i]] = (uint32_t)i;
}

constexpr int kBoxThreads = 256;
constexpr int kBoxBlocks = 256;

__global__ __launch_bounds__(kBoxThreads) void BBoxPartialKernel(int m, const uint32_t* __restrict__ idx,
                                                                 const double* __restrict__ x,
                                                                 const double* __restrict__ y,
                                                                 const double* __restrict__ z,
                                                                 double* __restrict__ part) {
  __shared__ double sh[6][kBoxThreads];
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int k = blockIdx.x * kBoxThreads + threadIdx.x; k < m; k += gridDim.x * kBoxThreads) {
    const uint32_t i = idx[k];
    const double p[3] = {x[i], y[i], z[i]};
    for (int d = 0; d < 3; ++d) {
      lo[d] = fmin(lo[d], p[d]);
      hi[d] = fmax(hi[d], p[d]);
    }
  }
  for (int d = 0; d < 3; ++d) {
    sh[d][threadIdx.x] = lo[d];
    sh[3 + d][threadIdx.x] = hi[d];
  }
  __syncthreads();
  for (int w = kBoxThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int d = 0; d < 3; ++d) {
        sh[d][threadIdx.x] = fmin(sh[d][threadIdx.x], sh[d][threadIdx.x + w]);
        sh[3 + d][threadIdx.x] = fmax(sh[3 + d][threadIdx.x], sh[3 + d][threadIdx.x + w]);
      }
    __syncthreads();
  }
  if (threadIdx.x < 6) part[blockIdx.x * 6 + threadIdx.x] = sh[threadIdx.x][0];
}

// Root cube of BarnesHutTree::Build (barnes-hut.cc:47-62).
__global__ void BHRootKernel(int blocks, const double* __restrict__ part, RootBox* __restrict__ root,
                             double* __restrict__ box_out) {
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int b = 0; b < blocks; ++b)
    for (int d = 0; d < 3; ++d) {
      lo[d] = fmin(lo[d], part[b * 6 + d]);
      hi[d] = fmax(hi[d], part[b * 6 + 3 + d]);
    }
  RootBox r;
  r.cx = 0.5 * (lo[0] + hi[0]);
  r.cy = 0.5 * (lo[1] + hi[1]);
  r.cz = 0.5 * (lo[2] + hi[2]);
  const double ext = fmax(fmax(hi[0] - lo[0], hi[1] - lo[1]), hi[2] - lo[2]);
  double half = 0.5 * ext * (1.0 + 1e-9);
  if (half < 1e-30) half = 1e-30;
  r.half = half;
  *root = r;
  for (int d = 0; d < 3; ++d) {
    box_out[d] = lo[d];
    box_out[3 + d] = hi[d];
  }
}

// Morton key by repeating the reference descent: octant test against the current centre, then
// move the centre by +-half/2 in FP64 (barnes-hut.cc:30-33,124-136).
__global__ void MortonKernel(int m, const uint32_t* __restrict__ idx, const double* __restrict__ x,
                             const double* __restrict__ y, const double* __restrict__ z,
                             const RootBox* __restrict__ root, uint64_t* __restrict__ keys,
                             uint32_t* __restrict__ vals) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= m) return;
  const uint32_t i = idx[k];
  const double px = x[i], py = y[i], pz = z[i];
  double cx = root->cx, cy = root->cy, cz = root->cz, h = root->half;
  uint64_t key = 0;
  for (int l = 0; l < kMaxLevel; ++l) {
    const int o = (px >= cx ? 1 : 0) | (py >= cy ? 2 : 0) | (pz >= cz ? 4 : 0);
    key = (key << 3) | (uint64_t)o;
    const double ch = 0.5 * h;
    cx = cx + ((o & 1) ? ch : -ch);
    cy = cy + ((o & 2) ? ch : -ch);
    cz = cz + ((o & 4) ? ch : -ch);
    h = ch;
  }
  keys[k] = key;
  vals[k] = i;
}

// ---- level-synchronous build ------------------------------------------------------------------
struct NodeArrays {  // BFS order
  int* b;
  int* e;
  int* depth;
  int* first_child;
  int* nchild;
  uint8_t* leaf;
  double* cx;
  double* cy;
  double* cz;
  double* half;
};

__device__ inline int Digit(uint64_t key, int level) { return (int)((key >> (3 * (kMaxLevel - 1 - level))) & 7u); }

__global__ void SplitCountKernel(int n0, int n1, NodeArrays nd, const uint64_t* __restrict__ keys, int leaf_cap,
                                 int* __restrict__ bounds, uint32_t* __restrict__ cnt) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int node = n0 + t;
  if (node >= n1) return;
  const int b = nd.b[node], e = nd.e[node], d = nd.depth[node];
  int* bd = bounds + 9 * t;
  if (e - b <= leaf_cap || d >= kMaxLevel) {
    nd.leaf[node] = 1;
    nd.nchild[node] = 0;
    nd.first_child[node] = -1;
    cnt[t] = 0;
    return;
  }
  nd.leaf[node] = 0;
  bd[0] = b;
  for (int o = 1; o < 8; ++o) {  // first position whose digit at level d is >= o
    int lo = bd[o - 1], hi = e;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (Digit(keys[mid], d) < o) lo = mid + 1;
      else hi = mid;
    }
    bd[o] = lo;
  }
  bd[8] = e;
  uint32_t c = 0;
  for (int o = 0; o < 8; ++o) c += bd[o + 1] > bd[o] ? 1u : 0u;
  cnt[t] = c;
}

__global__ void WriteChildrenKernel(int n0, int n1, int next_base, NodeArrays nd, const int* __restrict__ bounds,
                                    const uint32_t* __restrict__ off) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int node = n0 + t;
  if (node >= n1 || nd.leaf[node]) return;
  const int* bd = bounds + 9 * t;
  int c = next_base + (int)off[t];
  nd.first_child[node] = c;
  int k = 0;
  const double h = nd.half[node];
  const double ch = 0.5 * h;
  const double px = nd.cx[node], py = nd.cy[node], pz = nd.cz[node];
  for (int o = 0; o < 8; ++o) {
    if (bd[o + 1] == bd[o]) continue;
    const int ci = c + k++;
    nd.b[ci] = bd[o];
    nd.e[ci] = bd[o + 1];
    nd.depth[ci] = nd.depth[node] + 1;
    nd.cx[ci] = px + ((o & 1) ? ch : -ch);
    nd.cy[ci] = py + ((o & 2) ? ch : -ch);
    nd.cz[ci] = pz + ((o & 4) ? ch : -ch);
    nd.half[ci] = ch;
  }
  nd.nchild[node] = k;
}

// Leaves keep particle-index order (the reference's stable partition stops at the leaf).
__global__ void LeafSortKernel(int nn, NodeArrays nd, uint32_t* __restrict__ vals, uint32_t* __restrict__ over) {
  const int node = blockIdx.x * blockDim.x + threadIdx.x;
  if (node >= nn || !nd.leaf[node]) return;
  const int b = nd.b[node], e = nd.e[node];
  for (int k = b + 1; k < e; ++k) {
    const uint32_t v = vals[k];
    int j = k - 1;
    while (j >= b && vals[j] > v) {
      vals[j + 1] = vals[j];
      --j;
    }
    vals[j + 1] = v;
  }
  if (nd.depth[node] >= kMaxLevel && e - b > 8) atomicAdd(over, 1u);
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

// Dual monopoles (barnes-hut.cc:138-163): leaves sum q x over their particles in order; internal
// nodes combine their children's (charge, centre) in octant order; centre = moment * (1/charge).
__global__ void MomentKernel(int n0, int n1, NodeArrays nd, Moments mo, const uint32_t* __restrict__ vals,
                             const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ z,
                             const double* __restrict__ q) {
  const int node = n0 + blockIdx.x * blockDim.x + threadIdx.x;
  if (node >= n1) return;
  double qp = 0.0, qn = 0.0;
  double mpx = 0.0, mpy = 0.0, mpz = 0.0, mnx = 0.0, mny = 0.0, mnz = 0.0;
  if (nd.leaf[node]) {
    for (int k = nd.b[node]; k < nd.e[node]; ++k) {
      const uint32_t j = vals[k];
      const double qj = q[j];
      if (qj > 0.0) {
        qp = qp + qj;
        mpx = mpx + x[j] * qj;
        mpy = mpy + y[j] * qj;
        mpz = mpz + z[j] * qj;
      } else if (qj < 0.0) {
        const double a = -qj;
        qn = qn + a;
        mnx = mnx + x[j] * a;
        mny = mny + y[j] * a;
        mnz = mnz + z[j] * a;
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

__global__ void PreorderKeyKernel(int nn, NodeArrays nd, uint64_t* __restrict__ keys, uint32_t* __restrict__ ids) {
  const int node = blockIdx.x * blockDim.x + threadIdx.x;
  if (node >= nn) return;
  keys[node] = ((uint64_t)nd.b[node] << 5) | (uint64_t)nd.depth[node];
  ids[node] = (uint32_t)node;
}

// Walk-ready node record in DFS preorder.
struct WalkNode {
  double cx, cy, cz, half;     // geometric cube
  double pqx, pqy, pqz, qp;    // positive charge centre and charge
  double nqx, nqy, nqz, qn;    // negative charge centre and |charge|
  int b, e, skip, leaf;        // particle range in the sorted order, subtree skip pointer
};
struct WalkNodeF32 {  // FP32 hi/lo mirror of the monopoles in units (charge in e)
  float4 p_hi, p_lo, n_hi, n_lo;
};

__global__ void GatherPreorderKernel(int nn, NodeArrays nd, Moments mo, const uint64_t* __restrict__ skeys,
                                     const uint32_t* __restrict__ pre, WalkNode* __restrict__ out,
                                     WalkNodeF32* __restrict__ out32) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nn) return;
  const int node = (int)pre[p];
  WalkNode w;
  w.cx = nd.cx[node];
  w.cy = nd.cy[node];
  w.cz = nd.cz[node];
  w.half = nd.half[node];
  w.pqx = mo.cpx[node];
  w.pqy = mo.cpy[node];
  w.pqz = mo.cpz[node];
  w.qp = mo.qp[node];
  w.nqx = mo.cnx[node];
  w.nqy = mo.cny[node];
  w.nqz = mo.cnz[node];
  w.qn = mo.qn[node];
  w.b = nd.b[node];
  w.e = nd.e[node];
  w.leaf = nd.leaf[node];
  // skip = first preorder position whose key >= (e << 5)
  const uint64_t target = (uint64_t)w.e << 5;
  int lo = p + 1, hi = nn;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (skeys[mid] < target) lo = mid + 1;
    else hi = mid;
  }
  w.skip = lo;
  out[p] = w;
  WalkNodeF32 f;
  const double qs = 1.0 / kElementaryCharge;
  {
    const double xs = w.pqx * kUnitScale, ys = w.pqy * kUnitScale, zs = w.pqz * kUnitScale;
    const float xh = (float)xs, yh = (float)ys, zh = (float)zs;
    f.p_hi = make_float4(xh, yh, zh, (float)(w.qp * qs));
    f.p_lo = make_float4((float)(xs - xh), (float)(ys - yh), (float)(zs - zh), 0.f);
  }
  {
    const double xs = w.nqx * kUnitScale, ys = w.nqy * kUnitScale, zs = w.nqz * kUnitScale;
    const float xh = (float)xs, yh = (float)ys, zh = (float)zs;
    f.n_hi = make_float4(xh, yh, zh, (float)(-w.qn * qs));
    f.n_lo = make_float4((float)(xs - xh), (float)(ys - yh), (float)(zs - zh), 0.f);
  }
  out32[p] = f;
}

// ---- walk -------------------------------------------------------------------------------------
__device__ inline int WaveMin(int v) {
  for (int off = 32; off > 0; off >>= 1) v = min(v, __shfl_xor(v, off));
  return v;
}

__device__ __forceinline__ void PairF32Acc(float4 th, float4 tl, float4 sh, float4 sl, float& ax, float& ay,
                                           float& az) {
  const float dx = (th.x - sh.x) + (tl.x - sl.x);
  const float dy = (th.y - sh.y) + (tl.y - sl.y);
  const float dz = (th.z - sh.z) + (tl.z - sl.z);
  const float r2 = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, kEps2Units)));
  const float ri = __builtin_amdgcn_rsqf(r2);
  const float s = (sh.w * ri) * (ri * ri);
  ax = fmaf(s, dx, ax);
  ay = fmaf(s, dy, ay);
  az = fmaf(s, dz, az);
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

// MODE 0: FP32 force; MODE 1: FP64 force; MODE 2: FP64 potential (q_i phi_i into out_x).
template <int MODE>
__global__ __launch_bounds__(256) void BHWalkKernel(int m, int nn, const WalkNode* __restrict__ nodes,
                                                     const WalkNodeF32* __restrict__ nodes32,
                                                     const uint32_t* __restrict__ order,
                                                     const double* __restrict__ x, const double* __restrict__ y,
                                                     const double* __restrict__ z, const double* __restrict__ q,
                                                     const float4* __restrict__ src, double theta2,
                                                     double* __restrict__ out_x, double* __restrict__ out_y,
                                                     double* __restrict__ out_z, int64_t* __restrict__ counts) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = k < m;
  const uint32_t i = order[valid ? k : m - 1];
  const double px = x[i], py = y[i], pz = z[i], qi = q[i];
  const float4 th = src[2 * i], tl = src[2 * i + 1];
  // ComputeForce returns zero for a neutral target (barnes-hut.cc:168); the energy sum skips
  // neutral particles (energy_analysis.cc:64).
  const bool want = valid && qi != 0.0;
  int next = want ? 0 : nn;
  double ax = 0.0, ay = 0.0, az = 0.0;
  int64_t cnt = 0;
  int c = WaveMin(next);
  while (c < nn) {
    const WalkNode nd = nodes[c];
    const bool mine = next == c;
    bool opened = false;
    if (mine) {
      if (nd.qp == 0.0 && nd.qn == 0.0) {
        next = nd.skip;
      } else if (nd.leaf) {
        if (MODE == 0) {
          float fx = 0.f, fy = 0.f, fz = 0.f;
          for (int kk = nd.b; kk < nd.e; ++kk) {
            const uint32_t j = order[kk];
            if (j == i) continue;
            const float4 sh = src[2 * j];
            if (sh.w == 0.f) continue;
            PairF32Acc(th, tl, sh, src[2 * j + 1], fx, fy, fz);
            ++cnt;
          }
          ax += fx;
          ay += fy;
          az += fz;
        } else {
          for (int kk = nd.b; kk < nd.e; ++kk) {
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
            ++cnt;
          }
        }
        next = nd.skip;
      } else {
        const double dx = __dsub_rn(px, nd.cx), dy = __dsub_rn(py, nd.cy), dz = __dsub_rn(pz, nd.cz);
        const double d2 = Dot3(dx, dy, dz, dx, dy, dz);
        const double w = __dmul_rn(2.0, nd.half);
        if (__dmul_rn(w, w) < __dmul_rn(theta2, d2)) {  // accepted (barnes-hut.cc:201-215)
          if (MODE == 0) {
            const WalkNodeF32 f = nodes32[c];
            float fx = 0.f, fy = 0.f, fz = 0.f;
            if (nd.qp > 0.0) { PairF32Acc(th, tl, f.p_hi, f.p_lo, fx, fy, fz); ++cnt; }
            if (nd.qn > 0.0) { PairF32Acc(th, tl, f.n_hi, f.n_lo, fx, fy, fz); ++cnt; }
            ax += fx;
            ay += fy;
            az += fz;
          } else if (MODE == 1) {
            if (nd.qp > 0.0) { PairF64Acc(px, py, pz, qi, nd.pqx, nd.pqy, nd.pqz, nd.qp, ax, ay, az); ++cnt; }
            if (nd.qn > 0.0) { PairF64Acc(px, py, pz, qi, nd.nqx, nd.nqy, nd.nqz, -nd.qn, ax, ay, az); ++cnt; }
          } else {
            if (nd.qp > 0.0) {
              const double rx = __dsub_rn(px, nd.pqx), ry = __dsub_rn(py, nd.pqy), rz = __dsub_rn(pz, nd.pqz);
              const double d2p = Dot3(rx, ry, rz, rx, ry, rz);
              if (d2p >= kMinSquaredDistance) ax = __dadd_rn(ax, __ddiv_rn(nd.qp, sqrt(d2p)));
              ++cnt;
            }
            if (nd.qn > 0.0) {
              const double rx = __dsub_rn(px, nd.nqx), ry = __dsub_rn(py, nd.nqy), rz = __dsub_rn(pz, nd.nqz);
              const double d2n = Dot3(rx, ry, rz, rx, ry, rz);
              if (d2n >= kMinSquaredDistance) ax = __dsub_rn(ax, __ddiv_rn(nd.qn, sqrt(d2n)));
              ++cnt;
            }
          }
          next = nd.skip;
        } else {
          next = c + 1;
          opened = true;
        }
      }
    }
    c = __ballot(opened) ? c + 1 : WaveMin(next);
  }
  if (!valid) return;
  if (MODE == 0) {
    const double scale = qi != 0.0 ? kCoulomb * qi * kElementaryCharge * kInvUnit2 : 0.0;
    out_x[i] = scale * ax;
    out_y[i] = scale * ay;
    out_z[i] = scale * az;
  } else if (MODE == 1) {
    out_x[i] = ax;
    out_y[i] = ay;
    out_z[i] = az;
  } else {
    out_x[i] = __dmul_rn(qi, ax);
  }
  if (counts) counts[i] = cnt;
}

// ---- orchestration ------------------------------------------------------------------------------
struct BHTree {
  int m = 0;            // alive particles
  int nn = 0;           // nodes
  uint32_t* order = nullptr;  // sorted position -> particle index
  WalkNode* nodes = nullptr;
  WalkNodeF32* nodes32 = nullptr;
  int max_depth = 0;
  uint32_t deep_overflow = 0;
};

static int Alloc(am_ctx* ctx, const char* name, size_t bytes, void** p) { return Scratch(ctx, name, bytes, p); }

int BuildBHTree(am_store* s, int leaf_cap, BHTree* t) {
  am_ctx* ctx = s->ctx;
  hipStream_t st = ctx->stream;
  const int64_t n = s->n;
  int rc;
  void *alive, *off, *gathered, *part, *root, *box, *keys, *vals;
  if ((rc = Alloc(ctx, "bh_alive", (size_t)n * 4, &alive))) return rc;
  if ((rc = Alloc(ctx, "bh_off", (size_t)(n + 1) * 4, &off))) return rc;
  if ((rc = Alloc(ctx, "bh_gathered", (size_t)n * 4, &gathered))) return rc;
  if ((rc = Alloc(ctx, "bh_part", (size_t)kBoxBlocks * 6 * 8, &part))) return rc;
  if ((rc = Alloc(ctx, "bh_root", sizeof(RootBox) + 6 * 8, &root))) return rc;
  box = (char*)root + sizeof(RootBox);
  hipLaunchKernelGGL(AliveFlagKernel, dim3(Blocks(n, 256)), dim3(256), 0, st, n, s->flags, (uint32_t*)alive);
  AM_HIP(hipGetLastError());
  if ((rc = ExclusiveScan(ctx, (const uint32_t*)alive, (uint32_t*)off, n))) return rc;
  uint32_t m = 0;
  AM_HIP(hipMemcpyAsync(&m, (uint32_t*)off + n, 4, hipMemcpyDeviceToHost, st));
  AM_HIP(hipStreamSynchronize(st));
  t->m = (int)m;
  t->nn = 0;
  if (m == 0) return AM_OK;
  hipLaunchKernelGGL(CompactKernel, dim3(Blocks(n, 256)), dim3(256), 0, st, n, s->flags, (const uint32_t*)off,
                     (uint32_t*)gathered);
  const int bb = std::min(kBoxBlocks, Blocks(m, kBoxThreads));
  hipLaunchKernelGGL(BBoxPartialKernel, dim3(bb), dim3(kBoxThreads), 0, st, (int)m, (const uint32_t*)gathered, s->x,
                     s->y, s->z, (double*)part);
  hipLaunchKernelGGL(BHRootKernel, dim3(1), dim3(1), 0, st, bb, (const double*)part, (RootBox*)root, (double*)box);
  if ((rc = Alloc(ctx, "bh_keys", (size_t)m * 8, &keys))) return rc;
  if ((rc = Alloc(ctx, "bh_vals", (size_t)m * 4, &vals))) return rc;
  hipLaunchKernelGGL(MortonKernel, dim3(Blocks(m, 256)), dim3(256), 0, st, (int)m, (const uint32_t*)gathered, s->x,
                     s->y, s->z, (const RootBox*)root, (uint64_t*)keys, (uint32_t*)vals);
  AM_HIP(hipGetLastError());
  if ((rc = RadixSortPairs(ctx, (uint64_t*)keys, (uint32_t*)vals, m, 3 * kMaxLevel))) return rc;

  // Node arrays with a generous capacity; grown (with copy) when a level would overflow.
  size_t cap = (size_t)3 * m + 1024;
  std::vector<void*> arrays;
  NodeArrays nd;
  auto alloc_nodes = [&](size_t c) -> int {
    void* p;
    int r;
    const char* names[] = {"bh_nb", "bh_ne", "bh_nd", "bh_nfc", "bh_nnc", "bh_nleaf", "bh_ncx", "bh_ncy", "bh_ncz",
                           "bh_nhalf"};
    const size_t sz[] = {4, 4, 4, 4, 4, 1, 8, 8, 8, 8};
    void** dst[] = {(void**)&nd.b, (void**)&nd.e, (void**)&nd.depth, (void**)&nd.first_child, (void**)&nd.nchild,
                    (void**)&nd.leaf, (void**)&nd.cx, (void**)&nd.cy, (void**)&nd.cz, (void**)&nd.half};
    for (int a = 0; a < 10; ++a) {
      if ((r = Alloc(ctx, names[a], c * sz[a], &p))) return r;
      *dst[a] = p;
    }
    return AM_OK;
  };
  if ((rc = alloc_nodes(cap))) return rc;
  // root node (BFS index 0)
  {
    int zero = 0, mm = (int)m;
    AM_HIP(hipMemcpyAsync(nd.b, &zero, 4, hipMemcpyHostToDevice, st));
    AM_HIP(hipMemcpyAsync(nd.e, &mm, 4, hipMemcpyHostToDevice, st));
    AM_HIP(hipMemcpyAsync(nd.depth, &zero, 4, hipMemcpyHostToDevice, st));
    AM_HIP(hipMemcpyAsync(nd.cx, &((RootBox*)root)->cx, 8, hipMemcpyDeviceToDevice, st));
    AM_HIP(hipMemcpyAsync(nd.cy, &((RootBox*)root)->cy, 8, hipMemcpyDeviceToDevice, st));
    AM_HIP(hipMemcpyAsync(nd.cz, &((RootBox*)root)->cz, 8, hipMemcpyDeviceToDevice, st));
    AM_HIP(hipMemcpyAsync(nd.half, &((RootBox*)root)->half, 8, hipMemcpyDeviceToDevice, st));
    AM_HIP(hipStreamSynchronize(st));  // host ints above go out of scope
  }
  std::vector<int> level_start = {0, 1};
  int total = 1;
  for (int d = 0; d <= kMaxLevel; ++d) {
    const int n0 = level_start[d], n1 = level_start[d + 1];
    const int cnt_nodes = n1 - n0;
    if (cnt_nodes == 0) break;
    void *bounds, *cnt, *coff;
    if ((rc = Alloc(ctx, "bh_bounds", (size_t)cnt_nodes * 9 * 4, &bounds))) return rc;
    if ((rc = Alloc(ctx, "bh_cnt", (size_t)cnt_nodes * 4, &cnt))) return rc;
    if ((rc = Alloc(ctx, "bh_coff", (size_t)(cnt_nodes + 1) * 4, &coff))) return rc;
    hipLaunchKernelGGL(SplitCountKernel, dim3(Blocks(cnt_nodes, 256)), dim3(256), 0, st, n0, n1, nd,
                       (const uint64_t*)keys, leaf_cap, (int*)bounds, (uint32_t*)cnt);
    AM_HIP(hipGetLastError());
    if ((rc = ExclusiveScan(ctx, (const uint32_t*)cnt, (uint32_t*)coff, cnt_nodes))) return rc;
    uint32_t kids = 0;
    AM_HIP(hipMemcpyAsync(&kids, (uint32_t*)coff + cnt_nodes, 4, hipMemcpyDeviceToHost, st));
    AM_HIP(hipStreamSynchronize(st));
    if (kids == 0) break;
    if ((size_t)total + kids > cap) {  // grow with copy
      const size_t ncap = std::max(cap * 2, (size_t)total + kids + 1024);
      NodeArrays old = nd;
      std::vector<std::pair<void*, size_t>> saved;
      const size_t sz[] = {4, 4, 4, 4, 4, 1, 8, 8, 8, 8};
      void* olds[] = {old.b, old.e, old.depth, old.first_child, old.nchild, old.leaf, old.cx, old.cy, old.cz, old.half};
      for (int a = 0; a < 10; ++a) {
        void* tmp;
        AM_HIP(hipMallocAsync(&tmp, total * sz[a], st));
        AM_HIP(hipMemcpyAsync(tmp, olds[a], total * sz[a], hipMemcpyDeviceToDevice, st));
        saved.push_back({tmp, total * sz[a]});
      }
      if ((rc = alloc_nodes(ncap))) return rc;
      void* news[] = {nd.b, nd.e, nd.depth, nd.first_child, nd.nchild, nd.leaf, nd.cx, nd.cy, nd.cz, nd.half};
      for (int a = 0; a < 10; ++a) {
        AM_HIP(hipMemcpyAsync(news[a], saved[a].first, saved[a].second, hipMemcpyDeviceToDevice, st));
        AM_HIP(hipFreeAsync(saved[a].first, st));
      }
      cap = ncap;
    }
    hipLaunchKernelGGL(WriteChildrenKernel, dim3(Blocks(cnt_nodes, 256)), dim3(256), 0, st, n0, n1, total, nd,
                       (const int*)bounds, (const uint32_t*)coff);
    AM_HIP(hipGetLastError());
    total += (int)kids;
    level_start.push_back(total);
  }
  // Mark the deepest level (never split) as leaves.
  {
    const int n0 = level_start[level_start.size() - 2], n1 = level_start.back();
    if (n1 > n0) {
      void *bounds, *cnt;
      if ((rc = Alloc(ctx, "bh_bounds", (size_t)(n1 - n0) * 9 * 4, &bounds))) return rc;
      if ((rc = Alloc(ctx, "bh_cnt", (size_t)(n1 - n0) * 4, &cnt))) return rc;
      hipLaunchKernelGGL(SplitCountKernel, dim3(Blocks(n1 - n0, 256)), dim3(256), 0, st, n0, n1, nd,
                         (const uint64_t*)keys, (int)m + 1, (int*)bounds, (uint32_t*)cnt);
      AM_HIP(hipGetLastError());
    }
  }
  const int nn = total;
  const int nlevels = (int)level_start.size() - 1;
  void* over;
  if ((rc = Alloc(ctx, "bh_over", 4, &over))) return rc;
  AM_HIP(hipMemsetAsync(over, 0, 4, st));
  hipLaunchKernelGGL(LeafSortKernel, dim3(Blocks(nn, 256)), dim3(256), 0, st, nn, nd, (uint32_t*)vals,
                     (uint32_t*)over);
  AM_HIP(hipGetLastError());
  Moments mo;
  {
    const char* names[] = {"bh_qp", "bh_qn", "bh_cpx", "bh_cpy", "bh_cpz", "bh_cnx", "bh_cny", "bh_cnz"};
    double** dst[] = {&mo.qp, &mo.qn, &mo.cpx, &mo.cpy, &mo.cpz, &mo.cnx, &mo.cny, &mo.cnz};
    for (int a = 0; a < 8; ++a) {
      void* p;
      if ((rc = Alloc(ctx, names[a], (size_t)nn * 8, &p))) return rc;
      *dst[a] = (double*)p;
    }
  }
  for (int d = nlevels - 1; d >= 0; --d) {
    const int n0 = level_start[d], n1 = level_start[d + 1];
    if (n1 <= n0) continue;
    hipLaunchKernelGGL(MomentKernel, dim3(Blocks(n1 - n0, 256)), dim3(256), 0, st, n0, n1, nd, mo,
                       (const uint32_t*)vals, s->x, s->y, s->z, s->q);
    AM_HIP(hipGetLastError());
  }
  void *pkeys, *pids, *wn, *wn32;
  if ((rc = Alloc(ctx, "bh_pkeys", (size_t)nn * 8, &pkeys))) return rc;
  if ((rc = Alloc(ctx, "bh_pids", (size_t)nn * 4, &pids))) return rc;
  if ((rc = Alloc(ctx, "bh_walknodes", (size_t)nn * sizeof(WalkNode), &wn))) return rc;
  if ((rc = Alloc(ctx, "bh_walknodes32", (size_t)nn * sizeof(WalkNodeF32), &wn32))) return rc;
  hipLaunchKernelGGL(PreorderKeyKernel, dim3(Blocks(nn, 256)), dim3(256), 0, st, nn, nd, (uint64_t*)pkeys,
                     (uint32_t*)pids);
  AM_HIP(hipGetLastError());
  if ((rc = RadixSortPairs(ctx, (uint64_t*)pkeys, (uint32_t*)pids, nn, 40))) return rc;
  hipLaunchKernelGGL(GatherPreorderKernel, dim3(Blocks(nn, 256)), dim3(256), 0, st, nn, nd, mo,
                     (const uint64_t*)pkeys, (const uint32_t*)pids, (WalkNode*)wn, (WalkNodeF32*)wn32);
  AM_HIP(hipGetLastError());
  uint32_t ov = 0;
  AM_HIP(hipMemcpyAsync(&ov, over, 4, hipMemcpyDeviceToHost, st));
  AM_HIP(hipStreamSynchronize(st));
  t->nn = nn;
  t->order = (uint32_t*)vals;
  t->nodes = (WalkNode*)wn;
  t->nodes32 = (WalkNodeF32*)wn32;
  t->max_depth = nlevels - 1;
  t->deep_overflow = ov;
  return AM_OK;
}

static BHTree g_last_tree_info;  // diagnostics of the most recent build (host copy)

int BarnesHutWalk(am_store* s, double theta, int mode, double* ox, double* oy, double* oz, int64_t* counts) {
  BHTree t;
  int rc = BuildBHTree(s, 8, &t);
  if (rc) return rc;
  g_last_tree_info = t;
  if (t.m == 0) return AM_OK;
  am_ctx* ctx = s->ctx;
  const double theta2 = theta * theta;
  int slot;
  TimerBegin(ctx, mode == 0 ? "bh_walk_f32" : (mode == 1 ? "bh_walk_f64" : "bh_potential"), &slot);
  const dim3 grid(Blocks(t.m, 256)), block(256);
  if (mode == 0)
    hipLaunchKernelGGL(BHWalkKernel<0>, grid, block, 0, ctx->stream, t.m, t.nn, t.nodes, t.nodes32, t.order, s->x,
                       s->y, s->z, s->q, s->src, theta2, ox, oy, oz, counts);
  else if (mode == 1)
    hipLaunchKernelGGL(BHWalkKernel<1>, grid, block, 0, ctx->stream, t.m, t.nn, t.nodes, t.nodes32, t.order, s->x,
                       s->y, s->z, s->q, s->src, theta2, ox, oy, oz, counts);
  else
    hipLaunchKernelGGL(BHWalkKernel<2>, grid, block, 0, ctx->stream, t.m, t.nn, t.nodes, t.nodes32, t.order, s->x,
                       s->y, s->z, s->q, s->src, theta2, ox, oy, oz, counts);
  TimerEnd(ctx, slot);
  AM_HIP(hipGetLastError());
  return AM_OK;
}

__global__ void ZeroF64Kernel(int64_t n, double* __restrict__ a) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0.0;
}

int BarnesHutPotentialTerms(am_store* s, double theta, double* out) {
  hipLaunchKernelGGL(ZeroF64Kernel, dim3(Blocks(s->n, 256)), dim3(256), 0, s->ctx->stream, s->n, out);
  AM_HIP(hipGetLastError());
  return BarnesHutWalk(s, theta, 2, out, nullptr, nullptr, nullptr);
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
  if (n_alive) *n_alive = g_last_tree_info.m;
  if (n_nodes) *n_nodes = g_last_tree_info.nn;
  if (max_depth) *max_depth = g_last_tree_info.max_depth;
  if (deep_overflow) *deep_overflow = g_last_tree_info.deep_overflow;
  return AM_OK;
}

// Sorted particle order and preorder node ranges of the most recent build (host copies):
// order[n_alive] particle indices; begin/end/leaf[n_nodes] (any may be NULL).
int am_bh_tree_download(am_store* s, uint32_t* order, int32_t* begin, int32_t* end, uint8_t* leaf, double* qp,
                        double* qn) {
  if (!s) return Fail(AM_ERR_INVALID, "am_bh_tree_download: NULL store");
  am_ctx* ctx = s->ctx;
  const int m = g_last_tree_info.m, nn = g_last_tree_info.nn;
  if (order && m > 0)
    AM_HIP(hipMemcpyAsync(order, g_last_tree_info.order, (size_t)m * 4, hipMemcpyDeviceToHost, ctx->stream));
  if (nn > 0 && (begin || end || leaf || qp || qn)) {
    std::vector<WalkNode> w(nn);
    AM_HIP(hipMemcpyAsync(w.data(), g_last_tree_info.nodes, (size_t)nn * sizeof(WalkNode), hipMemcpyDeviceToHost,
                          ctx->stream));
    AM_HIP(hipStreamSynchronize(ctx->stream));
    for (int k = 0; k < nn; ++k) {
      if (begin) begin[k] = w[k].b;
      if (end) end[k] = w[k].e;
      if (leaf) leaf[k] = (uint8_t)w[k].leaf;
      if (qp) qp[k] = w[k].qp;
      if (qn) qn[k] = w[k].qn;
    }
  }
  AM_HIP(hipStreamSynchronize(ctx->stream));
  return AM_OK;
}

}  // extern "C"
