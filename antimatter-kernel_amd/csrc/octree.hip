// Count-capped octree in the reference's DFS preorder (see octree.hpp).
#include <algorithm>
#include <cmath>
#include <string>

#include "octree.hpp"

namespace am {

// Keeps alive particles with finite coordinates (non-finite ones stay out of every tree, as the
// direct solver skips them, coulomb_interaction.cc:55-56), and in mode 1 only charged ones.
__global__ void KeepFlagKernel(int64_t n, const uint8_t* __restrict__ flags, const double* __restrict__ q,
                               const double* __restrict__ x, const double* __restrict__ y,
                               const double* __restrict__ z, int mode, uint32_t* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const bool alive = !(flags[i] & kFlagAnnihilated) && Finite3(x[i], y[i], z[i]);
  out[i] = (alive && (mode == 0 || q[i] != 0.0)) ? 1u : 0u;
}

__global__ void CompactFlagKernel(int64_t n, const uint32_t* __restrict__ keep, const uint32_t* __restrict__ off,
                                  uint32_t* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && keep[i]) out[off[i]] = (uint32_t)i;
}

static std::string Name(const char* tag, const char* what) { return std::string(tag) + "_" + what; }

int GatherParticles(am_ctx* ctx, int64_t n, const uint8_t* flags, const double* q, const double* x, const double* y,
                    const double* z, int mode, uint32_t* out, int* m_out, const char* tag) {
  hipStream_t st = ctx->stream;
  void *keep, *off;
  int rc;
  if ((rc = Scratch(ctx, Name(tag, "keep").c_str(), (size_t)n * 4, &keep))) return rc;
  if ((rc = Scratch(ctx, Name(tag, "keepoff").c_str(), (size_t)(n + 1) * 4, &off))) return rc;
  hipLaunchKernelGGL(KeepFlagKernel, dim3(Blocks(n, 256)), dim3(256), 0, st, n, flags, q, x, y, z, mode,
                     (uint32_t*)keep);
  AM_HIP(hipGetLastError());
  if ((rc = ExclusiveScan(ctx, (const uint32_t*)keep, (uint32_t*)off, n))) return rc;
  hipLaunchKernelGGL(CompactFlagKernel, dim3(Blocks(n, 256)), dim3(256), 0, st, n, (const uint32_t*)keep,
                     (const uint32_t*)off, out);
  AM_HIP(hipGetLastError());
  uint32_t m = 0;
  AM_HIP(hipMemcpyAsync(&m, (uint32_t*)off + n, 4, hipMemcpyDeviceToHost, st));
  AM_HIP(hipStreamSynchronize(st));
  *m_out = (int)m;
  return AM_OK;
}

constexpr int kBoxThreads = 256;
constexpr int kBoxBlocks = 256;
static_assert(kBoxBlocks <= kBoxThreads, "BBoxFinalKernel takes one partial per thread");

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

// One workgroup of kBoxThreads: thread b takes block b's partial box, then a tree in LDS (min and
// max do not depend on the order).
__global__ __launch_bounds__(kBoxThreads) void BBoxFinalKernel(int blocks, const double* __restrict__ part,
                                                               double* __restrict__ box) {
  __shared__ double sh[6][kBoxThreads];
  const int t = threadIdx.x;
  for (int d = 0; d < 3; ++d) {
    sh[d][t] = t < blocks ? part[t * 6 + d] : INFINITY;
    sh[3 + d][t] = t < blocks ? part[t * 6 + 3 + d] : -INFINITY;
  }
  __syncthreads();
  for (int w = kBoxThreads / 2; w > 0; w >>= 1) {
    if (t < w)
      for (int d = 0; d < 3; ++d) {
        sh[d][t] = fmin(sh[d][t], sh[d][t + w]);
        sh[3 + d][t] = fmax(sh[3 + d][t], sh[3 + d][t + w]);
      }
    __syncthreads();
  }
  if (t < 6) box[t] = sh[t][0];
}

int GatheredBBox(am_ctx* ctx, const uint32_t* idx, int m, const double* x, const double* y, const double* z,
                 double* box_dev, const char* tag) {
  void* part;
  int rc = Scratch(ctx, Name(tag, "bboxpart").c_str(), (size_t)kBoxBlocks * 6 * 8, &part);
  if (rc) return rc;
  const int bb = std::min(kBoxBlocks, Blocks(m, kBoxThreads));
  hipLaunchKernelGGL(BBoxPartialKernel, dim3(bb), dim3(kBoxThreads), 0, ctx->stream, m, idx, x, y, z, (double*)part);
  AM_HIP(hipGetLastError());
  hipLaunchKernelGGL(BBoxFinalKernel, dim3(1), dim3(kBoxThreads), 0, ctx->stream, bb, (const double*)part, box_dev);
  AM_HIP(hipGetLastError());
  return AM_OK;
}

// Morton key by repeating the reference descent: octant test against the current centre, then
// move the centre by +-half/2 in FP64 (barnes-hut.cc:30-33,124-136; fast-multipole.cc:234-237,299-307).
__global__ void MortonKernel(int m, const uint32_t* __restrict__ idx, const double* __restrict__ x,
                             const double* __restrict__ y, const double* __restrict__ z,
                             const double* __restrict__ q, const RootBox* __restrict__ root,
                             uint64_t* __restrict__ keys, uint32_t* __restrict__ vals, double4* __restrict__ xq) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= m) return;
  const uint32_t i = idx[k];
  const double px = x[i], py = y[i], pz = z[i];
  xq[i] = make_double4(px, py, pz, q[i]);
  double cx = root->cx, cy = root->cy, cz = root->cz, h = root->half;
  uint64_t key = 0;
  for (int l = 0; l < kOctreeLevels; ++l) {
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

__device__ inline int Digit(uint64_t key, int level) {
  return (int)((key >> (3 * (kOctreeLevels - 1 - level))) & 7u);
}

// The octant starts of a node whose sorted keys are [b, e) (e > b): bd[o] = the first position whose
// digit at depth d is >= o (e if none), for o = 1..7; the digits are non-decreasing over the range.
// The seven lower-bound searches advance together, so each step issues seven independent loads and
// the chain of dependent loads is log2(e - b) long rather than seven searches long.  A finished
// search keeps its result; its load is clamped into the range and ignored.
__device__ inline void OctantStarts(const uint64_t* __restrict__ keys, int b, int e, int d, int (&bd)[9]) {
  int base[8], len[8];
#pragma unroll
  for (int o = 1; o < 8; ++o) {
    base[o] = b;
    len[o] = e - b;
  }
  for (int steps = 32 - __builtin_clz((unsigned)(e - b)); steps > 0; --steps) {
    uint64_t k[8];
#pragma unroll
    for (int o = 1; o < 8; ++o) k[o] = keys[min(base[o] + (len[o] >> 1), e - 1)];
#pragma unroll
    for (int o = 1; o < 8; ++o) {
      const int half = len[o] >> 1;
      const bool lt = Digit(k[o], d) < o;
      const bool live = len[o] > 0;
      base[o] = live && lt ? base[o] + half + 1 : base[o];
      len[o] = live ? (lt ? len[o] - half - 1 : half) : 0;
    }
  }
#pragma unroll
  for (int o = 1; o < 8; ++o) bd[o] = base[o];
}

// The root, and the first two level starts (depth 0 is node 0, depth 1 starts at node 1).
__global__ void RootNodeKernel(NodeArrays nd, const RootBox* __restrict__ root, int m, int* __restrict__ lvl) {
  nd.b[0] = 0;
  nd.e[0] = m;
  nd.depth[0] = 0;
  nd.cx[0] = root->cx;
  nd.cy[0] = root->cy;
  nd.cz[0] = root->cz;
  nd.half[0] = root->half;
  lvl[0] = 0;
  lvl[1] = 1;
}

// One level of the split, depth d (nodes [lvl[d], lvl[d + 1])), in one kernel.  A node holding more
// than leaf_cap particles above the deepest level splits into its non-empty octants (their key
// ranges found by binary search on the sorted keys); the level's children are numbered from
// lvl[d + 1] on, in node order and octant order within a node.  The child offsets come from a scan
// over tiles of 256 nodes: a tile takes the next tile id from a counter (so every tile it waits for
// is already running), publishes its child count, walks back over the earlier tiles' published
// counts to the first published prefix, and publishes its own prefix.  The last tile sets
// lvl[d + 2].  Children past cap are not written: flags[0] is set and the next levels are empty.
// A look-back that waits kLvSpinLimit polls gives up and sets flags[1] (as does a grid smaller than
// the level's tile count, which the launch bound rules out).
constexpr int kLvThreads = 256;
constexpr uint32_t kLvAgg = 1u << 30;
constexpr uint32_t kLvPre = 2u << 30;
constexpr uint32_t kLvVal = kLvAgg - 1;
constexpr int kLvSpinLimit = 1 << 22;

__global__ __launch_bounds__(kLvThreads) void OctreeLevelKernel(int d, int* __restrict__ lvl, NodeArrays nd,
                                                                const uint64_t* __restrict__ keys, int leaf_cap,
                                                                int64_t cap, uint32_t* __restrict__ counter,
                                                                uint32_t* __restrict__ status,
                                                                uint32_t* __restrict__ flags) {
  __shared__ uint32_t wsum[kLvThreads / 64];
  __shared__ uint32_t s_tile, s_excl;
  const int n0 = lvl[d], n1 = lvl[d + 1];
  const int count = n1 - n0;
  const int tiles = (count + kLvThreads - 1) / kLvThreads;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  if (count == 0) {
    if (blockIdx.x == 0 && t == 0) lvl[d + 2] = n1;
    return;
  }
  if (blockIdx.x == 0 && t == 0 && tiles > (int)gridDim.x) atomicOr(&flags[1], 2u);  // grid below the count
  if ((int)blockIdx.x >= tiles) return;
  if (t == 0) s_tile = atomicAdd(counter, 1u);
  __syncthreads();
  const int tile = (int)s_tile;
  const int node = n0 + tile * kLvThreads + t;
  int bd[9];
  uint32_t c = 0;
  bool split = false;
  if (node < n1) {
    const int b = nd.b[node], e = nd.e[node];
    split = e - b > leaf_cap && nd.depth[node] < kOctreeLevels;
    if (split) {
      bd[0] = b;
      OctantStarts(keys, b, e, d, bd);
      bd[8] = e;
      for (int o = 0; o < 8; ++o) c += bd[o + 1] > bd[o] ? 1u : 0u;
    }
  }
  // exclusive scan of the children over the tile (wave shuffles, one barrier)
  uint32_t x = c;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t y = (uint32_t)__shfl_up((int)x, off);
    if (lane >= off) x += y;
  }
  if (lane == 63) wsum[w] = x;
  __syncthreads();
  uint32_t local = x - c, total = 0;
  for (int i = 0; i < kLvThreads / 64; ++i) {
    if (i < w) local += wsum[i];
    total += wsum[i];
  }
  if (t == 0) {
    uint32_t excl = 0;
    uint32_t* mine = status + tile;
    if (tile == 0) {
      __hip_atomic_store(mine, kLvPre | total, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      __hip_atomic_store(mine, kLvAgg | total, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      int tt = tile - 1;
      int spins = 0;
      while (tt >= 0) {
        const uint32_t sw = __hip_atomic_load(status + tt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (sw & kLvPre) {
          excl += sw & kLvVal;
          break;
        }
        if (sw & kLvAgg) {
          excl += sw & kLvVal;
          --tt;
          continue;
        }
        if (++spins > kLvSpinLimit) {
          atomicOr(&flags[1], 1u);
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
      __hip_atomic_store(mine, kLvPre | (excl + total), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    s_excl = excl;
  }
  __syncthreads();
  const int64_t first = (int64_t)n1 + s_excl;  // the tile's first child id
  const bool over = first + total > cap;
  if (over && t == 0) atomicOr(&flags[0], 1u);
  if (node < n1) {
    if (!split) {
      nd.leaf[node] = 1;
      nd.nchild[node] = 0;
      nd.first_child[node] = -1;
    } else if (!over) {
      const int ci0 = (int)(first + local);
      nd.leaf[node] = 0;
      nd.first_child[node] = ci0;
      const double ch = 0.5 * nd.half[node];
      const double px = nd.cx[node], py = nd.cy[node], pz = nd.cz[node];
      const int dep = nd.depth[node] + 1;
      int k = 0;
      for (int o = 0; o < 8; ++o) {
        if (bd[o + 1] == bd[o]) continue;
        const int ci = ci0 + k++;
        nd.b[ci] = bd[o];
        nd.e[ci] = bd[o + 1];
        nd.depth[ci] = dep;
        nd.cx[ci] = px + ((o & 1) ? ch : -ch);
        nd.cy[ci] = py + ((o & 2) ? ch : -ch);
        nd.cz[ci] = pz + ((o & 4) ? ch : -ch);
        nd.half[ci] = ch;
      }
      nd.nchild[node] = k;
    }
  }
  if (tile == tiles - 1 && t == 0) lvl[d + 2] = over ? n1 : (int)(first + total);
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
  if (nd.depth[node] >= kOctreeLevels && e - b > 8) atomicAdd(over, 1u);
}

// The same order by G lanes per node (G >= the leaf capacity, a power of two <= 64): lane l of a
// leaf holding m <= G particles loads the l-th index and counts, by G shuffles within its group,
// the indices below it (they are distinct); that count is its position.  All loads of a group
// precede its stores, since the group is part of one wave.  A leaf longer than G (only at the depth
// limit) is sorted serially by its group's first lane.
template <int G>
__global__ __launch_bounds__(256) void LeafSortGroupKernel(int nn, NodeArrays nd, uint32_t* __restrict__ vals,
                                                           uint32_t* __restrict__ over) {
  static_assert(G >= 2 && G <= 64 && (G & (G - 1)) == 0, "group size");
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int node = (int)(t / G);
  const int l = (int)(threadIdx.x % G);
  const int gbase = (int)(threadIdx.x & 63) & ~(G - 1);
  const bool leaf = node < nn && nd.leaf[node];
  int b = 0, e = 0;
  if (leaf) {
    b = nd.b[node];
    e = nd.e[node];
  }
  const int m = e - b;
  const bool small = leaf && m <= G;
  const uint32_t v = small && l < m ? vals[b + l] : 0xffffffffu;
  int rank = 0;
#pragma unroll
  for (int j = 0; j < G; ++j) rank += __shfl(v, gbase + j) < v ? 1 : 0;
  if (small && l < m) vals[b + rank] = v;
  if (leaf && !small && l == 0) {
    for (int k = b + 1; k < e; ++k) {
      const uint32_t w = vals[k];
      int j = k - 1;
      while (j >= b && vals[j] > w) {
        vals[j + 1] = vals[j];
        --j;
      }
      vals[j + 1] = w;
    }
  }
  if (leaf && l == 0 && nd.depth[node] >= kOctreeLevels && m > 8) atomicAdd(over, 1u);
}

__global__ void ZeroU32Kernel(uint32_t* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

__global__ void ClearU32Kernel(uint32_t* __restrict__ p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0u;
}

__global__ void PreorderKeyKernel(int nn, NodeArrays nd, uint64_t* __restrict__ keys, uint32_t* __restrict__ ids) {
  const int node = blockIdx.x * blockDim.x + threadIdx.x;
  if (node >= nn) return;
  keys[node] = ((uint64_t)nd.b[node] << 5) | (uint64_t)nd.depth[node];
  ids[node] = (uint32_t)node;
}

__global__ void InversePermKernel(int nn, const uint32_t* __restrict__ pre, uint32_t* __restrict__ inv) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < nn) inv[pre[p]] = (uint32_t)p;
}

int BuildOctree(am_ctx* ctx, const uint32_t* particles, int m, const double* x, const double* y, const double* z,
                const double* q, const RootBox* root, int leaf_cap, const char* tag, Octree* t, double4* xq) {
  hipStream_t st = ctx->stream;
  int rc;
  t->m = m;
  t->nn = 0;
  t->level_start.clear();
  if (m == 0) return AM_OK;
  void *keys, *vals;
  if ((rc = Scratch(ctx, Name(tag, "keys").c_str(), (size_t)m * 8, &keys))) return rc;
  if ((rc = Scratch(ctx, Name(tag, "vals").c_str(), (size_t)m * 4, &vals))) return rc;
  hipLaunchKernelGGL(MortonKernel, dim3(Blocks(m, 256)), dim3(256), 0, st, m, particles, x, y, z, q, root,
                     (uint64_t*)keys, (uint32_t*)vals, xq);
  AM_HIP(hipGetLastError());
  if ((rc = RadixSortPairs(ctx, (uint64_t*)keys, (uint32_t*)vals, m, 3 * kOctreeLevels))) return rc;

  size_t cap = (size_t)3 * m + 1024;
  NodeArrays& nd = t->nd;
  const char* names[] = {"nb", "ne", "nd", "nfc", "nnc", "nleaf", "ncx", "ncy", "ncz", "nhalf"};
  const size_t sz[] = {4, 4, 4, 4, 4, 1, 8, 8, 8, 8};
  {  // a previous call's node arrays may be larger
    size_t held = SIZE_MAX;
    for (int a = 0; a < 10; ++a) held = std::min(held, ctx->scratch[Name(tag, names[a])].bytes / sz[a]);
    if (held != SIZE_MAX && held > cap) cap = std::min<size_t>(held, (size_t)kLvVal);
  }
  auto ptrs = [&]() {
    return std::vector<void**>{(void**)&nd.b, (void**)&nd.e, (void**)&nd.depth, (void**)&nd.first_child,
                               (void**)&nd.nchild, (void**)&nd.leaf, (void**)&nd.cx, (void**)&nd.cy,
                               (void**)&nd.cz, (void**)&nd.half};
  };
  {
    auto dst = ptrs();
    for (int a = 0; a < 10; ++a) {
      void* p;
      if ((rc = Scratch(ctx, Name(tag, names[a]).c_str(), cap * sz[a], &p))) return rc;
      *dst[a] = p;
    }
  }
  // The levels run back to back with no host round trip: level d's grid covers an upper bound of its
  // node count (at most 8 per parent, at most one per particle); its real count is read on the
  // device.  Level starts, the overflow and the look-back error come back once at the end; an
  // overflow rebuilds with larger node arrays.  Levels are launched up to one below the previous
  // build's deepest level; if that level still has nodes, the remaining levels follow after the
  // read-back.
  const int max_tiles = (m + kLvThreads - 1) / kLvThreads;
  const int64_t lv_words = 2 + (int64_t)(kOctreeLevels + 1) * (1 + max_tiles);  // flags, counters, status
  void *lvl, *lvw;
  if ((rc = Scratch(ctx, Name(tag, "lvl").c_str(), (kOctreeLevels + 3) * sizeof(int), &lvl))) return rc;
  if ((rc = Scratch(ctx, Name(tag, "lvwords").c_str(), (size_t)lv_words * 4, &lvw))) return rc;
  uint32_t* flags = (uint32_t*)lvw;
  uint32_t* counters = flags + 2;
  uint32_t* status = counters + kOctreeLevels + 1;
  int lv[kOctreeLevels + 3];
  const int hint = ctx->octree_depth[tag];  // 0 before the first build
  for (;;) {
    hipLaunchKernelGGL(ClearU32Kernel, dim3((unsigned)std::min<int64_t>(1024, (lv_words + 255) / 256)), dim3(256), 0,
                       st, (uint32_t*)lvw, lv_words);
    hipLaunchKernelGGL(RootNodeKernel, dim3(1), dim3(1), 0, st, nd, root, m, (int*)lvl);
    AM_HIP(hipGetLastError());
    int64_t ub = 1;
    int d = 0;
    int upto = hint > 0 ? std::min(hint + 1, kOctreeLevels) : kOctreeLevels;
    uint32_t fl[2];
    for (;;) {
      for (; d <= upto; ++d) {
        hipLaunchKernelGGL(OctreeLevelKernel, dim3((unsigned)((ub + kLvThreads - 1) / kLvThreads)), dim3(kLvThreads),
                           0, st, d, (int*)lvl, nd, (const uint64_t*)keys, leaf_cap, (int64_t)cap, counters + d,
                           status + (size_t)d * max_tiles, flags);
        ub = std::min<int64_t>(ub * 8, m);
      }
      AM_HIP(hipGetLastError());
      AM_HIP(hipMemcpyAsync(lv, lvl, sizeof(lv), hipMemcpyDeviceToHost, st));
      AM_HIP(hipMemcpyAsync(fl, flags, sizeof(fl), hipMemcpyDeviceToHost, st));
      AM_HIP(hipStreamSynchronize(st));
      if (fl[1]) return Fail(AM_ERR_HIP, "octree build: a look-back wait on the device gave up");
      // lv[0 .. upto + 2] are set: the levels 0 .. upto + 1.  Go on while level upto + 1 has nodes.
      if (fl[0] || upto == kOctreeLevels || lv[upto + 2] == lv[upto + 1]) break;
      upto = kOctreeLevels;
    }
    if (!fl[0]) break;
    if (cap >= (size_t)kLvVal) return Fail(AM_ERR_CAPACITY, "octree build: too many nodes");
    cap = std::min<size_t>(cap * 2, (size_t)kLvVal);  // rebuild with larger node arrays
    auto dst = ptrs();
    for (int a = 0; a < 10; ++a) {
      void* p;
      if ((rc = Scratch(ctx, Name(tag, names[a]).c_str(), cap * sz[a], &p))) return rc;
      *dst[a] = p;
    }
  }
  t->level_start = {0};
  for (int d = 0; d <= kOctreeLevels + 1 && lv[d + 1] > lv[d]; ++d) t->level_start.push_back(lv[d + 1]);
  ctx->octree_depth[tag] = (int)t->level_start.size() - 2;
  const int total = t->level_start.back();
  const int nn = total;
  void* over;
  if ((rc = Scratch(ctx, Name(tag, "over").c_str(), 4, &over))) return rc;
  hipLaunchKernelGGL(ZeroU32Kernel, dim3(1), dim3(64), 0, st, (uint32_t*)over, 1);
  // Leaves of at most 8 sort fastest on one thread each (14 against 17 us at N = 1M).
  if (leaf_cap > 8 && leaf_cap <= 32)
    hipLaunchKernelGGL(LeafSortGroupKernel<32>, dim3(Blocks((int64_t)nn * 32, 256)), dim3(256), 0, st, nn, nd,
                       (uint32_t*)vals, (uint32_t*)over);
  else
    hipLaunchKernelGGL(LeafSortKernel, dim3(Blocks(nn, 256)), dim3(256), 0, st, nn, nd, (uint32_t*)vals,
                       (uint32_t*)over);
  AM_HIP(hipGetLastError());
  void *pkeys, *pids, *inv;
  if ((rc = Scratch(ctx, Name(tag, "pkeys").c_str(), (size_t)nn * 8, &pkeys))) return rc;
  if ((rc = Scratch(ctx, Name(tag, "pids").c_str(), (size_t)nn * 4, &pids))) return rc;
  if ((rc = Scratch(ctx, Name(tag, "bfs2pre").c_str(), (size_t)nn * 4, &inv))) return rc;
  hipLaunchKernelGGL(PreorderKeyKernel, dim3(Blocks(nn, 256)), dim3(256), 0, st, nn, nd, (uint64_t*)pkeys,
                     (uint32_t*)pids);
  AM_HIP(hipGetLastError());
  if ((rc = RadixSortPairs(ctx, (uint64_t*)pkeys, (uint32_t*)pids, nn, 40))) return rc;
  hipLaunchKernelGGL(InversePermKernel, dim3(Blocks(nn, 256)), dim3(256), 0, st, nn, (const uint32_t*)pids,
                     (uint32_t*)inv);
  AM_HIP(hipGetLastError());
  uint32_t ov = 0, sort_err = 0;
  AM_HIP(hipMemcpyAsync(&ov, over, 4, hipMemcpyDeviceToHost, st));
  if (const uint32_t* w = SortErrorWord(ctx)) AM_HIP(hipMemcpyAsync(&sort_err, w, 4, hipMemcpyDeviceToHost, st));
  AM_HIP(hipStreamSynchronize(st));
  if ((rc = CheckSortErrorAfterSync(ctx, sort_err))) return rc;
  t->nn = nn;
  t->order = (uint32_t*)vals;
  t->pre = (uint32_t*)pids;
  t->bfs2pre = (uint32_t*)inv;
  t->pkeys = (uint64_t*)pkeys;
  t->deep_overflow = ov;
  return AM_OK;
}

}  // namespace am
