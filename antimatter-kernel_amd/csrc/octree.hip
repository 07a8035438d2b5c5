// Count-capped octree in the reference's DFS preorder (see octree.hpp).
#include <algorithm>
#include <cmath>
#include <string>

#include "octree.hpp"

namespace am {

__global__ void KeepFlagKernel(int64_t n, const uint8_t* __restrict__ flags, const double* __restrict__ q, int mode,
                               uint32_t* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const bool alive = !(flags[i] & kFlagAnnihilated);
  out[i] = (alive && (mode == 0 || q[i] != 0.0)) ? 1u : 0u;
}

__global__ void CompactFlagKernel(int64_t n, const uint32_t* __restrict__ keep, const uint32_t* __restrict__ off,
                                  uint32_t* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && keep[i]) out[off[i]] = (uint32_t)i;
}

static std::string Name(const char* tag, const char* what) { return std::string(tag) + "_" + what; }

int GatherParticles(am_ctx* ctx, int64_t n, const uint8_t* flags, const double* q, int mode, uint32_t* out,
                    int* m_out, const char* tag) {
  hipStream_t st = ctx->stream;
  void *keep, *off;
  int rc;
  if ((rc = Scratch(ctx, Name(tag, "keep").c_str(), (size_t)n * 4, &keep))) return rc;
  if ((rc = Scratch(ctx, Name(tag, "keepoff").c_str(), (size_t)(n + 1) * 4, &off))) return rc;
  hipLaunchKernelGGL(KeepFlagKernel, dim3(Blocks(n, 256)), dim3(256), 0, st, n, flags, q, mode, (uint32_t*)keep);
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

__global__ void BBoxFinalKernel(int blocks, const double* __restrict__ part, double* __restrict__ box) {
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int b = 0; b < blocks; ++b)
    for (int d = 0; d < 3; ++d) {
      lo[d] = fmin(lo[d], part[b * 6 + d]);
      hi[d] = fmax(hi[d], part[b * 6 + 3 + d]);
    }
  for (int d = 0; d < 3; ++d) {
    box[d] = lo[d];
    box[3 + d] = hi[d];
  }
}

int GatheredBBox(am_ctx* ctx, const uint32_t* idx, int m, const double* x, const double* y, const double* z,
                 double* box_dev, const char* tag) {
  void* part;
  int rc = Scratch(ctx, Name(tag, "bboxpart").c_str(), (size_t)kBoxBlocks * 6 * 8, &part);
  if (rc) return rc;
  const int bb = std::min(kBoxBlocks, Blocks(m, kBoxThreads));
  hipLaunchKernelGGL(BBoxPartialKernel, dim3(bb), dim3(kBoxThreads), 0, ctx->stream, m, idx, x, y, z, (double*)part);
  AM_HIP(hipGetLastError());
  hipLaunchKernelGGL(BBoxFinalKernel, dim3(1), dim3(1), 0, ctx->stream, bb, (const double*)part, box_dev);
  AM_HIP(hipGetLastError());
  return AM_OK;
}

// Morton key by repeating the reference descent: octant test against the current centre, then
// move the centre by +-half/2 in FP64 (barnes-hut.cc:30-33,124-136; fast-multipole.cc:234-237,299-307).
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

__global__ void RootNodeKernel(NodeArrays nd, const RootBox* __restrict__ root, int m) {
  nd.b[0] = 0;
  nd.e[0] = m;
  nd.depth[0] = 0;
  nd.cx[0] = root->cx;
  nd.cy[0] = root->cy;
  nd.cz[0] = root->cz;
  nd.half[0] = root->half;
}

__global__ void SplitCountKernel(int n0, int n1, NodeArrays nd, const uint64_t* __restrict__ keys, int leaf_cap,
                                 int* __restrict__ bounds, uint32_t* __restrict__ cnt) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int node = n0 + t;
  if (node >= n1) return;
  const int b = nd.b[node], e = nd.e[node], d = nd.depth[node];
  int* bd = bounds + 9 * t;
  if (e - b <= leaf_cap || d >= kOctreeLevels) {
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
  const int c = next_base + (int)off[t];
  nd.first_child[node] = c;
  int k = 0;
  const double ch = 0.5 * nd.half[node];
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
  if (nd.depth[node] >= kOctreeLevels && e - b > 8) atomicAdd(over, 1u);
}

__global__ void ZeroU32Kernel(uint32_t* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
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
                const RootBox* root, int leaf_cap, const char* tag, Octree* t) {
  hipStream_t st = ctx->stream;
  int rc;
  t->m = m;
  t->nn = 0;
  t->level_start.clear();
  if (m == 0) return AM_OK;
  void *keys, *vals;
  if ((rc = Scratch(ctx, Name(tag, "keys").c_str(), (size_t)m * 8, &keys))) return rc;
  if ((rc = Scratch(ctx, Name(tag, "vals").c_str(), (size_t)m * 4, &vals))) return rc;
  hipLaunchKernelGGL(MortonKernel, dim3(Blocks(m, 256)), dim3(256), 0, st, m, particles, x, y, z, root,
                     (uint64_t*)keys, (uint32_t*)vals);
  AM_HIP(hipGetLastError());
  if ((rc = RadixSortPairs(ctx, (uint64_t*)keys, (uint32_t*)vals, m, 3 * kOctreeLevels))) return rc;

  size_t cap = (size_t)3 * m + 1024;
  NodeArrays& nd = t->nd;
  const char* names[] = {"nb", "ne", "nd", "nfc", "nnc", "nleaf", "ncx", "ncy", "ncz", "nhalf"};
  const size_t sz[] = {4, 4, 4, 4, 4, 1, 8, 8, 8, 8};
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
  hipLaunchKernelGGL(RootNodeKernel, dim3(1), dim3(1), 0, st, nd, root, m);
  AM_HIP(hipGetLastError());
  t->level_start = {0, 1};
  int total = 1;
  for (int d = 0; d <= kOctreeLevels; ++d) {
    const int n0 = t->level_start[d], n1 = t->level_start[d + 1];
    const int cnt_nodes = n1 - n0;
    if (cnt_nodes == 0) break;
    void *bounds, *cnt, *coff;
    if ((rc = Scratch(ctx, Name(tag, "bounds").c_str(), (size_t)cnt_nodes * 9 * 4, &bounds))) return rc;
    if ((rc = Scratch(ctx, Name(tag, "cnt").c_str(), (size_t)cnt_nodes * 4, &cnt))) return rc;
    if ((rc = Scratch(ctx, Name(tag, "coff").c_str(), (size_t)(cnt_nodes + 1) * 4, &coff))) return rc;
    hipLaunchKernelGGL(SplitCountKernel, dim3(Blocks(cnt_nodes, 256)), dim3(256), 0, st, n0, n1, nd,
                       (const uint64_t*)keys, leaf_cap, (int*)bounds, (uint32_t*)cnt);
    AM_HIP(hipGetLastError());
    if ((rc = ExclusiveScan(ctx, (const uint32_t*)cnt, (uint32_t*)coff, cnt_nodes))) return rc;
    uint32_t kids = 0;
    AM_HIP(hipMemcpyAsync(&kids, (uint32_t*)coff + cnt_nodes, 4, hipMemcpyDeviceToHost, st));
    AM_HIP(hipStreamSynchronize(st));
    if (kids == 0) break;
    if ((size_t)total + kids > cap) {  // a previous call's buffers may already be large enough
      size_t held = SIZE_MAX;
      for (int a = 0; a < 10; ++a) held = std::min(held, ctx->scratch[Name(tag, names[a])].bytes / sz[a]);
      if ((size_t)total + kids <= held) cap = held;
    }
    if ((size_t)total + kids > cap) {  // grow, preserving the nodes built so far
      const size_t ncap = std::max(cap * 2, (size_t)total + kids + 1024);
      auto dst = ptrs();
      for (int a = 0; a < 10; ++a) {
        void* tmp;
        AM_HIP(hipMalloc(&tmp, (size_t)total * sz[a]));
        AM_HIP(hipMemcpyAsync(tmp, *dst[a], (size_t)total * sz[a], hipMemcpyDeviceToDevice, st));
        void* p;
        AM_HIP(hipStreamSynchronize(st));
        if ((rc = Scratch(ctx, Name(tag, names[a]).c_str(), ncap * sz[a], &p))) return rc;
        AM_HIP(hipMemcpyAsync(p, tmp, (size_t)total * sz[a], hipMemcpyDeviceToDevice, st));
        AM_HIP(hipStreamSynchronize(st));
        AM_HIP(hipFree(tmp));
        *dst[a] = p;
      }
      cap = ncap;
    }
    hipLaunchKernelGGL(WriteChildrenKernel, dim3(Blocks(cnt_nodes, 256)), dim3(256), 0, st, n0, n1, total, nd,
                       (const int*)bounds, (const uint32_t*)coff);
    AM_HIP(hipGetLastError());
    total += (int)kids;
    t->level_start.push_back(total);
  }
  const int nn = total;
  void* over;
  if ((rc = Scratch(ctx, Name(tag, "over").c_str(), 4, &over))) return rc;
  hipLaunchKernelGGL(ZeroU32Kernel, dim3(1), dim3(64), 0, st, (uint32_t*)over, 1);
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
  uint32_t ov = 0;
  AM_HIP(hipMemcpyAsync(&ov, over, 4, hipMemcpyDeviceToHost, st));
  AM_HIP(hipStreamSynchronize(st));
  t->nn = nn;
  t->order = (uint32_t*)vals;
  t->pre = (uint32_t*)pids;
  t->bfs2pre = (uint32_t*)inv;
  t->pkeys = (uint64_t*)pkeys;
  t->deep_overflow = ov;
  return AM_OK;
}

}  // namespace am
