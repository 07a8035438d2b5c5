// Cartesian-Taylor fast multipole method (src/interaction/fast-multipole.cc:17-713) on the device.
//
// Same method, same tree and same operators as the reference; the parallel structure is new:
//   * rescale to the unit cube (fast-multipole.cc:355-390) and the count-capped octree (leaf 32)
//     from Morton keys (octree.hpp), cells numbered in the reference's DFS preorder;
//   * radii bottom-up, P2M (one wave per leaf, one lane per multi-index), M2M bottom-up level by
//     level (one wave per parent), both in the reference's per-coefficient summation order;
//   * the serial dual-tree traversal (:484-526) becomes a level-synchronous expansion of (T, S)
//     pairs: each pair is accepted for M2L ((r_T + r_S)^2 < theta^2 d^2), kept for P2P (two
//     leaves) or split exactly as the reference splits it, so the M2L and P2P sets are the
//     reference's; the lists are then sorted by target cell;
//   * M2L: one wave per target cell; per source cell the derivative tensor D_alpha(t - c) of
//     1/r is built in LDS degree by degree with the reference recurrence (:207-232), then lane
//     gamma contracts L_gamma += (1/gamma!) sum_alpha (-1)^|alpha| M_alpha D_{alpha+gamma};
//   * L2L top-down level by level; L2P fused with P2P (one wave per leaf: 32 lanes per half of
//     the source list), unscaled to newtons (:663-674).
// All FP64.  Summation order differs from the reference only where the traversal order does
// (M2L and P2P list order), so results agree with it to rounding.
#include <algorithm>
#include <array>
#include <cmath>
#include <map>
#include <memory>
#include <vector>

#include "celltable.hpp"
#include "octree.hpp"

namespace am {

constexpr int kFmmMaxOrder = 8;
constexpr int kFmmMaxNm = (kFmmMaxOrder + 1) * (kFmmMaxOrder + 2) * (kFmmMaxOrder + 3) / 6;  // 165

// Device view of the expansion tables (FmmTables, fast-multipole.cc:41-178).
struct FmmTab {
  int P, nm, nd;
  const int* multi;      // [nd][3]
  const double* fact;    // [nd]
  const double* inv_fact;
  const double* sign;
  const int* rec;        // [nd][10]: axis, delta[3], minus1[3], minus2[3]
  const int* deg_start;  // [2P+2]: first index of each total degree
  const int* m2l;        // [nm][nm]
  const int* m2m_off;    // [nm+1] CSR by target coefficient a
  const int* m2m_bc;     // [..][2]: (b, c) in the reference list order
  const int* l2l_off;
  const int* l2l_bc;
  const int* grad;       // [nm][3]
  int n_m2m, n_l2l;      // entries of m2m_bc and l2l_bc (pairs)
};

struct HostTab {
  int P, nm, nd;
  std::vector<int> multi, rec, deg_start, m2l, m2m_off, m2m_bc, l2l_off, l2l_bc, grad;
  std::vector<double> fact, inv_fact, sign;
};

// Tables built on the host with the reference's definitions (fast-multipole.cc:78-177).
static HostTab MakeTables(int p) {
  HostTab t;
  t.P = p;
  const int size = 2 * p + 1;
  std::vector<int> lookup((size_t)size * size * size, -1);
  std::vector<std::array<int, 3>> multi;
  t.deg_start.clear();
  for (int deg = 0; deg <= 2 * p; ++deg) {
    t.deg_start.push_back((int)multi.size());
    for (int a = deg; a >= 0; --a)
      for (int b = deg - a; b >= 0; --b) {
        lookup[((size_t)a * size + b) * size + (deg - a - b)] = (int)multi.size();
        multi.push_back({a, b, deg - a - b});
      }
  }
  t.deg_start.push_back((int)multi.size());
  auto idx = [&](int a, int b, int c) {
    if (a < 0 || b < 0 || c < 0 || a >= size || b >= size || c >= size) return -1;
    return lookup[((size_t)a * size + b) * size + c];
  };
  t.nd = (int)multi.size();
  t.nm = (p + 1) * (p + 2) * (p + 3) / 6;
  for (auto& m : multi) t.multi.insert(t.multi.end(), m.begin(), m.end());
  std::vector<double> f(2 * p + 1, 1.0);
  for (int i = 1; i <= 2 * p; ++i) f[i] = f[i - 1] * i;
  for (int k = 0; k < t.nd; ++k) {
    const auto& m = multi[k];
    const double fk = f[m[0]] * f[m[1]] * f[m[2]];
    t.fact.push_back(fk);
    t.inv_fact.push_back(1.0 / fk);
    t.sign.push_back(((m[0] + m[1] + m[2]) % 2 == 0) ? 1.0 : -1.0);
  }
  t.rec.assign((size_t)t.nd * 10, -1);
  for (int k = 1; k < t.nd; ++k) {
    const auto& m = multi[k];
    int* r = &t.rec[(size_t)k * 10];
    r[0] = m[0] > 0 ? 0 : (m[1] > 0 ? 1 : 2);
    for (int j = 0; j < 3; ++j) {
      r[1 + j] = m[j];
      std::array<int, 3> q = m;
      if (m[j] >= 1) { q[j] -= 1; r[4 + j] = idx(q[0], q[1], q[2]); }
      if (m[j] >= 2) { q[j] -= 1; r[7 + j] = idx(q[0], q[1], q[2]); }
    }
  }
  t.m2l.resize((size_t)t.nm * t.nm);
  for (int g = 0; g < t.nm; ++g)
    for (int a = 0; a < t.nm; ++a)
      t.m2l[(size_t)g * t.nm + a] = idx(multi[a][0] + multi[g][0], multi[a][1] + multi[g][1], multi[a][2] + multi[g][2]);
  std::vector<std::vector<std::array<int, 2>>> m2m(t.nm), l2l(t.nm);
  for (int a = 0; a < t.nm; ++a)
    for (int b = 0; b < t.nm; ++b) {
      if (multi[b][0] <= multi[a][0] && multi[b][1] <= multi[a][1] && multi[b][2] <= multi[a][2])
        m2m[a].push_back({b, idx(multi[a][0] - multi[b][0], multi[a][1] - multi[b][1], multi[a][2] - multi[b][2])});
      if (multi[a][0] <= multi[b][0] && multi[a][1] <= multi[b][1] && multi[a][2] <= multi[b][2])
        l2l[a].push_back({b, idx(multi[b][0] - multi[a][0], multi[b][1] - multi[a][1], multi[b][2] - multi[a][2])});
    }
  t.m2m_off.push_back(0);
  t.l2l_off.push_back(0);
  for (int a = 0; a < t.nm; ++a) {
    for (auto& pr : m2m[a]) { t.m2m_bc.push_back(pr[0]); t.m2m_bc.push_back(pr[1]); }
    for (auto& pr : l2l[a]) { t.l2l_bc.push_back(pr[0]); t.l2l_bc.push_back(pr[1]); }
    t.m2m_off.push_back((int)t.m2m_bc.size() / 2);
    t.l2l_off.push_back((int)t.l2l_bc.size() / 2);
  }
  t.grad.resize((size_t)t.nm * 3);
  for (int g = 0; g < t.nm; ++g)
    for (int k = 0; k < 3; ++k) {
      std::array<int, 3> q = multi[g];
      t.grad[(size_t)g * 3 + k] = q[k] >= 1 ? (q[k] -= 1, idx(q[0], q[1], q[2])) : -1;
    }
  return t;
}

// Upload (cached per context and order) and return the device view.
static int DeviceTables(am_ctx* ctx, int p, FmmTab* out) {
  const HostTab t = MakeTables(p);
  std::vector<char> blob;
  auto put = [&](const void* src, size_t bytes) {
    size_t off = (blob.size() + 15) / 16 * 16;
    blob.resize(off + bytes);
    std::memcpy(blob.data() + off, src, bytes);
    return off;
  };
  const size_t o_multi = put(t.multi.data(), t.multi.size() * 4);
  const size_t o_fact = put(t.fact.data(), t.fact.size() * 8);
  const size_t o_ifact = put(t.inv_fact.data(), t.inv_fact.size() * 8);
  const size_t o_sign = put(t.sign.data(), t.sign.size() * 8);
  const size_t o_rec = put(t.rec.data(), t.rec.size() * 4);
  const size_t o_deg = put(t.deg_start.data(), t.deg_start.size() * 4);
  const size_t o_m2l = put(t.m2l.data(), t.m2l.size() * 4);
  const size_t o_m2mo = put(t.m2m_off.data(), t.m2m_off.size() * 4);
  const size_t o_m2mb = put(t.m2m_bc.data(), t.m2m_bc.size() * 4);
  const size_t o_l2lo = put(t.l2l_off.data(), t.l2l_off.size() * 4);
  const size_t o_l2lb = put(t.l2l_bc.data(), t.l2l_bc.size() * 4);
  const size_t o_grad = put(t.grad.data(), t.grad.size() * 4);
  const std::string name = "fmm_tables_" + std::to_string(p);
  const bool fresh = ctx->scratch.find(name) == ctx->scratch.end();
  void* dev;
  int rc = Scratch(ctx, name.c_str(), blob.size(), &dev);
  if (rc) return rc;
  if (fresh) {
    AM_HIP(hipMemcpyAsync(dev, blob.data(), blob.size(), hipMemcpyHostToDevice, ctx->stream));
    AM_HIP(hipStreamSynchronize(ctx->stream));
  }
  char* d = (char*)dev;
  *out = FmmTab{t.P, t.nm, t.nd, (const int*)(d + o_multi), (const double*)(d + o_fact), (const double*)(d + o_ifact),
                (const double*)(d + o_sign), (const int*)(d + o_rec), (const int*)(d + o_deg),
                (const int*)(d + o_m2l), (const int*)(d + o_m2mo), (const int*)(d + o_m2mb),
                (const int*)(d + o_l2lo), (const int*)(d + o_l2lb), (const int*)(d + o_grad),
                (int)(t.m2m_bc.size() / 2), (int)(t.l2l_bc.size() / 2)};
  return AM_OK;
}

struct FmmFrame {
  double mid[3];
  double length;
  double force_scale;  // k_C / L^2
  double min_sq;       // 1e-30 / L^2
};

// Unit-cube rescaling (fast-multipole.cc:355-390).
__global__ void FrameKernel(const double* __restrict__ box, FmmFrame* __restrict__ fr, RootBox* __restrict__ root) {
  const double lo[3] = {box[0], box[1], box[2]}, hi[3] = {box[3], box[4], box[5]};
  const double ext = fmax(fmax(hi[0] - lo[0], hi[1] - lo[1]), hi[2] - lo[2]);
  const double L = fmax(ext, 1e-15);
  FmmFrame f;
  for (int d = 0; d < 3; ++d) f.mid[d] = 0.5 * (lo[d] + hi[d]);
  f.length = L;
  f.force_scale = kCoulomb / (L * L);
  f.min_sq = kMinSquaredDistance / (L * L);
  *fr = f;
  RootBox r;
  r.cx = 0.0;
  r.cy = 0.0;
  r.cz = 0.0;
  r.half = 0.5 * (1.0 + 1e-9);
  *root = r;
}

__global__ void RescaleKernel(int m, const uint32_t* __restrict__ idx, const double* __restrict__ x,
                              const double* __restrict__ y, const double* __restrict__ z,
                              const FmmFrame* __restrict__ fr, double* __restrict__ rx, double* __restrict__ ry,
                              double* __restrict__ rz) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= m) return;
  const uint32_t i = idx[k];
  const double L = fr->length;
  rx[i] = (x[i] - fr->mid[0]) / L;
  ry[i] = (y[i] - fr->mid[1]) / L;
  rz[i] = (z[i] - fr->mid[2]) / L;
}

// Tree-ordered particle copies (fast-multipole.cc:401-410).  xq: the rescaled (x, y, z, q) by
// particle index, one record each (written by the Morton pass).
__global__ void SortedCopyKernel(int m, const uint32_t* __restrict__ order, const double4* __restrict__ xq,
                                 double4* __restrict__ sp) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= m) return;
  sp[k] = xq[order[k]];
}

// FP32 hi/lo source mirror in tree order (for the FP32 near field).
__global__ void SortedSrcKernel(int m, const uint32_t* __restrict__ order, const float4* __restrict__ src,
                                float4* __restrict__ out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= m) return;
  const uint32_t i = order[k];
  out[2 * k] = src[2 * i];
  out[2 * k + 1] = src[2 * i + 1];
}

struct Cell {  // preorder cell record
  double cx, cy, cz, radius;
  double half;  // cube half width (the root's times 2^-depth, exactly)
  int b, e, leaf, nchild;
  int child[8];
};

// Radii bottom-up (fast-multipole.cc:412-438), in BFS order.
__global__ void RadiusKernel(int n0, int n1, NodeArrays nd, const double4* __restrict__ sp, double* __restrict__ rad) {
  const int node = n0 + blockIdx.x * blockDim.x + threadIdx.x;
  if (node >= n1) return;
  const double cx = nd.cx[node], cy = nd.cy[node], cz = nd.cz[node];
  if (nd.leaf[node]) {
    double r2 = 0.0;
    for (int k = nd.b[node]; k < nd.e[node]; ++k) {
      const double dx = sp[k].x - cx, dy = sp[k].y - cy, dz = sp[k].z - cz;
      r2 = fmax(r2, dx * dx + dy * dy + dz * dz);
    }
    rad[node] = sqrt(r2);
  } else {
    double r = 0.0;
    const int c0 = nd.first_child[node];
    for (int c = c0; c < c0 + nd.nchild[node]; ++c) {
      const double dx = nd.cx[c] - cx, dy = nd.cy[c] - cy, dz = nd.cz[c] - cz;
      r = fmax(r, sqrt(dx * dx + dy * dy + dz * dz) + rad[c]);
    }
    rad[node] = r;
  }
}

__global__ void GatherCellsKernel(int nn, NodeArrays nd, const uint32_t* __restrict__ pre,
                                  const uint32_t* __restrict__ bfs2pre, const double* __restrict__ rad,
                                  Cell* __restrict__ cells, double4* __restrict__ geo, double4* __restrict__ trv,
                                  uint32_t* __restrict__ tinfo) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nn) return;
  const int node = (int)pre[p];
  Cell c;
  c.cx = nd.cx[node];
  c.cy = nd.cy[node];
  c.cz = nd.cz[node];
  c.radius = rad[node];
  c.half = nd.half[node];
  c.b = nd.b[node];
  c.e = nd.e[node];
  c.leaf = nd.leaf[node];
  c.nchild = nd.nchild[node];
  for (int k = 0; k < 8; ++k) c.child[k] = k < c.nchild ? (int)bfs2pre[nd.first_child[node] + k] : -1;
  cells[p] = c;
  geo[p] = make_double4(c.cx, c.cy, c.cz, c.half);
  trv[p] = make_double4(c.cx, c.cy, c.cz, c.radius);
  tinfo[p] = (c.leaf ? 1u : 0u) | ((uint32_t)c.nchild << 1);
}

__device__ inline double Monomial(const FmmTab& t, int k, double x, double y, double z) {
  const int a = t.multi[3 * k], b = t.multi[3 * k + 1], c = t.multi[3 * k + 2];
  double px = 1.0, py = 1.0, pz = 1.0;
  for (int i = 0; i < a; ++i) px = px * x;
  for (int i = 0; i < b; ++i) py = py * y;
  for (int i = 0; i < c; ++i) pz = pz * z;
  return px * py * pz;
}

// Monomial x^a y^b z^c of multi-index k from a copy of the multi table (e.g. in LDS), with the
// same products as Monomial.
__device__ inline double MonomialAt(const int* __restrict__ multi, int k, double x, double y, double z) {
  const int a = multi[3 * k], b = multi[3 * k + 1], c = multi[3 * k + 2];
  double px = 1.0, py = 1.0, pz = 1.0;
  for (int i = 0; i < a; ++i) px = px * x;
  for (int i = 0; i < b; ++i) py = py * y;
  for (int i = 0; i < c; ++i) pz = pz * z;
  return px * py * pz;
}

// P2M (fast-multipole.cc:443-461): one wave per cell, lane = coefficient.
__global__ __launch_bounds__(64) void P2MKernel(int nn, const Cell* __restrict__ cells, const double4* __restrict__ sp,
                                                FmmTab t, double* __restrict__ M) {
  __shared__ int s_multi[3 * kFmmMaxNm];
  const int p = blockIdx.x;
  const Cell c = cells[p];
  const int a = threadIdx.x;
  if (!c.leaf) {
    for (int k = a; k < t.nm; k += 64) M[(size_t)p * t.nm + k] = 0.0;
    return;
  }
  for (int k = a; k < 3 * t.nm; k += 64) s_multi[k] = t.multi[k];
  __syncthreads();
  for (int k = a; k < t.nm; k += 64) {
    const double ifk = t.inv_fact[k];
    double acc = 0.0;
    for (int j = c.b; j < c.e; ++j) {
      const double4 s = sp[j];
      if (s.w == 0.0) continue;
      acc = acc + s.w * MonomialAt(s_multi, k, s.x - c.cx, s.y - c.cy, s.z - c.cz) * ifk;
    }
    M[(size_t)p * t.nm + k] = acc;
  }
}

// Dynamic LDS of M2MKernel and L2LKernel: the tables their inner loops read, staged once per
// workgroup (multi, fact, inv_fact of the first nm indices, the CSR offsets and the (b, c) pairs).
inline size_t TransferLds(const FmmTab& t, bool l2l) {
  return sizeof(double) * 2 * (size_t)t.nm +
         sizeof(int) * (3 * (size_t)t.nm + (size_t)t.nm + 1 + 2 * (size_t)(l2l ? t.n_l2l : t.n_m2m));
}
struct TransferTabs {
  double *fact, *ifact;
  int *multi, *off, *bc;
};
__device__ inline TransferTabs StageTransferTabs(const FmmTab& t, bool l2l) {
  extern __shared__ double tlds[];
  TransferTabs r;
  r.fact = tlds;
  r.ifact = tlds + t.nm;
  r.multi = reinterpret_cast<int*>(tlds + 2 * t.nm);
  r.off = r.multi + 3 * t.nm;
  r.bc = r.off + t.nm + 1;
  const int* off = l2l ? t.l2l_off : t.m2m_off;
  const int* bc = l2l ? t.l2l_bc : t.m2m_bc;
  const int nbc = 2 * (l2l ? t.n_l2l : t.n_m2m);
  for (int k = threadIdx.x; k < t.nm; k += blockDim.x) {
    r.fact[k] = t.fact[k];
    r.ifact[k] = t.inv_fact[k];
  }
  for (int k = threadIdx.x; k < 3 * t.nm; k += blockDim.x) r.multi[k] = t.multi[k];
  for (int k = threadIdx.x; k <= t.nm; k += blockDim.x) r.off[k] = off[k];
  for (int k = threadIdx.x; k < nbc; k += blockDim.x) r.bc[k] = bc[k];
  return r;
}

// M2M (fast-multipole.cc:463-482) for the parents of one BFS level.
__global__ __launch_bounds__(64) void M2MKernel(int n0, NodeArrays nd, const uint32_t* __restrict__ bfs2pre,
                                                FmmTab t, double* __restrict__ M) {
  __shared__ double mono[kFmmMaxNm];
  const int node = n0 + blockIdx.x;
  if (nd.leaf[node]) return;
  const TransferTabs tt = StageTransferTabs(t, false);
  const int p = (int)bfs2pre[node];
  const int c0 = nd.first_child[node];
  double acc[3] = {0.0, 0.0, 0.0};  // up to 3 coefficients per lane (nm <= 165 < 192)
  for (int k = 0; k < 3; ++k) {
    const int a = threadIdx.x + 64 * k;
    if (a < t.nm) acc[k] = M[(size_t)p * t.nm + a];
  }
  for (int c = c0; c < c0 + nd.nchild[node]; ++c) {
    const double dx = nd.cx[c] - nd.cx[node], dy = nd.cy[c] - nd.cy[node], dz = nd.cz[c] - nd.cz[node];
    __syncthreads();
    for (int k = threadIdx.x; k < t.nm; k += 64) mono[k] = MonomialAt(tt.multi, k, dx, dy, dz);
    __syncthreads();
    const double* mc = M + (size_t)bfs2pre[c] * t.nm;
    for (int k = 0; k < 3; ++k) {
      const int a = threadIdx.x + 64 * k;
      if (a >= t.nm) continue;
      double v = acc[k];
      for (int q = tt.off[a]; q < tt.off[a + 1]; ++q) {
        const int b = tt.bc[2 * q], cc = tt.bc[2 * q + 1];
        v = v + mc[b] * mono[cc] * tt.ifact[cc];
      }
      acc[k] = v;
    }
  }
  for (int k = 0; k < 3; ++k) {
    const int a = threadIdx.x + 64 * k;
    if (a < t.nm) M[(size_t)p * t.nm + a] = acc[k];
  }
}

// One expansion step of the dual-tree traversal (fast-multipole.cc:494-525): pair i of the frontier
// is accepted for M2L, kept for P2P, or split.  The tile kernels below classify each 1024-pair tile's
// pairs and count its three outputs, scan the tile counts, and write every output at its tile's
// offset plus a workgroup scan.  The decision reads the traversal records trv[c] = (centre, radius)
// and tinfo[c] = leaf | nchild << 1 (72 bytes per pair instead of two whole cell records); children
// come from the cell record only for split pairs.
__device__ inline int TraverseKind(const double4 T, const double4 S, uint32_t ti, uint32_t si, double theta2) {
  const double dx = T.x - S.x, dy = T.y - S.y, dz = T.z - S.z;
  const double dist2 = dx * dx + dy * dy + dz * dz;
  const double rs = T.w + S.w;
  if (rs * rs < theta2 * dist2) return 0;  // M2L
  const bool t_leaf = ti & 1u, s_leaf = si & 1u;
  if (t_leaf && s_leaf) return 1;  // P2P
  return (s_leaf || (!t_leaf && T.w >= S.w)) ? 2 : 3;  // split the target / the source
}

// One traversal step over the frontier pairs in tiles of kTrvTile pairs (item u of thread t is pair
// tile * kTrvTile + u * kTrvThreads + t, so loads are coalesced).  Each pair yields an M2L pair, a
// P2P pair, or the pairs of the split cell's children; the three outputs keep the frontier order.
constexpr int kTrvThreads = 256;
constexpr int kTrvItems = 4;
constexpr int kTrvTile = kTrvThreads * kTrvItems;

__device__ inline uint32_t TraverseNew(int kind, uint32_t ti, uint32_t si) {
  return kind == 2 ? (ti >> 1) : (kind == 3 ? (si >> 1) : 0u);
}

// Block-wide exclusive scan of three counts in thread order (blocks of THREADS); returns the block
// totals in tot.  Two barriers; wsum may be reused after the second.
template <int THREADS = kTrvThreads>
__device__ inline void TraverseBlockScan(uint32_t (&c)[3], uint32_t (&ex)[3], uint32_t (&tot)[3],
                                         uint32_t (*wsum)[THREADS / 64]) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint32_t x[3] = {c[0], c[1], c[2]};
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const uint32_t y = (uint32_t)__shfl_up((int)x[k], off);
      if (lane >= off) x[k] += y;
    }
  }
  if (lane == 63)
#pragma unroll
    for (int k = 0; k < 3; ++k) wsum[k][w] = x[k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    uint32_t before = 0, all = 0;
    for (int v = 0; v < THREADS / 64; ++v) {
      if (v < w) before += wsum[k][v];
      all += wsum[k][v];
    }
    ex[k] = before + x[k] - c[k];
    tot[k] = all;
  }
  __syncthreads();
}

// Classifies the tile's pairs (kind per pair, kept for the write) and counts its outputs.
__global__ __launch_bounds__(kTrvThreads) void TraverseTileCountKernel(int np, const uint2* __restrict__ pairs,
                                                                       const double4* __restrict__ trv,
                                                                       const uint32_t* __restrict__ tinfo,
                                                                       double theta2, uint8_t* __restrict__ kinds,
                                                                       uint32_t* __restrict__ tile_cnt) {
  __shared__ uint32_t wsum[3][kTrvThreads / 64];
  uint32_t c[3] = {0u, 0u, 0u};
#pragma unroll
  for (int u = 0; u < kTrvItems; ++u) {
    const int64_t i = (int64_t)blockIdx.x * kTrvTile + u * kTrvThreads + threadIdx.x;
    if (i < np) {
      const uint2 pr = pairs[i];
      const uint32_t ti = tinfo[pr.x], si = tinfo[pr.y];
      const int kind = TraverseKind(trv[pr.x], trv[pr.y], ti, si, theta2);
      kinds[i] = (uint8_t)kind;
      c[0] += kind == 0 ? 1u : 0u;
      c[1] += kind == 1 ? 1u : 0u;
      c[2] += TraverseNew(kind, ti, si);
    }
  }
  uint32_t ex[3], tot[3];
  TraverseBlockScan(c, ex, tot, wsum);
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < 3; ++k) tile_cnt[3 * (size_t)blockIdx.x + k] = tot[k];
}

// Exclusive scans over the tiles of each output's counts (one block of kTrvScanThreads); tot = the
// three totals.
constexpr int kTrvScanThreads = 1024;
__global__ __launch_bounds__(kTrvScanThreads) void TraverseTileScanKernel(int64_t ntiles,
                                                                          const uint32_t* __restrict__ tile_cnt,
                                                                          uint32_t* __restrict__ tile_off,
                                                                          uint32_t* __restrict__ tot) {
  __shared__ uint32_t wsum[3][kTrvScanThreads / 64];
  uint32_t run[3] = {0u, 0u, 0u};
  for (int64_t t0 = 0; t0 < ntiles; t0 += kTrvScanThreads) {
    const int64_t t = t0 + threadIdx.x;
    uint32_t c[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) c[k] = t < ntiles ? tile_cnt[3 * t + k] : 0u;
    uint32_t ex[3], all[3];
    TraverseBlockScan<kTrvScanThreads>(c, ex, all, wsum);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      if (t < ntiles) tile_off[3 * t + k] = run[k] + ex[k];
      run[k] += all[k];
    }
  }
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < 3; ++k) tot[k] = run[k];
}

// Writes the tile's outputs at the tile's offsets plus the exclusive counts of the earlier pairs of
// the tile, item round by item round (the frontier order).
__global__ __launch_bounds__(kTrvThreads) void TraverseTileWriteKernel(int np, const uint2* __restrict__ pairs,
                                                                       const uint8_t* __restrict__ kinds,
                                                                       const uint32_t* __restrict__ tinfo,
                                                                       const Cell* __restrict__ cells,
                                                                       const uint32_t* __restrict__ tile_off,
                                                                       uint64_t* __restrict__ m2l_out,
                                                                       uint64_t* __restrict__ p2p_out,
                                                                       uint2* __restrict__ next) {
  __shared__ uint32_t wsum[3][kTrvThreads / 64];
  uint32_t base[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) base[k] = tile_off[3 * (size_t)blockIdx.x + k];
#pragma unroll 1
  for (int u = 0; u < kTrvItems; ++u) {
    const int64_t i = (int64_t)blockIdx.x * kTrvTile + u * kTrvThreads + threadIdx.x;
    int kind = -1;
    uint2 pr = make_uint2(0u, 0u);
    uint32_t c[3] = {0u, 0u, 0u};
    if (i < np) {
      pr = pairs[i];
      kind = kinds[i];
      c[0] = kind == 0 ? 1u : 0u;
      c[1] = kind == 1 ? 1u : 0u;
      if (kind >= 2) c[2] = TraverseNew(kind, tinfo[pr.x], tinfo[pr.y]);
    }
    uint32_t ex[3], tot[3];
    TraverseBlockScan(c, ex, tot, wsum);
    const uint64_t key = ((uint64_t)pr.x << 32) | pr.y;
    if (kind == 0) {
      m2l_out[base[0] + ex[0]] = key;
    } else if (kind == 1) {
      p2p_out[base[1] + ex[1]] = key;
    } else if (kind >= 2) {
      uint2* out = next + base[2] + ex[2];
      if (kind == 2) {
        const Cell& T = cells[pr.x];
        for (int k = 0; k < T.nchild; ++k) out[k] = make_uint2((uint32_t)T.child[k], pr.y);
      } else {
        const Cell& S = cells[pr.y];
        for (int k = 0; k < S.nchild; ++k) out[k] = make_uint2(pr.x, (uint32_t)S.child[k]);
      }
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) base[k] += tot[k];
  }
}

// CSR offsets of target cells in a pair list (T << 32 | S) grouped by ascending T.
__global__ void PairOffsetsKernel(int nn, const uint64_t* __restrict__ keys, int64_t np, int64_t* __restrict__ off) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t > nn) return;
  const uint64_t k = (uint64_t)t << 32;
  int64_t lo = 0, hi = np;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (keys[mid] < k) lo = mid + 1;
    else hi = mid;
  }
  off[t] = lo;
}

// ---- M2L (fast-multipole.cc:528-555) by displacement class ---------------------------------------
// Cell centres sit on the octree lattice, so the displacement between the two cells of an M2L pair
// is an integer multiple q of the finer cell's half width h_f = root_half * 2^-d.  The key (d, q)
// fixes the displacement exactly, and the derivative tensor of 1/r depends on nothing else.  Pairs
// are classed by key in an open-addressing table, each class's tensor is built once, and the
// per-pair work is the contraction sum_alpha (-1)^|alpha| M_alpha D_{alpha+gamma}.  A pair whose key
// does not fit (|q| >= 2^18) or that finds no slot within the probe limit builds its own tensor.
constexpr int kQBits = 19;
constexpr int64_t kQOff = (int64_t)1 << (kQBits - 1);
constexpr uint64_t kQMask = ((uint64_t)1 << kQBits) - 1;
constexpr uint64_t kClassEmpty = ~0ull;
constexpr uint32_t kNoClass = 0xffffffffu;
constexpr int kClassProbes = 64;
constexpr int kM2LWaves = 4;     // waves per workgroup, one target cell each
constexpr int kM2LLdsOrder = 5;  // orders whose index tables fit in LDS beside the wave slices

__device__ inline void WaveSync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Key (d: 5 bits, q: 3 x 19 bits) of the displacement T - S; false if it does not fit.  T and S are
// (centre, half width) records.
__device__ inline bool M2LClassKey(const double4 T, const double4 S, int root_exp, uint64_t* key) {
  const double hf = fmin(T.w, S.w);
  const int d = root_exp - ilogb(hf);
  if (d < 0 || d > 31) return false;
  const double qd[3] = {(T.x - S.x) / hf, (T.y - S.y) / hf, (T.z - S.z) / hf};
  uint64_t k = (uint64_t)d;
  for (int a = 0; a < 3; ++a) {
    const double r = rint(qd[a]);
    if (!(fabs(r) < (double)kQOff) || fabs(qd[a] - r) > 1e-6) return false;
    k = (k << kQBits) | (uint64_t)((int64_t)r + kQOff);
  }
  *key = k;
  return true;
}

// The displacement of a key: q * h_f, exact.
__device__ inline void M2LClassDisplacement(uint64_t key, double root_half, double xs[3]) {
  const int d = (int)(key >> (3 * kQBits));
  const double hf = ldexp(root_half, -d);
  xs[0] = (double)((int64_t)((key >> (2 * kQBits)) & kQMask) - kQOff) * hf;
  xs[1] = (double)((int64_t)((key >> kQBits) & kQMask) - kQOff) * hf;
  xs[2] = (double)((int64_t)(key & kQMask) - kQOff) * hf;
}

// Empties the table and zeroes the unclassed-pair count.
__global__ void ClassClearKernel(uint64_t size, uint64_t* __restrict__ hkeys, uint32_t* __restrict__ n_uncl) {
  const uint64_t h = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (h < size) hkeys[h] = kClassEmpty;
  if (h == 0) *n_uncl = 0u;
}

// Inserts every pair's key; pair_slot[q] = its table slot, or kNoClass (counted in *n_uncl).  geo
// holds each cell's (centre, half width): 32 bytes instead of the whole cell record per lookup.
// max_probes = 0 leaves every pair unclassed (a diagnostic of the unclassed paths).
__global__ void M2LClassInsertKernel(int64_t np, const uint64_t* __restrict__ pairs, const double4* __restrict__ geo,
                                     uint64_t* __restrict__ hkeys, uint64_t mask, int max_probes,
                                     uint32_t* __restrict__ pair_slot, uint32_t* __restrict__ n_uncl) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= np) return;
  const uint64_t pk = pairs[q];
  const int root_exp = ilogb(geo[0].w);
  uint64_t key;
  uint32_t slot = kNoClass;
  if (M2LClassKey(geo[pk >> 32], geo[(uint32_t)pk], root_exp, &key)) {
    uint64_t h = CellHashMix(key) & mask;
    for (int probe = 0; probe < max_probes; ++probe) {
      const uint64_t cur = hkeys[h];
      if (cur == key) {
        slot = (uint32_t)h;
        break;
      }
      if (cur == kClassEmpty) {
        const unsigned long long prev =
            atomicCAS(reinterpret_cast<unsigned long long*>(hkeys + h), (unsigned long long)kClassEmpty,
                      (unsigned long long)key);
        if (prev == kClassEmpty || prev == key) {
          slot = (uint32_t)h;
          break;
        }
      }
      h = (h + 1) & mask;
    }
  }
  pair_slot[q] = slot;
  if (slot == kNoClass) atomicAdd(n_uncl, 1u);
}

__global__ void ClassOccupiedKernel(uint64_t size, const uint64_t* __restrict__ hkeys, uint32_t* __restrict__ occ) {
  const uint64_t h = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (h < size) occ[h] = hkeys[h] != kClassEmpty ? 1u : 0u;
}

// ids[h] (exclusive scan of occupancy) numbers the classes in slot order; cls_key[id] = key.
__global__ void ClassListKernel(uint64_t size, const uint64_t* __restrict__ hkeys, const uint32_t* __restrict__ ids,
                                uint64_t* __restrict__ cls_key) {
  const uint64_t h = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (h < size && hkeys[h] != kClassEmpty) cls_key[ids[h]] = hkeys[h];
}

// Class id of every pair (kNoClass for unclassed pairs).
__global__ void PairClassKernel(int64_t np, const uint32_t* __restrict__ pair_slot, const uint32_t* __restrict__ ids,
                                uint32_t* __restrict__ pair_cls) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= np) return;
  const uint32_t slot = pair_slot[q];
  pair_cls[q] = slot == kNoClass ? kNoClass : ids[slot];
}

// Derivative tensor D_k (k < nd) of 1/r at xs into the wave's LDS slice, degree by degree (the
// recurrence of fast-multipole.cc:207-232).
__device__ inline void DerivTensorWave(const double xs[3], const FmmTab& t, const int* __restrict__ rec,
                                       double* __restrict__ Dw) {
  const int lane = threadIdx.x & 63;
  const double inv_r2 = 1.0 / (xs[0] * xs[0] + xs[1] * xs[1] + xs[2] * xs[2]);
  if (lane == 0) Dw[0] = sqrt(inv_r2);
  for (int g = 1; g <= 2 * t.P; ++g) {
    WaveSync();
    for (int k = t.deg_start[g] + lane; k < t.deg_start[g + 1]; k += 64) {
      const int* rr = rec + 10 * k;
      const int i = rr[0];
      const double xi = i == 0 ? xs[0] : (i == 1 ? xs[1] : xs[2]);
      double sum = (2 * rr[1 + i] - 1) * xi * Dw[rr[4 + i]];
      if (rr[7 + i] >= 0) sum += (rr[1 + i] - 1) * (rr[1 + i] - 1) * Dw[rr[7 + i]];
      for (int j = 0; j < 3; ++j) {
        if (j == i || rr[1 + j] == 0) continue;
        sum += 2.0 * rr[1 + j] * xs[j] * Dw[rr[4 + j]];
        if (rr[7 + j] >= 0) sum += rr[1 + j] * (rr[1 + j] - 1) * Dw[rr[7 + j]];
      }
      Dw[k] = -sum * inv_r2;
    }
  }
  WaveSync();
}

// One wave per class: the tensor of the class's displacement into Dc[class][nd].
__global__ __launch_bounds__(64) void ClassTensorKernel(int n_cls, const uint64_t* __restrict__ cls_key,
                                                        const Cell* __restrict__ cells, FmmTab t,
                                                        double* __restrict__ Dc) {
  extern __shared__ double lds[];
  const int c = blockIdx.x;
  if (c >= n_cls) return;
  double xs[3];
  M2LClassDisplacement(cls_key[c], cells[0].half, xs);
  DerivTensorWave(xs, t, t.rec, lds);
  for (int k = threadIdx.x; k < t.nd; k += 64) Dc[(size_t)c * t.nd + k] = lds[k];
}

// One wave per target cell (kM2LWaves per workgroup), sources of its list in order.  Per pair the
// class tensor (or, unclassed, the pair's own) and the multipole are staged in the wave's LDS slice
// and lane gamma contracts; the next pair's tensor and multipole are fetched into registers while
// the current one is contracted.  Per-target sums are in list order: deterministic.
// DPL: tensor entries per lane, ceil(nd / 64) rounded up to 3, 5, 8 or 16 (orders 4, 5, 6, 8).
constexpr int kMPerLane = (kFmmMaxNm + 63) / 64;  // 3
constexpr int kRegNm = 36;  // register contraction for nm <= 35 (orders <= 4), padded to 4

__device__ __forceinline__ double ReadLaneF64(double v, int src_lane) {
  const uint64_t u = __double_as_longlong(v);
  const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)u, src_lane);
  const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(u >> 32), src_lane);
  return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

template <int DPL, bool REG>
__global__ __launch_bounds__(64 * kM2LWaves) void M2LKernel(int nn, const Cell* __restrict__ cells,
                                                             const uint64_t* __restrict__ list,
                                                             const int64_t* __restrict__ off,
                                                             const uint32_t* __restrict__ pair_cls,
                                                             const double* __restrict__ Dc, FmmTab t,
                                                             const double* __restrict__ M, double* __restrict__ L) {
  // Dynamic LDS: per wave D[nd] and Ms[nm] (doubles); sign[nm]; then, for orders <= kM2LLdsOrder,
  // the m2l index table [nm][nm] and the recurrence table [nd][10] (ints).
  extern __shared__ double lds[];
  const int nm = t.nm, nd = t.nd;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  double* s_sign = lds + kM2LWaves * (nd + nm);
  int* s_m2l = (int*)(s_sign + nm);
  int* s_rec = s_m2l + nm * nm;
  const bool lds_tab = t.P <= kM2LLdsOrder;
  if (lds_tab) {
    for (int k = threadIdx.x; k < nm * nm; k += 64 * kM2LWaves) s_m2l[k] = t.m2l[k];
    for (int k = threadIdx.x; k < nd * 10; k += 64 * kM2LWaves) s_rec[k] = t.rec[k];
  }
  for (int k = threadIdx.x; k < nm; k += 64 * kM2LWaves) s_sign[k] = t.sign[k];
  __syncthreads();  // the only workgroup barrier: waves run independently from here
  const int p = blockIdx.x * kM2LWaves + wave;
  if (p >= nn) return;
  const int* m2l = lds_tab ? s_m2l : t.m2l;
  const int* rec = lds_tab ? s_rec : t.rec;
  double* Dw = lds + wave * (nd + nm);
  double* Mw = Dw + nd;
  const Cell T = cells[p];
  double acc[3] = {0.0, 0.0, 0.0};
  const int64_t b = off[p], e = off[p + 1];
  double dreg[DPL], mreg[kMPerLane];
  int ixr[REG ? kRegNm : 1];
  if (REG)
#pragma unroll
    for (int al = 0; al < (REG ? kRegNm : 1); ++al) ixr[al] = (lane < nm && al < nm) ? m2l[lane * nm + al] : 0;
  double msig = 0.0;  // REG: this lane's signed multipole coefficient of the current pair
  int s_cur = 0;
  uint32_t c_cur = kNoClass;
  auto fetch = [&](int64_t q, int& s, uint32_t& c) {
    s = (int)(uint32_t)list[q];
    c = pair_cls[q];
#pragma unroll
    for (int r = 0; r < kMPerLane; ++r) {
      const int k = lane + 64 * r;
      mreg[r] = k < nm ? M[(size_t)s * nm + k] : 0.0;
    }
    if (c != kNoClass) {
      const double* src = Dc + (size_t)c * nd;
#pragma unroll
      for (int r = 0; r < DPL; ++r) {
        const int k = lane + 64 * r;
        if (k < nd) dreg[r] = src[k];
      }
    }
  };
  if (b < e) fetch(b, s_cur, c_cur);
  for (int64_t q = b; q < e; ++q) {
    if (REG) {
      msig = lane < nm ? s_sign[lane] * mreg[0] : 0.0;
    } else {
#pragma unroll
      for (int r = 0; r < kMPerLane; ++r) {
        const int k = lane + 64 * r;
        if (k < nm) Mw[k] = s_sign[k] * mreg[r];
      }
    }
    if (c_cur != kNoClass) {
#pragma unroll
      for (int r = 0; r < DPL; ++r) {
        const int k = lane + 64 * r;
        if (k < nd) Dw[k] = dreg[r];
      }
      WaveSync();
    } else {
      const Cell& S = cells[s_cur];
      const double xs[3] = {T.cx - S.cx, T.cy - S.cy, T.cz - S.cz};
      DerivTensorWave(xs, t, rec, Dw);
    }
    if (q + 1 < e) fetch(q + 1, s_cur, c_cur);  // in flight during the contraction
    if (REG) {
      // nm <= 35: lane g holds its index row in registers; multipole coefficient al is broadcast
      // from lane al (after the sign) with readlane, so only the tensor gather touches LDS.
      if (lane < nm) {
        double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
#pragma unroll
        for (int al = 0; al < kRegNm; al += 4) {
          if (al + 0 < nm) a0 += ReadLaneF64(msig, al + 0) * Dw[ixr[al + 0]];
          if (al + 1 < nm) a1 += ReadLaneF64(msig, al + 1) * Dw[ixr[al + 1]];
          if (al + 2 < nm) a2 += ReadLaneF64(msig, al + 2) * Dw[ixr[al + 2]];
          if (al + 3 < nm) a3 += ReadLaneF64(msig, al + 3) * Dw[ixr[al + 3]];
        }
        acc[0] += ((a0 + a1) + (a2 + a3)) * t.inv_fact[lane];
      }
      WaveSync();
      continue;
    }
    for (int k = 0; k < 3; ++k) {
      const int g = lane + 64 * k;
      if (g >= nm) break;
      const int* ix = m2l + g * nm;
      double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
      int al = 0;
      for (; al + 4 <= nm; al += 4) {
        a0 += Mw[al] * Dw[ix[al]];
        a1 += Mw[al + 1] * Dw[ix[al + 1]];
        a2 += Mw[al + 2] * Dw[ix[al + 2]];
        a3 += Mw[al + 3] * Dw[ix[al + 3]];
      }
      for (; al < nm; ++al) a0 += Mw[al] * Dw[ix[al]];
      acc[k] += ((a0 + a1) + (a2 + a3)) * t.inv_fact[g];
    }
    WaveSync();  // the slice is rewritten by the next pair
  }
  for (int k = 0; k < 3; ++k) {
    const int g = lane + 64 * k;
    if (g < nm) L[(size_t)p * nm + g] = acc[k];
  }
}

// FP32 tensors of the classes (the FP32-precision path).
__global__ void TensorToF32Kernel(int64_t n, const double* __restrict__ Dc, float* __restrict__ Dc32) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) Dc32[k] = (float)Dc[k];
}

// ---- M2L on MFMA (the FP32-precision path, orders <= 5) -------------------------------------------
// For displacement class c the M2L is one fixed matrix: L_gamma += sum_alpha A_c[gamma][alpha] M_alpha
// with A_c[gamma][alpha] = (-1)^|alpha| D_{alpha+gamma} / gamma!.  The sign is folded into the FP32
// multipoles and 1/gamma! is applied in FP64 to each target's sum, so the product matrix is the plain
// gather D_{alpha+gamma} from the class tensor.  Pairs of one class run as dense products
// D_c [M'_S1 .. M'_S16] on v_mfma_f32_16x16x4f32, up to 16 pairs (one per column) per product.
//
// Accumulation without per-pair rows: the pairs are sorted by (target group, class), a group being
// kM2LGroupTargets consecutive BFS ids (cells of one level in Morton order, so neighbouring targets
// share most of their classes: at N = 1M about 85% of the 16 columns of a product are filled).  Two
// pairs of one class never share a target (cell centres are distinct, so a target and a displacement
// fix the source), so the columns of a product belong to distinct targets of the group.  A
// workgroup takes one item (a contiguous piece of a group's pair list), keeps the group's locals in
// LDS in FP64, and writes them as the item's partial; a reduction adds a group's partials in item
// order.  The waves of a workgroup take contiguous eighths of the item and add their products into
// the LDS locals in a fixed turn order (round r: wave 0, 1, .., 7, a wave without a product passing
// its turn), so every sum has a fixed order and the result is deterministic.  A product's FP32
// result for one column depends on that column alone.  Unclassed pairs get classes of their own.
// Operand layout (16x16x4): lane l = (j, h) with j = l%16, h = l/16 holds A[j][h] and B[h][j];
// the result D[4h + v][j], v = 0..3.
typedef float f32x4v __attribute__((ext_vector_type(4)));
constexpr int kM2LGroupTargets = 256;  // target cells per group (consecutive BFS ids)
constexpr int kM2LGroupBits = 8;
constexpr int kM2LMfmaMaxOrder = 5;  // orders of the matrix-core M2L (nm <= 56)
constexpr int kM2LWavesP4 = 16;     // waves per workgroup of the product kernel (orders <= 4)
constexpr int kM2LWavesP5 = 8;      // (order 5: more registers per wave)

// FP32 signed multipoles in the B-operand order of the products: cell S's row of 4 nkp floats holds
// (-1)^|al| M_al at h * nkp + ks for al = 4 ks + h (0 past nm); nkp is the k-step count rounded up
// to 4, so lane (j, h) reads its values as aligned 16-byte loads.
__global__ void MultipoleToF32Kernel(int nn, int nm, int nkp, const double* __restrict__ M,
                                     const double* __restrict__ sign, float* __restrict__ M32) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (int64_t)nn * 4 * nkp) return;
  const int64_t S = k / (4 * nkp);
  const int e = (int)(k - S * 4 * nkp), h = e / nkp, ks = e - h * nkp, al = 4 * ks + h;
  M32[k] = al < nm ? (float)(sign[al] * M[S * nm + al]) : 0.f;
}

// Unclassed pairs (pair_slot kNoClass) get class ids n_cls + u, u their rank in list order.
__global__ void UnclassedFlagKernel(int64_t np, const uint32_t* __restrict__ pair_slot, uint32_t* __restrict__ flag) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < np) flag[q] = pair_slot[q] == kNoClass ? 1u : 0u;
}

__global__ void UnclassedListKernel(int64_t np, const uint32_t* __restrict__ pair_slot,
                                    const uint32_t* __restrict__ rank, uint32_t* __restrict__ ulist) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < np && pair_slot[q] == kNoClass) ulist[rank[q]] = (uint32_t)q;
}

// One wave per unclassed pair: the FP32 tensor of its own displacement.
__global__ __launch_bounds__(64) void UnclassedTensorKernel(int n_uncl, const uint32_t* __restrict__ ulist,
                                                            const uint64_t* __restrict__ list,
                                                            const Cell* __restrict__ cells, FmmTab t,
                                                            float* __restrict__ Dc32) {
  extern __shared__ double lds[];
  const int u = blockIdx.x;
  if (u >= n_uncl) return;
  const uint64_t pk = list[ulist[u]];
  const Cell& T = cells[(int)(pk >> 32)];
  const Cell& S = cells[(uint32_t)pk];
  const double xs[3] = {T.cx - S.cx, T.cy - S.cy, T.cz - S.cz};
  DerivTensorWave(xs, t, t.rec, lds);
  for (int k = threadIdx.x; k < t.nd; k += 64) Dc32[(size_t)u * t.nd + k] = (float)lds[k];
}

// Bit layout of the sorted pair keys: (group, class, target within the group) above the source
// cell when all fit in 64 bits ("packed", the source then rides in the key), else (group, class,
// target) with the source as the sort value.
struct M2LKeyLayout {
  int packed;
  int sbits, cls_shift, gshift;
  uint64_t smask, cls_mask;
};

__global__ void M2LGroupKeyKernel(int64_t np, const uint64_t* __restrict__ list, const uint32_t* __restrict__ pre,
                                  const uint32_t* __restrict__ pair_slot, const uint32_t* __restrict__ ids,
                                  const uint32_t* __restrict__ urank, uint32_t n_cls, M2LKeyLayout kl,
                                  uint64_t* __restrict__ keys, uint32_t* __restrict__ vals) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= np) return;
  const uint64_t pk = list[q];
  const uint32_t S = (uint32_t)pk;
  const uint32_t bfs = pre[(uint32_t)(pk >> 32)];
  const uint32_t slot = pair_slot[q];
  const uint64_t cls = slot == kNoClass ? (uint64_t)n_cls + urank[q] : (uint64_t)ids[slot];
  const uint64_t g = bfs >> kM2LGroupBits, loc = bfs & (kM2LGroupTargets - 1);
  const uint64_t hi = (g << kl.gshift) | (cls << kl.cls_shift) | (loc << (kl.cls_shift - kM2LGroupBits));
  if (kl.packed) {
    keys[q] = hi | S;
  } else {
    keys[q] = hi;
    vals[q] = S;
  }
}

// goff[g] = first sorted position of group g (g = 0 .. n_groups).
__global__ void GroupOffsetsKernel(int n_groups, const uint64_t* __restrict__ keys, int64_t np, int gshift,
                                   int64_t* __restrict__ goff) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g > n_groups) return;
  int64_t lo = 0, hi = np;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if ((keys[mid] >> gshift) < (uint64_t)g) lo = mid + 1;
    else hi = mid;
  }
  goff[g] = lo;
}

struct M2LItem {
  int64_t b, e;  // sorted positions
  int g, pad;    // group
};

// ---- exact, order-independent accumulation of the M2L products -------------------------------------
// A product's FP32 result x for coefficient gamma of a target (|gamma| = n) is bounded by
//     b_n = sum_a Dk_c(n + a) Ma_S(a),
// Dk_c(k) = max |D| over the class tensor's entries of degree k, Ma_S(a) = sum over |alpha| = a of
// |M'_alpha(S)| (the FP32 multipoles).  With B = 2 n_prod(group) max_pairs(b_n) per (group, n) --
// n_prod, the group's product count, bounds any target's pair count, and 2 covers the FP32 rounding
// of the product -- every x is rounded to the grid 2^eg with 2^(eg + 52) > B: (x + C) - C with
// C = 1.5 2^(eg + 52).  Sums of grid multiples below 2^(eg + 53) are exact in FP64, so the atomic
// LDS adds, and the reduction over a group's items, give the same bits in any order.  The rounding
// error, at most 2^(eg - 1) = B 2^-53 per product, is far below the FP32 error of the product.

// Ma[S][a] (a = 0 .. P): the per-degree L1 norms of each cell's FP32 multipoles.
__global__ void CellNormKernel(int nn, FmmTab t, const double* __restrict__ M, float* __restrict__ ma) {
  const int S = blockIdx.x * blockDim.x + threadIdx.x;
  if (S >= nn) return;
  for (int a = 0; a <= t.P; ++a) {
    float s = 0.f;
    for (int k = t.deg_start[a]; k < t.deg_start[a + 1]; ++k) s += fabsf((float)M[(size_t)S * t.nm + k]);
    ma[(size_t)S * (t.P + 1) + a] = s;
  }
}

// Dk[c][k] (k = 0 .. 2P) for every class tensor (one wave per class).
__global__ __launch_bounds__(64) void ClassDegMaxKernel(int n_tens, const float* __restrict__ Dc32, FmmTab t,
                                                        float* __restrict__ dk) {
  const int c = blockIdx.x, lane = threadIdx.x;
  if (c >= n_tens) return;
  for (int d = 0; d <= 2 * t.P; ++d) {
    float m = 0.f;
    for (int k = t.deg_start[d] + lane; k < t.deg_start[d + 1]; k += 64) m = fmaxf(m, fabsf(Dc32[(size_t)c * t.nd + k]));
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
    if (lane == 0) dk[(size_t)c * (2 * t.P + 1) + d] = m;
  }
}

// Workgroup b's item: consecutive items go to one XCD (workgroups are dealt to the 8 XCDs in turn),
// so each XCD works through a contiguous range of groups and its L2 holds their sources.
__device__ inline int XcdItem(int b, int n) {
  const int q = n / 8, r = n % 8, x = b % 8, i = b / 8;
  return x < r ? x * (q + 1) + i : r * (q + 1) + (x - r) * q + i;
}

// Slice w of W of an item: the sorted positions that wave w of the item's workgroup takes.
__device__ inline void M2LSlice(const M2LItem& it, int w, int W, int64_t& sb, int64_t& se) {
  const int64_t span = it.e - it.b;
  sb = it.b + span * w / W;
  se = it.b + span * (w + 1) / W;
}

// Products of a slice: its sorted pairs cut into runs of one class, each cut into pieces of at most
// 16 (a product has one pair per column).  One wave per slice, 64 positions per step.  Counting
// pass (start == nullptr): the slice's product count into cnt[s].  Writing pass: product p (numbered
// from off[s]) gets start[p] (its first position), cls[p] and col[16 p + c] = (source << 8 | target
// within the group) for its column c; start[off[s + 1]] = the slice's end, so every product's
// length is start[p + 1] - start[p].
// The writing pass also takes the group's bound max_pairs(b_n) into gmax[g][n] (float bits, atomic
// max: the same in any order).
__global__ __launch_bounds__(64) void M2LProductKernel(int n_slices, int W, const M2LItem* __restrict__ items,
                                                       const uint64_t* __restrict__ keys,
                                                       const uint32_t* __restrict__ vals, M2LKeyLayout kl,
                                                       const uint32_t* __restrict__ off, uint32_t* __restrict__ cnt,
                                                       uint32_t* __restrict__ start, uint32_t* __restrict__ cls,
                                                       uint32_t* __restrict__ col, int P,
                                                       const float* __restrict__ ma,
                                                       const float* __restrict__ dk, uint32_t* __restrict__ gmax) {
  const int s = blockIdx.x, lane = threadIdx.x;
  if (s >= n_slices) return;
  int64_t sb, se;
  const M2LItem itm = items[s / W];
  M2LSlice(itm, s % W, W, sb, se);
  float bmax[kM2LMfmaMaxOrder + 1];  // (indices are compile-time: the loops are unrolled)
#pragma unroll
  for (int n = 0; n <= kM2LMfmaMaxOrder; ++n) bmax[n] = 0.f;
  const bool write = start != nullptr;
  uint32_t np = write ? off[s] : 0u;  // products before the window
  int64_t rs = sb, hp = sb;           // start of the run and of the product that the window continues
  uint32_t prev = 0xffffffffu;        // class of the position before the window
  const uint64_t incl = lane == 63 ? ~0ull : ((2ull << lane) - 1);
  // keys come in blocks of four steps (256 positions), each block loaded while the one before it
  // is processed
  uint64_t kq[4], kn[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) kq[q] = sb + 64 * q + lane < se ? keys[sb + 64 * q + lane] : 0ull;
  for (int64_t wb4 = sb; wb4 < se; wb4 += 256) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int64_t en = wb4 + 256 + 64 * q + lane;
      kn[q] = en < se ? keys[en] : 0ull;
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int64_t wb = wb4 + 64 * q;
      if (wb >= se) break;
      const int64_t e = wb + lane;
      const bool valid = e < se;
      const uint64_t k = kq[q];
      const uint32_t c = (uint32_t)((k >> kl.cls_shift) & kl.cls_mask);
      uint32_t pc = (uint32_t)__shfl_up((int)c, 1);
      if (lane == 0) pc = prev;
      const uint64_t runs = __ballot(valid && (e == sb || c != pc));
      const uint64_t rb = runs & incl;
      const int64_t myrs = rb ? wb + 63 - __builtin_clzll(rb) : rs;
      const bool head = valid && ((e - myrs) & 15) == 0;
      const uint64_t heads = __ballot(head);
      const uint64_t hb = heads & incl;
      const int64_t myhp = hb ? wb + 63 - __builtin_clzll(hb) : hp;
      if (write && valid) {
        const uint32_t p = np + (uint32_t)__popcll(hb) - 1u;  // hb == 0: the product of the previous step
        const uint32_t S = kl.packed ? (uint32_t)(k & kl.smask) : vals[e];
        const uint32_t loc = (uint32_t)((k >> (kl.cls_shift - kM2LGroupBits)) & (kM2LGroupTargets - 1));
        col[16 * (size_t)p + (size_t)(e - myhp)] = (S << 8) | loc;
        const float* dkc = dk + (size_t)c * (2 * P + 1);
        const float* mas = ma + (size_t)S * (P + 1);
        float dv[2 * kM2LMfmaMaxOrder + 1], mv[kM2LMfmaMaxOrder + 1];
#pragma unroll
        for (int d = 0; d <= 2 * kM2LMfmaMaxOrder; ++d) dv[d] = d <= 2 * P ? dkc[d] : 0.f;
#pragma unroll
        for (int a = 0; a <= kM2LMfmaMaxOrder; ++a) mv[a] = a <= P ? mas[a] : 0.f;
#pragma unroll
        for (int n = 0; n <= kM2LMfmaMaxOrder; ++n) {
          float bsum = 0.f;
#pragma unroll
          for (int a = 0; a <= kM2LMfmaMaxOrder; ++a) bsum += dv[n + a] * mv[a];
          bmax[n] = fmaxf(bmax[n], bsum);
        }
        if (head) {
          start[p] = (uint32_t)e;
          cls[p] = c;
        }
      }
      np += (uint32_t)__popcll(heads);
      hp = (int64_t)__shfl(myhp, 63);  // (lanes past se keep the values of the last valid lane)
      rs = (int64_t)__shfl(myrs, 63);
      prev = __shfl(c, 63);
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) kq[q] = kn[q];
  }
  if (lane == 0) {
    if (write) start[off[s + 1]] = (uint32_t)se;
    else cnt[s] = np;
  }
  if (write) {
#pragma unroll
    for (int n = 0; n <= kM2LMfmaMaxOrder; ++n) {
      float m = bmax[n];
      for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
      if (lane == 0 && n <= P && m > 0.f) atomicMax(&gmax[(size_t)itm.g * (P + 1) + n], __float_as_uint(m));
    }
  }
}

// LDS bytes of M2LGroupKernel<NRB, NKS, NDL, W>: the group's locals [G][nm | 1] (doubles; an odd
// row stride spreads the 16 targets of a product's adds over the LDS banks), the
// per-lane fragment index table [NRB NKS][64], per wave the FP32 class tensor Df[nd + 1] (the
// whole rounded up to 8 bytes), the grid constants [nm] (doubles).
static size_t M2LGroupLds(int nm, int nd, int nrb, int nks, int waves) {
  return sizeof(double) * (size_t)kM2LGroupTargets * (nm | 1) + sizeof(int) * (size_t)nrb * nks * 64 +
         sizeof(float) * (size_t)((waves * (nd + 1) + 1) & ~1) + sizeof(double) * (size_t)nm;
}

__global__ void M2LZeroKernel(int64_t n, uint32_t* __restrict__ p) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

template <int NRB, int NKS, int NDL, int W>
__global__ __launch_bounds__(64 * W) void M2LGroupKernel(
    const M2LItem* __restrict__ items, const int* __restrict__ gitem, const uint32_t* __restrict__ slice_off,
    const uint32_t* __restrict__ pstart, const uint32_t* __restrict__ pcls, const uint32_t* __restrict__ pcol,
    const float* __restrict__ Dc32, FmmTab t, const float* __restrict__ M32, const uint32_t* __restrict__ gmax,
    double* __restrict__ partial) {
  constexpr int NKP = (NKS + 3) & ~3;
  constexpr int G = kM2LGroupTargets;
  extern __shared__ double lds[];
  const int nm = t.nm, nd = t.nd;
  double* acc = lds;
  const int as = nm | 1;  // row stride of the locals
  int* s_aidx = reinterpret_cast<int*>(acc + G * as);
  float* s_df = reinterpret_cast<float*>(s_aidx + NRB * NKS * 64);
  double* s_grid = reinterpret_cast<double*>(s_df + ((W * (nd + 1) + 1) & ~1));  // C per coefficient
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int j = lane & 15, h = lane >> 4;
  const int item = XcdItem(blockIdx.x, gridDim.x);
  for (int e = threadIdx.x; e < G * as; e += 64 * W) acc[e] = 0.0;
  {
    const int grp = items[item].g;
    const double n_prod = (double)(slice_off[gitem[grp + 1] * W] - slice_off[gitem[grp] * W]);
    for (int g = threadIdx.x; g < nm; g += 64 * W) {
      const int n = t.multi[3 * g] + t.multi[3 * g + 1] + t.multi[3 * g + 2];
      const double B = 2.0 * n_prod * (double)__uint_as_float(gmax[(size_t)grp * (t.P + 1) + n]);
      s_grid[g] = B > 0.0 ? 1.5 * ldexp(1.0, ilogb(B) + 1) : 0.0;  // 1.5 2^(eg + 52), eg = ilogb(B) + 1 - 52
    }
  }
  // Fragment (rb, ks) of lane (j, h) is D[m2l[16 rb + j][4 ks + h]]; outside the nm x nm matrix it
  // is Df[nd], which holds 0.
  for (int e = threadIdx.x; e < NRB * NKS * 64; e += 64 * W) {
    const int rk = e >> 6, l = e & 63, rb = rk / NKS, ks = rk - rb * NKS;
    const int g = rb * 16 + (l & 15), al = ks * 4 + (l >> 4);
    s_aidx[e] = (g < nm && al < nm) ? t.m2l[g * nm + al] : nd;
  }
  float* Df = s_df + wave * (nd + 1);
  if (lane == 0) Df[nd] = 0.f;
  __syncthreads();
  // this wave's products
  const int slice = item * W + wave;
  const int pb = (int)slice_off[slice], pe = (int)slice_off[slice + 1];
  // A product's columns into the locals: rounded to the grid, then LDS atomic adds, exact and so
  // independent of the order in which the waves make them.
  auto add_columns = [&](const f32x4v (&a)[NRB], int loc, bool present) {
    double* row = acc + loc * as;
#pragma unroll
    for (int rb = 0; rb < NRB; ++rb)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int g = 16 * rb + 4 * h + v;
        if (present && g < nm) {
          const double C = s_grid[g];
          const double x = ((double)a[rb][v] + C) - C;
          __hip_atomic_fetch_add(row + g, x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        }
      }
  };
  // Software pipeline over three slots, the stage written out three times with the slots rotated so
  // that no register a load may still be writing is copied.  At stage k: product k+2's table entry
  // (start, next start, class, column word) is loaded; product k+1's multipole rows and class tensor
  // are loaded from its entry (loaded a stage earlier); product k-1's result is added; product k
  // runs on the matrix cores into the same accumulators.  Every stage issues the same loads in the
  // same order (positions past the wave's last product are clamped to it), so the compiler's
  // in-order waits are exact.
  struct Slot {
    uint32_t st0, st1, cl, cw;  // table entry: start, next start, class, column word
    float bv[NKS], dv[NDL];
  };
  f32x4v acc_r[NRB];
  auto load_entry = [&](int k, Slot& s) {
    const int kk = k < pe ? k : pe - 1;
    s.st0 = pstart[kk];
    s.st1 = pstart[kk + 1];
    s.cl = pcls[kk];
    s.cw = pcol[16 * (size_t)kk + j];
  };
  auto load_operands = [&](Slot& s) {  // the lane's multipole values of column j's source, the tensor
    const bool present = j < (int)(s.st1 - s.st0);
    const uint32_t S = present ? (s.cw >> 8) : 0u;  // cell 0 for an empty column (never added)
    const float* mrow = M32 + (size_t)S * (4 * NKP) + h * NKP;
#pragma unroll
    for (int i = 0; i < NKS / 4; ++i) {
      const f32x4v v = reinterpret_cast<const f32x4v*>(mrow)[i];
      s.bv[4 * i] = v[0];
      s.bv[4 * i + 1] = v[1];
      s.bv[4 * i + 2] = v[2];
      s.bv[4 * i + 3] = v[3];
    }
#pragma unroll
    for (int i = (NKS / 4) * 4; i < NKS; ++i) s.bv[i] = mrow[i];
    const uint32_t c = (uint32_t)__builtin_amdgcn_readfirstlane((int)s.cl);
#pragma unroll
    for (int i = 0; i < NDL; ++i) {
      const int k = lane + 64 * i;  // values past nd are loaded from nd - 1 and never used
      s.dv[i] = Dc32[(size_t)c * nd + (k < nd ? k : nd - 1)];
    }
  };
  float afrag[NRB][NKS];
  uint32_t cur = 0xffffffffu;
  int locB = 0;
  bool presB = false, pending = false;
  Slot s0, s1, s2;
#pragma unroll
  for (int rb = 0; rb < NRB; ++rb) acc_r[rb] = f32x4v{0.f, 0.f, 0.f, 0.f};
  if (pb < pe) {
    load_entry(pb, s0);
    load_entry(pb + 1, s1);
    load_operands(s0);
  }
  int k = pb;
  auto stage = [&](Slot& c, Slot& n, Slot& f) -> bool {
    if (k >= pe) {  // no product k: the last one (k-1) is added and the wave is done
      if (pending) add_columns(acc_r, locB, presB);
      pending = false;
      return false;
    }
    load_entry(k + 2, f);
    load_operands(n);
    if (pending) add_columns(acc_r, locB, presB);  // product k-1's columns into the locals
    const uint32_t ccls = (uint32_t)__builtin_amdgcn_readfirstlane((int)c.cl);
    if (ccls != cur) {  // the class's fragments
      WaveSync();
#pragma unroll
      for (int i = 0; i < NDL; ++i)
        if (lane + 64 * i < nd) Df[lane + 64 * i] = c.dv[i];
      WaveSync();
#pragma unroll
      for (int rb = 0; rb < NRB; ++rb)
#pragma unroll
        for (int ks = 0; ks < NKS; ++ks) afrag[rb][ks] = Df[s_aidx[(rb * NKS + ks) * 64 + lane]];
      cur = ccls;
    }
#pragma unroll
    for (int rb = 0; rb < NRB; ++rb) acc_r[rb] = f32x4v{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < NKS; ++ks)
#pragma unroll
      for (int rb = 0; rb < NRB; ++rb)
        acc_r[rb] = __builtin_amdgcn_mfma_f32_16x16x4f32(afrag[rb][ks], c.bv[ks], acc_r[rb], 0, 0, 0);
    locB = (int)(c.cw & (G - 1));
    presB = j < (int)(c.st1 - c.st0);
    pending = true;
    ++k;
    return true;
  };
  while (stage(s0, s1, s2) && stage(s1, s2, s0) && stage(s2, s0, s1)) {
  }
  __syncthreads();
  double* out = partial + (size_t)item * G * nm;
  for (int e = threadIdx.x; e < G * nm; e += 64 * W) {
    const int loc = e / nm;
    out[e] = acc[loc * as + (e - loc * nm)];
  }
}

// L of each target cell: its group's item partials added in item order, times 1/gamma!.
__global__ void M2LPartialReduceKernel(int nn, int nm, const int* __restrict__ gitem,
                                       const double* __restrict__ partial, const double* __restrict__ inv_fact,
                                       const uint32_t* __restrict__ bfs2pre, double* __restrict__ L) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)nn * nm) return;
  const int bfs = (int)(idx / nm), gm = (int)(idx - (int64_t)bfs * nm);
  const int g = bfs >> kM2LGroupBits, loc = bfs & (kM2LGroupTargets - 1);
  double s = 0.0;
  for (int it = gitem[g]; it < gitem[g + 1]; ++it) s += partial[((size_t)it * kM2LGroupTargets + loc) * nm + gm];
  L[(size_t)bfs2pre[bfs] * nm + gm] = s * inv_fact[gm];
}

// L2L (fast-multipole.cc:557-578) for the parents of one BFS level.
__global__ __launch_bounds__(64) void L2LKernel(int n0, NodeArrays nd, const uint32_t* __restrict__ bfs2pre, FmmTab t,
                                                double* __restrict__ L) {
  __shared__ double mono[kFmmMaxNm];
  __shared__ double lp[kFmmMaxNm];
  const int node = n0 + blockIdx.x;
  if (nd.leaf[node]) return;
  const TransferTabs tt = StageTransferTabs(t, true);
  const int p = (int)bfs2pre[node];
  for (int k = threadIdx.x; k < t.nm; k += 64) lp[k] = L[(size_t)p * t.nm + k];
  const int c0 = nd.first_child[node];
  for (int c = c0; c < c0 + nd.nchild[node]; ++c) {
    const double dx = nd.cx[c] - nd.cx[node], dy = nd.cy[c] - nd.cy[node], dz = nd.cz[c] - nd.cz[node];
    __syncthreads();
    for (int k = threadIdx.x; k < t.nm; k += 64) mono[k] = MonomialAt(tt.multi, k, dx, dy, dz);
    __syncthreads();
    double* lc = L + (size_t)bfs2pre[c] * t.nm;
    for (int a = threadIdx.x; a < t.nm; a += 64) {
      double v = lc[a];
      for (int q = tt.off[a]; q < tt.off[a + 1]; ++q) {
        const int bb = tt.bc[2 * q], cc = tt.bc[2 * q + 1];
        v += lp[bb] * tt.fact[bb] * tt.ifact[a] * tt.ifact[cc] * mono[cc];
      }
      lc[a] = v;
    }
  }
}

// FP64 L2P + P2P (fast-multipole.cc:580-661) and unscaling (:663-674), also the potential
// energy: one wave per leaf; lane l owns particle b + (l & 31) of the current 32-particle chunk
// and the sources of parity l >> 5.
__global__ __launch_bounds__(64) void L2PP2PKernel(int nn, const Cell* __restrict__ cells,
                                                   const uint64_t* __restrict__ list,
                                                   const int64_t* __restrict__ off, FmmTab t,
                                                   const double* __restrict__ L, const double4* __restrict__ sp,
                                                   const uint32_t* __restrict__ order,
                                                   const FmmFrame* __restrict__ fr, int want_forces,
                                                   double* __restrict__ fx, double* __restrict__ fy,
                                                   double* __restrict__ fz, double* __restrict__ phi_out) {
  __shared__ double lt[kFmmMaxNm];
  const int p = blockIdx.x;
  const Cell T = cells[p];
  if (!T.leaf) return;
  for (int k = threadIdx.x; k < t.nm; k += 64) lt[k] = L[(size_t)p * t.nm + k];
  __syncthreads();
  const int lane = threadIdx.x & 31, half = threadIdx.x >> 5;
  const double min_sq = fr->min_sq;
  const int64_t lb = off[p], le = off[p + 1];
  for (int base = T.b; base < T.e; base += 32) {
    const int i = base + lane;
    const bool valid = i < T.e;
    const double4 si = sp[valid ? i : T.b];
    double pot = 0.0, ex = 0.0, ey = 0.0, ez = 0.0;
    if (valid && half == 0) {  // L2P
      const double x = si.x - T.cx, y = si.y - T.cy, z = si.z - T.cz;
      if (!want_forces)
        for (int g = 0; g < t.nm; ++g) pot += lt[g] * Monomial(t, g, x, y, z);
      if (want_forces) {
        double gx = 0.0, gy = 0.0, gz = 0.0;
        for (int g = 1; g < t.nm; ++g) {
          const int* gi = t.grad + 3 * g;
          const int* mg = t.multi + 3 * g;
          if (gi[0] >= 0) gx += mg[0] * lt[g] * Monomial(t, gi[0], x, y, z);
          if (gi[1] >= 0) gy += mg[1] * lt[g] * Monomial(t, gi[1], x, y, z);
          if (gi[2] >= 0) gz += mg[2] * lt[g] * Monomial(t, gi[2], x, y, z);
        }
        ex = -gx;
        ey = -gy;
        ez = -gz;
      }
    }
    if (valid) {
      for (int64_t q = lb; q < le; ++q) {
        const Cell& S = cells[(int)(uint32_t)list[q]];
        for (int j = S.b + half; j < S.e; j += 2) {
          if (j == i) continue;
          const double4 sj = sp[j];
          if (sj.w == 0.0) continue;
          const double rx = si.x - sj.x, ry = si.y - sj.y, rz = si.z - sj.z;
          const double r2 = rx * rx + ry * ry + rz * rz;
          if (r2 < min_sq) continue;
          const double inv_r = 1.0 / sqrt(r2);
          if (want_forces) {
            const double f = sj.w * inv_r / r2;
            ex += rx * f;
            ey += ry * f;
            ez += rz * f;
          } else {
            pot += sj.w * inv_r;
          }
        }
      }
    }
    ex += __shfl_down(ex, 32);
    ey += __shfl_down(ey, 32);
    ez += __shfl_down(ez, 32);
    pot += __shfl_down(pot, 32);
    if (valid && half == 0) {
      const uint32_t o = order[i];
      if (want_forces) {
        const double s = fr->force_scale * si.w;
        fx[o] = s * ex;
        fy[o] = s * ey;
        fz[o] = s * ez;
      } else {
        phi_out[i] = si.w * pot;  // q_i phi_i in tree order
      }
    }
  }
}

// L2P gradient for order P (fast-multipole.cc:663-674): the local expansion is
// phi(x) = sum_g L_g x^g over the multi-indices of total degree <= P in the table order (degree
// ascending, then a and b descending), and grad phi is summed over g in that order with the same
// products as the runtime-order loop above; unrolled, with the coefficients read as wave-uniform
// scalars.
template <int P>
__device__ __forceinline__ void L2PGrad(const double* __restrict__ L, double x, double y, double z, double& gx,
                                        double& gy, double& gz) {
  double px[P + 1], py[P + 1], pz[P + 1];
  px[0] = py[0] = pz[0] = 1.0;
#pragma unroll
  for (int k = 1; k <= P; ++k) {
    px[k] = px[k - 1] * x;
    py[k] = py[k - 1] * y;
    pz[k] = pz[k - 1] * z;
  }
  gx = gy = gz = 0.0;
  int g = 0;
#pragma unroll
  for (int deg = 0; deg <= P; ++deg)
#pragma unroll
    for (int a = deg; a >= 0; --a)
#pragma unroll
      for (int b = deg - a; b >= 0; --b) {
        const int c = deg - a - b;
        const double l = L[g];
        if (a > 0) gx += (double)a * l * (px[a - 1] * py[b] * pz[c]);
        if (b > 0) gy += (double)b * l * (px[a] * py[b - 1] * pz[c]);
        if (c > 0) gz += (double)c * l * (px[a] * py[b] * pz[c - 1]);
        ++g;
      }
}

// L2P forces for the FP32 path (fast-multipole.cc:663-674): one wave per leaf, kP2PWaves leaves
// per workgroup, lane = particle of the leaf (at most 32).  Writes F = -q grad(phi_local) in
// newtons; P2PF32Kernel then adds the near field.
constexpr int kP2PWaves = 4;

template <int P>
__global__ __launch_bounds__(64 * kP2PWaves) void L2PForceKernel(int nn, const Cell* __restrict__ cells, int nm,
                                                                 const double* __restrict__ L,
                                                                 const double4* __restrict__ sp,
                                                                 const uint32_t* __restrict__ order,
                                                                 const FmmFrame* __restrict__ fr,
                                                                 double* __restrict__ fx, double* __restrict__ fy,
                                                                 double* __restrict__ fz) {
  const int p = blockIdx.x * kP2PWaves + (threadIdx.x >> 6);
  if (p >= nn) return;
  const Cell& T = cells[p];
  if (!T.leaf) return;
  for (int i = T.b + (threadIdx.x & 63); i < T.e; i += 64) {
    const double4 si = sp[i];
    double gx, gy, gz;
    L2PGrad<P>(L + (size_t)p * nm, si.x - T.cx, si.y - T.cy, si.z - T.cz, gx, gy, gz);
    const double s = fr->force_scale * si.w;
    const uint32_t o = order[i];
    fx[o] = s * -gx;
    fy[o] = s * -gy;
    fz[o] = s * -gz;
  }
}

// FP32 near field (the P2P of fast-multipole.cc:663-674 in the hi/lo pair form), added to the
// L2P forces.  The work is cut into items so that no leaf with a long near list holds one wave
// for the whole kernel: item = (target leaf p, batch of <= 32 of its targets, a run of segments
// of <= kP2PSeg cells of its near list), at most kP2PItemsPerBatch items per batch (so the
// partials stay bounded when the lists are very long).  P2PItemKernel sums one item per wave into
// a partial per target; P2PReduceKernel adds the partials of a target in list order, so the
// result does not depend on scheduling.
//   * A segment's source cells are contiguous ranges of the tree-ordered particles.  They are
//     staged in LDS as (first particle, exclusive prefix count), so the segment reads as one
//     flattened source sequence: position k maps to a particle by a binary search in the prefix
//     counts, and 64 consecutive positions are mostly consecutive particles.
//   * A batch of nt targets takes nt_pad = the next power of two >= max(nt, 4) lanes; the
//     64 / nt_pad groups of such lanes split every 64-source chunk (group s takes sources s,
//     s + 64 / nt_pad, ...).  A chunk is staged in the wave's LDS slice while the next one is
//     fetched.
//   * Per chunk the pair sum is FP32 and is added into FP64; the lane groups are folded by a fixed
//     xor tree.  A source equal to the target contributes exactly 0 (softened coincident pair),
//     and so do the charge-0 padding entries.
constexpr int kP2PSeg = 64;
constexpr int kP2PItemsPerBatch = 16;

// Segments of a near list of np cells, and segments per item.
__host__ __device__ inline int P2PSegments(int64_t np) { return (int)((np + kP2PSeg - 1) / kP2PSeg); }
__host__ __device__ inline int P2PSegsPerItem(int nseg) {
  return (nseg + kP2PItemsPerBatch - 1) / kP2PItemsPerBatch;
}

// One staged chunk against a batch of NPAD target lanes: group grp takes sources grp + r * 64 / NPAD
// (a fixed trip count, unrolled, so the LDS reads go out ahead of the arithmetic).
template <int NPAD>
__device__ __forceinline__ void P2PChunk(const float4* __restrict__ hi, const float4* __restrict__ lo, int grp,
                                         const float4 th, const float4 tl, float& px, float& py, float& pz) {
  constexpr int kGroups = 64 / NPAD;
#pragma unroll 4
  for (int r = 0; r < NPAD; ++r) PairF32(th, tl, hi[grp + r * kGroups], lo[grp + r * kGroups], px, py, pz);
}

struct P2PItem {
  uint32_t leaf;
  uint16_t batch, part;  // segments [part * G, (part + 1) * G) of the list, G = P2PSegsPerItem
};

// Items per cell: (target batches) x (items per batch); 0 for internal cells.
__global__ void P2PItemCountKernel(int nn, const Cell* __restrict__ cells, const int64_t* __restrict__ off,
                                   uint32_t* __restrict__ cnt) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nn) return;
  const Cell& T = cells[p];
  const int nseg = P2PSegments(off[p + 1] - off[p]);
  const int g = P2PSegsPerItem(nseg);
  cnt[p] = T.leaf ? (uint32_t)((T.e - T.b + 31) / 32) * (uint32_t)((nseg + g - 1) / g) : 0u;
}

// Items of cell p at base[p]: batch-major, then in list order.
__global__ void P2PItemListKernel(int nn, const Cell* __restrict__ cells, const int64_t* __restrict__ off,
                                  const uint32_t* __restrict__ base, const uint32_t* __restrict__ cnt,
                                  P2PItem* __restrict__ items) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nn || cnt[p] == 0) return;
  const int nseg = P2PSegments(off[p + 1] - off[p]);
  const int g = P2PSegsPerItem(nseg), ni = (nseg + g - 1) / g;
  const int nb = (int)(cnt[p] / (uint32_t)ni);
  P2PItem* out = items + base[p];
  for (int b = 0; b < nb; ++b)
    for (int k = 0; k < ni; ++k) out[b * ni + k] = P2PItem{(uint32_t)p, (uint16_t)b, (uint16_t)k};
}

// part[item][slot] = (sum_x, sum_y, sum_z) over the item's sources, e/unit^2, slot < batch size.
__global__ __launch_bounds__(64 * kP2PWaves) void P2PItemKernel(
    uint32_t n_items, const P2PItem* __restrict__ items, const Cell* __restrict__ cells,
    const uint64_t* __restrict__ list, const int64_t* __restrict__ off, const float4* __restrict__ sp32,
    double* __restrict__ part) {
  __shared__ uint32_t s_first[kP2PWaves][kP2PSeg], s_pre[kP2PWaves][kP2PSeg + 1];
  __shared__ float4 s_hi[kP2PWaves][64], s_lo[kP2PWaves][64];
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const uint32_t it = blockIdx.x * kP2PWaves + w;
  if (it >= n_items) return;  // the wave's own exit: no workgroup barrier below
  const P2PItem item = items[it];
  const Cell& T = cells[item.leaf];
  const int tb = T.b + 32 * item.batch;
  const int nt = T.e - tb < 32 ? T.e - tb : 32;
  const int sh = nt <= 4 ? 2 : (nt <= 8 ? 3 : (nt <= 16 ? 4 : 5));  // log2 nt_pad
  const int slot = l & ((1 << sh) - 1), grp = l >> sh;
  const bool valid = slot < nt;
  const int i = tb + (valid ? slot : 0);
  const float4 th = sp32[2 * i], tl = sp32[2 * i + 1];
  const int64_t lb = off[item.leaf], le = off[item.leaf + 1];
  const int gper = P2PSegsPerItem(P2PSegments(le - lb));
  const int64_t s0 = lb + (int64_t)kP2PSeg * gper * item.part;
  const int64_t s1 = s0 + (int64_t)kP2PSeg * gper < le ? s0 + (int64_t)kP2PSeg * gper : le;
  double ax = 0.0, ay = 0.0, az = 0.0;
  for (int64_t seg = s0; seg < s1; seg += kP2PSeg) {
    const int ns = (int)((s1 - seg) < kP2PSeg ? (s1 - seg) : kP2PSeg);
    uint32_t first = 0, cnt = 0;
    if (l < ns) {
      const Cell& S = cells[(uint32_t)list[seg + l]];
      first = (uint32_t)S.b;
      cnt = (uint32_t)(S.e - S.b);
    }
    uint32_t incl = cnt;  // inclusive scan over the wave
    for (int d = 1; d < 64; d <<= 1) {
      const uint32_t y = __shfl_up(incl, d);
      if (l >= d) incl += y;
    }
    const uint32_t run = __shfl(incl, 63);
    WaveSync();  // the previous segment's tables have been read
    s_first[w][l] = first;
    s_pre[w][l] = incl - cnt;  // entries past the segment hold run
    if (l == 0) s_pre[w][kP2PSeg] = run;
    WaveSync();
    // particle at flattened position k < run: the last q with s_pre[q] <= k
    auto fetch = [&](uint32_t k, float4& h, float4& lo) {
      h = make_float4(0.f, 0.f, 0.f, 0.f);
      lo = h;
      if (k < run) {
        int a = 0;
#pragma unroll
        for (int step = kP2PSeg / 2; step > 0; step >>= 1)
          if (s_pre[w][a + step] <= k) a += step;
        const uint32_t j = s_first[w][a] + (k - s_pre[w][a]);
        h = sp32[2 * j];
        lo = sp32[2 * j + 1];
      }
    };
    float4 nh, nl;
    fetch((uint32_t)l, nh, nl);
    for (uint32_t k0 = 0; k0 < run; k0 += 64) {
      WaveSync();  // the previous chunk has been read
      s_hi[w][l] = nh;
      s_lo[w][l] = nl;
      WaveSync();
      fetch(k0 + 64 + (uint32_t)l, nh, nl);
      float px = 0.f, py = 0.f, pz = 0.f;
      switch (sh) {
        case 2: P2PChunk<4>(s_hi[w], s_lo[w], grp, th, tl, px, py, pz); break;
        case 3: P2PChunk<8>(s_hi[w], s_lo[w], grp, th, tl, px, py, pz); break;
        case 4: P2PChunk<16>(s_hi[w], s_lo[w], grp, th, tl, px, py, pz); break;
        default: P2PChunk<32>(s_hi[w], s_lo[w], grp, th, tl, px, py, pz); break;
      }
      ax += (double)px;
      ay += (double)py;
      az += (double)pz;
    }
  }
  for (int d = 1 << sh; d < 64; d <<= 1) {
    ax += __shfl_xor(ax, d);
    ay += __shfl_xor(ay, d);
    az += __shfl_xor(az, d);
  }
  if (grp == 0 && valid) {
    double* o = part + ((size_t)it * 32 + slot) * 3;
    o[0] = ax;
    o[1] = ay;
    o[2] = az;
  }
}

// F_i += k_C q_i e 2^40 * (sum over the item partials of i's batch, in list order).  One thread
// per target slot of every leaf batch.
__global__ __launch_bounds__(64 * kP2PWaves) void P2PReduceKernel(
    int nn, const Cell* __restrict__ cells, const int64_t* __restrict__ off, const uint32_t* __restrict__ base,
    const uint32_t* __restrict__ cnt, const double* __restrict__ part, const double4* __restrict__ sp,
    const uint32_t* __restrict__ order, double* __restrict__ fx, double* __restrict__ fy, double* __restrict__ fz) {
  const int p = blockIdx.x * kP2PWaves + (threadIdx.x >> 6);
  if (p >= nn || cnt[p] == 0) return;
  const Cell& T = cells[p];
  const int nseg = P2PSegments(off[p + 1] - off[p]);
  const int g = P2PSegsPerItem(nseg), ni = (nseg + g - 1) / g;
  for (int i = T.b + (threadIdx.x & 63); i < T.e; i += 64) {
    const int b = (i - T.b) / 32, slot = (i - T.b) % 32;
    double sx = 0.0, sy = 0.0, sz = 0.0;
    for (int k = 0; k < ni; ++k) {
      const double* o = part + (((size_t)base[p] + (size_t)b * ni + k) * 32 + slot) * 3;
      sx += o[0];
      sy += o[1];
      sz += o[2];
    }
    const double s32 = kCoulomb * sp[i].w * kElementaryCharge * kInvUnit2;
    const uint32_t o = order[i];
    fx[o] += s32 * sx;
    fy[o] += s32 * sy;
    fz[o] += s32 * sz;
  }
}

static void LaunchL2PForce(int order, hipStream_t st, int nn, const Cell* cells, int nm, const double* L,
                           const double4* sp, const uint32_t* ord, const FmmFrame* fr, double* fx, double* fy,
                           double* fz) {
  const dim3 grid((nn + kP2PWaves - 1) / kP2PWaves), block(64 * kP2PWaves);
#define AM_L2P_CASE(P)                                                                                          \
  case P:                                                                                                       \
    hipLaunchKernelGGL(L2PForceKernel<P>, grid, block, 0, st, nn, cells, nm, L, sp, ord, fr, fx, fy, fz); \
    break;
  switch (order) {
    AM_L2P_CASE(1)
    AM_L2P_CASE(2)
    AM_L2P_CASE(3)
    AM_L2P_CASE(4)
    AM_L2P_CASE(5)
    AM_L2P_CASE(6)
    AM_L2P_CASE(7)
    AM_L2P_CASE(8)
  }
#undef AM_L2P_CASE
}

// Neutral particles (fast-multipole.cc:690-691) and particles with a non-finite coordinate (left out of
// the tree) get zero force.
__global__ void ZeroNeutralForceKernel(int64_t n, const uint8_t* __restrict__ flags, const double* __restrict__ q,
                                       const double* __restrict__ x, const double* __restrict__ y,
                                       const double* __restrict__ z, double* __restrict__ fx,
                                       double* __restrict__ fy, double* __restrict__ fz) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || (flags[i] & kFlagAnnihilated) || (q[i] != 0.0 && Finite3(x[i], y[i], z[i]))) return;
  fx[i] = 0.0;
  fy[i] = 0.0;
  fz[i] = 0.0;
}


// want_forces: forces into the store (sources = alive, q != 0); otherwise the potential energy
// (sources = all alive particles, fast-multipole.cc:335-340 via energy_analysis.cc:71-81).
int FastMultipoleRun(am_store* s, int order, double theta, bool want_forces, double* pe, int precision = AM_FP64) {
  am_ctx* ctx = s->ctx;
  hipStream_t st = ctx->stream;
  const int64_t n = s->n;
  int rc;
  order = std::min(std::max(order, 1), kFmmMaxOrder);
  FmmTab tab;
  if ((rc = DeviceTables(ctx, order, &tab))) return rc;
  if (want_forces) {
    hipLaunchKernelGGL(ZeroNeutralForceKernel, dim3(Blocks(n, 256)), dim3(256), 0, st, n, s->flags, s->q, s->x,
                       s->y, s->z, s->fx, s->fy, s->fz);
    AM_HIP(hipGetLastError());
  }
  void *gathered, *frame, *rx, *ry, *rz;
  if ((rc = Scratch(ctx, "fmm_gathered", (size_t)n * 4, &gathered))) return rc;
  if ((rc = Scratch(ctx, "fmm_frame", sizeof(FmmFrame) + sizeof(RootBox) + 6 * 8, &frame))) return rc;
  RootBox* root = (RootBox*)((char*)frame + sizeof(FmmFrame));
  double* box = (double*)((char*)root + sizeof(RootBox));
  int m = 0;
  if ((rc = GatherParticles(ctx, n, s->flags, s->q, s->x, s->y, s->z, want_forces ? 1 : 0, (uint32_t*)gathered, &m, "fmm"))) return rc;
  if (pe) *pe = 0.0;
  s->fmm_stats = FmmRunStats();
  if (m == 0) return AM_OK;
  if ((rc = GatheredBBox(ctx, (const uint32_t*)gathered, m, s->x, s->y, s->z, box, "fmm"))) return rc;
  hipLaunchKernelGGL(FrameKernel, dim3(1), dim3(1), 0, st, (const double*)box, (FmmFrame*)frame, root);
  if ((rc = Scratch(ctx, "fmm_rx", (size_t)n * 8, &rx))) return rc;
  if ((rc = Scratch(ctx, "fmm_ry", (size_t)n * 8, &ry))) return rc;
  if ((rc = Scratch(ctx, "fmm_rz", (size_t)n * 8, &rz))) return rc;
  hipLaunchKernelGGL(RescaleKernel, dim3(Blocks(m, 256)), dim3(256), 0, st, m, (const uint32_t*)gathered, s->x, s->y,
                     s->z, (const FmmFrame*)frame, (double*)rx, (double*)ry, (double*)rz);
  AM_HIP(hipGetLastError());
  Octree o;
  void* xq;
  if ((rc = Scratch(ctx, "fmm_xq", (size_t)n * sizeof(double4), &xq))) return rc;
  if ((rc = BuildOctree(ctx, (const uint32_t*)gathered, m, (const double*)rx, (const double*)ry, (const double*)rz,
                        s->q, root, 32, "fmm", &o, (double4*)xq)))
    return rc;
  const int nn = o.nn;
  const int nlevels = (int)o.level_start.size() - 1;
  void *sp, *rad, *cells, *M, *L;
  if ((rc = Scratch(ctx, "fmm_sp", (size_t)m * sizeof(double4), &sp))) return rc;
  if ((rc = Scratch(ctx, "fmm_rad", (size_t)nn * 8, &rad))) return rc;
  if ((rc = Scratch(ctx, "fmm_cells", (size_t)nn * sizeof(Cell), &cells))) return rc;
  void *geo, *trv, *tinfo;
  if ((rc = Scratch(ctx, "fmm_cellgeo", (size_t)nn * sizeof(double4), &geo))) return rc;
  if ((rc = Scratch(ctx, "fmm_celltrv", (size_t)nn * sizeof(double4), &trv))) return rc;
  if ((rc = Scratch(ctx, "fmm_cellinfo", (size_t)nn * 4, &tinfo))) return rc;
  if ((rc = Scratch(ctx, "fmm_M", (size_t)nn * tab.nm * 8, &M))) return rc;
  if ((rc = Scratch(ctx, "fmm_L", (size_t)nn * tab.nm * 8, &L))) return rc;
  hipLaunchKernelGGL(SortedCopyKernel, dim3(Blocks(m, 256)), dim3(256), 0, st, m, (const uint32_t*)o.order,
                     (const double4*)xq, (double4*)sp);
  for (int d = nlevels - 1; d >= 0; --d) {
    const int n0 = o.level_start[d], n1 = o.level_start[d + 1];
    hipLaunchKernelGGL(RadiusKernel, dim3(Blocks(n1 - n0, 256)), dim3(256), 0, st, n0, n1, o.nd,
                       (const double4*)sp, (double*)rad);
  }
  hipLaunchKernelGGL(GatherCellsKernel, dim3(Blocks(nn, 256)), dim3(256), 0, st, nn, o.nd, (const uint32_t*)o.pre,
                     (const uint32_t*)o.bfs2pre, (const double*)rad, (Cell*)cells, (double4*)geo, (double4*)trv,
                     (uint32_t*)tinfo);
  AM_HIP(hipGetLastError());
  int slot;
  TimerBegin(ctx, "fmm_upward", &slot);
  hipLaunchKernelGGL(P2MKernel, dim3(nn), dim3(64), 0, st, nn, (const Cell*)cells, (const double4*)sp, tab, (double*)M);
  for (int d = nlevels - 2; d >= 0; --d) {
    const int n0 = o.level_start[d], n1 = o.level_start[d + 1];
    hipLaunchKernelGGL(M2MKernel, dim3(n1 - n0), dim3(64), TransferLds(tab, false), st, n0, o.nd,
                       (const uint32_t*)o.bfs2pre, tab,
                       (double*)M);
  }
  TimerEnd(ctx, slot);
  AM_HIP(hipGetLastError());

  // Dual-tree traversal, level-synchronous.
  const double theta2 = theta * theta;
  int64_t n_m2l = 0, n_p2p = 0;
  size_t cap_m2l = (size_t)64 * nn + 1024, cap_p2p = (size_t)32 * nn + 1024;
  void *m2l, *p2p;
  if ((rc = Scratch(ctx, "fmm_m2l", cap_m2l * 8, &m2l))) return rc;
  if ((rc = Scratch(ctx, "fmm_p2p", cap_p2p * 8, &p2p))) return rc;
  void *fa, *fb;
  size_t cap_fa = (size_t)8 * nn + 64, cap_fb = cap_fa;
  if ((rc = Scratch(ctx, "fmm_front_a", cap_fa * 8, &fa))) return rc;
  if ((rc = Scratch(ctx, "fmm_front_b", cap_fb * 8, &fb))) return rc;
  {
    const uint2 rootpair = make_uint2(0u, 0u);
    AM_HIP(hipMemcpyAsync(fa, &rootpair, 8, hipMemcpyHostToDevice, st));
    AM_HIP(hipStreamSynchronize(st));
  }
  int64_t np = 1;
  TimerBegin(ctx, "fmm_traverse", &slot);
  while (np > 0) {
    const int64_t ntiles = (np + kTrvTile - 1) / kTrvTile;
    void *kinds, *tcnt, *toff, *tot_dev;
    if ((rc = Scratch(ctx, "fmm_trav_kind", (size_t)np, &kinds))) return rc;
    if ((rc = Scratch(ctx, "fmm_trav_tcnt", (size_t)ntiles * 3 * 4, &tcnt))) return rc;
    if ((rc = Scratch(ctx, "fmm_trav_toff", (size_t)ntiles * 3 * 4, &toff))) return rc;
    if ((rc = Scratch(ctx, "fmm_trav_tot", 16, &tot_dev))) return rc;
    hipLaunchKernelGGL(TraverseTileCountKernel, dim3((unsigned)ntiles), dim3(kTrvThreads), 0, st, (int)np,
                       (const uint2*)fa, (const double4*)trv, (const uint32_t*)tinfo, theta2, (uint8_t*)kinds,
                       (uint32_t*)tcnt);
    hipLaunchKernelGGL(TraverseTileScanKernel, dim3(1), dim3(kTrvScanThreads), 0, st, ntiles, (const uint32_t*)tcnt,
                       (uint32_t*)toff, (uint32_t*)tot_dev);
    AM_HIP(hipGetLastError());
    uint32_t tot[3];
    AM_HIP(hipMemcpyAsync(tot, tot_dev, 12, hipMemcpyDeviceToHost, st));
    AM_HIP(hipStreamSynchronize(st));
    // grow outputs (preserving contents)
    auto grow = [&](const char* name, void** buf, size_t& cap, size_t need, size_t used, size_t elem) -> int {
      if (need <= cap) return AM_OK;
      const size_t held = ctx->scratch[name].bytes / elem;  // a previous call's buffer may be larger
      if (need <= held) {
        cap = held;
        return AM_OK;
      }
      const size_t ncap = std::max(need, cap * 2);
      void* tmp = nullptr;
      if (used) {
        AM_HIP(hipMalloc(&tmp, used * elem));
        AM_HIP(hipMemcpyAsync(tmp, *buf, used * elem, hipMemcpyDeviceToDevice, st));
        AM_HIP(hipStreamSynchronize(st));
      }
      int r = Scratch(ctx, name, ncap * elem, buf);
      if (r) return r;
      if (used) {
        AM_HIP(hipMemcpyAsync(*buf, tmp, used * elem, hipMemcpyDeviceToDevice, st));
        AM_HIP(hipStreamSynchronize(st));
        AM_HIP(hipFree(tmp));
      }
      cap = ncap;
      return AM_OK;
    };
    if ((rc = grow("fmm_m2l", &m2l, cap_m2l, (size_t)n_m2l + tot[0], (size_t)n_m2l, 8))) return rc;
    if ((rc = grow("fmm_p2p", &p2p, cap_p2p, (size_t)n_p2p + tot[1], (size_t)n_p2p, 8))) return rc;
    if ((rc = grow("fmm_front_b", &fb, cap_fb, tot[2], 0, 8))) return rc;
    hipLaunchKernelGGL(TraverseTileWriteKernel, dim3((unsigned)ntiles), dim3(kTrvThreads), 0, st, (int)np,
                       (const uint2*)fa, (const uint8_t*)kinds, (const uint32_t*)tinfo, (const Cell*)cells,
                       (const uint32_t*)toff, (uint64_t*)m2l + n_m2l, (uint64_t*)p2p + n_p2p, (uint2*)fb);
    AM_HIP(hipGetLastError());
    n_m2l += tot[0];
    n_p2p += tot[1];
    np = tot[2];
    std::swap(fa, fb);
    std::swap(cap_fa, cap_fb);
    // keep the scratch names pointing at the right buffers
    std::swap(ctx->scratch["fmm_front_a"], ctx->scratch["fmm_front_b"]);
  }
  // Both lists are grouped by target cell for the kernels that walk them per target: a stable sort of
  // the (T << 32 | S) pairs on the cb bits of T.  The MFMA M2L sorts its pairs by (target group,
  // class) instead, straight from the traversal's order.
  const bool mfma_plan = precision == AM_FP32 && want_forces && tab.nm <= 56;
  int cb = 1;
  while ((1 << cb) < nn) ++cb;
  void *off_m2l = nullptr, *off_p2p;
  bool m2l_sorted = false;
  auto sort_m2l = [&]() -> int {
    int r;
    if ((r = RadixSortKeys(ctx, (uint64_t*)m2l, n_m2l, 32, cb))) return r;
    if ((r = Scratch(ctx, "fmm_off_m2l", (size_t)(nn + 1) * 8, &off_m2l))) return r;
    hipLaunchKernelGGL(PairOffsetsKernel, dim3(Blocks(nn + 1, 256)), dim3(256), 0, st, nn, (const uint64_t*)m2l,
                       n_m2l, (int64_t*)off_m2l);
    AM_HIP(hipGetLastError());
    m2l_sorted = true;
    return AM_OK;
  };
  if (!mfma_plan && (rc = sort_m2l())) return rc;
  if ((rc = RadixSortKeys(ctx, (uint64_t*)p2p, n_p2p, 32, cb))) return rc;
  if ((rc = Scratch(ctx, "fmm_off_p2p", (size_t)(nn + 1) * 8, &off_p2p))) return rc;
  hipLaunchKernelGGL(PairOffsetsKernel, dim3(Blocks(nn + 1, 256)), dim3(256), 0, st, nn, (const uint64_t*)p2p, n_p2p,
                     (int64_t*)off_p2p);
  TimerEnd(ctx, slot);
  AM_HIP(hipGetLastError());

  uint32_t* m2l_counts = nullptr;  // the unclassed-pair count
  TimerBegin(ctx, "fmm_m2l", &slot);
  {
    // classes of the M2L pairs (table of at most 2^22 slots, load factor <= 1/2 for up to 2^21 keys)
    uint64_t tsize = 1024;
    while (tsize < 2 * (uint64_t)std::max<int64_t>(n_m2l, 1) && tsize < ((uint64_t)1 << 22)) tsize <<= 1;
    void *hk, *pslot, *occ, *ids, *cnt;
    if ((rc = Scratch(ctx, "fmm_cls_keys", tsize * 8, &hk))) return rc;
    if ((rc = Scratch(ctx, "fmm_pair_slot", (size_t)std::max<int64_t>(n_m2l, 1) * 4, &pslot))) return rc;
    if ((rc = Scratch(ctx, "fmm_cls_occ", tsize * 4, &occ))) return rc;
    if ((rc = Scratch(ctx, "fmm_cls_ids", (tsize + 1) * 4, &ids))) return rc;
    if ((rc = Scratch(ctx, "fmm_m2l_counts", 4, &cnt))) return rc;
    m2l_counts = (uint32_t*)cnt;
    hipLaunchKernelGGL(ClassClearKernel, dim3(Blocks((int64_t)tsize, 256)), dim3(256), 0, st, tsize, (uint64_t*)hk,
                       m2l_counts);
    auto insert = [&]() {
      if (n_m2l > 0)
        hipLaunchKernelGGL(M2LClassInsertKernel, dim3(Blocks(n_m2l, 256)), dim3(256), 0, st, n_m2l,
                           (const uint64_t*)m2l, (const double4*)geo, (uint64_t*)hk, tsize - 1,
                           ctx->fmm_classes ? kClassProbes : 0, (uint32_t*)pslot, m2l_counts);
    };
    insert();
    hipLaunchKernelGGL(ClassOccupiedKernel, dim3(Blocks((int64_t)tsize, 256)), dim3(256), 0, st, tsize,
                       (const uint64_t*)hk, (uint32_t*)occ);
    AM_HIP(hipGetLastError());
    if ((rc = ExclusiveScan(ctx, (const uint32_t*)occ, (uint32_t*)ids, (int64_t)tsize))) return rc;
    uint32_t n_cls = 0, sort_err = 0, n_uncl = 0;  // (the traversal's list sorts are checked here too)
    AM_HIP(hipMemcpyAsync(&n_cls, (uint32_t*)ids + tsize, 4, hipMemcpyDeviceToHost, st));
    AM_HIP(hipMemcpyAsync(&n_uncl, m2l_counts, 4, hipMemcpyDeviceToHost, st));
    if (const uint32_t* w = SortErrorWord(ctx)) AM_HIP(hipMemcpyAsync(&sort_err, w, 4, hipMemcpyDeviceToHost, st));
    AM_HIP(hipStreamSynchronize(st));
    if ((rc = CheckSortErrorAfterSync(ctx, sort_err))) return rc;
    void *ckey, *Dc;
    if ((rc = Scratch(ctx, "fmm_cls_key", (size_t)std::max<uint32_t>(n_cls, 1) * 8, &ckey))) return rc;
    if ((rc = Scratch(ctx, "fmm_cls_tensor", (size_t)std::max<uint32_t>(n_cls, 1) * tab.nd * 8, &Dc))) return rc;
    hipLaunchKernelGGL(ClassListKernel, dim3(Blocks((int64_t)tsize, 256)), dim3(256), 0, st, tsize,
                       (const uint64_t*)hk, (const uint32_t*)ids, (uint64_t*)ckey);
    if (n_cls > 0)
      hipLaunchKernelGGL(ClassTensorKernel, dim3(n_cls), dim3(64), sizeof(double) * tab.nd, st, (int)n_cls,
                         (const uint64_t*)ckey, (const Cell*)cells, tab, (double*)Dc);
    AM_HIP(hipGetLastError());
    // The MFMA path gives each unclassed pair a tensor of its own, up to a bound; past it (or with
    // no pairs) the FP64 contraction runs.
    const bool mfma = mfma_plan && n_m2l > 0 && (int64_t)n_uncl <= ctx->fmm_max_unclassed;
    if (mfma) {
      const int nrb = tab.nm <= 36 ? 3 : 4, nks = tab.nm <= 36 ? 9 : 14;  // the instance below
      const int nkp = (nks + 3) & ~3;
      const int64_t n_tens = (int64_t)n_cls + n_uncl;
      void *Dc32, *M32, *keys, *vals = nullptr, *urank = nullptr;
      if ((rc = Scratch(ctx, "fmm_cls_tensor32", (size_t)std::max<int64_t>(n_tens, 1) * tab.nd * 4, &Dc32))) return rc;
      if ((rc = Scratch(ctx, "fmm_M32", (size_t)nn * 4 * nkp * 4, &M32))) return rc;
      const int64_t nt = (int64_t)n_cls * tab.nd;
      if (nt > 0)
        hipLaunchKernelGGL(TensorToF32Kernel, dim3(Blocks(nt, 256)), dim3(256), 0, st, nt, (const double*)Dc,
                           (float*)Dc32);
      if (n_uncl > 0) {
        void *flag, *ulist;
        if ((rc = Scratch(ctx, "fmm_uncl_flag", (size_t)n_m2l * 4, &flag))) return rc;
        if ((rc = Scratch(ctx, "fmm_uncl_rank", (size_t)(n_m2l + 1) * 4, &urank))) return rc;
        if ((rc = Scratch(ctx, "fmm_uncl_list", (size_t)n_uncl * 4, &ulist))) return rc;
        hipLaunchKernelGGL(UnclassedFlagKernel, dim3(Blocks(n_m2l, 256)), dim3(256), 0, st, n_m2l,
                           (const uint32_t*)pslot, (uint32_t*)flag);
        AM_HIP(hipGetLastError());
        if ((rc = ExclusiveScan(ctx, (const uint32_t*)flag, (uint32_t*)urank, n_m2l))) return rc;
        hipLaunchKernelGGL(UnclassedListKernel, dim3(Blocks(n_m2l, 256)), dim3(256), 0, st, n_m2l,
                           (const uint32_t*)pslot, (const uint32_t*)urank, (uint32_t*)ulist);
        hipLaunchKernelGGL(UnclassedTensorKernel, dim3(n_uncl), dim3(64), sizeof(double) * tab.nd, st, (int)n_uncl,
                           (const uint32_t*)ulist, (const uint64_t*)m2l, (const Cell*)cells, tab,
                           (float*)Dc32 + (size_t)n_cls * tab.nd);
        AM_HIP(hipGetLastError());
      }
      hipLaunchKernelGGL(MultipoleToF32Kernel, dim3(Blocks((int64_t)nn * 4 * nkp, 256)), dim3(256), 0, st, nn,
                         tab.nm, nkp, (const double*)M, tab.sign, (float*)M32);
      // sort keys: (group, class, target within the group[, source])
      const int n_groups = (nn + kM2LGroupTargets - 1) >> kM2LGroupBits;
      int cbits = 1, gbits = 1;
      while (((int64_t)1 << cbits) < n_tens) ++cbits;
      while ((1 << gbits) < n_groups) ++gbits;
      M2LKeyLayout kl;
      kl.packed = cb + kM2LGroupBits + cbits + gbits <= 64 && n_m2l < ((int64_t)1 << 30) - 1;
      kl.sbits = cb;
      kl.smask = ((uint64_t)1 << cb) - 1;
      kl.cls_shift = (kl.packed ? cb : 0) + kM2LGroupBits;
      kl.gshift = kl.cls_shift + cbits;
      kl.cls_mask = ((uint64_t)1 << cbits) - 1;
      if ((rc = Scratch(ctx, "fmm_m2l_keys", (size_t)n_m2l * 8, &keys))) return rc;
      if (!kl.packed && (rc = Scratch(ctx, "fmm_m2l_vals", (size_t)n_m2l * 4, &vals))) return rc;
      hipLaunchKernelGGL(M2LGroupKeyKernel, dim3(Blocks(n_m2l, 256)), dim3(256), 0, st, n_m2l, (const uint64_t*)m2l,
                         (const uint32_t*)o.pre, (const uint32_t*)pslot, (const uint32_t*)ids,
                         (const uint32_t*)urank, n_cls, kl, (uint64_t*)keys, (uint32_t*)vals);
      AM_HIP(hipGetLastError());
      // a stable sort on (group, class): within a run the traversal's order is kept
      if (kl.packed) rc = RadixSortKeys(ctx, (uint64_t*)keys, n_m2l, kl.cls_shift, cbits + gbits);
      else rc = RadixSortPairs(ctx, (uint64_t*)keys, (uint32_t*)vals, n_m2l, kl.gshift + gbits);
      if (rc) return rc;
      void* goff_dev;
      if ((rc = Scratch(ctx, "fmm_m2l_goff", (size_t)(n_groups + 1) * 8, &goff_dev))) return rc;
      hipLaunchKernelGGL(GroupOffsetsKernel, dim3(Blocks(n_groups + 1, 256)), dim3(256), 0, st, n_groups,
                         (const uint64_t*)keys, n_m2l, kl.gshift, (int64_t*)goff_dev);
      AM_HIP(hipGetLastError());
      std::vector<int64_t> goff(n_groups + 1);
      AM_HIP(hipMemcpyAsync(goff.data(), goff_dev, goff.size() * 8, hipMemcpyDeviceToHost, st));
      if (const uint32_t* w = SortErrorWord(ctx)) AM_HIP(hipMemcpyAsync(&sort_err, w, 4, hipMemcpyDeviceToHost, st));
      AM_HIP(hipStreamSynchronize(st));
      if ((rc = CheckSortErrorAfterSync(ctx, sort_err))) return rc;
      // Items: each group's pair list in pieces of about n_m2l / (4 CUs) pairs, so that every CU
      // takes several workgroups; every group has at least one item (its locals are written even
      // when it has no pairs).
      const int64_t piece = std::max<int64_t>(4096, n_m2l / (4 * std::max(ctx->num_cus, 1)));
      std::vector<M2LItem> items;
      std::vector<int> gitem(n_groups + 1);
      for (int g = 0; g < n_groups; ++g) {
        gitem[g] = (int)items.size();
        const int64_t b = goff[g], c = goff[g + 1] - b, k = std::max<int64_t>(1, (c + piece - 1) / piece);
        for (int64_t i = 0; i < k; ++i) items.push_back(M2LItem{b + c * i / k, b + c * (i + 1) / k, g, 0});
      }
      gitem[n_groups] = (int)items.size();
      const int n_items = (int)items.size();
      void *items_dev, *gitem_dev, *partial;
      if ((rc = Scratch(ctx, "fmm_m2l_items", (size_t)n_items * sizeof(M2LItem), &items_dev))) return rc;
      if ((rc = Scratch(ctx, "fmm_m2l_gitem", gitem.size() * 4, &gitem_dev))) return rc;
      if ((rc = Scratch(ctx, "fmm_m2l_partial", (size_t)n_items * kM2LGroupTargets * tab.nm * 8, &partial))) return rc;
      AM_HIP(hipMemcpyAsync(items_dev, items.data(), (size_t)n_items * sizeof(M2LItem), hipMemcpyHostToDevice, st));
      AM_HIP(hipMemcpyAsync(gitem_dev, gitem.data(), gitem.size() * 4, hipMemcpyHostToDevice, st));
      AM_HIP(hipStreamSynchronize(st));  // the host vectors go out of scope below
      // Products: each wave's slice of an item cut into runs of one class of at most 16 pairs,
      // written as a table (a counting pass, a scan, a writing pass), so the product kernel reads
      // its operands' addresses without a chain of dependent loads.
      const int waves = tab.nm <= 36 ? kM2LWavesP4 : kM2LWavesP5;  // of the instance below
      const int n_slices = n_items * waves;
      // the bounds of the exact accumulation
      void *mnorm, *dmax, *gmax;
      if ((rc = Scratch(ctx, "fmm_m2l_mnorm", (size_t)nn * (tab.P + 1) * 4, &mnorm))) return rc;
      if ((rc = Scratch(ctx, "fmm_m2l_dmax", (size_t)std::max<int64_t>(n_tens, 1) * (2 * tab.P + 1) * 4, &dmax)))
        return rc;
      if ((rc = Scratch(ctx, "fmm_m2l_gmax", (size_t)n_groups * (tab.P + 1) * 4, &gmax))) return rc;
      hipLaunchKernelGGL(CellNormKernel, dim3(Blocks(nn, 256)), dim3(256), 0, st, nn, tab, (const double*)M,
                         (float*)mnorm);
      hipLaunchKernelGGL(ClassDegMaxKernel, dim3((unsigned)n_tens), dim3(64), 0, st, (int)n_tens, (const float*)Dc32,
                         tab, (float*)dmax);
      hipLaunchKernelGGL(M2LZeroKernel, dim3(Blocks((int64_t)n_groups * (tab.P + 1), 256)), dim3(256), 0, st,
                         (int64_t)n_groups * (tab.P + 1), (uint32_t*)gmax);
      AM_HIP(hipGetLastError());
      void *scnt, *soff, *pstart, *pcls, *pcol;
      if ((rc = Scratch(ctx, "fmm_m2l_slice_cnt", (size_t)n_slices * 4, &scnt))) return rc;
      if ((rc = Scratch(ctx, "fmm_m2l_slice_off", (size_t)(n_slices + 1) * 4, &soff))) return rc;
      hipLaunchKernelGGL(M2LProductKernel, dim3(n_slices), dim3(64), 0, st, n_slices, waves,
                         (const M2LItem*)items_dev, (const uint64_t*)keys, (const uint32_t*)vals, kl,
                         (const uint32_t*)nullptr, (uint32_t*)scnt, (uint32_t*)nullptr, (uint32_t*)nullptr,
                         (uint32_t*)nullptr, tab.P, (const float*)mnorm, (const float*)dmax, (uint32_t*)gmax);
      AM_HIP(hipGetLastError());
      if ((rc = ExclusiveScan(ctx, (const uint32_t*)scnt, (uint32_t*)soff, n_slices))) return rc;
      uint32_t n_prod = 0;
      AM_HIP(hipMemcpyAsync(&n_prod, (uint32_t*)soff + n_slices, 4, hipMemcpyDeviceToHost, st));
      AM_HIP(hipStreamSynchronize(st));
      if ((rc = Scratch(ctx, "fmm_m2l_pstart", (size_t)(n_prod + 1) * 4, &pstart))) return rc;
      if ((rc = Scratch(ctx, "fmm_m2l_pcls", (size_t)std::max<uint32_t>(n_prod, 1) * 4, &pcls))) return rc;
      if ((rc = Scratch(ctx, "fmm_m2l_pcol", (size_t)std::max<uint32_t>(n_prod, 1) * 64, &pcol))) return rc;
      hipLaunchKernelGGL(M2LProductKernel, dim3(n_slices), dim3(64), 0, st, n_slices, waves,
                         (const M2LItem*)items_dev, (const uint64_t*)keys, (const uint32_t*)vals, kl,
                         (const uint32_t*)soff, (uint32_t*)nullptr, (uint32_t*)pstart, (uint32_t*)pcls,
                         (uint32_t*)pcol, tab.P, (const float*)mnorm, (const float*)dmax, (uint32_t*)gmax);
      AM_HIP(hipGetLastError());
      const size_t lds = M2LGroupLds(tab.nm, tab.nd, nrb, nks, waves);
      auto launch_group = [&](auto kernel) -> int {
        AM_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)lds));
        hipLaunchKernelGGL(kernel, dim3(n_items), dim3(64 * waves), lds, st, (const M2LItem*)items_dev,
                           (const int*)gitem_dev, (const uint32_t*)soff, (const uint32_t*)pstart,
                           (const uint32_t*)pcls, (const uint32_t*)pcol, (const float*)Dc32, tab, (const float*)M32,
                           (const uint32_t*)gmax, (double*)partial);
        AM_HIP(hipGetLastError());
        return AM_OK;
      };
      int mslot;
      TimerBegin(ctx, "fmm_m2l_mfma", &mslot);
      if (tab.nm <= 36) rc = launch_group(M2LGroupKernel<3, 9, 3, kM2LWavesP4>);
      else rc = launch_group(M2LGroupKernel<4, 14, 5, kM2LWavesP5>);
      TimerEnd(ctx, mslot);
      if (rc) return rc;
      hipLaunchKernelGGL(M2LPartialReduceKernel, dim3(Blocks((int64_t)nn * tab.nm, 256)), dim3(256), 0, st, nn,
                         tab.nm, (const int*)gitem_dev, (const double*)partial, tab.inv_fact,
                         (const uint32_t*)o.bfs2pre, (double*)L);
      AM_HIP(hipGetLastError());
      s->fmm_stats.m2l_items = n_items;
    } else {
      if (!m2l_sorted) {  // the list in target order, and its slots again in that order
        if ((rc = sort_m2l())) return rc;
        insert();
      }
      const size_t m2l_lds =
          sizeof(double) * ((size_t)kM2LWaves * (tab.nd + tab.nm) + tab.nm) +
          (tab.P <= kM2LLdsOrder ? sizeof(int) * ((size_t)tab.nm * tab.nm + (size_t)tab.nd * 10) : 0);
      if (n_m2l > 0)
        hipLaunchKernelGGL(PairClassKernel, dim3(Blocks(n_m2l, 256)), dim3(256), 0, st, n_m2l, (const uint32_t*)pslot,
                           (const uint32_t*)ids, (uint32_t*)pslot);  // in place: slot -> class id
      const int dpl = (tab.nd + 63) / 64;
      auto launch = [&](auto kernel) {
        hipLaunchKernelGGL(kernel, dim3(Blocks(nn, kM2LWaves)), dim3(64 * kM2LWaves), m2l_lds, st, nn,
                           (const Cell*)cells, (const uint64_t*)m2l, (const int64_t*)off_m2l, (const uint32_t*)pslot,
                           (const double*)Dc, tab, (const double*)M, (double*)L);
      };
      if (tab.nm <= kRegNm - 1 && dpl <= 3) launch(M2LKernel<3, true>);
      else if (dpl <= 3) launch(M2LKernel<3, false>);
      else if (dpl <= 5) launch(M2LKernel<5, false>);
      else if (dpl <= 8) launch(M2LKernel<8, false>);
      else launch(M2LKernel<16, false>);
    }
    s->fmm_stats.m2l_classes = n_cls;
    s->fmm_stats.m2l_unclassed = n_uncl;
  }
  TimerEnd(ctx, slot);
  AM_HIP(hipGetLastError());
  TimerBegin(ctx, "fmm_downward", &slot);
  for (int d = 0; d < nlevels - 1; ++d) {
    const int n0 = o.level_start[d], n1 = o.level_start[d + 1];
    hipLaunchKernelGGL(L2LKernel, dim3(n1 - n0), dim3(64), TransferLds(tab, true), st, n0, o.nd,
                       (const uint32_t*)o.bfs2pre, tab,
                       (double*)L);
  }
  TimerEnd(ctx, slot);
  AM_HIP(hipGetLastError());
  void* sp32 = nullptr;
  if (want_forces && precision == AM_FP32) {
    if ((rc = Scratch(ctx, "fmm_sp32", (size_t)m * 2 * sizeof(float4), &sp32))) return rc;
    hipLaunchKernelGGL(SortedSrcKernel, dim3(Blocks(m, 256)), dim3(256), 0, st, m, (const uint32_t*)o.order, s->src,
                       (float4*)sp32);
    AM_HIP(hipGetLastError());
  }
  void* phi = nullptr;
  if (!want_forces) {
    if ((rc = Scratch(ctx, "fmm_phi", (size_t)m * 8, &phi))) return rc;
  }
  TimerBegin(ctx, "fmm_l2p_p2p", &slot);
  if (sp32) {
    void *icnt, *ibase, *items, *part;
    if ((rc = Scratch(ctx, "fmm_p2p_icnt", (size_t)nn * 4, &icnt))) return rc;
    if ((rc = Scratch(ctx, "fmm_p2p_ibase", (size_t)(nn + 1) * 4, &ibase))) return rc;
    hipLaunchKernelGGL(P2PItemCountKernel, dim3(Blocks(nn, 256)), dim3(256), 0, st, nn, (const Cell*)cells,
                       (const int64_t*)off_p2p, (uint32_t*)icnt);
    AM_HIP(hipGetLastError());
    if ((rc = ExclusiveScan(ctx, (const uint32_t*)icnt, (uint32_t*)ibase, nn))) return rc;
    uint32_t n_items = 0, sort_err = 0;
    AM_HIP(hipMemcpyAsync(&n_items, (uint32_t*)ibase + nn, 4, hipMemcpyDeviceToHost, st));
    if (const uint32_t* w = SortErrorWord(ctx)) AM_HIP(hipMemcpyAsync(&sort_err, w, 4, hipMemcpyDeviceToHost, st));
    AM_HIP(hipStreamSynchronize(st));
    if ((rc = CheckSortErrorAfterSync(ctx, sort_err))) return rc;
    if ((rc = Scratch(ctx, "fmm_p2p_items", (size_t)std::max<uint32_t>(n_items, 1) * sizeof(P2PItem), &items)))
      return rc;
    if ((rc = Scratch(ctx, "fmm_p2p_part", (size_t)std::max<uint32_t>(n_items, 1) * 32 * 3 * 8, &part))) return rc;
    hipLaunchKernelGGL(P2PItemListKernel, dim3(Blocks(nn, 256)), dim3(256), 0, st, nn, (const Cell*)cells,
                       (const int64_t*)off_p2p, (const uint32_t*)ibase, (const uint32_t*)icnt, (P2PItem*)items);
    AM_HIP(hipGetLastError());
    LaunchL2PForce(order, st, nn, (const Cell*)cells, tab.nm, (const double*)L, (const double4*)sp,
                   (const uint32_t*)o.order, (const FmmFrame*)frame, s->fx, s->fy, s->fz);
    AM_HIP(hipGetLastError());
    if (n_items > 0)
      hipLaunchKernelGGL(P2PItemKernel, dim3((n_items + kP2PWaves - 1) / kP2PWaves), dim3(64 * kP2PWaves), 0, st,
                         n_items, (const P2PItem*)items, (const Cell*)cells, (const uint64_t*)p2p,
                         (const int64_t*)off_p2p, (const float4*)sp32, (double*)part);
    AM_HIP(hipGetLastError());
    hipLaunchKernelGGL(P2PReduceKernel, dim3((nn + kP2PWaves - 1) / kP2PWaves), dim3(64 * kP2PWaves), 0, st, nn,
                       (const Cell*)cells, (const int64_t*)off_p2p, (const uint32_t*)ibase, (const uint32_t*)icnt,
                       (const double*)part, (const double4*)sp, (const uint32_t*)o.order, s->fx, s->fy, s->fz);
  } else
    hipLaunchKernelGGL(L2PP2PKernel, dim3(nn), dim3(64), 0, st, nn, (const Cell*)cells, (const uint64_t*)p2p,
                       (const int64_t*)off_p2p, tab, (const double*)L, (const double4*)sp, (const uint32_t*)o.order,
                       (const FmmFrame*)frame, want_forces ? 1 : 0, s->fx, s->fy, s->fz, (double*)phi);
  TimerEnd(ctx, slot);
  AM_HIP(hipGetLastError());
  s->fmm_stats.cells = nn;
  s->fmm_stats.m2l_pairs = n_m2l;
  s->fmm_stats.p2p_cell_pairs = n_p2p;
  s->fmm_stats.m = m;
  s->fmm_stats.max_depth = nlevels - 1;
  if (!want_forces) {
    double sum = 0.0;
    if ((rc = ReduceSum(ctx, (const double*)phi, m, &sum))) return rc;
    FmmFrame fh;
    AM_HIP(hipMemcpyAsync(&fh, frame, sizeof(fh), hipMemcpyDeviceToHost, st));
    AM_HIP(hipStreamSynchronize(st));
    *pe = 0.5 * kCoulomb / fh.length * sum;
  }
  return AM_OK;
}

int FastMultipolePotentialEnergy(am_store* s, int order, double theta, double* pe) {
  return FastMultipoleRun(s, order, theta, false, pe);
}

}  // namespace am

using namespace am;

extern "C" {

int am_compute_fast_multipole_coulomb_forces(am_store* s, int order, double theta) {
  if (!s) return Fail(AM_ERR_INVALID, "am_compute_fast_multipole_coulomb_forces: NULL store");
  if (!(theta >= 0.0)) return Fail(AM_ERR_INVALID, "fast_multipole_theta must be >= 0");
  if (s->n == 0) return AM_OK;
  if (s->n > (int64_t)1 << 30) return Fail(AM_ERR_CAPACITY, "FMM: n too large");
  AM_HIP(hipSetDevice(s->ctx->device));
  return FastMultipoleRun(s, order, theta, true, nullptr);
}

// FMM with a choice of arithmetic: AM_FP64 (the reference arithmetic) or AM_FP32 (P2P on hi/lo FP32
// coordinates with FP64 accumulation per work item; for orders <= 5 the M2L contraction on FP32
// matrix cores; the tree, P2M, M2M, L2L and L2P stay FP64).
int am_compute_fast_multipole_coulomb_forces_ex(am_store* s, int order, double theta, int precision) {
  if (!s) return Fail(AM_ERR_INVALID, "am_compute_fast_multipole_coulomb_forces_ex: NULL store");
  if (precision != AM_FP32 && precision != AM_FP64) return Fail(AM_ERR_INVALID, "unknown precision");
  if (!(theta >= 0.0)) return Fail(AM_ERR_INVALID, "fast_multipole_theta must be >= 0");
  if (s->n == 0) return AM_OK;
  if (s->n > (int64_t)1 << 30) return Fail(AM_ERR_CAPACITY, "FMM: n too large");
  AM_HIP(hipSetDevice(s->ctx->device));
  return FastMultipoleRun(s, order, theta, true, nullptr, precision);
}

// FMM statistics of the most recent run: sources, cells, M2L cell pairs, P2P cell pairs, depth.
int am_fmm_stats(am_store* s, int64_t* n_sources, int64_t* n_cells, int64_t* m2l_pairs, int64_t* p2p_cell_pairs,
                 int32_t* max_depth, int64_t* m2l_classes) {
  if (!s) return Fail(AM_ERR_INVALID, "am_fmm_stats: NULL store");
  if (n_sources) *n_sources = s->fmm_stats.m;
  if (n_cells) *n_cells = s->fmm_stats.cells;
  if (m2l_pairs) *m2l_pairs = s->fmm_stats.m2l_pairs;
  if (p2p_cell_pairs) *p2p_cell_pairs = s->fmm_stats.p2p_cell_pairs;
  if (max_depth) *max_depth = s->fmm_stats.max_depth;
  if (m2l_classes) *m2l_classes = s->fmm_stats.m2l_classes;
  return AM_OK;
}

int am_fmm_m2l_stats(am_store* s, int64_t* unclassed, int64_t* items) {
  if (!s) return Fail(AM_ERR_INVALID, "am_fmm_m2l_stats: NULL store");
  if (unclassed) *unclassed = s->fmm_stats.m2l_unclassed;
  if (items) *items = s->fmm_stats.m2l_items;
  return AM_OK;
}

int am_debug_fmm_m2l_options(am_ctx* ctx, int classes, int64_t max_unclassed) {
  if (!ctx) return Fail(AM_ERR_INVALID, "am_debug_fmm_m2l_options: NULL context");
  ctx->fmm_classes = classes != 0;
  ctx->fmm_max_unclassed = max_unclassed;
  return AM_OK;
}

}  // extern "C"
