// Stable LSD radix sort of (uint64 key, uint32 value) pairs and an exclusive prefix scan.
//
// Sort: 8-bit digits, one pass per digit of the key's significant bits.  Each pass:
//   1. RadixHistKernel: per-tile (2048 keys) digit histogram in LDS -> hist[digit][tile].
//   2. ExclusiveScan over hist (digit-major), giving every (digit, tile) its global base.
//   3. RadixScatterKernel: the tile is processed in 8 rounds of 256 keys in input order; inside a
//      wave the lanes with equal digits are found with 8 ballots (one per digit bit), so a key's
//      rank = base[digit] + keys of that digit in earlier waves of the round + earlier lanes of
//      its own wave.  Input order is preserved within a digit: the sort is stable.
// Scan: see ExclusiveScan below.
#include <algorithm>

#include "am_internal.hpp"

namespace am {

constexpr int kSortThreads = 256;
constexpr int kSortRounds = 8;
constexpr int kSortTile = kSortThreads * kSortRounds;  // 2048 keys per workgroup

__global__ __launch_bounds__(kSortThreads) void RadixHistKernel(const uint64_t* __restrict__ keys, int64_t n,
                                                                 int shift, uint32_t* __restrict__ hist,
                                                                 int num_tiles) {
  __shared__ uint32_t h[256];
  const int t = threadIdx.x;
  h[t] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * kSortTile;
  for (int r = 0; r < kSortRounds; ++r) {
    const int64_t idx = base + r * kSortThreads + t;
    if (idx < n) atomicAdd(&h[(keys[idx] >> shift) & 255u], 1u);
  }
  __syncthreads();
  hist[(size_t)t * num_tiles + blockIdx.x] = h[t];
}

__global__ __launch_bounds__(kSortThreads) void RadixScatterKernel(
    const uint64_t* __restrict__ kin, const uint32_t* __restrict__ vin, uint64_t* __restrict__ kout,
    uint32_t* __restrict__ vout, int64_t n, int shift, const uint32_t* __restrict__ offsets, int num_tiles) {
  __shared__ uint32_t base[256];
  __shared__ uint32_t wcount[kSortThreads / 64][256];
  const int t = threadIdx.x;
  const int wave = t >> 6;
  const int lane = t & 63;
  base[t] = offsets[(size_t)t * num_tiles + blockIdx.x];
  for (int w = 0; w < kSortThreads / 64; ++w) wcount[w][t] = 0;
  __syncthreads();
  const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  const int64_t tile0 = (int64_t)blockIdx.x * kSortTile;
  for (int r = 0; r < kSortRounds; ++r) {
    const int64_t idx = tile0 + r * kSortThreads + t;
    const bool valid = idx < n;
    const uint64_t key = valid ? kin[idx] : 0ull;
    const uint32_t val = valid ? vin[idx] : 0u;
    const uint32_t d = (uint32_t)(key >> shift) & 255u;
    uint64_t mask = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const bool bit = (d >> b) & 1u;
      const uint64_t bits = __ballot(bit);
      mask &= bit ? bits : ~bits;
    }
    const uint32_t rank = (uint32_t)__popcll(mask & lt_mask);
    if (valid && rank == 0) wcount[wave][d] = (uint32_t)__popcll(mask);
    __syncthreads();
    if (valid) {
      uint32_t pre = 0;
      for (int w = 0; w < wave; ++w) pre += wcount[w][d];
      const uint32_t pos = base[d] + pre + rank;
      kout[pos] = key;
      vout[pos] = val;
    }
    __syncthreads();
    uint32_t tot = 0;
    for (int w = 0; w < kSortThreads / 64; ++w) {
      tot += wcount[w][t];
      wcount[w][t] = 0;
    }
    base[t] += tot;
    __syncthreads();
  }
}

// Exclusive scan: tiles of 4096 counts.  (1) tile sums with coalesced loads; (2) one workgroup
// scans the tile sums; (3) each tile is staged in LDS (coalesced), every thread scans 16
// consecutive counts, the 256 thread totals are scanned, and the tile is stored coalesced.
constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr int kScanTileN = kScanThreads * kScanItems;
constexpr int kScanSumsThreads = 1024;

__device__ inline int PadIdx(int i) { return i + (i >> 4); }  // one pad word per 16: no bank conflicts

__device__ inline uint32_t BlockExclusiveScan256(uint32_t v, uint32_t* sh, uint32_t* total) {
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int off = 1; off < kScanThreads; off <<= 1) {
    const uint32_t a = t >= off ? sh[t - off] : 0u;
    __syncthreads();
    sh[t] += a;
    __syncthreads();
  }
  const uint32_t incl = sh[t];
  *total = sh[kScanThreads - 1];
  __syncthreads();
  return incl - v;
}

__global__ __launch_bounds__(kScanThreads) void ScanTileSumKernel(const uint32_t* __restrict__ in, int64_t n,
                                                                  uint32_t* __restrict__ sums) {
  __shared__ uint32_t sh[kScanThreads];
  const int64_t base = (int64_t)blockIdx.x * kScanTileN;
  uint32_t s = 0;
  for (int k = 0; k < kScanItems; ++k) {
    const int64_t i = base + k * kScanThreads + threadIdx.x;
    if (i < n) s += in[i];
  }
  uint32_t tot;
  BlockExclusiveScan256(s, sh, &tot);
  if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

__global__ __launch_bounds__(kScanSumsThreads) void ScanSumsKernel(const uint32_t* __restrict__ in,
                                                                   uint32_t* __restrict__ out, int64_t n) {
  __shared__ uint32_t sums[kScanSumsThreads];
  const int t = threadIdx.x;
  const int64_t seg = (n + kScanSumsThreads - 1) / kScanSumsThreads;
  const int64_t b = std::min<int64_t>(n, (int64_t)t * seg);
  const int64_t e = std::min<int64_t>(n, b + seg);
  uint32_t s = 0;
  for (int64_t k = b; k < e; ++k) s += in[k];
  sums[t] = s;
  __syncthreads();
  for (int off = 1; off < kScanSumsThreads; off <<= 1) {
    const uint32_t v = t >= off ? sums[t - off] : 0u;
    __syncthreads();
    sums[t] += v;
    __syncthreads();
  }
  uint32_t run = t > 0 ? sums[t - 1] : 0u;
  for (int64_t k = b; k < e; ++k) {
    const uint32_t v = in[k];
    out[k] = run;
    run += v;
  }
  if (t == kScanSumsThreads - 1) out[n] = sums[kScanSumsThreads - 1];
}

__global__ __launch_bounds__(kScanThreads) void ScanTileApplyKernel(const uint32_t* __restrict__ in, int64_t n,
                                                                    const uint32_t* __restrict__ tile_off,
                                                                    int num_tiles, uint32_t* __restrict__ out) {
  __shared__ uint32_t tile[kScanTileN + kScanTileN / 16];
  __shared__ uint32_t sh[kScanThreads];
  const int t = threadIdx.x;
  const int64_t base = (int64_t)blockIdx.x * kScanTileN;
  for (int k = 0; k < kScanItems; ++k) {
    const int li = k * kScanThreads + t;
    const int64_t i = base + li;
    tile[PadIdx(li)] = i < n ? in[i] : 0u;
  }
  __syncthreads();
  uint32_t s = 0;
  for (int k = 0; k < kScanItems; ++k) s += tile[PadIdx(t * kScanItems + k)];
  uint32_t tot;
  uint32_t run = BlockExclusiveScan256(s, sh, &tot) + tile_off[blockIdx.x];
  for (int k = 0; k < kScanItems; ++k) {
    const int li = PadIdx(t * kScanItems + k);
    const uint32_t v = tile[li];
    tile[li] = run;
    run += v;
  }
  __syncthreads();
  for (int k = 0; k < kScanItems; ++k) {
    const int li = k * kScanThreads + t;
    const int64_t i = base + li;
    if (i < n) out[i] = tile[PadIdx(li)];
  }
  if (blockIdx.x == 0 && t == 0) out[n] = tile_off[num_tiles];
}

int ExclusiveScan(am_ctx* ctx, const uint32_t* in, uint32_t* out, int64_t n) {
  if (n <= 0) {
    hipLaunchKernelGGL(ScanSumsKernel, dim3(1), dim3(kScanSumsThreads), 0, ctx->stream, in, out, (int64_t)0);
    AM_HIP(hipGetLastError());
    return AM_OK;
  }
  const int64_t tiles = (n + kScanTileN - 1) / kScanTileN;
  if (tiles > 0x7fffffff) return Fail(AM_ERR_CAPACITY, "scan: n too large");
  void* buf;
  int rc = Scratch(ctx, "scan_tiles", (size_t)(2 * tiles + 1) * 4, &buf);
  if (rc) return rc;
  uint32_t* sums = (uint32_t*)buf;
  uint32_t* offs = sums + tiles;
  hipLaunchKernelGGL(ScanTileSumKernel, dim3((unsigned)tiles), dim3(kScanThreads), 0, ctx->stream, in, n, sums);
  AM_HIP(hipGetLastError());
  hipLaunchKernelGGL(ScanSumsKernel, dim3(1), dim3(kScanSumsThreads), 0, ctx->stream, (const uint32_t*)sums, offs,
                     tiles);
  AM_HIP(hipGetLastError());
  hipLaunchKernelGGL(ScanTileApplyKernel, dim3((unsigned)tiles), dim3(kScanThreads), 0, ctx->stream, in, n,
                     (const uint32_t*)offs, (int)tiles, out);
  AM_HIP(hipGetLastError());
  return AM_OK;
}

// Sizes every scratch buffer a scan of n counts and a sort of n keys will ask for, so that a
// later call allocates nothing (required before stream capture).
int ReserveSortScratch(am_ctx* ctx, int64_t n) {
  const int64_t num_tiles = (n + kSortTile - 1) / kSortTile;
  const int64_t hist = 256 * num_tiles;
  const int64_t scan_n = std::max<int64_t>(n, hist);
  const int64_t scan_tiles = (scan_n + kScanTileN - 1) / kScanTileN;
  void* p;
  int rc;
  if ((rc = Scratch(ctx, "sort_keys", (size_t)n * 8, &p))) return rc;
  if ((rc = Scratch(ctx, "sort_vals", (size_t)n * 4, &p))) return rc;
  if ((rc = Scratch(ctx, "sort_hist", (size_t)hist * 4, &p))) return rc;
  if ((rc = Scratch(ctx, "sort_offs", (size_t)(hist + 1) * 4, &p))) return rc;
  if ((rc = Scratch(ctx, "scan_tiles", (size_t)(2 * scan_tiles + 1) * 4, &p))) return rc;
  return AM_OK;
}

int RadixSortPairs(am_ctx* ctx, uint64_t* keys, uint32_t* vals, int64_t n, int key_bits) {
  if (n <= 1) return AM_OK;
  if (n > 0x7fffffffLL) return Fail(AM_ERR_CAPACITY, "radix sort: n too large");
  const int num_tiles = (int)((n + kSortTile - 1) / kSortTile);
  void *kbuf, *vbuf, *hist, *offs;
  int rc;
  if ((rc = Scratch(ctx, "sort_keys", (size_t)n * 8, &kbuf))) return rc;
  if ((rc = Scratch(ctx, "sort_vals", (size_t)n * 4, &vbuf))) return rc;
  if ((rc = Scratch(ctx, "sort_hist", (size_t)256 * num_tiles * 4, &hist))) return rc;
  if ((rc = Scratch(ctx, "sort_offs", ((size_t)256 * num_tiles + 1) * 4, &offs))) return rc;
  uint64_t* ka = keys;
  uint32_t* va = vals;
  uint64_t* kb = (uint64_t*)kbuf;
  uint32_t* vb = (uint32_t*)vbuf;
  const int passes = (std::max(1, std::min(64, key_bits)) + 7) / 8;
  for (int p = 0; p < passes; ++p) {
    const int shift = 8 * p;
    hipLaunchKernelGGL(RadixHistKernel, dim3(num_tiles), dim3(kSortThreads), 0, ctx->stream, ka, n, shift,
                       (uint32_t*)hist, num_tiles);
    AM_HIP(hipGetLastError());
    if ((rc = ExclusiveScan(ctx, (const uint32_t*)hist, (uint32_t*)offs, (int64_t)256 * num_tiles))) return rc;
    hipLaunchKernelGGL(RadixScatterKernel, dim3(num_tiles), dim3(kSortThreads), 0, ctx->stream, ka, va, kb, vb, n,
                       shift, (const uint32_t*)offs, num_tiles);
    AM_HIP(hipGetLastError());
    std::swap(ka, kb);
    std::swap(va, vb);
  }
  if (ka != keys) {
    AM_HIP(hipMemcpyAsync(keys, ka, (size_t)n * 8, hipMemcpyDeviceToDevice, ctx->stream));
    AM_HIP(hipMemcpyAsync(vals, va, (size_t)n * 4, hipMemcpyDeviceToDevice, ctx->stream));
  }
  return AM_OK;
}

}  // namespace am

using namespace am;

extern "C" {

// Diagnostics: the device scan and radix sort on host buffers (for tests).
int am_debug_exclusive_scan(am_ctx* ctx, const uint32_t* in, int64_t n, uint32_t* out) {
  if (!ctx || n < 0 || (n > 0 && (!in || !out))) return Fail(AM_ERR_INVALID, "am_debug_exclusive_scan: bad args");
  void *din, *dout;
  int rc;
  if ((rc = Scratch(ctx, "dbg_in", (size_t)(n + 1) * 4, &din))) return rc;
  if ((rc = Scratch(ctx, "dbg_out", (size_t)(n + 1) * 4, &dout))) return rc;
  if (n) AM_HIP(hipMemcpyAsync(din, in, (size_t)n * 4, hipMemcpyHostToDevice, ctx->stream));
  if ((rc = ExclusiveScan(ctx, (const uint32_t*)din, (uint32_t*)dout, n))) return rc;
  AM_HIP(hipMemcpyAsync(out, dout, (size_t)(n + 1) * 4, hipMemcpyDeviceToHost, ctx->stream));
  AM_HIP(hipStreamSynchronize(ctx->stream));
  return AM_OK;
}

int am_debug_radix_sort(am_ctx* ctx, uint64_t* keys, uint32_t* vals, int64_t n, int key_bits) {
  if (!ctx || n < 0 || (n > 0 && (!keys || !vals))) return Fail(AM_ERR_INVALID, "am_debug_radix_sort: bad args");
  if (n == 0) return AM_OK;
  void *dk, *dv;
  int rc;
  if ((rc = Scratch(ctx, "dbg_keys", (size_t)n * 8, &dk))) return rc;
  if ((rc = Scratch(ctx, "dbg_vals", (size_t)n * 4, &dv))) return rc;
  AM_HIP(hipMemcpyAsync(dk, keys, (size_t)n * 8, hipMemcpyHostToDevice, ctx->stream));
  AM_HIP(hipMemcpyAsync(dv, vals, (size_t)n * 4, hipMemcpyHostToDevice, ctx->stream));
  if ((rc = RadixSortPairs(ctx, (uint64_t*)dk, (uint32_t*)dv, n, key_bits))) return rc;
  AM_HIP(hipMemcpyAsync(keys, dk, (size_t)n * 8, hipMemcpyDeviceToHost, ctx->stream));
  AM_HIP(hipMemcpyAsync(vals, dv, (size_t)n * 4, hipMemcpyDeviceToHost, ctx->stream));
  AM_HIP(hipStreamSynchronize(ctx->stream));
  return AM_OK;
}

}  // extern "C"
