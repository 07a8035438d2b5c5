// Stable LSD radix sort of (uint64 key, uint32 value) pairs and an exclusive prefix scan.
//
// Sort: 8-bit digits, one pass per digit of the key's significant bits.  Each pass:
//   1. RadixHistKernel: per-tile (2048 keys) digit histogram in LDS -> hist[digit][tile].
//   2. ExclusiveScan over hist (digit-major), giving every (digit, tile) its global base.
//   3. RadixScatterKernel: the tile is processed in 8 rounds of 256 keys in input order; inside a
//      wave the lanes with equal digits are found with 8 ballots (one per digit bit), so a key's
//      rank = base[digit] + keys of that digit in earlier waves of the round + earlier lanes of
//      its own wave.  Input order is preserved within a digit: the sort is stable.
// Scan: one workgroup of 1024 threads; each thread owns a contiguous segment.
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

constexpr int kScanThreads = 1024;

__global__ __launch_bounds__(kScanThreads) void ExclusiveScanKernel(const uint32_t* __restrict__ in,
                                                                     uint32_t* __restrict__ out, int64_t n) {
  __shared__ uint32_t sums[kScanThreads];
  const int t = threadIdx.x;
  const int64_t seg = (n + kScanThreads - 1) / kScanThreads;
  const int64_t b = std::min<int64_t>(n, (int64_t)t * seg);
  const int64_t e = std::min<int64_t>(n, b + seg);
  uint32_t s = 0;
  for (int64_t k = b; k < e; ++k) s += in[k];
  sums[t] = s;
  __syncthreads();
  // Hillis-Steele inclusive scan of the segment sums.
  for (int off = 1; off < kScanThreads; off <<= 1) {
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
  if (t == kScanThreads - 1) out[n] = sums[kScanThreads - 1];
}

int ExclusiveScan(am_ctx* ctx, const uint32_t* in, uint32_t* out, int64_t n) {
  hipLaunchKernelGGL(ExclusiveScanKernel, dim3(1), dim3(kScanThreads), 0, ctx->stream, in, out, n);
  AM_HIP(hipGetLastError());
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
