// Stable LSD radix sort of (uint64 key, uint32 value) pairs and an exclusive prefix scan.
// (Below 2^30 keys the single-pass-per-digit sort near the end of this file is used; the tiled
// passes described here remain for larger inputs.)
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

// ---- Single-pass-per-digit LSD radix sort ----------------------------------------------------------
// One read of the keys gives every pass's global digit histogram (OsHistKernel) and one small kernel
// their exclusive scans (OsDigitBaseKernel).  Each pass is then one kernel over tiles of 4096 keys
// (OsScatterKernel):
//   * a workgroup takes the next tile id from a counter, so every tile it may wait for belongs to a
//     workgroup that is already running;
//   * wave w ranks the tile's w-th quarter (1024 consecutive keys, 16 rounds of 64) by digit with 8
//     ballots per round and wave-private digit counters, so ranking needs no workgroup barrier;
//     quarters in wave order keep the sort stable;
//   * the tile publishes its digit counts, then walks back over the earlier tiles' published counts
//     (and stops at the first published prefix) to find how many keys of each digit come before it,
//     and publishes its own prefix;
//   * the keys are staged in LDS in digit order and written out in runs of consecutive positions.
// The look-back waits at most kOsSpinLimit polls per tile and digit; a wait that long sets the sticky
// error word (SortErrorWord) instead of hanging, and no write ever goes past n.
constexpr int kOsThreads = 256;
constexpr int kOsWaves = kOsThreads / 64;
constexpr int kOsItems = 16;
constexpr int kOsTile = kOsThreads * kOsItems;  // 4096 keys
constexpr uint32_t kOsAgg = 1u << 30;           // status word: the tile's own count
constexpr uint32_t kOsPre = 2u << 30;           // status word: the inclusive prefix through the tile
constexpr uint32_t kOsVal = kOsAgg - 1;
constexpr int kOsSpinLimit = 1 << 22;
constexpr int kOsMaxPasses = 8;

// Exclusive scan of one value per thread of a 256-thread workgroup (wave shuffles, one barrier);
// wsum is [kOsWaves] LDS words.
__device__ inline uint32_t OsBlockExclusiveScan(uint32_t v, uint32_t* wsum) {
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  uint32_t x = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t y = (uint32_t)__shfl_up((int)x, off);
    if (lane >= off) x += y;
  }
  if (lane == 63) wsum[w] = x;
  __syncthreads();
  uint32_t pre = 0;
  for (int i = 0; i < w; ++i) pre += wsum[i];
  return pre + x - v;
}

__global__ void OsClearKernel(uint32_t* __restrict__ p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0u;
}

// One tile of kOsTile keys per workgroup; every pass's digit counts added into ghist[pass][256].
// Keys often arrive grouped (lists in traversal order, narrow key ranges), and same-address LDS
// atomics serialise, so each run of equal digits over consecutive lanes of a load is added once,
// by its first lane.
// Digit mask of pass p of a sort on key_bits bits: 8 bits, fewer in the last pass.
__device__ __host__ inline uint32_t OsDigitMask(int p, int key_bits) {
  const int b = key_bits - 8 * p;
  return b >= 8 ? 255u : (1u << b) - 1u;
}

__global__ __launch_bounds__(kOsThreads) void OsHistKernel(const uint64_t* __restrict__ keys, int64_t n, int shift0,
                                                           int key_bits, uint32_t* __restrict__ ghist) {
  const int passes = (key_bits + 7) / 8;
  __shared__ uint32_t h[kOsMaxPasses][256];
  const int t = threadIdx.x, lane = t & 63;
  for (int p = 0; p < passes; ++p) h[p][t] = 0u;
  __syncthreads();
  const int64_t tile0 = (int64_t)blockIdx.x * kOsTile;
  uint64_t key[kOsItems];
#pragma unroll
  for (int r = 0; r < kOsItems; ++r) {
    const int64_t idx = tile0 + r * kOsThreads + t;
    key[r] = idx < n ? keys[idx] : 0ull;
  }
#pragma unroll
  for (int r = 0; r < kOsItems; ++r) {
    const bool valid = tile0 + r * kOsThreads + t < n;
    const int nv = __popcll(__ballot(valid));  // the valid lanes are lanes 0 .. nv - 1
    const uint32_t plo = (uint32_t)__shfl_up((int)(uint32_t)key[r], 1);
    const uint32_t phi = (uint32_t)__shfl_up((int)(uint32_t)(key[r] >> 32), 1);
    const uint64_t prev = ((uint64_t)phi << 32) | plo;
    for (int p = 0; p < passes; ++p) {
      const int sh = shift0 + 8 * p;
      const uint32_t dm = OsDigitMask(p, key_bits);
      const uint32_t d = (uint32_t)(key[r] >> sh) & dm;
      const bool head = valid && (lane == 0 || d != ((uint32_t)(prev >> sh) & dm));
      const uint64_t heads = __ballot(head);
      if (head) {
        const uint64_t above = lane == 63 ? 0ull : heads >> (lane + 1);
        const uint32_t len = above ? (uint32_t)__builtin_ctzll(above) + 1u : (uint32_t)(nv - lane);
        atomicAdd(&h[p][d], len);
      }
    }
  }
  __syncthreads();
  for (int p = 0; p < passes; ++p)
    if (h[p][t]) atomicAdd(&ghist[p * 256 + t], h[p][t]);
}

// gbase[pass][d] = keys of a smaller digit in that pass (one workgroup per pass).
__global__ __launch_bounds__(kOsThreads) void OsDigitBaseKernel(const uint32_t* __restrict__ ghist,
                                                                uint32_t* __restrict__ gbase) {
  __shared__ uint32_t wsum[kOsWaves];
  const int p = blockIdx.x, t = threadIdx.x;
  gbase[p * 256 + t] = OsBlockExclusiveScan(ghist[p * 256 + t], wsum);
}

template <bool VALS>  // values travel with the keys
__global__ __launch_bounds__(kOsThreads) void OsScatterKernel(
    const uint64_t* __restrict__ kin, const uint32_t* __restrict__ vin, uint64_t* __restrict__ kout,
    uint32_t* __restrict__ vout, int64_t n, int shift, uint32_t dmask, const uint32_t* __restrict__ gbase,
    uint32_t* __restrict__ counter, uint32_t* __restrict__ status, uint32_t* __restrict__ err) {
  __shared__ uint64_t sk[kOsTile];           // keys, then values, in digit order
  __shared__ uint32_t wcnt[kOsWaves][256];  // per-wave digit counts, then per-wave digit offsets
  __shared__ uint32_t loc[256], glob[256], wsum[kOsWaves];
  __shared__ uint32_t s_tile;
  const int t = threadIdx.x, w = t >> 6, lane = t & 63;
  if (t == 0) s_tile = atomicAdd(counter, 1u);
  for (int i = 0; i < kOsWaves; ++i) wcnt[i][t] = 0u;
  __syncthreads();
  const uint32_t tile = s_tile;
  const int64_t tile0 = (int64_t)tile * kOsTile;
  const int64_t q0 = tile0 + (int64_t)w * (kOsTile / kOsWaves);  // this wave's quarter
  uint64_t key[kOsItems];
  uint32_t val[kOsItems], rk[kOsItems];
#pragma unroll
  for (int r = 0; r < kOsItems; ++r) {
    const int64_t idx = q0 + r * 64 + lane;
    key[r] = idx < n ? kin[idx] : 0ull;
    val[r] = (VALS && idx < n) ? vin[idx] : 0u;
  }
  const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  uint32_t* mycnt = wcnt[w];
#pragma unroll
  for (int r = 0; r < kOsItems; ++r) {
    const bool valid = q0 + r * 64 + lane < n;
    const uint32_t d = (uint32_t)(key[r] >> shift) & dmask;
    uint64_t mask = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const bool bit = (d >> b) & 1u;
      const uint64_t bits = __ballot(bit);
      mask &= bit ? bits : ~bits;
    }
    const uint32_t before = mycnt[d];
    rk[r] = before + (uint32_t)__popcll(mask & lt);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (valid && (mask & lt) == 0) mycnt[d] = before + (uint32_t)__popcll(mask);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
  __syncthreads();
  // digit t: the tile's count, and each wave's offset among the tile's keys of digit t
  uint32_t c = 0;
  for (int i = 0; i < kOsWaves; ++i) {
    const uint32_t ci = wcnt[i][t];
    wcnt[i][t] = c;
    c += ci;
  }
  // keys of digit t in the earlier tiles: thread t looks back over their status words
  uint32_t excl = 0;
  uint32_t* mine = status + (size_t)tile * 256 + t;
  if (tile == 0) {
    __hip_atomic_store(mine, kOsPre | c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  } else {
    __hip_atomic_store(mine, kOsAgg | c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    int64_t tt = (int64_t)tile - 1;
    int spins = 0;
    while (tt >= 0) {
      const uint32_t sw =
          __hip_atomic_load(status + (size_t)tt * 256 + t, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (sw & kOsPre) {
        excl += sw & kOsVal;
        break;
      }
      if (sw & kOsAgg) {
        excl += sw & kOsVal;
        --tt;
        continue;
      }
      if (++spins > kOsSpinLimit) {
        atomicOr(err, 1u);
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    __hip_atomic_store(mine, kOsPre | (excl + c), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  const uint32_t lbase = OsBlockExclusiveScan(c, wsum);
  loc[t] = lbase;
  glob[t] = gbase[t] + excl;
  __syncthreads();
  // stage the keys in digit order and write runs of consecutive positions, then the values the same
  // way (through the same LDS), so that the tile needs 32 KB of LDS
  uint32_t sp[kOsItems];
#pragma unroll
  for (int r = 0; r < kOsItems; ++r) {
    const uint32_t d = (uint32_t)(key[r] >> shift) & dmask;
    sp[r] = loc[d] + wcnt[w][d] + rk[r];
    if (q0 + r * 64 + lane < n) sk[sp[r]] = key[r];
  }
  __syncthreads();
  const int tile_n = (int)std::min<int64_t>(kOsTile, n - tile0);
  uint32_t dst[kOsItems];
#pragma unroll
  for (int r = 0; r < kOsItems; ++r) {
    const int i = r * kOsThreads + t;
    dst[r] = 0xffffffffu;
    if (i < tile_n) {
      const uint64_t k = sk[i];
      const uint32_t d = (uint32_t)(k >> shift) & dmask;
      dst[r] = glob[d] + (uint32_t)i - loc[d];
      if (dst[r] < (uint64_t)n) kout[dst[r]] = k;
    }
  }
  if (VALS) {
    __syncthreads();
    uint32_t* sv = reinterpret_cast<uint32_t*>(sk);
#pragma unroll
    for (int r = 0; r < kOsItems; ++r)
      if (q0 + r * 64 + lane < n) sv[sp[r]] = val[r];
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kOsItems; ++r) {
      const int i = r * kOsThreads + t;
      if (i < tile_n && dst[r] < (uint64_t)n) vout[dst[r]] = sv[i];
    }
  }
}

// Words of the single-pass sort's scratch (cleared by every sort): histograms [passes][256], digit
// bases [passes][256], tile counters [passes], then the status words [passes][tiles][256].  The
// look-back error word is kept apart (SortErrorWord); no sort clears it, only its report does.
static int64_t OsMetaWords(int passes) { return (int64_t)passes * 512 + passes; }

__global__ void OsZeroWordKernel(uint32_t* w) { *w = 0u; }

uint32_t* SortErrorWord(am_ctx* ctx) {
  Buffer& b = ctx->scratch["sort_err"];
  if (!b.ptr) {
    void* p;
    if (Scratch(ctx, "sort_err", 4, &p)) return nullptr;
    hipLaunchKernelGGL(OsZeroWordKernel, dim3(1), dim3(1), 0, ctx->stream, (uint32_t*)p);
  }
  return (uint32_t*)b.ptr;
}

int CheckSortErrorAfterSync(am_ctx* ctx, uint32_t word) {
  if (!word) return AM_OK;
  // Reported once: the word is cleared (in stream order) so later calls are not failed for it.
  if (const uint32_t* w = SortErrorWord(ctx))
    hipLaunchKernelGGL(OsZeroWordKernel, dim3(1), dim3(1), 0, ctx->stream, const_cast<uint32_t*>(w));
  return Fail(AM_ERR_HIP, "radix sort: a look-back wait on the device gave up");
}

int CheckSortError(am_ctx* ctx) {
  uint32_t word = 0;
  if (const uint32_t* w = SortErrorWord(ctx))
    AM_HIP(hipMemcpyAsync(&word, w, 4, hipMemcpyDeviceToHost, ctx->stream));
  AM_HIP(hipStreamSynchronize(ctx->stream));
  return CheckSortErrorAfterSync(ctx, word);
}
static int64_t OsWords(int64_t n, int passes) {
  return OsMetaWords(passes) + (int64_t)passes * ((n + kOsTile - 1) / kOsTile) * 256;
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
  if ((rc = Scratch(ctx, "sort_onesweep", (size_t)OsWords(n, kOsMaxPasses) * 4, &p))) return rc;
  if (!SortErrorWord(ctx)) return Fail(AM_ERR_HIP, "radix sort: error word");
  return AM_OK;
}

// The single-pass-per-digit sort of keys (and vals, when given) on bits [bit0, bit0 + 8 passes);
// n < kOsVal.  The result ends in keys / vals.
static int OneSweepSort(am_ctx* ctx, uint64_t* keys, uint32_t* vals, int64_t n, int bit0, int key_bits) {
  const int passes = (key_bits + 7) / 8;
  void *kbuf, *vbuf = nullptr, *os;
  int rc;
  if ((rc = Scratch(ctx, "sort_keys", (size_t)n * 8, &kbuf))) return rc;
  if (vals && (rc = Scratch(ctx, "sort_vals", (size_t)n * 4, &vbuf))) return rc;
  const int64_t tiles = (n + kOsTile - 1) / kOsTile;
  const int64_t words = OsWords(n, passes);
  if ((rc = Scratch(ctx, "sort_onesweep", (size_t)words * 4, &os))) return rc;
  uint32_t* err = SortErrorWord(ctx);
  if (!err) return Fail(AM_ERR_HIP, "radix sort: error word");
  uint32_t* ghist = (uint32_t*)os;
  uint32_t* gbase = ghist + passes * 256;
  uint32_t* counters = gbase + passes * 256;
  uint32_t* status = ghist + OsMetaWords(passes);
  hipLaunchKernelGGL(OsClearKernel, dim3((unsigned)std::min<int64_t>(2048, (words + 255) / 256)), dim3(256), 0,
                     ctx->stream, (uint32_t*)os, words);
  hipLaunchKernelGGL(OsHistKernel, dim3((unsigned)tiles), dim3(kOsThreads), 0, ctx->stream, keys, n, bit0, key_bits,
                     ghist);
  hipLaunchKernelGGL(OsDigitBaseKernel, dim3(passes), dim3(kOsThreads), 0, ctx->stream, (const uint32_t*)ghist, gbase);
  AM_HIP(hipGetLastError());
  uint64_t *ka = keys, *kb = (uint64_t*)kbuf;
  uint32_t *va = vals, *vb = (uint32_t*)vbuf;
  for (int p = 0; p < passes; ++p) {
    const uint32_t dm = OsDigitMask(p, key_bits);
    if (vals)
      hipLaunchKernelGGL(OsScatterKernel<true>, dim3((unsigned)tiles), dim3(kOsThreads), 0, ctx->stream, ka, va, kb,
                         vb, n, bit0 + 8 * p, dm, gbase + 256 * p, counters + p, status + (size_t)p * tiles * 256, err);
    else
      hipLaunchKernelGGL(OsScatterKernel<false>, dim3((unsigned)tiles), dim3(kOsThreads), 0, ctx->stream, ka,
                         (const uint32_t*)nullptr, kb, (uint32_t*)nullptr, n, bit0 + 8 * p, dm, gbase + 256 * p,
                         counters + p, status + (size_t)p * tiles * 256, err);
    AM_HIP(hipGetLastError());
    std::swap(ka, kb);
    std::swap(va, vb);
  }
  if (ka != keys) {
    AM_HIP(hipMemcpyAsync(keys, ka, (size_t)n * 8, hipMemcpyDeviceToDevice, ctx->stream));
    if (vals) AM_HIP(hipMemcpyAsync(vals, va, (size_t)n * 4, hipMemcpyDeviceToDevice, ctx->stream));
  }
  return AM_OK;
}

int RadixSortPairs(am_ctx* ctx, uint64_t* keys, uint32_t* vals, int64_t n, int key_bits) {
  if (n <= 1) return AM_OK;
  if (n > 0x7fffffffLL) return Fail(AM_ERR_CAPACITY, "radix sort: n too large");
  const int passes = (std::max(1, std::min(64, key_bits)) + 7) / 8;
  if (n < (int64_t)kOsVal) return OneSweepSort(ctx, keys, vals, n, 0, std::max(1, std::min(64, key_bits)));
  // per-tile histograms and a scan per pass (the status words of the single-pass sort count up to
  // kOsVal)
  void *kbuf, *vbuf, *hist, *offs;
  int rc;
  if ((rc = Scratch(ctx, "sort_keys", (size_t)n * 8, &kbuf))) return rc;
  if ((rc = Scratch(ctx, "sort_vals", (size_t)n * 4, &vbuf))) return rc;
  const int num_tiles = (int)((n + kSortTile - 1) / kSortTile);
  if ((rc = Scratch(ctx, "sort_hist", (size_t)256 * num_tiles * 4, &hist))) return rc;
  if ((rc = Scratch(ctx, "sort_offs", ((size_t)256 * num_tiles + 1) * 4, &offs))) return rc;
  uint64_t *ka = keys, *kb = (uint64_t*)kbuf;
  uint32_t *va = vals, *vb = (uint32_t*)vbuf;
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

int RadixSortKeys(am_ctx* ctx, uint64_t* keys, int64_t n, int bit0, int key_bits) {
  if (n <= 1) return AM_OK;
  if (n >= (int64_t)kOsVal) return Fail(AM_ERR_CAPACITY, "radix sort of keys: n too large");
  if (bit0 < 0 || bit0 + key_bits > 64 || key_bits < 1) return Fail(AM_ERR_INVALID, "radix sort of keys: bad bit range");
  return OneSweepSort(ctx, keys, nullptr, n, bit0, key_bits);
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
  uint32_t err = 0;
  if (const uint32_t* w = SortErrorWord(ctx)) AM_HIP(hipMemcpyAsync(&err, w, 4, hipMemcpyDeviceToHost, ctx->stream));
  AM_HIP(hipStreamSynchronize(ctx->stream));
  return CheckSortErrorAfterSync(ctx, err);
}

int am_debug_radix_sort_keys(am_ctx* ctx, uint64_t* keys, int64_t n, int bit0, int key_bits) {
  if (!ctx || n < 0 || (n > 0 && !keys)) return Fail(AM_ERR_INVALID, "am_debug_radix_sort_keys: bad args");
  if (n == 0) return AM_OK;
  void* dk;
  int rc;
  if ((rc = Scratch(ctx, "dbg_keys", (size_t)n * 8, &dk))) return rc;
  AM_HIP(hipMemcpyAsync(dk, keys, (size_t)n * 8, hipMemcpyHostToDevice, ctx->stream));
  if ((rc = RadixSortKeys(ctx, (uint64_t*)dk, n, bit0, key_bits))) return rc;
  AM_HIP(hipMemcpyAsync(keys, dk, (size_t)n * 8, hipMemcpyDeviceToHost, ctx->stream));
  uint32_t err = 0;
  if (const uint32_t* w = SortErrorWord(ctx)) AM_HIP(hipMemcpyAsync(&err, w, 4, hipMemcpyDeviceToHost, ctx->stream));
  AM_HIP(hipStreamSynchronize(ctx->stream));
  return CheckSortErrorAfterSync(ctx, err);
}

}  // extern "C"
