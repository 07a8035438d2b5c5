// Cell table: open addressing from a cell key to the (start, count) of its run in a sorted key
// array.  It is built once after the radix sort (one insert per run, load factor <= 1/2); a
// stencil query is then one or two independent probes instead of a binary search over all n keys.
// The runs keep the sorted order, so iterating a cell visits its particles exactly as before.
#pragma once

#include "am_internal.hpp"

namespace am {

constexpr uint64_t kCellEmpty = ~0ull;  // also the key of particles in no cell (never inserted)

struct CellTable {
  const uint64_t* keys = nullptr;
  const uint2* range = nullptr;  // (start, count) in the sorted order
  uint64_t mask = 0;
};

// MurmurHash3's 64-bit finaliser (public domain).
__host__ __device__ inline uint64_t CellHashMix(uint64_t k) {
  k ^= k >> 33;
  k *= 0xff51afd7ed558ccdull;
  k ^= k >> 33;
  k *= 0xc4ceb9fe1a85ec53ull;
  k ^= k >> 33;
  return k;
}

// (start, count) of the cell, (0, 0) if empty.  Terminates: at most half the slots are used.
__device__ inline uint2 CellLookup(const CellTable& t, uint64_t key) {
  uint64_t h = CellHashMix(key) & t.mask;
  for (;;) {
    const uint64_t k = t.keys[h];
    if (k == key) return t.range[h];
    if (k == kCellEmpty) return make_uint2(0u, 0u);
    h = (h + 1) & t.mask;
  }
}

// Entries for n sorted keys: a power of two >= 2n.
inline uint64_t CellTableSize(int64_t n) {
  uint64_t s = 1024;
  while (s < 2 * (uint64_t)n) s <<= 1;
  return s;
}

// Clears the table (fill kernel, safe under stream capture) and inserts every run of the sorted
// keys skeys[0..n) other than kCellEmpty.  hkeys / hrange hold `size` entries.
int LaunchCellTable(hipStream_t st, const uint64_t* skeys, int64_t n, uint64_t* hkeys, uint2* hrange, uint64_t size);

}  // namespace am
