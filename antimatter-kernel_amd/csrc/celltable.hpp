// Cell table: open addressing from a cell key to the (start, count) of its run in a sorted key
// array.  It is built once after the radix sort (one insert per run, load factor <= 1/2); a
// stencil query is then one or two independent probes instead of a binary search over all n keys.
// The runs keep the sorted order, so iterating a cell visits its particles exactly as before.
#pragma once

#include "am_internal.hpp"

namespace am {

constexpr uint64_t kCellEmpty = ~0ull;  // also the key of particles in no cell (never inserted)

// One table slot: the cell key and its run, in one 16-byte record, so that an insert and a hit
// touch one line.
struct __align__(16) CellSlot {
  uint64_t key;    // kCellEmpty if free
  uint32_t start;  // run in the sorted order
  uint32_t count;
};

struct CellTable {
  const CellSlot* slots = nullptr;
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
    const CellSlot s = t.slots[h];
    if (s.key == key) return make_uint2(s.start, s.count);
    if (s.key == kCellEmpty) return make_uint2(0u, 0u);
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
// keys skeys[0..n) other than kCellEmpty.  slots holds `size` entries.
int LaunchCellTable(hipStream_t st, const uint64_t* skeys, int64_t n, CellSlot* slots, uint64_t size);

}  // namespace am
