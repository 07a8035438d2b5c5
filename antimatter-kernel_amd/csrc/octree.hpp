// Count-capped octree shared by the Barnes-Hut and FMM solvers.
//
// The reference builds both trees by recursive, stable, in-place octant partitioning
// (barnes-hut.cc:68-164, fast-multipole.cc:257-311): a node splits while it holds more than the
// leaf capacity, children are visited in octant order 0..7 (bit 0 = x >= c_x, bit 1 = y, bit 2 = z)
// and every child centre is its parent centre +- half/2 in FP64.  BuildOctree reproduces that tree
// from 63-bit Morton keys (21 levels of the same FP64 comparisons), a stable radix sort and a
// level-synchronous split, and numbers the nodes in the reference's DFS preorder.
#pragma once

#include <vector>

#include "am_internal.hpp"

namespace am {

constexpr int kOctreeLevels = 21;  // 63-bit keys; the reference allows 32 levels

struct RootBox {
  double cx, cy, cz, half;
};

struct NodeArrays {  // BFS order (level by level, children of a node contiguous in octant order)
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

struct Octree {
  int m = 0;                    // particles in the tree
  int nn = 0;                   // nodes
  uint32_t* order = nullptr;    // sorted position -> particle index (leaves in index order)
  NodeArrays nd{};              // BFS arrays
  std::vector<int> level_start; // BFS id range of depth d: [level_start[d], level_start[d+1])
  uint32_t* pre = nullptr;      // preorder position -> BFS id
  uint32_t* bfs2pre = nullptr;  // BFS id -> preorder position
  uint64_t* pkeys = nullptr;    // sorted preorder keys (first particle << 5 | depth)
  uint32_t deep_overflow = 0;   // depth-21 leaves holding more than 8 particles
};

// particles: m indices (ascending) into x/y/z/q; root: device pointer; tag prefixes scratch names.
// xq (indexed like x) receives (x, y, z, q) of the m particles as one record each, so that later
// tree-order gathers read one line per particle.
int BuildOctree(am_ctx* ctx, const uint32_t* particles, int m, const double* x, const double* y, const double* z,
                const double* q, const RootBox* root, int leaf_cap, const char* tag, Octree* t, double4* xq);

// Compacts the indices i < n of alive particles with finite coordinates into out (ascending);
// returns count.  mode 0: all of them; mode 1: those with q != 0.
int GatherParticles(am_ctx* ctx, int64_t n, const uint8_t* flags, const double* q, const double* x, const double* y,
                    const double* z, int mode, uint32_t* out, int* m_out, const char* tag);

// Bounding box of the gathered particles: box_dev receives lo[3], hi[3] (device).
int GatheredBBox(am_ctx* ctx, const uint32_t* idx, int m, const double* x, const double* y, const double* z,
                 double* box_dev, const char* tag);

}  // namespace am
