// Barnes-Hut tree records shared by the walk (bh.hip) and the learned correction (mlp.hip).
#pragma once

#include "octree.hpp"

namespace am {

// Learned-correction features of a node (preorder), per sign s: the charge centre's offset from the
// cube centre in half-widths (3), the normalised second moments S_s / (Q_s h^2) (6) and Q_s / e.
struct FeatNode {
  float fp[10];
  float fn[10];
  float h;  // half width in units of 2^-20 m
  float pad;
};

// Flags of a walk record.
constexpr uint32_t kNodeTwig = 1;    // internal node whose children are all leaves
constexpr uint32_t kNodeEmpty = 2;   // no charge of either sign: the walk skips it
constexpr uint32_t kNodeHasPos = 4;  // qp > 0
constexpr uint32_t kNodeHasNeg = 8;  // qn > 0
constexpr uint32_t kNodeLeaf = 16;

// A node as the walk reads it, 128 bytes in DFS preorder.  The first 80 bytes (WalkHot) are what
// an FP32 visit needs; the walk fetches them with two wide scalar loads ahead of the visit (the
// successor skip(c) while c is being visited).  The rest (WalkCold) holds the FP64 charges, read by
// the FP64 walks, and the FP64 cube, read by the exact acceptance test near its threshold.
struct WalkNode {
  float chx, chy, chz;      // cube centre in units of 2^-20 m, FP32 hi parts
  // The acceptance test's left side W = (2 half)^2 in units^2 bracketed for the FP32 test:
  // w2hi >= W (1 + 2^-16) and w2lo <= W (1 - 2^-16) (FP32 roundings of W (1 +- 2^-15)).
  float w2hi;
  float clx, cly, clz;      // lo parts: centre = hi + lo to ~2^-48 relative
  uint32_t flags;           // kNodeTwig | kNodeEmpty | kNodeHasPos | kNodeHasNeg | kNodeLeaf
  int b, e, skip;           // particle range in the sorted order, subtree skip pointer
  float w2lo;
  // Charge centre minus cube centre in units (|offset| <= half width, so FP32 keeps it to 2^-24 of
  // the half width), charge in e; an empty class is (0, 0, 0, charge 0).
  float4 p_rel, n_rel;
  double qp, qn;            // positive charge and |negative charge|
  double cx, cy, cz, half;  // geometric cube in FP64
};
struct WalkHot {  // the first 80 bytes of a WalkNode
  float chx, chy, chz, w2hi, clx, cly, clz;
  uint32_t flags;
  int b, e, skip;
  float w2lo;
  float4 p_rel, n_rel;
};
struct WalkCold {  // the last 48 bytes of a WalkNode
  double qp, qn;
  double cx, cy, cz, half;
};
static_assert(sizeof(WalkNode) == 128 && sizeof(WalkHot) == 80 && sizeof(WalkCold) == 48, "walk record layout");

struct WalkNodeAux {  // read by the FP64 walks' accepted nodes and by the learned correction
  float4 p_hi, p_lo, n_hi, n_lo;  // FP32 hi/lo charge centres in units, charge in e
  double pqx, pqy, pqz;           // FP64 charge centres
  double nqx, nqy, nqz;
};

struct BHTree {
  Octree tree;
  int m = 0;            // alive particles
  int nn = 0;           // nodes
  uint32_t* order = nullptr;  // sorted position -> particle index
  WalkNode* nodes = nullptr;
  WalkNodeAux* nodes32 = nullptr;
  const RootBox* root = nullptr;  // device: the root cube (the FP32 acceptance test's range check)
  FeatNode* feat = nullptr;  // only when built with features
  float4* ssrc = nullptr;    // FP32 hi/lo mirror in tree order (2 per sorted position)
  int max_depth = 0;
  uint32_t deep_overflow = 0;
};

// Builds the tree of the store's alive particles (leaf capacity leaf_cap); with_features adds the
// per-node features of the learned correction.
int BuildBHTree(am_store* s, int leaf_cap, BHTree* t, bool with_features = false);

// FP32 Barnes-Hut force walk over the targets tsel[0..n_targets) (sorted positions; NULL: all),
// writing forces of those targets; with rec != NULL the accepted nodes of target k are recorded in
// rec[k * rec_cap + c] (c < rec_cap) and their number in rec_cnt[k].
int LaunchBHWalkF32(am_store* s, const BHTree& t, double theta, const uint32_t* tsel, int n_targets, double* ox,
                    double* oy, double* oz, uint32_t* rec, int rec_cap, uint32_t* rec_cnt);



}  // namespace am
