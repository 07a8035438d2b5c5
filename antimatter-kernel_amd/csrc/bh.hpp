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

struct WalkNode {
  // The first 64 bytes are what the walk reads at every visit (one scalar fetch).
  float chx, chy, chz;  // cube centre in units of 2^-20 m, FP32 hi parts
  float w2;             // (2 half)^2 in units^2, rounded to FP32: the acceptance test's left side
  float clx, cly, clz;  // lo parts: centre = hi + lo to ~2^-48 relative
  int twig;             // internal node whose children are all leaves
  double qp, qn;        // positive charge and |negative charge|
  int b, e, skip, leaf; // particle range in the sorted order, subtree skip pointer
  double pqx, pqy, pqz; // positive charge centre
  double nqx, nqy, nqz; // negative charge centre
  double cx, cy, cz, half;  // geometric cube in FP64 (the exact acceptance test)
};
struct WalkNodeF32 {  // FP32 monopoles, charge in e
  // Charge centre minus cube centre in units of 2^-20 m (|offset| <= half width, so FP32 keeps it to
  // 2^-24 of the half width); an empty class is (0, 0, 0, charge 0).  The walk reads these 32 bytes.
  float4 p_rel, n_rel;
  float4 p_hi, p_lo, n_hi, n_lo;  // hi/lo charge centres (learned-correction residuals)
};

struct BHTree {
  Octree tree;
  int m = 0;            // alive particles
  int nn = 0;           // nodes
  uint32_t* order = nullptr;  // sorted position -> particle index
  WalkNode* nodes = nullptr;
  WalkNodeF32* nodes32 = nullptr;
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
