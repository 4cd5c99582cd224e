// plan.hpp -- the host side of an FMM on one rank: domain, adaptive octree, box index,
// leaf numbering, interaction lists and the points in leaf order.
//
// This is the one-rank image of what fmm-exec builds before an evaluation
// (FmmBuilder::build, fmm-exec/src/fmm.rs:1853-1935): the domain
// (compute_global_bounding_box, octree/src/octree.rs:1278), the level-16 keys of sources
// then targets, the tree refined by distinct keys and 2:1 balanced
// (octree/src/octree/implementation.rs:682-716, 1014-1056), the plan's box index and
// lists (fmm-plan/src/index.rs, interaction_manager.rs) and the leaf-scaled points
// (fmm-exec/src/geometry.rs:291).
#pragma once

#include <array>
#include <cstdint>
#include <vector>

#include "common.hpp"

namespace fmmh {

// CSR rows: row r spans ent[off[r] .. off[r+1]); aux carries an offset index (V) or is
// empty.
struct Csr {
  std::vector<uint32_t> off;
  std::vector<uint32_t> ent;
  std::vector<uint16_t> aux;
  uint32_t rows() const { return off.empty() ? 0 : static_cast<uint32_t>(off.size() - 1); }
  uint64_t size() const { return off.empty() ? 0 : off.back(); }
};

struct PlanOptions {
  int64_t leaf_points = 64;  // max_fine_keys: refine while a box holds more distinct keys
  int max_level = kDeepest;
  // Near rule with named crossovers (fmm-plan/src/near.rs; NearRuleChoice::Crossovers):
  // a W pair converts to P2P where its source leaf holds fewer than near_w sources, an X
  // pair where its target leaf holds fewer than near_x targets.  Negative: off.
  double near_w = -1.0;
  double near_x = -1.0;
};

struct Plan {
  PlanOptions options;
  // Domain: the bounding box [xmin, ymin, zmin, xmax, ymax, zmax], its lower corner a and
  // side w (fmm-exec/src/geometry.rs:120-146).
  std::array<double, 6> box{};
  std::array<double, 3> lower{};
  double side = 0.0;
  int nlevels = 0;

  // Box index: per level the held keys in Morton order (fmm-plan/src/index.rs:1-40).
  std::vector<std::vector<uint64_t>> keys;
  std::vector<std::vector<uint32_t>> box_leaf;     // leaf index or kNone
  std::vector<std::vector<uint32_t>> first_child;  // index of child 0 on level + 1 or kNone
  std::vector<std::vector<uint32_t>> parent;       // box index of the parent on level - 1

  // Leaves by (level, key): leaves of level l are leaf_off[l] .. leaf_off[l+1].
  std::vector<uint64_t> leaf_key;
  std::vector<uint32_t> leaf_box;
  std::vector<uint32_t> leaf_off;

  // Lists.  near[l], w[l]: one row per leaf of level l; near entries are leaf indices
  // ascending (U list and the leaf itself), w entries box indices on level l + 1.
  // v[l], x[l]: one row per box of level l; v entries are box indices on level l with
  // their offset index in aux, ascending in it; x entries leaf indices ascending.
  std::vector<Csr> near, w, v, x;

  // Points.  src_start/tgt_start: CSR of the points of each leaf in leaf order;
  // src_order/tgt_order: caller index of the point at each leaf-order position;
  // us/ut: leaf-scaled coordinates (3 per point) in leaf order, f64.
  int64_t ns = 0, nt = 0;
  bool same_points = false;  // targets are the sources (one array given for both)
  std::vector<uint32_t> src_start, tgt_start, src_order, tgt_order;
  std::vector<double> us, ut;

  // Radius of each leaf in domain units (side / 2^(level + 1); per cloud in a forest).
  std::vector<double> leaf_radius;

  // During a build only: the leaves in Morton order and the leaf index of each.
  std::vector<uint64_t> morton_leaves;
  std::vector<uint32_t> morton_to_leaf;

  uint32_t nleaves() const { return static_cast<uint32_t>(leaf_key.size()); }
  uint32_t nboxes(int level) const { return static_cast<uint32_t>(keys[level].size()); }
  double radius(int level) const;  // r_l = w / 2^(l+1)
};

// Builds the plan of sources and targets (3 doubles per point).
Plan build_plan(const double *sources, int64_t ns, const double *targets, int64_t nt,
                const PlanOptions &options);
// The same build on the GPU for its per-point stages (plan_device.hip): the bounding box,
// the fine keys, their sort and distinct set, and the points' leaves, order and leaf
// coordinates; the tree and lists from the distinct keys on the host.  The same plan.
Plan build_plan_device(int device, const double *sources, int64_t ns, const double *targets, int64_t nt,
                       const PlanOptions &options);

// The stages of a build: the checked settings, the domain of the points' bounding box
// [lo, hi], the fine (level-16) key of a point, the leaves, boxes and lists from the sorted
// distinct fine keys (leaving morton_leaves / morton_to_leaf for the points), a point's
// leaf-scaled coordinates in a leaf, and the near rule and leaf radii once the points are
// placed.
Plan plan_begin(int64_t ns, int64_t nt, bool same_points, const PlanOptions &options);
void plan_domain(Plan &plan, const double lo[3], const double hi[3]);
uint64_t plan_fine_key(const Plan &plan, const double *x);
void plan_tree(Plan &plan, std::vector<uint64_t> fine);
void plan_leaf_coordinates(const Plan &plan, const double *x, uint32_t leaf, double *u);
void plan_finish(Plan &plan);

// A forest: independent clouds (each its own domain and tree, sources = targets) as one
// plan whose levels hold every cloud's boxes and leaves, cloud after cloud; lists never
// cross clouds and caller indices run over the clouds in order (counts[c] points each).
Plan build_forest(const double *points, const int64_t *counts, int nclouds, const PlanOptions &options);

// Tables (CONVENTIONS §3.12): column-major (p+1)^2 x (p+1)^2 matrices, entry (i, k) at
// i + k n; M2M and L2L by child octant (8 each), M2L by offset index (316).
// Rotation M2L tables: per offset index (316) the forward multipole rotation F(Q) and the
// backward local rotation G(Q^T), per degree n a column-major (2n+1)^2 block (degrees in
// order, rotation_block_size(p) values per offset); per distinct |d|^2 (dist2, indexed by
// dist_of[offset]) the coaxial M2L, per order i a column-major (p+1-i)^2 block [j-i][n-i]
// (coaxial_block_size(p) values per distance).  T(d) = G A F equals the dense M2L table.
int rotation_block_size(int p);
// Every degree's block (column-major, degrees in order) of the coefficient rotation for the
// rotation matrix R (row-major): multipoles (local = false) or locals (local = true).
void rotation_blocks_of(int p, const double R[9], bool local, double *out);
int coaxial_block_size(int p);
void build_rotation_tables(int p, std::vector<double> &fwd, std::vector<double> &bwd, std::vector<double> &coax,
                           std::vector<uint16_t> &dist_of, std::vector<double> &dist2);

void build_tables(int p, std::vector<double> &m2m, std::vector<double> &l2l,
                  std::vector<double> &m2l);

}  // namespace fmmh
