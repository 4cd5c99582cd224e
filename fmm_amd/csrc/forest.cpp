// forest.cpp -- many independent clouds as one plan (batched evaluation).
//
// Each cloud gets its own tree from build_plan (its own domain, refinement and lists,
// exactly as if it were evaluated alone).  The forest stacks them level by level: level
// l holds the boxes of cloud 0, then cloud 1, ...; the leaves of level l likewise; every
// index in the lists, parents and children is shifted to the stacked numbering.  Lists
// never cross clouds, frames are formed from each cloud's own integer cells, and each
// leaf keeps its cloud's radius for the output pass, so one set of launches evaluates
// every cloud with each cloud's own arithmetic.
#include <algorithm>
#include <cstdint>
#include <vector>

#include "plan.hpp"

namespace fmmh {

Plan build_forest(const double *points, const int64_t *counts, int nclouds, const PlanOptions &options) {
  require(nclouds >= 1, "build_forest: no clouds");
  std::vector<Plan> clouds(nclouds);
  std::vector<int64_t> first(nclouds + 1, 0);
  for (int c = 0; c < nclouds; ++c) {
    require(counts[c] >= 1, "build_forest: an empty cloud");
    first[c + 1] = first[c] + counts[c];
  }
  require(first[nclouds] < (int64_t(1) << 31), "build_forest: too many points", kTooLarge);
#pragma omp parallel for schedule(dynamic, 1)
  for (int c = 0; c < nclouds; ++c) {
    const double *p = points + 3 * first[c];
    clouds[c] = build_plan(p, counts[c], p, counts[c], options);
  }

  Plan f;
  f.options = options;
  f.box = clouds[0].box;
  f.lower = clouds[0].lower;
  f.side = clouds[0].side;
  f.ns = f.nt = first[nclouds];
  f.same_points = true;
  int L = 0;
  for (const Plan &p : clouds) L = std::max(L, p.nlevels);
  f.nlevels = L;

  // box bases per cloud and level, leaf bases per cloud and level
  std::vector<std::vector<uint32_t>> bb(nclouds, std::vector<uint32_t>(L + 1, 0));
  std::vector<std::vector<uint32_t>> lb(nclouds, std::vector<uint32_t>(L, 0));
  f.leaf_off.assign(L + 1, 0);
  for (int l = 0; l < L; ++l) {
    uint32_t nb = 0, nl = 0;
    for (int c = 0; c < nclouds; ++c) {
      const Plan &p = clouds[c];
      bb[c][l] = nb;
      nb += l < p.nlevels ? p.nboxes(l) : 0;
      nl += l < p.nlevels ? p.leaf_off[l + 1] - p.leaf_off[l] : 0;
    }
    f.leaf_off[l + 1] = f.leaf_off[l] + nl;
    uint32_t at = f.leaf_off[l];
    for (int c = 0; c < nclouds; ++c) {
      const Plan &p = clouds[c];
      lb[c][l] = at;
      at += l < p.nlevels ? p.leaf_off[l + 1] - p.leaf_off[l] : 0;
    }
  }
  // global leaf index of cloud c's leaf i
  auto leaf_of = [&](int c, uint32_t i) -> uint32_t {
    const Plan &p = clouds[c];
    const int l = int(std::upper_bound(p.leaf_off.begin(), p.leaf_off.end(), i) - p.leaf_off.begin()) - 1;
    return lb[c][l] + (i - p.leaf_off[l]);
  };
  auto shift = [](uint32_t v, uint32_t by) { return v == kNone ? kNone : v + by; };

  f.keys.resize(L);
  f.box_leaf.resize(L);
  f.first_child.resize(L);
  f.parent.resize(L);
  f.near.resize(L);
  f.w.resize(L);
  f.v.resize(L);
  f.x.resize(L);
  const uint32_t nleaves = f.leaf_off[L];
  f.leaf_key.resize(nleaves);
  f.leaf_box.resize(nleaves);
  f.leaf_radius.resize(nleaves);
  for (int l = 0; l < L; ++l) {
    Csr &near = f.near[l], &w = f.w[l], &v = f.v[l], &x = f.x[l];
    near.off = {0};
    w.off = {0};
    v.off = {0};
    x.off = {0};
    for (int c = 0; c < nclouds; ++c) {
      const Plan &p = clouds[c];
      if (l >= p.nlevels) continue;
      for (uint32_t b = 0; b < p.nboxes(l); ++b) {
        f.keys[l].push_back(p.keys[l][b]);
        const uint32_t bl = p.box_leaf[l][b];
        f.box_leaf[l].push_back(bl == kNone ? kNone : leaf_of(c, bl));
        f.first_child[l].push_back(shift(p.first_child[l][b], l + 1 < L ? bb[c][l + 1] : 0));
        f.parent[l].push_back(shift(p.parent[l][b], l >= 1 ? bb[c][l - 1] : 0));
      }
      for (uint32_t i = p.leaf_off[l]; i < p.leaf_off[l + 1]; ++i) {
        const uint32_t g = leaf_of(c, i);
        f.leaf_key[g] = p.leaf_key[i];
        f.leaf_box[g] = p.leaf_box[i] + bb[c][l];
        f.leaf_radius[g] = p.leaf_radius[i];
      }
      // near and W: one row per leaf of the level
      const Csr &pn = p.near[l], &pw = p.w[l];
      for (uint32_t r = 0; r < pn.rows(); ++r) {
        for (uint32_t e = pn.off[r]; e < pn.off[r + 1]; ++e) near.ent.push_back(leaf_of(c, pn.ent[e]));
        near.off.push_back(static_cast<uint32_t>(near.ent.size()));
      }
      for (uint32_t r = 0; r < pw.rows(); ++r) {
        for (uint32_t e = pw.off[r]; e < pw.off[r + 1]; ++e) w.ent.push_back(pw.ent[e] + bb[c][l + 1]);
        w.off.push_back(static_cast<uint32_t>(w.ent.size()));
      }
      // V and X: one row per box of the level
      const Csr &pv = p.v[l], &px = p.x[l];
      for (uint32_t r = 0; r < pv.rows(); ++r) {
        for (uint32_t e = pv.off[r]; e < pv.off[r + 1]; ++e) {
          v.ent.push_back(pv.ent[e] + bb[c][l]);
          v.aux.push_back(pv.aux[e]);
        }
        v.off.push_back(static_cast<uint32_t>(v.ent.size()));
      }
      for (uint32_t r = 0; r < px.rows(); ++r) {
        for (uint32_t e = px.off[r]; e < px.off[r + 1]; ++e) x.ent.push_back(leaf_of(c, px.ent[e]));
        x.off.push_back(static_cast<uint32_t>(x.ent.size()));
      }
    }
  }

  // points in the stacked leaf order; caller indices run over the clouds
  auto points_of = [&](bool src) {
    std::vector<uint32_t> start(nleaves + 1, 0), order;
    std::vector<double> u;
    order.reserve(f.ns);
    u.reserve(3 * f.ns);
    // inverse of leaf_of: (cloud, local leaf) of every global leaf
    std::vector<std::pair<int, uint32_t>> owner(nleaves);
    for (int c = 0; c < nclouds; ++c)
      for (uint32_t i = 0; i < clouds[c].nleaves(); ++i) owner[leaf_of(c, i)] = {c, i};
    for (uint32_t g = 0; g < nleaves; ++g) {
      const Plan &p = clouds[owner[g].first];
      const uint32_t i = owner[g].second;
      const std::vector<uint32_t> &st = src ? p.src_start : p.tgt_start;
      const std::vector<uint32_t> &od = src ? p.src_order : p.tgt_order;
      const std::vector<double> &pu = src ? p.us : p.ut;
      for (uint32_t k = st[i]; k < st[i + 1]; ++k) {
        order.push_back(static_cast<uint32_t>(first[owner[g].first] + od[k]));
        u.insert(u.end(), {pu[3 * k], pu[3 * k + 1], pu[3 * k + 2]});
      }
      start[g + 1] = static_cast<uint32_t>(order.size());
    }
    if (src) {
      f.src_start.swap(start);
      f.src_order.swap(order);
      f.us.swap(u);
    } else {
      f.tgt_start.swap(start);
      f.tgt_order.swap(order);
      f.ut.swap(u);
    }
  };
  points_of(true);
  points_of(false);
  return f;
}

}  // namespace fmmh
