// plan.cpp -- octree, box index, interaction lists and leaf-ordered points on one rank.
//
// The tree follows the reference's rules (references in plan.hpp); the algorithms are
// this library's own: keys are sorted once and refined over contiguous key ranges, the
// 2:1 balance ripples level by level from the deepest interior boxes up (every interior
// box's 26 same-level neighbour cells must exist, the condition the reference's closure
// enforces in octree/src/octree/implementation.rs:1080-1120), and the lists are
// resolved through one open-addressing table of all held keys.
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>

#include "plan.hpp"

namespace fmmh {

double Plan::radius(int level) const { return side / std::ldexp(1.0, level + 1); }

namespace {

// Open-addressing map key -> (box index, is_leaf) over every held key.
struct KeyMap {
  std::vector<uint64_t> slot_key;
  std::vector<uint32_t> slot_val;  // box index; bit 31 set for a leaf
  uint64_t mask = 0;

  void init(size_t count) {
    size_t cap = 16;
    while (cap < 2 * count + 16) cap <<= 1;
    slot_key.assign(cap, ~0ULL);
    slot_val.assign(cap, 0);
    mask = cap - 1;
  }
  static uint64_t hash(uint64_t k) {
    k ^= k >> 33;
    k *= 0xff51afd7ed558ccdULL;
    k ^= k >> 33;
    return k;
  }
  void insert(uint64_t k, uint32_t v) {
    uint64_t h = hash(k) & mask;
    while (slot_key[h] != ~0ULL && slot_key[h] != k) h = (h + 1) & mask;
    slot_key[h] = k;
    slot_val[h] = v;
  }
  // returns false when absent
  bool get(uint64_t k, uint32_t &v) const {
    uint64_t h = hash(k) & mask;
    while (slot_key[h] != ~0ULL) {
      if (slot_key[h] == k) {
        v = slot_val[h];
        return true;
      }
      h = (h + 1) & mask;
    }
    return false;
  }
};

constexpr uint32_t kLeafBit = 0x80000000u;

// morton::from_physical_point (octree/src/morton.rs:255-273) at level 16: the reference
// coordinate (x - min) / (max - min) per axis, scaled by 2^16, truncated (negative and
// NaN to 0 as Rust's saturating cast) and clamped to the last cell.
inline uint64_t fine_key(const double *x, const std::array<double, 6> &box) {
  uint64_t idx[3];
  for (int k = 0; k < 3; ++k) {
    const double ref = (x[k] - box[k]) / (box[k + 3] - box[k]);
    const double v = ref * 65536.0;
    uint64_t i = (v > 0.0) ? static_cast<uint64_t>(v) : 0;
    idx[k] = i < 65535 ? i : 65535;
  }
  return make_key(kDeepest, idx[0], idx[1], idx[2]);
}

void refine(const uint64_t *keys, size_t n, uint64_t box, int max_level, int64_t max_keys,
            std::vector<uint64_t> &out) {
  const int level = key_level(box);
  if (level < max_level && static_cast<int64_t>(n) > max_keys) {
    const int shift = kLevelBits + 3 * (kDeepest - (level + 1));
    size_t at = 0;
    for (int o = 0; o < 8; ++o) {
      const uint64_t c = child_key(box, o);
      const uint64_t prefix = c >> shift;
      const size_t start = at;
      // keys are sorted: the child's keys are one contiguous run
      size_t lo = at, hi = n;
      while (lo < hi) {
        const size_t mid = (lo + hi) / 2;
        if ((keys[mid] >> shift) <= prefix)
          lo = mid + 1;
        else
          hi = mid;
      }
      at = lo;
      refine(keys + start, at - start, c, max_level, max_keys, out);
    }
  } else {
    out.push_back(box);
  }
}

// 2:1 balance: process interior boxes from the deepest interior level up; every
// same-level neighbour cell of an interior box must be held, so a coarser leaf covering
// one is split down to it (creating interior boxes only on coarser levels, handled later).
std::vector<uint64_t> ripple_balance(std::vector<uint64_t> leaves) {
  int deepest = 0;
  for (uint64_t k : leaves) deepest = std::max(deepest, key_level(k));
  if (deepest == 0) return leaves;
  // held keys: leaf flag
  KeyMap held;
  held.init(leaves.size() * 4 + 64);
  std::vector<std::vector<uint64_t>> interior(kDeepest + 1);
  auto add_interior = [&](uint64_t k) {
    held.insert(k, 0);
    interior[key_level(k)].push_back(k);
  };
  for (uint64_t k : leaves) {
    held.insert(k, kLeafBit);
    uint64_t a = k;
    while (key_level(a) > 0) {
      a = parent_key(a);
      uint32_t v;
      if (held.get(a, v)) break;
      add_interior(a);
    }
  }
  // the map must grow if many splits happen: rebuild with more room when half full
  size_t count = 0;
  for (auto &lv : interior) count += lv.size();
  count += leaves.size();
  auto grow = [&]() {
    if (2 * (count + 64) < held.slot_key.size()) return;
    KeyMap bigger;
    bigger.init(held.slot_key.size());
    for (size_t s = 0; s < held.slot_key.size(); ++s)
      if (held.slot_key[s] != ~0ULL) bigger.insert(held.slot_key[s], held.slot_val[s]);
    held = std::move(bigger);
  };
  for (int j = deepest - 1; j >= 1; --j) {
    const std::vector<uint64_t> level_boxes = interior[j];  // stable during this level
    for (uint64_t P : level_boxes) {
      const Cell c = decode(P);
      const int64_t size = int64_t(1) << j;
      for (int dx = -1; dx <= 1; ++dx)
        for (int dy = -1; dy <= 1; ++dy)
          for (int dz = -1; dz <= 1; ++dz) {
            if (!dx && !dy && !dz) continue;
            const int64_t x = c.i + dx, y = c.j + dy, z = c.k + dz;
            if (x < 0 || y < 0 || z < 0 || x >= size || y >= size || z >= size) continue;
            const uint64_t N = make_key(j, x, y, z);
            uint32_t v;
            if (held.get(N, v)) continue;
            // the covering leaf is the nearest held ancestor
            uint64_t L = N;
            do {
              L = parent_key(L);
            } while (!held.get(L, v));
            // split L toward N
            while (key_level(L) < j) {
              held.insert(L, 0);
              interior[key_level(L)].push_back(L);
              const int lv = key_level(L) + 1;
              const uint64_t next = ancestor_at(N, lv);
              for (int o = 0; o < 8; ++o) {
                const uint64_t ch = child_key(L, o);
                held.insert(ch, kLeafBit);
                ++count;
              }
              grow();
              L = next;
            }
          }
    }
  }
  std::vector<uint64_t> out;
  out.reserve(held.slot_key.size() / 2);
  for (size_t s = 0; s < held.slot_key.size(); ++s)
    if (held.slot_key[s] != ~0ULL && (held.slot_val[s] & kLeafBit)) out.push_back(held.slot_key[s]);
  std::sort(out.begin(), out.end());
  return out;
}

void leaf_coordinates(const double *x, uint64_t leaf, const Plan &plan, double *u) {
  // fmm-exec/src/geometry.rs:291-298: fl(fl(x - a) 2^(l+1) / w - (2i + 1))
  const Cell c = decode(leaf);
  const double scale = std::ldexp(1.0, c.level + 1);
  const int64_t idx[3] = {c.i, c.j, c.k};
  for (int k = 0; k < 3; ++k) {
    const double d = x[k] - plan.lower[k];
    const double s = d * scale;
    u[k] = s / plan.side - static_cast<double>(2 * idx[k] + 1);
  }
}

void assign_points(const double *pts, int64_t n, const Plan &plan,
                   const std::vector<uint64_t> &leaves_morton,
                   const std::vector<uint32_t> &morton_to_leaf, std::vector<uint32_t> &start,
                   std::vector<uint32_t> &order, std::vector<double> &u) {
  const uint32_t nl = plan.nleaves();
  std::vector<uint32_t> leaf_of(n);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    const uint64_t k = fine_key(pts + 3 * i, plan.box);
    const size_t pos = std::upper_bound(leaves_morton.begin(), leaves_morton.end(), k) - leaves_morton.begin();
    leaf_of[i] = morton_to_leaf[pos - 1];
  }
  start.assign(nl + 1, 0);
  for (int64_t i = 0; i < n; ++i) start[leaf_of[i] + 1]++;
  for (uint32_t l = 0; l < nl; ++l) start[l + 1] += start[l];
  std::vector<uint32_t> pos(start.begin(), start.end() - 1);
  order.resize(n);
  for (int64_t i = 0; i < n; ++i) order[pos[leaf_of[i]]++] = static_cast<uint32_t>(i);
  u.resize(3 * n);
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t l = 0; l < nl; ++l)
    for (uint32_t j = start[l]; j < start[l + 1]; ++j)
      leaf_coordinates(pts + 3 * order[j], plan.leaf_key[l], plan, u.data() + 3 * j);
}

// The near rule with named crossovers (fmm-plan/src/near.rs:15-40): a W entry of a leaf
// that is itself a leaf joins the near row when it holds fewer than near_w sources; the
// X entries of a leaf join its near row when it holds fewer than near_x targets.  Near
// rows keep leaf (level, key) order; the converted entries leave their W and X rows.
void apply_near_rule(Plan &plan) {
  const double nw = plan.options.near_w, nx = plan.options.near_x;
  if (nw < 0 && nx < 0) return;
  for (int l = 0; l < plan.nlevels; ++l) {
    const uint32_t l0 = plan.leaf_off[l], nll = plan.leaf_off[l + 1] - l0;
    Csr &N = plan.near[l];
    Csr &W = plan.w[l];
    Csr &X = plan.x[l];
    Csr nn, nwc, nxc;
    nn.off.assign(nll + 1, 0);
    nwc.off.assign(nll + 1, 0);
    std::vector<bool> drop_x(plan.nboxes(l), false);
    for (uint32_t r = 0; r < nll; ++r) {
      const uint32_t lf = l0 + r;
      std::vector<uint32_t> row(N.ent.begin() + N.off[r], N.ent.begin() + N.off[r + 1]);
      for (uint32_t e = W.off[r]; e < W.off[r + 1]; ++e) {
        const uint32_t sb = W.ent[e];
        const uint32_t sl = l + 1 < plan.nlevels ? plan.box_leaf[l + 1][sb] : kNone;
        if (nw >= 0 && sl != kNone && double(plan.src_start[sl + 1] - plan.src_start[sl]) < nw)
          row.push_back(sl);
        else
          nwc.ent.push_back(sb);
      }
      const uint32_t b = plan.leaf_box[lf];
      if (nx >= 0 && double(plan.tgt_start[lf + 1] - plan.tgt_start[lf]) < nx) {
        for (uint32_t e = X.off[b]; e < X.off[b + 1]; ++e) row.push_back(X.ent[e]);
        drop_x[b] = true;
      }
      std::sort(row.begin(), row.end());
      nn.ent.insert(nn.ent.end(), row.begin(), row.end());
      nn.off[r + 1] = static_cast<uint32_t>(nn.ent.size());
      nwc.off[r + 1] = static_cast<uint32_t>(nwc.ent.size());
    }
    const uint32_t nb = plan.nboxes(l);
    nxc.off.assign(nb + 1, 0);
    for (uint32_t b = 0; b < nb; ++b) {
      if (!drop_x[b]) nxc.ent.insert(nxc.ent.end(), X.ent.begin() + X.off[b], X.ent.begin() + X.off[b + 1]);
      nxc.off[b + 1] = static_cast<uint32_t>(nxc.ent.size());
    }
    N = std::move(nn);
    W = std::move(nwc);
    X = std::move(nxc);
  }
}

}  // namespace

Plan plan_begin(int64_t ns, int64_t nt, bool same_points, const PlanOptions &options) {
  require(ns >= 0 && nt >= 0 && ns + nt > 0, "build_plan: no points");
  require(ns + nt < (int64_t(1) << 32), "build_plan: too many points", kTooLarge);
  require(options.leaf_points >= 1, "build_plan: leaf_points must be positive");
  Plan plan;
  plan.options = options;
  plan.ns = ns;
  plan.nt = nt;
  plan.same_points = same_points;
  return plan;
}

void plan_domain(Plan &plan, const double lo[3], const double hi[3]) {
  double diam[3], mean[3];
  for (int k = 0; k < 3; ++k) {
    diam[k] = hi[k] - lo[k];
    mean[k] = lo[k] + 0.5 * diam[k];
  }
  double maxd = std::max(std::max(diam[0], diam[1]), diam[2]);
  require(maxd > 0.0 && std::isfinite(maxd), "build_plan: the points span nothing");
  maxd = maxd * (1.0 + 1.0 / 65536.0);
  for (int k = 0; k < 3; ++k) {
    plan.box[k] = mean[k] - 0.5 * maxd;
    plan.box[k + 3] = mean[k] + 0.5 * maxd;
    plan.lower[k] = plan.box[k];
  }
  const double sides[3] = {plan.box[3] - plan.box[0], plan.box[4] - plan.box[1], plan.box[5] - plan.box[2]};
  plan.side = std::max(std::max(sides[0], sides[1]), sides[2]);
}

uint64_t plan_fine_key(const Plan &plan, const double *x) { return fine_key(x, plan.box); }

void plan_tree(Plan &plan, std::vector<uint64_t> fine) {
  const PlanOptions &options = plan.options;
  // Refinement from the root, then 2:1 balance.
  std::vector<uint64_t> leaves;
  refine(fine.data(), fine.size(), 0ULL, std::min(options.max_level, kDeepest), options.leaf_points, leaves);
  std::vector<uint64_t>().swap(fine);
  leaves = ripple_balance(std::move(leaves));

  // Box index: leaves and all their ancestors, per level in Morton order.
  int deepest = 0;
  for (uint64_t k : leaves) deepest = std::max(deepest, key_level(k));
  plan.nlevels = deepest + 1;
  const int L = plan.nlevels;
  plan.keys.assign(L, {});
  for (uint64_t k : leaves) {
    plan.keys[key_level(k)].push_back(k);
    // a leaf's parent is interior; child 0 adds its parent, which is enough since every
    // interior box has all eight children held
    if (key_level(k) > 0 && octant_of(k) == 0) {
      uint64_t a = parent_key(k);
      plan.keys[key_level(a)].push_back(a);
    }
  }
  // interior boxes whose child 0 is itself interior: add ancestors level by level
  for (int l = L - 1; l >= 1; --l) {
    std::vector<uint64_t> add;
    for (uint64_t k : plan.keys[l])
      if (octant_of(k) == 0) add.push_back(parent_key(k));
    auto &dst = plan.keys[l - 1];
    dst.insert(dst.end(), add.begin(), add.end());
    std::sort(dst.begin(), dst.end());
    dst.erase(std::unique(dst.begin(), dst.end()), dst.end());
  }
  std::sort(plan.keys[L - 1].begin(), plan.keys[L - 1].end());
  if (plan.keys[0].empty()) plan.keys[0].push_back(0ULL);

  // Leaves by (level, key).
  plan.leaf_key = leaves;
  std::stable_sort(plan.leaf_key.begin(), plan.leaf_key.end(), [](uint64_t a, uint64_t b) {
    const int la = key_level(a), lb = key_level(b);
    return la != lb ? la < lb : a < b;
  });
  const uint32_t nl = plan.nleaves();
  plan.leaf_box.resize(nl);
  plan.leaf_off.assign(L + 1, 0);
  plan.box_leaf.assign(L, {});
  plan.first_child.assign(L, {});
  plan.parent.assign(L, {});
  KeyMap map;
  size_t total = 0;
  for (int l = 0; l < L; ++l) total += plan.keys[l].size();
  map.init(total);
  for (int l = 0; l < L; ++l) {
    plan.box_leaf[l].assign(plan.keys[l].size(), kNone);
    for (uint32_t b = 0; b < plan.keys[l].size(); ++b) map.insert(plan.keys[l][b], b);
  }
  for (uint32_t i = 0; i < nl; ++i) {
    const uint64_t k = plan.leaf_key[i];
    const int l = key_level(k);
    uint32_t b = 0;
    map.get(k, b);
    plan.leaf_box[i] = b;
    plan.box_leaf[l][b] = i;
    plan.leaf_off[l + 1]++;
    map.insert(k, b | kLeafBit);
  }
  for (int l = 0; l < L; ++l) plan.leaf_off[l + 1] += plan.leaf_off[l];
  for (int l = 0; l < L; ++l) {
    const uint32_t nb = plan.nboxes(l);
    plan.first_child[l].assign(nb, kNone);
    plan.parent[l].assign(nb, kNone);
    for (uint32_t b = 0; b < nb; ++b) {
      const uint64_t k = plan.keys[l][b];
      uint32_t v;
      if (plan.box_leaf[l][b] == kNone && l + 1 < L && map.get(child_key(k, 0), v))
        plan.first_child[l][b] = v & ~kLeafBit;
      if (l > 0 && map.get(parent_key(k), v)) plan.parent[l][b] = v & ~kLeafBit;
    }
  }
  auto held = [&](uint64_t k, uint32_t &box, bool &leaf) {
    uint32_t v;
    if (!map.get(k, v)) return false;
    box = v & ~kLeafBit;
    leaf = (v & kLeafBit) != 0;
    return true;
  };

  // Interaction lists (fmm-plan/src/interaction_manager.rs:10-28).
  plan.near.assign(L, {});
  plan.w.assign(L, {});
  plan.v.assign(L, {});
  plan.x.assign(L, {});
  for (int l = 0; l < L; ++l) {
    const uint32_t nb = plan.nboxes(l);
    const uint32_t l0 = plan.leaf_off[l], nll = plan.leaf_off[l + 1] - l0;
    std::vector<std::vector<std::pair<uint16_t, uint32_t>>> vrows(nb);
    std::vector<std::vector<uint64_t>> xrows(nb), urows(nll), wrows(nll);
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t bi = 0; bi < nb; ++bi) {
      const uint64_t B = plan.keys[l][bi];
      const Cell cb = decode(B);
      if (l >= 1) {
        const uint64_t P = parent_key(B);
        const Cell cp = decode(P);
        const int64_t size = int64_t(1) << (l - 1);
        for (int dx = -1; dx <= 1; ++dx)
          for (int dy = -1; dy <= 1; ++dy)
            for (int dz = -1; dz <= 1; ++dz) {
              if (!dx && !dy && !dz) continue;
              const int64_t x = cp.i + dx, y = cp.j + dy, z = cp.k + dz;
              if (x < 0 || y < 0 || z < 0 || x >= size || y >= size || z >= size) continue;
              const uint64_t N = make_key(l - 1, x, y, z);
              uint32_t nbx;
              bool nleaf;
              if (!held(N, nbx, nleaf)) continue;
              if (nleaf) {
                if (!adjacent(N, B)) xrows[bi].push_back(N);
                continue;
              }
              for (int o = 0; o < 8; ++o) {
                const int64_t ci = 2 * x + ((o >> 2) & 1), cj = 2 * y + ((o >> 1) & 1), ck = 2 * z + (o & 1);
                const int64_t ddx = cb.i - ci, ddy = cb.j - cj, ddz = cb.k - ck;
                if (std::llabs(ddx) <= 1 && std::llabs(ddy) <= 1 && std::llabs(ddz) <= 1) continue;
                uint32_t cbx;
                bool cleaf;
                held(make_key(l, ci, cj, ck), cbx, cleaf);
                vrows[bi].push_back({static_cast<uint16_t>(offset_index((int)ddx, (int)ddy, (int)ddz)), cbx});
              }
            }
        std::sort(vrows[bi].begin(), vrows[bi].end());
        std::sort(xrows[bi].begin(), xrows[bi].end());
      }
      const uint32_t leaf = plan.box_leaf[l][bi];
      if (leaf == kNone) continue;
      const uint32_t r = leaf - l0;
      auto &U = urows[r];
      auto &W = wrows[r];
      U.push_back(B);
      const int64_t size = int64_t(1) << l;
      for (int dx = -1; dx <= 1; ++dx)
        for (int dy = -1; dy <= 1; ++dy)
          for (int dz = -1; dz <= 1; ++dz) {
            if (!dx && !dy && !dz) continue;
            const int64_t x = cb.i + dx, y = cb.j + dy, z = cb.k + dz;
            if (x < 0 || y < 0 || z < 0 || x >= size || y >= size || z >= size) continue;
            const uint64_t N = make_key(l, x, y, z);
            uint32_t nbx;
            bool nleaf;
            if (held(N, nbx, nleaf)) {
              if (nleaf) {
                U.push_back(N);
                continue;
              }
              for (int o = 0; o < 8; ++o) {
                const uint64_t C = child_key(N, o);
                uint32_t cbx;
                bool cleaf;
                held(C, cbx, cleaf);
                if (adjacent(C, B)) {
                  // 2:1 balance makes every child of a neighbour that touches a leaf a leaf
                  if (!cleaf) throw Error(kInvalidArgument, "tree is not 2:1 balanced (U)");
                  U.push_back(C);
                } else {
                  W.push_back(C);
                }
              }
            } else {
              // covered by a coarser leaf (one level up under 2:1 balance)
              uint64_t A = N;
              bool aleaf = false;
              uint32_t abx;
              do {
                A = parent_key(A);
              } while (!held(A, abx, aleaf));
              if (aleaf && adjacent(A, B)) U.push_back(A);
            }
          }
      std::sort(U.begin(), U.end(), [](uint64_t a, uint64_t b) {
        const int la = key_level(a), lb = key_level(b);
        return la != lb ? la < lb : a < b;
      });
      U.erase(std::unique(U.begin(), U.end()), U.end());
      std::sort(W.begin(), W.end());
    }
    auto &V = plan.v[l];
    auto &X = plan.x[l];
    V.off.assign(nb + 1, 0);
    X.off.assign(nb + 1, 0);
    for (uint32_t b = 0; b < nb; ++b) {
      V.off[b + 1] = V.off[b] + static_cast<uint32_t>(vrows[b].size());
      X.off[b + 1] = X.off[b] + static_cast<uint32_t>(xrows[b].size());
    }
    V.ent.resize(V.off[nb]);
    V.aux.resize(V.off[nb]);
    X.ent.resize(X.off[nb]);
    for (uint32_t b = 0; b < nb; ++b) {
      for (size_t e = 0; e < vrows[b].size(); ++e) {
        V.aux[V.off[b] + e] = vrows[b][e].first;
        V.ent[V.off[b] + e] = vrows[b][e].second;
      }
      for (size_t e = 0; e < xrows[b].size(); ++e) {
        uint32_t bx;
        bool lf;
        held(xrows[b][e], bx, lf);
        X.ent[X.off[b] + e] = plan.box_leaf[l - 1][bx];
      }
    }
    auto &N = plan.near[l];
    auto &Wc = plan.w[l];
    N.off.assign(nll + 1, 0);
    Wc.off.assign(nll + 1, 0);
    for (uint32_t r = 0; r < nll; ++r) {
      N.off[r + 1] = N.off[r] + static_cast<uint32_t>(urows[r].size());
      Wc.off[r + 1] = Wc.off[r] + static_cast<uint32_t>(wrows[r].size());
    }
    N.ent.resize(N.off[nll]);
    Wc.ent.resize(Wc.off[nll]);
    for (uint32_t r = 0; r < nll; ++r) {
      for (size_t e = 0; e < urows[r].size(); ++e) {
        uint32_t bx;
        bool lf;
        held(urows[r][e], bx, lf);
        N.ent[N.off[r] + e] = plan.box_leaf[key_level(urows[r][e])][bx];
      }
      for (size_t e = 0; e < wrows[r].size(); ++e) {
        uint32_t bx;
        bool lf;
        held(wrows[r][e], bx, lf);
        Wc.ent[Wc.off[r] + e] = bx;
      }
    }
  }

  // the leaves in Morton order and their leaf indices, for the points' assignment
  plan.morton_leaves = leaves;  // sorted by key
  plan.morton_to_leaf.resize(leaves.size());
  for (size_t i = 0; i < leaves.size(); ++i) {
    uint32_t b;
    bool lf;
    held(leaves[i], b, lf);
    plan.morton_to_leaf[i] = plan.box_leaf[key_level(leaves[i])][b];
  }
}

void plan_finish(Plan &plan) {
  apply_near_rule(plan);
  plan.leaf_radius.resize(plan.nleaves());
  for (int l = 0; l < plan.nlevels; ++l)
    for (uint32_t i = plan.leaf_off[l]; i < plan.leaf_off[l + 1]; ++i) plan.leaf_radius[i] = plan.radius(l);
  std::vector<uint64_t>().swap(plan.morton_leaves);
  std::vector<uint32_t>().swap(plan.morton_to_leaf);
}

void plan_leaf_coordinates(const Plan &plan, const double *x, uint32_t leaf, double *u) {
  leaf_coordinates(x, plan.leaf_key[leaf], plan, u);
}

Plan build_plan(const double *sources, int64_t ns, const double *targets, int64_t nt,
                const PlanOptions &options) {
  Plan plan = plan_begin(ns, nt, sources == targets && ns == nt, options);
  const int64_t n = ns + nt;
  auto point = [&](int64_t i) { return i < ns ? sources + 3 * i : targets + 3 * (i - ns); };
  // Domain: compute_global_bounding_box (octree/src/octree.rs:1278-1369).
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int64_t i = 0; i < n; ++i)
    for (int k = 0; k < 3; ++k) {
      const double v = point(i)[k];
      require(std::isfinite(v), "build_plan: a coordinate is not finite");
      lo[k] = std::min(lo[k], v);
      hi[k] = std::max(hi[k], v);
    }
  plan_domain(plan, lo, hi);
  // Distinct level-16 keys of sources then targets.
  std::vector<uint64_t> fine(n);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) fine[i] = fine_key(point(i), plan.box);
  std::sort(fine.begin(), fine.end());
  fine.erase(std::unique(fine.begin(), fine.end()), fine.end());
  plan_tree(plan, std::move(fine));
  // Points in leaf order; the leaf of a point is the last leaf at or before its fine key
  // in Morton order (leaves are disjoint and cover the domain).
  assign_points(sources, ns, plan, plan.morton_leaves, plan.morton_to_leaf, plan.src_start, plan.src_order, plan.us);
  assign_points(targets, nt, plan, plan.morton_leaves, plan.morton_to_leaf, plan.tgt_start, plan.tgt_order, plan.ut);
  plan_finish(plan);
  return plan;
}

}  // namespace fmmh
