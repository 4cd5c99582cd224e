// operator.hip -- the device operator (see operator.hpp).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <array>
#include <cstring>
#include <map>

#include "kernels.hpp"
#include "operator.hpp"
#include "pwtables.hpp"
#include "schedule.hpp"
#include "table_cache.hpp"

namespace fmmh {

void hip_check(hipError_t e, const char *what) {
  if (e == hipSuccess) return;
  const Status s = (e == hipErrorOutOfMemory) ? kOutOfMemory : kDeviceError;
  throw Error(s, std::string(what) + ": " + hipGetErrorString(e));
}

DeviceOperator::DeviceOperator(const Plan &plan, const OperatorOptions &options) : opt_(options) {
  require(opt_.p >= 0 && opt_.p <= dev::kMaxDegree, "the device path supports 0 <= p <= 15", kUnsupportedLayout);
  require(opt_.precision == kPlainF64 || opt_.precision == kMixed, "unknown precision", kUnsupportedPrecision);
  require(opt_.m2l == kM2lDense || opt_.m2l == kM2lRotation || opt_.m2l == kM2lPlaneWave, "unknown M2L method",
          kUnsupportedLayout);
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) throw Error(kNoDevice, "no HIP device");
  require(opt_.device >= 0 && opt_.device < ndev, "device ordinal out of range", kNoDevice);
  hip_check(hipSetDevice(opt_.device), "hipSetDevice");
  hip_check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
  hip_check(hipEventCreateWithFlags(&order_event_, hipEventDisableTiming), "hipEventCreate");
  hip_check(dev::upload_constants(), "upload constants");

  nlevels_ = plan.nlevels;
  nc_ = ncoef(opt_.p);
  ns_ = plan.ns;
  nt_ = plan.nt;
  same_points_ = plan.same_points;
  nleaves_ = plan.nleaves();
  leaf_off_ = plan.leaf_off;
  box_base_.assign(nlevels_ + 1, 0);
  for (int l = 0; l < nlevels_; ++l) box_base_[l + 1] = box_base_[l] + plan.nboxes(l);
  nbox_total_ = box_base_[nlevels_];
  auto gid = [&](int level, uint32_t b) { return box_base_[level] + b; };

  // cells of leaves and boxes
  std::vector<uint32_t> leaf_cell(4 * size_t(nleaves_)), box_cell(4 * nbox_total_);
  for (uint32_t i = 0; i < nleaves_; ++i) {
    const Cell c = decode(plan.leaf_key[i]);
    leaf_cell[4 * i] = c.level;
    leaf_cell[4 * i + 1] = uint32_t(c.i);
    leaf_cell[4 * i + 2] = uint32_t(c.j);
    leaf_cell[4 * i + 3] = uint32_t(c.k);
  }
  for (int l = 0; l < nlevels_; ++l)
    for (uint32_t b = 0; b < plan.nboxes(l); ++b) {
      const Cell c = decode(plan.keys[l][b]);
      const size_t g = gid(l, b);
      box_cell[4 * g] = c.level;
      box_cell[4 * g + 1] = uint32_t(c.i);
      box_cell[4 * g + 2] = uint32_t(c.j);
      box_cell[4 * g + 3] = uint32_t(c.k);
    }
  leaf_cell_.upload(leaf_cell);
  box_cell_.upload(box_cell);
  leaf_r_.upload(plan.leaf_radius);
  src_start_.upload(plan.src_start);
  tgt_start_.upload(plan.tgt_start);
  src_order_.upload(plan.src_order);
  tgt_order_.upload(plan.tgt_order);

  // near rows of every leaf, concatenated in leaf order
  std::vector<uint32_t> near_off(nleaves_ + 1, 0), near_ent;
  for (int l = 0; l < nlevels_; ++l) {
    const Csr &N = plan.near[l];
    for (uint32_t r = 0; r < N.rows(); ++r) {
      const uint32_t lf = plan.leaf_off[l] + r;
      near_off[lf + 1] = N.off[r + 1] - N.off[r];
      for (uint32_t e = N.off[r]; e < N.off[r + 1]; ++e) {
        near_ent.push_back(N.ent[e]);
        const uint32_t sl = N.ent[e];
        p2p_pairs_ += uint64_t(plan.src_start[sl + 1] - plan.src_start[sl]) *
                      (plan.tgt_start[lf + 1] - plan.tgt_start[lf]);
      }
    }
  }
  for (uint32_t i = 0; i < nleaves_; ++i) near_off[i + 1] += near_off[i];
  near_off_.upload(near_off);
  near_ent_.upload(near_ent);

  // points in leaf order
  std::vector<double> sx(ns_), sy(ns_), sz(ns_), tx(nt_), ty(nt_), tz(nt_);
  for (int64_t i = 0; i < ns_; ++i) {
    sx[i] = plan.us[3 * i];
    sy[i] = plan.us[3 * i + 1];
    sz[i] = plan.us[3 * i + 2];
  }
  for (int64_t i = 0; i < nt_; ++i) {
    tx[i] = plan.ut[3 * i];
    ty[i] = plan.ut[3 * i + 1];
    tz[i] = plan.ut[3 * i + 2];
  }
  sx_.upload(sx);
  sy_.upload(sy);
  sz_.upload(sz);
  tx_.upload(tx);
  ty_.upload(ty);
  tz_.upload(tz);
  q_.alloc(ns_);
  q_in_.alloc(ns_);
  if (opt_.precision == kMixed && ns_ > 0) {
    std::vector<int32_t> irec(4 * size_t(ns_));
    dev::build_source_records(sx.data(), sy.data(), sz.data(), size_t(ns_), irec.data());
    std::vector<float> rec(irec.size());
    std::memcpy(rec.data(), irec.data(), irec.size() * sizeof(int32_t));
    src_rec_.upload(rec);
    // the fixed-point separations assume |x'| <= 7 half-widths for every near pair (2:1
    // balance), in either leaf's frame; check the plan's near rows once
    for (int l = 0; l < plan.nlevels; ++l) {
      const Csr &nr = plan.near[l];
      for (uint32_t r = 0; r + 1 < uint32_t(nr.off.size()); ++r) {
        const Cell t = decode(plan.leaf_key[plan.leaf_off[l] + r]);
        for (uint32_t e = nr.off[r]; e < nr.off[r + 1]; ++e) {
          const Cell c = decode(plan.leaf_key[nr.ent[e]]);
          const double ct[3] = {double(t.i), double(t.j), double(t.k)}, cs[3] = {double(c.i), double(c.j), double(c.k)};
          // max over the axes of |c''| + rho: the targets in the source leaf's half-widths
          // (the uniform-source form) and the sources in the target leaf's (the staged form)
          double reach = 0.0;
          for (int k = 0; k < 3; ++k) {
            const double dt = (2.0 * ct[k] + 1.0) * std::ldexp(1.0, -int(t.level));
            const double ds = (2.0 * cs[k] + 1.0) * std::ldexp(1.0, -int(c.level));
            reach = std::max(reach, std::fabs(dt - ds) * std::ldexp(1.0, int(c.level)) +
                                        std::ldexp(1.0, int(c.level) - int(t.level)));
            reach = std::max(reach, std::fabs(dt - ds) * std::ldexp(1.0, int(t.level)) +
                                        std::ldexp(1.0, int(t.level) - int(c.level)));
          }
          require(reach <= 7.0, "P2P: a near pair spans more than 7 half-widths (the tree is not 2:1 balanced)",
                  kUnsupportedLayout);
        }
      }
    }
  }

  // expansions and outputs (the clustered-mode scratch rows are sized in build_schedules)
  // 64 doubles past the last box: the per-M-tile M2M reads a source expansion as four
  // runs of 4 ceil(nc / 16) values (at most 15 past its end) and masks the tail
  mult_.alloc(nbox_total_ * nc_ + 64);
  phi_hat_.alloc(nt_);
  grad_hat_.alloc(opt_.gradients ? 3 * nt_ : 0);
  phi_.alloc(nt_);
  grad_.alloc(opt_.gradients ? 3 * nt_ : 0);

  // tables, swizzled for the MFMA contraction
  {
    std::vector<double> m2m, l2l, m2l;
    cached_tables(opt_.p, m2m, l2l, m2l);
    const size_t n2 = size_t(nc_) * nc_, sw = dev::swizzled_table_size(opt_.p);
    auto swz = [&](const std::vector<double> &t, int count, DevBuf<double> &dst) {
      std::vector<double> out(sw * count);
      for (int i = 0; i < count; ++i) dev::swizzle_table(opt_.p, t.data() + i * n2, out.data() + i * sw);
      dst.upload(out);
    };
    swz(m2m, 8, t_m2m_);
    swz(l2l, 8, t_l2l_);
    if (nlevels_ > 2 && opt_.m2l == kM2lDense) {
      if (opt_.precision == kMixed) {
        const size_t sw16 = dev::swizzled_table_size_bf16(opt_.p);
        std::vector<uint16_t> out(sw16 * 316);
        for (int i = 0; i < 316; ++i) dev::swizzle_table_bf16(opt_.p, m2l.data() + i * n2, out.data() + i * sw16);
        t_m2l16_.upload(out);
      } else {
        swz(m2l, 316, t_m2l_);
      }
    }
  }

  // leaf-kernel rows
  {
    std::vector<uint32_t> id, cell, off{0}, ent;
    p2m_level_.assign(nlevels_ + 1, 0);
    for (int l = 0; l < nlevels_; ++l) {
      for (uint32_t lf = plan.leaf_off[l]; lf < plan.leaf_off[l + 1]; ++lf) {
        if (plan.src_start[lf + 1] == plan.src_start[lf]) continue;
        id.push_back(gid(l, plan.leaf_box[lf]));
        cell.insert(cell.end(), {leaf_cell[4 * lf], leaf_cell[4 * lf + 1], leaf_cell[4 * lf + 2], leaf_cell[4 * lf + 3]});
        ent.push_back(lf);
        off.push_back(static_cast<uint32_t>(ent.size()));
      }
      p2m_level_[l + 1] = static_cast<uint32_t>(id.size());
    }
    upload_rows(p2m_rows_, id, cell, off, ent);
  }
  {
    std::vector<uint32_t> id, cell, off{0}, ent;
    l2p_level_.assign(nlevels_ + 1, 0);
    for (int l = 0; l < nlevels_; ++l) {
      for (uint32_t lf = plan.leaf_off[l]; lf < plan.leaf_off[l + 1]; ++lf) {
        if (plan.tgt_start[lf + 1] == plan.tgt_start[lf]) continue;
        id.push_back(lf);
        ent.push_back(gid(l, plan.leaf_box[lf]));
        off.push_back(static_cast<uint32_t>(ent.size()));
      }
      l2p_level_[l + 1] = static_cast<uint32_t>(id.size());
    }
    upload_rows(l2p_rows_, id, cell, off, ent);
  }
  {
    std::vector<uint32_t> id, cell, off{0}, ent;
    m2p_level_.assign(nlevels_ + 1, 0);
    for (int l = 0; l < nlevels_; ++l) {
      const Csr &W = plan.w[l];
      for (uint32_t r = 0; r < W.rows(); ++r) {
        const uint32_t lf = plan.leaf_off[l] + r;
        if (W.off[r + 1] == W.off[r] || plan.tgt_start[lf + 1] == plan.tgt_start[lf]) continue;
        id.push_back(lf);
        for (uint32_t e = W.off[r]; e < W.off[r + 1]; ++e) ent.push_back(gid(l + 1, W.ent[e]));
        off.push_back(static_cast<uint32_t>(ent.size()));
      }
      m2p_level_[l + 1] = static_cast<uint32_t>(id.size());
    }
    upload_rows(m2p_rows_, id, cell, off, ent);
  }
  {
    std::vector<uint32_t> id, cell, off{0}, ent;
    p2l_level_.assign(nlevels_ + 1, 0);
    for (int l = 0; l < nlevels_; ++l) {
      const Csr &X = plan.x[l];
      for (uint32_t b = 0; l >= 1 && b < X.rows(); ++b) {
        if (X.off[b + 1] == X.off[b]) continue;
        const size_t g = gid(l, b);
        id.push_back(uint32_t(g));
        cell.insert(cell.end(), {box_cell[4 * g], box_cell[4 * g + 1], box_cell[4 * g + 2], box_cell[4 * g + 3]});
        for (uint32_t e = X.off[b]; e < X.off[b + 1]; ++e) ent.push_back(X.ent[e]);
        off.push_back(static_cast<uint32_t>(ent.size()));
      }
      p2l_level_[l + 1] = static_cast<uint32_t>(id.size());
    }
    upload_rows(p2l_rows_, id, cell, off, ent);
  }
  build_schedules(plan);
  if (opt_.m2l == kM2lRotation) build_rotation(plan);
  if (opt_.m2l == kM2lPlaneWave) build_plane_wave(plan);
  loc_.alloc(nbox_total_ * nc_);
  if (opt_.precision == kMixed)
    slots32_.alloc(nslots_ * nc_);
  else
    slots_.alloc(nslots_ * nc_);
  if (opt_.precision == kMixed) mult16_.alloc(nbox_total_ * dev::split_expansion_size(opt_.p));
  if (opt_.precision == kMixed && near_ent_.count > 0) {
    // the near entries whose pairs the fixed-point P2P must test (one pass at build)
    near_close_.alloc(near_ent_.count);
    dev::p2p_close_flags(stream_, nleaves_, near_off_.ptr, near_ent_.ptr, leaf_cell_.ptr, src_start_.ptr,
                         tgt_start_.ptr, src_rec_.ptr, tx_.ptr, ty_.ptr, tz_.ptr, near_close_.ptr);
    hip_check(hipStreamSynchronize(stream_), "P2P close-pair flags");
  }
  hip_check(hipDeviceSynchronize(), "operator build");
}

DeviceOperator::~DeviceOperator() {
  if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
  if (graph_) (void)hipGraphDestroy(graph_);
  for (hipEvent_t e : events_) (void)hipEventDestroy(e);
  if (order_event_) (void)hipEventDestroy(order_event_);
  if (stream_) (void)hipStreamDestroy(stream_);
}

void DeviceOperator::wait_for(hipStream_t producer) {
  hip_check(hipEventRecord(order_event_, producer), "hipEventRecord");
  hip_check(hipStreamWaitEvent(stream_, order_event_, 0), "hipStreamWaitEvent");
}

void DeviceOperator::set_timing(bool enable) {
  sync();
  timing_ = enable;
  if (enable && events_.empty()) {
    events_.resize(1024);
    for (auto &e : events_) hip_check(hipEventCreate(&e), "hipEventCreate");
  }
  marks_.clear();
  marks_.reserve(events_.size() / 2);
  events_used_ = 0;
  for (int k = 0; k < 16; ++k) {
    kind_ms_[k] = 0.0;
    kind_launches_[k] = 0;
  }
}

template <class F>
void DeviceOperator::timed(int kind, F &&f) {
  if (!timing_ || events_used_ + 2 > events_.size()) {
    f();
    return;
  }
  Mark m{kind, events_[events_used_], events_[events_used_ + 1]};
  events_used_ += 2;
  hip_check(hipEventRecord(m.a, stream_), "hipEventRecord");
  f();
  hip_check(hipEventRecord(m.b, stream_), "hipEventRecord");
  marks_.push_back(m);
}

void DeviceOperator::collect_timings() {
  for (const Mark &m : marks_) {
    float ms = 0.0f;
    hip_check(hipEventElapsedTime(&ms, m.a, m.b), "hipEventElapsedTime");
    kind_ms_[m.kind] += ms;
    kind_launches_[m.kind] += 1;
  }
  marks_.clear();
  events_used_ = 0;
}

void DeviceOperator::kind_times(double *ms, int64_t *launches) {
  sync();
  for (int k = 0; k < kNumKinds; ++k) {
    if (ms) ms[k] = kind_ms_[k];
    if (launches) launches[k] = kind_launches_[k];
  }
}

void DeviceOperator::build_schedules(const Plan &plan) {
  auto gid = [&](int level, uint32_t b) { return box_base_[level] + b; };
  // rank of offset d for targets of octant o: (parent-level neighbour D, source octant o').
  // The plan's V offsets are d = target - source, so the source child sits at ob - d
  // (child units, from the target's parent's lower corner): D = floor((ob - d) / 2) + 1
  // per axis in 0..2 and o' = (ob - d) mod 2, for every offset a V list can hold.
  for (int oc = 0; oc < 8; ++oc) {
    const int ob[3] = {(oc >> 2) & 1, (oc >> 1) & 1, oc & 1};
    for (int d = 0; d < 316; ++d) {
      int dv[3];
      offset_of_index(d, dv);
      int D = 0, op = 0;
      for (int a = 0; a < 3; ++a) {
        const int s = ob[a] - dv[a];
        const int bit = ((s % 2) + 2) % 2;
        D = 3 * D + ((s - bit) / 2 + 1);
        op = 2 * op + bit;
      }
      m2l_rank_[oc][d] = static_cast<uint16_t>(8 * D + op);
    }
  }
  const int C = dev::COLS;
  m2m_.resize(nlevels_);
  l2l_.resize(nlevels_);
  for (int l = 0; l < nlevels_; ++l) {
    // M2M into the parents of level l, by octant (fmm-plan/src/operator.rs:256-262)
    if (l + 1 < nlevels_) {
      // tasks of 16 parents, one wave per child octant (dev::m2m_split)
      const int width = 16;
      ScheduleBuilder sb(width);
      const int C = width;
      std::vector<uint32_t> cols;
      std::vector<std::vector<std::pair<uint16_t, uint32_t>>> lists;
      for (uint32_t b = 0; b < plan.nboxes(l); ++b) {
        const uint32_t fc = plan.first_child[l][b];
        if (fc == kNone) continue;
        cols.push_back(gid(l, b));
        std::vector<std::pair<uint16_t, uint32_t>> li;
        for (int o = 0; o < 8; ++o) li.push_back({uint16_t(o), gid(l + 1, fc + o)});
        lists.push_back(std::move(li));
        if (int(cols.size()) == C) {
          sb.add_task(cols, lists);
          cols.clear();
          lists.clear();
        }
      }
      if (!cols.empty()) sb.add_task(cols, lists);
      sb.finish(m2m_[l]);
    }
    if (l >= 1) {
      // L2L into the boxes of level l, grouped by octant so a task shares one table;
      // 16 columns per task (waves split the rows) on small levels, 64 while 128-column
      // tasks would fill fewer than four rounds of the GPU, else 128
      const uint32_t nb = plan.nboxes(l);
      const int tuned = l < kTuneLevels ? opt_.tuning.l2l_width[l] : 0;
      require(tuned == 0 || tuned == 16 || tuned == 64 || tuned == dev::COLS, "tuning: L2L width 16, 64 or 128");
      const int width = tuned ? tuned : nb < 64u * dev::COLS ? 16 : (nb < 4u * 512u * dev::COLS ? 64 : dev::COLS);
      ScheduleBuilder sb(width);
      const int C = width;
      for (int o = 0; o < 8; ++o) {
        std::vector<uint32_t> cols;
        std::vector<std::vector<std::pair<uint16_t, uint32_t>>> lists;
        for (uint32_t b = 0; b < plan.nboxes(l); ++b) {
          if (octant_of(plan.keys[l][b]) != o) continue;
          cols.push_back(gid(l, b));
          lists.push_back({{uint16_t(o), gid(l - 1, plan.parent[l][b])}});
          if (int(cols.size()) == C) {
            sb.add_task(cols, lists);
            cols.clear();
            lists.clear();
          }
        }
        if (!cols.empty()) sb.add_task(cols, lists);
      }
      sb.finish(l2l_[l]);
    }
  }
  // M2L of every level in one schedule.  Dense levels: tiles of COLS targets of one
  // octant, whose groups are the offsets their parity allows.  Sparse levels (most of
  // a tile's column slots without a source) and small ones: clustered mode (below),
  // tasks writing scratch rows, then an ordered reduction per target.
  ScheduleBuilder all;
  std::vector<uint32_t> red_id, red_cell, red_off{0}, red_ent;
  m2l_level_task_.assign(nlevels_ + 1, 0);
  m2l_red_level_.assign(nlevels_ + 1, 0);
  nslots_ = 0;
  for (int l = 0; l < nlevels_; ++l) {
    if (l >= 2 && opt_.m2l != kM2lDense) {
      // the rotation and plane-wave M2L have their own rows (build_rotation); count the pairs only
      all.pairs += plan.v[l].size();
    } else if (l >= 2) {
      const Csr &V = plan.v[l];
      // Tasks of the eight octant classes interleaved (task k of class 0, of class 1, ...),
      // and each task's groups ordered by (parent-level neighbour D, source octant o'),
      // d = 2 D + o' - o: the k-th tasks of the eight classes hold the children of the
      // same parents, so at the same group position they read the same source boxes,
      // and the launch keeps those eight tasks on one XCD (one L2).
      ScheduleBuilder tile;
      std::vector<std::vector<uint32_t>> cls_cols(8);
      std::vector<std::vector<std::vector<std::pair<uint16_t, uint32_t>>>> cls_lists(8);
      for (uint32_t b = 0; b < plan.nboxes(l); ++b) {
        if (V.off[b + 1] == V.off[b]) continue;
        const int o = octant_of(plan.keys[l][b]);
        cls_cols[o].push_back(gid(l, b));
        std::vector<std::pair<uint16_t, uint32_t>> li;
        for (uint32_t e = V.off[b]; e < V.off[b + 1]; ++e) li.push_back({V.aux[e], gid(l, V.ent[e])});
        cls_lists[o].push_back(std::move(li));
      }
      size_t most = 0;
      for (int o = 0; o < 8; ++o) most = std::max(most, cls_cols[o].size());
      for (size_t at = 0; at < most; at += size_t(C))
        for (int o = 0; o < 8; ++o) {
          if (at >= cls_cols[o].size()) continue;
          const size_t end = std::min(cls_cols[o].size(), at + size_t(C));
          std::vector<uint32_t> cols(cls_cols[o].begin() + at, cls_cols[o].begin() + end);
          std::vector<std::vector<std::pair<uint16_t, uint32_t>>> lists(cls_lists[o].begin() + at,
                                                                          cls_lists[o].begin() + end);
          tile.add_task(cols, lists, m2l_rank_[o].data());
        }
      // Box mode needs both a dense level and enough tasks to fill the GPU: a task walks
      // all of its ~189 groups in sequence, so a level of a few dozen tasks would be one
      // long chain per CU.
      constexpr uint32_t kMinBoxTasks = 256;
      const int tuned = l < kTuneLevels ? opt_.tuning.m2l_mode[l] : 0;
      require(tuned >= 0 && tuned <= 2, "tuning: M2L mode 0, 1 or 2");
      const bool box = tile.pairs == 0 || (tuned ? tuned == 1
                                                 : (2 * tile.pairs >= tile.slots && tile.ntasks() >= kMinBoxTasks));
      if (l < kTuneLevels) m2l_mode_taken_[l] = tile.pairs == 0 ? 0 : box ? 1 : 2;
      if (box) {
        all.append(tile);
      } else {
        // Sparse level (few of a target's possible offsets present), or a small one: clustered mode.
        // A target's V entries fall into the children of its parent's neighbours D;
        // each (target, D) gets a scratch slot, and a task holds the slots of one
        // (target octant o, D) class, whose columns share the same offsets d = 2 D + o'
        // - o (at most 8 groups, one per source octant o').  The slots of a target are
        // then added in ascending D, so every local adds its V terms in rank order.
        // Classes of the same D and chunk position are adjacent for all eight o (the
        // children of the same parents, reading the sources of the same neighbour cells
        // on one XCD).
        struct Member {
          uint32_t slot;
          std::vector<std::pair<uint16_t, uint32_t>> list;
        };
        std::vector<std::vector<Member>> cls(8 * 27);  // [D * 8 + o]
        for (uint32_t b = 0; b < plan.nboxes(l); ++b) {
          if (V.off[b + 1] == V.off[b]) continue;
          const int o = octant_of(plan.keys[l][b]);
          std::map<int, std::vector<std::pair<uint16_t, uint32_t>>> by_D;
          for (uint32_t e = V.off[b]; e < V.off[b + 1]; ++e) {
            const int D = m2l_rank_[o][V.aux[e]] / 8;
            require(D >= 0 && D < 27, "M2L schedule: V offset outside the parent's neighbours", kUnsupportedLayout);
            by_D[D].push_back({V.aux[e], gid(l, V.ent[e])});
          }
          red_id.push_back(gid(l, b));
          for (auto &kv : by_D) {
            const uint32_t slot = static_cast<uint32_t>(nbox_total_ + nslots_++);
            red_ent.push_back(slot);
            cls[size_t(kv.first) * 8 + o].push_back({slot, std::move(kv.second)});
          }
          red_off.push_back(static_cast<uint32_t>(red_ent.size()));
        }
        ScheduleBuilder clustered;
        for (int D = 0; D < 27; ++D) {
          size_t most = 0;
          for (int o = 0; o < 8; ++o) most = std::max(most, cls[size_t(D) * 8 + o].size());
          for (size_t at = 0; at < most; at += size_t(C))
            for (int o = 0; o < 8; ++o) {
              const auto &mem = cls[size_t(D) * 8 + o];
              if (at >= mem.size()) continue;
              std::vector<uint32_t> cols;
              std::vector<std::vector<std::pair<uint16_t, uint32_t>>> lists;
              for (size_t k = at; k < mem.size() && k < at + size_t(C); ++k) {
                cols.push_back(mem[k].slot);
                lists.push_back(mem[k].list);
              }
              clustered.add_task(cols, lists, m2l_rank_[o].data());
            }
        }
        all.append(clustered);
      }
    }
    m2l_level_task_[l + 1] = all.ntasks();
    m2l_red_level_[l + 1] = static_cast<uint32_t>(red_id.size());
  }
  all.finish(m2l_);
  upload_rows(m2l_red_, red_id, red_cell, red_off, red_ent);
}

// Plane-wave M2L by merge and shift (fmm-kernels/src/plane_wave.rs:1-60,
// fmm-tables/src/plane_wave/geometry.rs): the quadrature of degree p, the accumulators and
// shifts (the same for every parent), and per level the parents of its boxes with their
// 216 slots and 8 children, read off the V lists: every V entry of a child is a slot of
// its parent, and every pair (slot, child) that is not adjacent is a V entry when the slot
// is held (checked here).
void DeviceOperator::build_plane_wave(const Plan &plan) {
  const PlaneWave pw = ::fmmh::build_plane_wave(opt_.p, 0);
  require(pw.nodes <= dev::kPwMaxNodes, "plane wave: too many quadrature nodes", kUnsupportedLayout);
  const size_t lds = dev::pw_local_lds(opt_.p, pw.H, pw.nodes);
  require(lds <= 160 * 1024, "plane wave: the quadrature does not fit in LDS", kUnsupportedLayout);
  hip_check(dev::pw_allow_lds(lds), "plane wave LDS");
  pw_first_.upload(pw.first);
  pw_nalpha_.upload(pw.nalpha);
  pw_fourier_.upload(pw.fourier);
  pw_opow_.upload(pw.opow);
  pw_ipow_.upload(pw.ipow);
  pw_hw_.upload(pw.hw);
  pw_hc_.upload(pw.hc);
  pw_hs_.upload(pw.hs);
  pw_hk_.upload(pw.hk);
  pw_fwd_.upload(pw.fwd);
  pw_bwd_.upload(pw.bwd);
  // the (node, order) sums of the local step, by decreasing alpha count
  std::vector<int> gitem;
  for (int k = 0; k < pw.nodes; ++k)
    for (int m = 0; m < pw.fourier[k]; ++m) gitem.push_back(k * (opt_.p + 1) + m);
  std::stable_sort(gitem.begin(), gitem.end(),
                   [&](int x, int y) { return pw.nalpha[x / (opt_.p + 1)] > pw.nalpha[y / (opt_.p + 1)]; });
  pw_gitem_.upload(gitem);
  // the merge's shifts: one table per distinct source shift, the child shifts
  const size_t H2 = 2 * size_t(pw.H);
  const PwMergeGeometry &geo = plane_wave_merge_geometry();
  static_assert(kPwAccStride == dev::kPwAccStride, "one accumulator stride");
  std::vector<double> src_shift(geo.shift_d.size() * H2), child_shift(size_t(kPwDirections) * kPwChildren * H2);
  for (size_t i = 0; i < geo.shift_d.size(); ++i)
    plane_wave_shift(pw, geo.shift_dir[i], geo.shift_d[i].data(), &src_shift[i * H2]);
  for (int dir = 0; dir < kPwDirections; ++dir)
    for (int c = 0; c < kPwChildren; ++c) {
      const double d[3] = {((c >> 2) & 1) - 0.5, ((c >> 1) & 1) - 0.5, (c & 1) - 0.5};
      plane_wave_shift(pw, dir, d, &child_shift[(size_t(kPwChildren) * dir + c) * H2]);
    }
  pw_acc_first_.upload(geo.acc_first);
  pw_acc_mask_.upload(geo.acc_mask);
  pw_src_shift_.upload(src_shift);
  pw_child_shift_.upload(child_shift);
  // parents by the level of their children, children in box order
  const PwMergeLists lists = plane_wave_merge_lists(plan, box_base_, std::vector<char>(nlevels_, 1));
  pw_par_level_ = lists.par_level;
  pw_tgt_level_ = lists.tgt_level;
  pw_pe_first_.upload(lists.pe_first);
  pw_pe_.upload(lists.pe);
  pw_children_.upload(lists.children);
  pw_targets_.upload(lists.targets);
  const uint32_t box0 = nlevels_ > 2 ? static_cast<uint32_t>(box_base_[2]) : 0;
  pw_box0_ = box0;
  pw_nbox_ = nlevels_ > 2 ? static_cast<uint32_t>(box_base_[nlevels_] - box_base_[2]) : 0;
  pw_waves_.alloc(size_t(pw_nbox_) * 12 * pw.H);
  pw_inc_.alloc(size_t(pw_nbox_) * 12 * pw.H);
  pw_ = dev::PwArgs{opt_.p,         pw.nodes,     pw.H,       rotation_block_size(opt_.p), pw_first_.ptr,
                    pw_nalpha_.ptr, pw_fourier_.ptr, pw_opow_.ptr, pw_ipow_.ptr,              pw_hw_.ptr,
                    pw_hc_.ptr,     pw_hs_.ptr,   pw_hk_.ptr, pw_fwd_.ptr,                 pw_bwd_.ptr,
                    static_cast<int>(gitem.size()), pw_gitem_.ptr};
  pw_merge_ = dev::PwMerge{pw_acc_first_.ptr, pw_acc_mask_.ptr, pw_src_shift_.ptr, pw_child_shift_.ptr};
}

void DeviceOperator::build_rotation(const Plan &plan) {
  if (opt_.m2l == kM2lRotation) {
  std::vector<double> fwd, bwd, coax, dist2;
  std::vector<uint16_t> dist_of;
  build_rotation_tables(opt_.p, fwd, bwd, coax, dist_of, dist2);
  rot_rb_ = rotation_block_size(opt_.p);
  rot_cb_ = coaxial_block_size(opt_.p);
  if (opt_.precision == kMixed) {
    rot_fwd32_.upload(std::vector<float>(fwd.begin(), fwd.end()));
    rot_bwd32_.upload(std::vector<float>(bwd.begin(), bwd.end()));
    rot_coax32_.upload(std::vector<float>(coax.begin(), coax.end()));
  } else {
    rot_fwd_.upload(fwd);
    rot_bwd_.upload(bwd);
    rot_coax_.upload(coax);
  }
  rot_dist_.upload(dist_of);
  }
  // rows: target boxes with V entries, level by level; entries in row order
  std::vector<uint32_t> box, off{0}, src;
  std::vector<uint16_t> ofs;
  rot_level_.assign(nlevels_ + 1, 0);
  for (int l = 0; l < nlevels_; ++l) {
    const Csr &V = plan.v[l];
    for (uint32_t b = 0; l >= 2 && b < plan.nboxes(l); ++b) {
      if (V.off[b + 1] == V.off[b]) continue;
      box.push_back(static_cast<uint32_t>(box_base_[l] + b));
      for (uint32_t e = V.off[b]; e < V.off[b + 1]; ++e) {
        src.push_back(static_cast<uint32_t>(box_base_[l] + V.ent[e]));
        ofs.push_back(V.aux[e]);
      }
      off.push_back(static_cast<uint32_t>(src.size()));
    }
    rot_level_[l + 1] = static_cast<uint32_t>(box.size());
  }
  rot_box_.upload(box);
  rot_off_.upload(off);
  rot_src_.upload(src);
  rot_offset_.upload(ofs);
  if (opt_.m2l != kM2lRotation) return;
  // pairs by level, then offset, then row: blocks of at most pb pairs of one offset
  const bool mixed = opt_.precision == kMixed;
  const int pb = mixed ? dev::rotation_pairs_per_block<float>(opt_.p, rot_rb_, rot_cb_)
                       : dev::rotation_pairs_per_block<double>(opt_.p, rot_rb_, rot_cb_);
  hip_check(mixed ? dev::rotation_allow_lds<float>(opt_.p, rot_rb_, rot_cb_)
                  : dev::rotation_allow_lds<double>(opt_.p, rot_rb_, rot_cb_),
            "rotation M2L LDS");
  std::vector<uint32_t> blk_first, pair_src, pair_of(src.size());
  std::vector<uint16_t> blk_offset;
  rot_blk_level_.assign(nlevels_ + 1, 0);
  uint32_t npairs = 0;
  for (int l = 0; l < nlevels_; ++l) {
    std::vector<std::vector<uint32_t>> by_t(316);
    for (uint32_t r = rot_level_[l]; r < rot_level_[l + 1]; ++r)
      for (uint32_t e = off[r]; e < off[r + 1]; ++e) by_t[ofs[e]].push_back(e);
    for (int t = 0; t < 316; ++t)
      for (size_t k = 0; k < by_t[t].size(); ++k) {
        if (k % size_t(pb) == 0) {
          blk_first.push_back(npairs);
          blk_offset.push_back(static_cast<uint16_t>(t));
        }
        const uint32_t e = by_t[t][k];
        pair_src.push_back(src[e]);
        pair_of[e] = npairs++;
      }
    rot_blk_level_[l + 1] = static_cast<uint32_t>(blk_offset.size());
  }
  blk_first.push_back(npairs);
  rot_blk_first_.upload(blk_first);
  rot_blk_offset_.upload(blk_offset);
  rot_pair_src_.upload(pair_src);
  // chunks of whole blocks within each level, at most chunk_pairs pairs of scratch rows
  const size_t chunk_pairs =
      std::max<size_t>(size_t(pb), dev::kRotScratchBytes / (size_t(nc_) * (mixed ? sizeof(float) : sizeof(double))));
  rot_chunk_blk_.clear();
  rot_chunk_level_.assign(nlevels_ + 1, 0);
  for (int l = 0; l < nlevels_; ++l) {
    rot_chunk_level_[l] = static_cast<uint32_t>(rot_chunk_blk_.size());
    for (uint32_t bk = rot_blk_level_[l]; bk < rot_blk_level_[l + 1]; ++bk)
      if (bk == rot_blk_level_[l] || blk_first[bk + 1] - blk_first[rot_chunk_blk_.back()] > chunk_pairs)
        rot_chunk_blk_.push_back(bk);
  }
  rot_chunk_level_[nlevels_] = static_cast<uint32_t>(rot_chunk_blk_.size());
  const size_t nchunks = rot_chunk_blk_.size();
  rot_chunk_blk_.push_back(static_cast<uint32_t>(blk_offset.size()));
  rot_chunk_q_.assign(nchunks + 1, npairs);
  for (size_t c = 0; c < nchunks; ++c) rot_chunk_q_[c] = blk_first[rot_chunk_blk_[c]];
  // per chunk, its targets' rows (level by level, row order) with that chunk's pairs in V
  // order, relative to the chunk's first pair
  std::vector<std::vector<uint32_t>> cid(nchunks), coff(nchunks), cent(nchunks);
  for (size_t c = 0; c < nchunks; ++c) coff[c].push_back(0);
  for (uint32_t r = 0; r + 1 < off.size(); ++r) {
    std::map<size_t, std::vector<uint32_t>> by_chunk;
    for (uint32_t e = off[r]; e < off[r + 1]; ++e) {
      const uint32_t q = pair_of[e];
      const size_t c = size_t(std::upper_bound(rot_chunk_q_.begin(), rot_chunk_q_.begin() + nchunks, q) -
                              rot_chunk_q_.begin()) - 1;
      by_chunk[c].push_back(q - rot_chunk_q_[c]);
    }
    for (const auto &kv : by_chunk) {
      cid[kv.first].push_back(box[r]);
      cent[kv.first].insert(cent[kv.first].end(), kv.second.begin(), kv.second.end());
      coff[kv.first].push_back(static_cast<uint32_t>(cent[kv.first].size()));
    }
  }
  std::vector<uint32_t> red_id, red_off{0}, red_ent;
  rot_chunk_row_.assign(1, 0);
  size_t most = 0;
  for (size_t c = 0; c < nchunks; ++c) {
    const uint32_t base = static_cast<uint32_t>(red_ent.size());
    red_id.insert(red_id.end(), cid[c].begin(), cid[c].end());
    for (size_t k = 1; k < coff[c].size(); ++k) red_off.push_back(base + coff[c][k]);
    red_ent.insert(red_ent.end(), cent[c].begin(), cent[c].end());
    rot_chunk_row_.push_back(static_cast<uint32_t>(red_id.size()));
    most = std::max<size_t>(most, rot_chunk_q_[c + 1] - rot_chunk_q_[c]);
  }
  upload_rows(rot_red_, red_id, {}, red_off, red_ent);
  if (mixed) {
    rot_slots32_.alloc(most * nc_);
    mult32_.alloc(nbox_total_ * nc_);
  } else {
    rot_slots_.alloc(most * nc_);
  }
}

size_t DeviceOperator::device_bytes() const {
  return mult_.bytes() + loc_.bytes() + slots_.bytes() + slots32_.bytes() + t_m2l_.bytes() + t_m2m_.bytes() +
         t_l2l_.bytes() + sx_.bytes() * 8 + src_rec_.bytes() + rot_slots_.bytes() + rot_slots32_.bytes() +
         mult32_.bytes() + pw_waves_.bytes() + pw_inc_.bytes() + pw_src_shift_.bytes() + pw_child_shift_.bytes();
}

void DeviceOperator::launch_contract(const Schedule &s, uint32_t task_first, uint32_t ntasks, const double *X,
                                     double *Y, const DevBuf<double> &tables, uint32_t xchunk) {
  dev::contract_f64(stream_, opt_.p, s.width, ntasks, s.task_group_off.ptr + task_first, s.group_table.ptr,
                    s.group_src.ptr, s.col_out.ptr + size_t(task_first) * s.width, tables.ptr, X, Y,
                    static_cast<uint32_t>(nbox_total_), slots_.ptr, xchunk);
}

// M2L of the levels level_first .. level_last: the contraction tasks, then the ordered
// reduction of the clustered-mode levels.
void DeviceOperator::launch_m2l(int level_first, int level_last) {
  if (opt_.m2l == kM2lPlaneWave) {
    const uint32_t q0 = pw_par_level_[level_first], q1 = pw_par_level_[level_last + 1];
    if (q1 == q0) return;
    // the waves of the boxes of these levels, the parents' merge and shift into their
    // children's incoming waves, then the children's locals
    const uint32_t b0 = box_base_[level_first], b1 = box_base_[level_last + 1];
    const uint32_t t0 = pw_tgt_level_[level_first], t1 = pw_tgt_level_[level_last + 1];
    dev::pw_outgoing(stream_, pw_, b1 - b0, b0, mult_.ptr, pw_waves_.ptr + size_t(b0 - pw_box0_) * 12 * pw_.H);
    timed(kKindM2lContract, [&] {
      dev::pw_merge(stream_, pw_, pw_merge_, q1 - q0, pw_pe_first_.ptr + size_t(q0) * 6 * dev::kPwAccStride,
                    pw_pe_.ptr, pw_children_.ptr + size_t(q0) * kPwChildren, pw_box0_, pw_waves_.ptr, pw_inc_.ptr);
    });
    dev::pw_local(stream_, pw_, t1 - t0, pw_targets_.ptr + t0, pw_box0_, pw_inc_.ptr, loc_.ptr);
    return;
  }
  if (opt_.m2l == kM2lRotation) {
    if (opt_.precision == kMixed) {
      // the multipoles of these levels in f32, read by every pair that uses them
      const size_t x0 = box_base_[level_first], x1 = box_base_[level_last + 1];
      dev::to_f32(stream_, (x1 - x0) * nc_, mult_.ptr + x0 * nc_, mult32_.ptr + x0 * nc_);
    }
    // chunk by chunk: the pairs' rows, then each target adds that chunk's rows in the
    // order of its V list
    for (uint32_t c = rot_chunk_level_[level_first]; c < rot_chunk_level_[level_last + 1]; ++c) {
      const uint32_t c0 = rot_chunk_blk_[c], c1 = rot_chunk_blk_[c + 1], q0 = rot_chunk_q_[c];
      timed(kKindM2lContract, [&] {
        if (opt_.precision == kMixed)
          dev::m2l_rotation<float>(stream_, opt_.p, c1 - c0, q0, rot_blk_first_.ptr + c0, rot_blk_offset_.ptr + c0,
                                   rot_pair_src_.ptr, rot_dist_.ptr, rot_fwd32_.ptr, rot_bwd32_.ptr, rot_coax32_.ptr,
                                   rot_rb_, rot_cb_, mult32_.ptr, rot_slots32_.ptr);
        else
          dev::m2l_rotation<double>(stream_, opt_.p, c1 - c0, q0, rot_blk_first_.ptr + c0, rot_blk_offset_.ptr + c0,
                                    rot_pair_src_.ptr, rot_dist_.ptr, rot_fwd_.ptr, rot_bwd_.ptr, rot_coax_.ptr,
                                    rot_rb_, rot_cb_, mult_.ptr, rot_slots_.ptr);
      });
      const uint32_t w0 = rot_chunk_row_[c], w1 = rot_chunk_row_[c + 1];
      dev::reduce_slots(stream_, opt_.p, w1 - w0, rot_red_.row_id.ptr + w0, rot_red_.row_off.ptr + w0,
                        rot_red_.row_ent.ptr, 0, rot_slots_.ptr,
                        opt_.precision == kMixed ? rot_slots32_.ptr : nullptr, loc_.ptr);
    }
    return;
  }
  const uint32_t t0 = m2l_level_task_[level_first], t1 = m2l_level_task_[level_last + 1];
  if (t1 > t0) {
    if (opt_.precision == kMixed) {
      const size_t b0 = box_base_[level_first], b1 = box_base_[level_last + 1];
      const size_t sz = dev::split_expansion_size(opt_.p);
      dev::split_bf16(stream_, opt_.p, b1 - b0, mult_.ptr + b0 * nc_, mult16_.ptr + b0 * sz);
      timed(kKindM2lContract, [&] {
        dev::contract_bf16(stream_, opt_.p, t1 - t0, m2l_.task_group_off.ptr + t0, m2l_.group_table.ptr,
                           m2l_.group_src.ptr, m2l_.col_out.ptr + size_t(t0) * dev::COLS, t_m2l16_.ptr,
                           mult16_.ptr, loc_.ptr, static_cast<uint32_t>(nbox_total_), slots32_.ptr, kM2lXcdRun,
                           opt_.tuning.m2l_stage);
      });
    } else {
      timed(kKindM2lContract, [&] { launch_contract(m2l_, t0, t1 - t0, mult_.ptr, loc_.ptr, t_m2l_, kM2lXcdRun); });
    }
  }
  const uint32_t r0 = m2l_red_level_[level_first], r1 = m2l_red_level_[level_last + 1];
  if (r1 > r0)
    dev::reduce_slots(stream_, opt_.p, r1 - r0, m2l_red_.row_id.ptr + r0, m2l_red_.row_off.ptr + r0,
                      m2l_red_.row_ent.ptr, static_cast<uint32_t>(nbox_total_), slots_.ptr, slots32_.ptr, loc_.ptr);
}

void DeviceOperator::launch_p2l(int level_first, int level_last) {
  const uint32_t r0 = p2l_level_[level_first], r1 = p2l_level_[level_last + 1];
  dev::expand(stream_, opt_.p, 1, r1 - r0, p2l_rows_.row_id.ptr + r0, p2l_rows_.row_cell.ptr + 4 * size_t(r0),
              p2l_rows_.row_off.ptr + r0, p2l_rows_.row_ent.ptr, leaf_cell_.ptr, src_start_.ptr, sx_.ptr, sy_.ptr,
              sz_.ptr, q_.ptr, loc_.ptr);
}

void DeviceOperator::load_charges(const double *charges, bool on_device) {
  if (ns_ == 0) return;
  const double *src = charges;
  if (!on_device) {
    hip_check(hipMemcpyAsync(q_in_.ptr, charges, ns_ * sizeof(double), hipMemcpyHostToDevice, stream_), "charges");
    src = q_in_.ptr;
  }
  dev::gather_charges(stream_, src, src_order_.ptr, ns_, q_.ptr, src_rec_.count ? src_rec_.ptr : nullptr);
}

void DeviceOperator::reset() {
  dev::fill(stream_, mult_.ptr, mult_.count, 0.0);
  dev::fill(stream_, loc_.ptr, loc_.count, 0.0);
  dev::fill(stream_, phi_hat_.ptr, phi_hat_.count, 0.0);
  dev::fill(stream_, grad_hat_.ptr, grad_hat_.count, 0.0);
}

void DeviceOperator::p2m(int level) {
  if (level < 0 || level >= nlevels_) return;
  timed(0, [&] {
    const uint32_t r0 = p2m_level_[level], r1 = p2m_level_[level + 1];
    dev::expand(stream_, opt_.p, 0, r1 - r0, p2m_rows_.row_id.ptr + r0, p2m_rows_.row_cell.ptr + 4 * size_t(r0),
                p2m_rows_.row_off.ptr + r0, p2m_rows_.row_ent.ptr, leaf_cell_.ptr, src_start_.ptr, sx_.ptr, sy_.ptr,
                sz_.ptr, q_.ptr, mult_.ptr);
  });
}

void DeviceOperator::m2m(int level) {
  if (level < 0 || level + 1 >= nlevels_) return;
  timed(1, [&] {
    const Schedule &sc = m2m_[level];
    dev::m2m_split(stream_, opt_.p, sc.ntasks, sc.task_group_off.ptr, sc.group_table.ptr, sc.group_src.ptr,
                   sc.col_out.ptr, t_m2m_.ptr, mult_.ptr, mult_.ptr,
                   level < kTuneLevels ? opt_.tuning.m2m_form[level] : 0);
  });
}

void DeviceOperator::m2l(int level) {
  if (level < 2 || level >= nlevels_) return;
  timed(3, [&] {
    launch_m2l(level, level);
  });
}

void DeviceOperator::p2l(int level) {
  if (level < 1 || level >= nlevels_) return;
  timed(4, [&] { launch_p2l(level, level); });
}

void DeviceOperator::l2l(int level) {
  if (level < 1 || level >= nlevels_) return;
  timed(2, [&] {
    launch_contract(l2l_[level], 0, l2l_[level].ntasks, loc_.ptr, loc_.ptr, t_l2l_);
  });
}

void DeviceOperator::l2p(int level) {
  if (level < 0 || level >= nlevels_) return;
  timed(5, [&] {
    const uint32_t r0 = l2p_level_[level], r1 = l2p_level_[level + 1];
    dev::evaluate(stream_, opt_.p, 0, opt_.gradients, r1 - r0, l2p_rows_.row_id.ptr + r0, l2p_rows_.row_off.ptr + r0,
                  l2p_rows_.row_ent.ptr, box_cell_.ptr, leaf_cell_.ptr, tgt_start_.ptr, tx_.ptr, ty_.ptr, tz_.ptr,
                  loc_.ptr, phi_hat_.ptr, grad_hat_.ptr);
  });
}

void DeviceOperator::m2p(int level) {
  if (level < 0 || level >= nlevels_) return;
  timed(6, [&] {
    const uint32_t r0 = m2p_level_[level], r1 = m2p_level_[level + 1];
    dev::evaluate(stream_, opt_.p, 1, opt_.gradients, r1 - r0, m2p_rows_.row_id.ptr + r0, m2p_rows_.row_off.ptr + r0,
                  m2p_rows_.row_ent.ptr, box_cell_.ptr, leaf_cell_.ptr, tgt_start_.ptr, tx_.ptr, ty_.ptr, tz_.ptr,
                  mult_.ptr, phi_hat_.ptr, grad_hat_.ptr);
  });
}

void DeviceOperator::launch_p2p(uint32_t leaf_first, uint32_t count) {
  const uint32_t umin = uint32_t(p2p_uniform_min());
  if (opt_.precision == kMixed)
    dev::p2p<float>(stream_, leaf_first, count, opt_.gradients, umin, near_off_.ptr, near_ent_.ptr, leaf_cell_.ptr,
                    src_start_.ptr, tgt_start_.ptr, sx_.ptr, sy_.ptr, sz_.ptr, q_.ptr, src_rec_.ptr,
                    near_close_.ptr, tx_.ptr, ty_.ptr, tz_.ptr, phi_hat_.ptr, grad_hat_.ptr);
  else
    dev::p2p<double>(stream_, leaf_first, count, opt_.gradients, umin, near_off_.ptr, near_ent_.ptr,
                     leaf_cell_.ptr, src_start_.ptr, tgt_start_.ptr, sx_.ptr, sy_.ptr, sz_.ptr, q_.ptr, nullptr,
                     nullptr, tx_.ptr, ty_.ptr, tz_.ptr, phi_hat_.ptr, grad_hat_.ptr);
}

void DeviceOperator::p2p(int level) {
  if (level < 0 || level >= nlevels_) return;
  timed(7, [&] {
    const uint32_t l0 = leaf_off_[level], l1 = leaf_off_[level + 1];
    launch_p2p(l0, l1 - l0);
  });
}

void DeviceOperator::evaluate() {
  const int L = nlevels_ - 1;
  // upward: P2M of every leaf (leaves have no M2M rows), then M2M from the deepest
  // parents up (the per-level order p2m(l), m2m(l-1) gives the same sums)
  timed(0, [&] {
    dev::expand(stream_, opt_.p, 0, p2m_rows_.rows, p2m_rows_.row_id.ptr, p2m_rows_.row_cell.ptr,
                p2m_rows_.row_off.ptr, p2m_rows_.row_ent.ptr, leaf_cell_.ptr, src_start_.ptr, sx_.ptr, sy_.ptr, sz_.ptr,
                q_.ptr, mult_.ptr);
  });
  for (int l = L - 1; l >= 0; --l) m2m(l);
  // downward: M2L and P2L read only multipoles and points, so every level's M2L is one
  // launch and every level's P2L another; then the L2L chain from the root down.  Each
  // local receives its M2L, P2L and L2L terms, each kind in the order of its row.
  if (L >= 2) timed(3, [&] { launch_m2l(2, L); });
  if (L >= 1) timed(4, [&] { launch_p2l(1, L); });
  for (int l = 1; l <= L; ++l) l2l(l);
  // leaves: each kind over every level at once; a target still receives L2P, M2P and
  // P2P contributions in that order
  timed(5, [&] {
    dev::evaluate(stream_, opt_.p, 0, opt_.gradients, l2p_rows_.rows, l2p_rows_.row_id.ptr, l2p_rows_.row_off.ptr,
                  l2p_rows_.row_ent.ptr, box_cell_.ptr, leaf_cell_.ptr, tgt_start_.ptr, tx_.ptr, ty_.ptr, tz_.ptr,
                  loc_.ptr, phi_hat_.ptr, grad_hat_.ptr);
  });
  timed(6, [&] {
    dev::evaluate(stream_, opt_.p, 1, opt_.gradients, m2p_rows_.rows, m2p_rows_.row_id.ptr, m2p_rows_.row_off.ptr,
                  m2p_rows_.row_ent.ptr, box_cell_.ptr, leaf_cell_.ptr, tgt_start_.ptr, tx_.ptr, ty_.ptr, tz_.ptr,
                  mult_.ptr, phi_hat_.ptr, grad_hat_.ptr);
  });
  timed(7, [&] { launch_p2p(0, nleaves_); });
}

void DeviceOperator::evaluate_graph() {
  require(!timing_, "evaluate_graph: kind timing must be off (events are not captured)", kUnsupportedLayout);
  if (!graph_exec_) {
    // every buffer is fixed at build and every clear is a kernel, so the captured
    // launches stay valid for the operator's lifetime
    hip_check(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    reset();
    evaluate();
    hip_check(hipStreamEndCapture(stream_, &graph_), "hipStreamEndCapture");
    hip_check(hipGraphInstantiate(&graph_exec_, graph_, nullptr, nullptr, 0), "hipGraphInstantiate");
  }
  hip_check(hipGraphLaunch(graph_exec_, stream_), "hipGraphLaunch");
}

void DeviceOperator::read_output(double *phi, double *grad, bool on_device) {
  dev::output_pass(stream_, nleaves_, tgt_start_.ptr, leaf_r_.ptr, tgt_order_.ptr, phi_hat_.ptr,
                   grad_hat_.ptr, on_device ? phi : phi_.ptr, on_device ? grad : grad_.ptr,
                   opt_.gradients && grad != nullptr);
  if (!on_device) {
    hip_check(hipMemcpyAsync(phi, phi_.ptr, nt_ * sizeof(double), hipMemcpyDeviceToHost, stream_), "download phi");
    if (opt_.gradients && grad)
      hip_check(hipMemcpyAsync(grad, grad_.ptr, 3 * nt_ * sizeof(double), hipMemcpyDeviceToHost, stream_),
                "download grad");
  }
  sync();
}

void DeviceOperator::vjp(const double *q, const double *g, const double *field, bool on_device, double *grad_q,
                         double *grad_pos) {
  require(same_points_, "vjp: the adjoint is implemented for sources = targets", kUnsupportedLayout);
  require(opt_.gradients, "vjp: the operator must compute gradients", kUnsupportedLayout);
  const size_t n = size_t(nt_);
  if (n == 0) return;
  const double *dq = q, *dg = g, *df = field;
  double *oq = grad_q, *op = grad_pos;
  if (!on_device) {
    if (vjp_in_.count == 0) vjp_in_.alloc(5 * n);
    if (vjp_out_.count == 0) vjp_out_.alloc(4 * n);
    hip_check(hipMemcpyAsync(vjp_in_.ptr, q, n * sizeof(double), hipMemcpyHostToDevice, stream_), "vjp q");
    hip_check(hipMemcpyAsync(vjp_in_.ptr + n, g, n * sizeof(double), hipMemcpyHostToDevice, stream_), "vjp g");
    if (grad_pos)
      hip_check(hipMemcpyAsync(vjp_in_.ptr + 2 * n, field, 3 * n * sizeof(double), hipMemcpyHostToDevice, stream_),
                "vjp field");
    dq = vjp_in_.ptr;
    dg = vjp_in_.ptr + n;
    df = vjp_in_.ptr + 2 * n;
    oq = vjp_out_.ptr;
    op = grad_pos ? vjp_out_.ptr + n : nullptr;
  }
  // the adjoint field: charges g at the same points, potential and gradient
  load_charges(dg, true);
  reset();
  evaluate();
  dev::output_pass(stream_, nleaves_, tgt_start_.ptr, leaf_r_.ptr, tgt_order_.ptr, phi_hat_.ptr,
                   grad_hat_.ptr, phi_.ptr, grad_.ptr, 1);
  dev::vjp_combine(stream_, n, dq, dg, phi_.ptr, grad_.ptr, df, oq, op);
  if (!on_device) {
    hip_check(hipMemcpyAsync(grad_q, oq, n * sizeof(double), hipMemcpyDeviceToHost, stream_), "vjp grad_q");
    if (grad_pos)
      hip_check(hipMemcpyAsync(grad_pos, op, 3 * n * sizeof(double), hipMemcpyDeviceToHost, stream_),
                "vjp grad_pos");
  }
  sync();
}

void DeviceOperator::read_expansions(bool locals, int level, double *out) {
  require(level >= 0 && level < nlevels_, "read_expansions: level out of range");
  sync();
  const size_t n = size_t(box_base_[level + 1] - box_base_[level]) * nc_;
  const double *src = (locals ? loc_.ptr : mult_.ptr) + size_t(box_base_[level]) * nc_;
  hip_check(hipMemcpy(out, src, n * sizeof(double), hipMemcpyDeviceToHost), "read_expansions");
}

void DeviceOperator::sync() {
  hip_check(hipStreamSynchronize(stream_), "evaluation");
  hip_check(hipGetLastError(), "launch");
  if (!marks_.empty()) collect_timings();
}

void direct_sum(int device, const double *sources, const double *charges, int64_t ns, const double *targets,
                int64_t nt, double *phi, double *grad, bool on_device) {
  require(ns >= 0 && nt >= 0, "direct_sum: negative size");
  if (nt == 0) return;
  hip_check(hipSetDevice(device), "hipSetDevice");
  hipStream_t s = nullptr;
  hip_check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
  const int nch = dev::direct_chunks(ns, nt);
  DevBuf<double> part, in, out;
  part.alloc(size_t(nch) * nt * 8);
  const double *src = sources, *q = charges, *tgt = targets;
  double *ph = phi, *gr = grad;
  if (!on_device) {
    in.alloc(4 * size_t(ns) + 3 * size_t(nt));
    out.alloc(4 * size_t(nt));
    hip_check(hipMemcpyAsync(in.ptr, sources, 3 * ns * sizeof(double), hipMemcpyHostToDevice, s), "sources");
    hip_check(hipMemcpyAsync(in.ptr + 3 * ns, charges, ns * sizeof(double), hipMemcpyHostToDevice, s), "charges");
    hip_check(hipMemcpyAsync(in.ptr + 4 * ns, targets, 3 * nt * sizeof(double), hipMemcpyHostToDevice, s), "targets");
    src = in.ptr;
    q = in.ptr + 3 * ns;
    tgt = in.ptr + 4 * ns;
    ph = out.ptr;
    gr = out.ptr + nt;
  }
  dev::direct_sum(s, src, q, ns, tgt, nt, grad != nullptr, part.ptr, nch, ph, grad ? gr : nullptr);
  if (!on_device) {
    hip_check(hipMemcpyAsync(phi, ph, nt * sizeof(double), hipMemcpyDeviceToHost, s), "phi");
    if (grad) hip_check(hipMemcpyAsync(grad, gr, 3 * nt * sizeof(double), hipMemcpyDeviceToHost, s), "grad");
  }
  const hipError_t e = hipStreamSynchronize(s);
  (void)hipStreamDestroy(s);
  hip_check(e, "direct_sum");
}

}  // namespace fmmh
