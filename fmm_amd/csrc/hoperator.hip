// hoperator.hip -- the Helmholtz device operator (see hoperator.hpp).
//
// Box b holds three coefficient planes per expansion (vector ids 3b, 3b + 1, 3b + 2: real
// part, imaginary part, negated imaginary part).  A complex table T = Tr + i Ti is stored as
// the real tables Tr and Ti, and T x becomes, per output box, two columns of the grouped
// contraction with two groups each:
//   real column:  Tr <- plane 0 of the source,  Ti <- plane 2 (-Im)
//   imag column:  Tr <- plane 1 (Im),           Ti <- plane 0
// so every column of every group holds a source (no sign per column, no empty slots).  The
// negated planes are written by a small kernel once a level's expansions are complete.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <map>

#include "hkernels.hpp"
#include "hoperator.hpp"
#include "hpwtables.hpp"
#include "htables.hpp"
#include "pwtables.hpp"
#include "kernels.hpp"
#include "schedule.hpp"

namespace fmmh {

namespace {

// rank of offset d among the M2L groups of a target of octant o: (parent-level neighbour,
// source octant), as the Laplace operator orders them for L2 reuse; the plan's offsets
// are target minus source, so the source child sits at ob - d from the parent's corner
uint16_t m2l_rank(int oc, int d) {
  const int ob[3] = {(oc >> 2) & 1, (oc >> 1) & 1, oc & 1};
  int dv[3];
  offset_of_index(d, dv);
  int D = 0, op = 0;
  for (int a = 0; a < 3; ++a) {
    const int s = ob[a] - dv[a];
    const int bit = ((s % 2) + 2) % 2;
    D = 3 * D + ((s - bit) / 2 + 1);
    op = 2 * op + bit;
  }
  return static_cast<uint16_t>(8 * D + op);
}

bool all_zero(const double *v, size_t n) {
  for (size_t i = 0; i < n; ++i)
    if (v[i] != 0.0) return false;
  return true;
}

}  // namespace

HelmholtzOperator::HelmholtzOperator(const Plan &plan, const HelmholtzOptions &options) : opt_(options) {
  require(opt_.p >= 0 && opt_.p <= dev::kMaxDegree, "the device path supports 0 <= p <= 15", kUnsupportedLayout);
  require(opt_.precision == kPlainF64 || opt_.precision == kMixed, "unknown precision", kUnsupportedPrecision);
  require(opt_.m2l == kM2lDense || opt_.m2l == kM2lRotation || opt_.m2l == kM2lPlaneWave,
          "Helmholtz: M2L dense, rotation or plane wave", kUnsupportedLayout);
  require(std::isfinite(opt_.kre) && std::isfinite(opt_.kim) && (opt_.kre != 0.0 || opt_.kim != 0.0),
          "Helmholtz: the wavenumber must be finite and nonzero");
  require(opt_.kre >= 0.0 && opt_.kim >= 0.0, "Helmholtz: Re k >= 0 and Im k >= 0 (outgoing, not growing)");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) throw Error(kNoDevice, "no HIP device");
  require(opt_.device >= 0 && opt_.device < ndev, "device ordinal out of range", kNoDevice);
  hip_check(hipSetDevice(opt_.device), "hipSetDevice");
  hip_check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
  hip_check(hipEventCreateWithFlags(&order_event_, hipEventDisableTiming), "hipEventCreate");

  const int p = opt_.p;
  nlevels_ = plan.nlevels;
  nc_ = ncoef(p);
  ns_ = plan.ns;
  nt_ = plan.nt;
  nleaves_ = plan.nleaves();
  box_base_.assign(nlevels_ + 1, 0);
  for (int l = 0; l < nlevels_; ++l) box_base_[l + 1] = box_base_[l] + plan.nboxes(l);
  nbox_total_ = box_base_[nlevels_];
  require(3 * nbox_total_ < size_t(kNone), "Helmholtz: too many boxes", kTooLarge);
  auto gid = [&](int level, uint32_t b) { return box_base_[level] + b; };
  // one domain: every leaf's radius is its level's (a forest's clouds are not one domain)
  for (int l = 0; l < nlevels_; ++l)
    for (uint32_t lf = plan.leaf_off[l]; lf < plan.leaf_off[l + 1]; ++lf)
      require(plan.leaf_radius[lf] == plan.radius(l), "Helmholtz: a plan of one domain is required",
              kUnsupportedLayout);

  // level constants in each level's units
  const cplx k(opt_.kre, opt_.kim);
  std::vector<cplx> kh(nlevels_);
  std::vector<double> hlev(4 * size_t(std::max(nlevels_, 1)), 0.0);
  for (int l = 0; l < nlevels_; ++l) {
    kh[l] = k * plan.radius(l);
    hlev[4 * l] = kh[l].real();
    hlev[4 * l + 1] = kh[l].imag();
    hlev[4 * l + 2] = h_scale(kh[l]);
    hlev[4 * l + 3] = h_beta(kh[l]);
  }
  hlev_.upload(hlev);
  std::vector<double> norm;
  h_norm_factors(p, norm);
  norm_.upload(norm);

  // cells, points, near rows (as the Laplace operator)
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
  {
    std::vector<double> a(ns_), b(ns_), c(ns_);
    for (int64_t i = 0; i < ns_; ++i) {
      a[i] = plan.us[3 * i];
      b[i] = plan.us[3 * i + 1];
      c[i] = plan.us[3 * i + 2];
    }
    sx_.upload(a);
    sy_.upload(b);
    sz_.upload(c);
    if (opt_.precision == kMixed && ns_ > 0) {
      std::vector<float> rec(8 * size_t(ns_));
      dev::h_source_records(a.data(), b.data(), c.data(), size_t(ns_), rec.data());
      src_rec_.upload(rec);
    }
    a.resize(nt_);
    b.resize(nt_);
    c.resize(nt_);
    for (int64_t i = 0; i < nt_; ++i) {
      a[i] = plan.ut[3 * i];
      b[i] = plan.ut[3 * i + 1];
      c[i] = plan.ut[3 * i + 2];
    }
    tx_.upload(a);
    ty_.upload(b);
    tz_.upload(c);
  }
  qr_.alloc(ns_);
  qi_.alloc(ns_);
  q_in_.alloc(2 * ns_);
  // 64 doubles past the last row: the M2M's per-tile form reads a source row as four runs
  // of 4 ceil(nc / 16) values (at most 15 past its end) and masks the tail
  mult_.alloc(3 * nbox_total_ * nc_ + 64);
  loc_.alloc(3 * nbox_total_ * nc_);
  phi_hat_.alloc(2 * nt_);
  grad_hat_.alloc(opt_.gradients ? 6 * nt_ : 0);
  phi_.alloc(2 * nt_);
  grad_.alloc(opt_.gradients ? 6 * nt_ : 0);

  // leaf-kernel rows
  {
    std::vector<uint32_t> id, cell, off{0}, ent;
    for (int l = 0; l < nlevels_; ++l)
      for (uint32_t lf = plan.leaf_off[l]; lf < plan.leaf_off[l + 1]; ++lf) {
        if (plan.src_start[lf + 1] == plan.src_start[lf]) continue;
        id.push_back(gid(l, plan.leaf_box[lf]));
        cell.insert(cell.end(), {leaf_cell[4 * lf], leaf_cell[4 * lf + 1], leaf_cell[4 * lf + 2], leaf_cell[4 * lf + 3]});
        ent.push_back(lf);
        off.push_back(static_cast<uint32_t>(ent.size()));
      }
    upload_rows(p2m_rows_, id, cell, off, ent);
  }
  {
    std::vector<uint32_t> id, cell, off{0}, ent;
    for (int l = 0; l < nlevels_; ++l)
      for (uint32_t lf = plan.leaf_off[l]; lf < plan.leaf_off[l + 1]; ++lf) {
        if (plan.tgt_start[lf + 1] == plan.tgt_start[lf]) continue;
        id.push_back(lf);
        ent.push_back(gid(l, plan.leaf_box[lf]));
        off.push_back(static_cast<uint32_t>(ent.size()));
      }
    upload_rows(l2p_rows_, id, cell, off, ent);
  }
  {
    std::vector<uint32_t> id, cell, off{0}, ent;
    for (int l = 0; l < nlevels_; ++l) {
      const Csr &W = plan.w[l];
      for (uint32_t r = 0; r < W.rows(); ++r) {
        const uint32_t lf = plan.leaf_off[l] + r;
        if (W.off[r + 1] == W.off[r] || plan.tgt_start[lf + 1] == plan.tgt_start[lf]) continue;
        id.push_back(lf);
        for (uint32_t e = W.off[r]; e < W.off[r + 1]; ++e) ent.push_back(gid(l + 1, W.ent[e]));
        off.push_back(static_cast<uint32_t>(ent.size()));
      }
    }
    upload_rows(m2p_rows_, id, cell, off, ent);
  }
  {
    std::vector<uint32_t> id, cell, off{0}, ent;
    for (int l = 1; l < nlevels_; ++l) {
      const Csr &X = plan.x[l];
      for (uint32_t b = 0; b < X.rows(); ++b) {
        if (X.off[b + 1] == X.off[b]) continue;
        const size_t g = gid(l, b);
        id.push_back(uint32_t(g));
        cell.insert(cell.end(), {box_cell[4 * g], box_cell[4 * g + 1], box_cell[4 * g + 2], box_cell[4 * g + 3]});
        for (uint32_t e = X.off[b]; e < X.off[b + 1]; ++e) ent.push_back(X.ent[e]);
        off.push_back(static_cast<uint32_t>(ent.size()));
      }
    }
    upload_rows(p2l_rows_, id, cell, off, ent);
  }

  // tables and schedules
  const size_t n2 = size_t(nc_) * nc_, sw = dev::swizzled_table_size(p);
  auto swizzle_planes = [&](const std::vector<double> &t, size_t count, std::vector<double> &out) {
    const size_t base = out.size();
    out.resize(base + 2 * count * sw);
    for (size_t i = 0; i < 2 * count; ++i) dev::swizzle_table(p, t.data() + i * n2, out.data() + base + i * sw);
  };
  // M2M and L2L: per parent level l, 8 octants x (re, im), ids 16 l + 2 o + plane
  {
    std::vector<double> m2m_sw, l2l_sw;
    m2m_.resize(nlevels_);
    l2l_.resize(nlevels_);
    m2m_complex_ = false;
    for (int l = 0; l + 1 < nlevels_; ++l) {
      std::vector<double> tm, tl;
      h_m2m_tables(p, kh[l], tm);
      h_l2l_tables(p, kh[l], tl);
      for (int o = 0; o < 8; ++o)
        if (!all_zero(tm.data() + (2 * o + 1) * n2, n2) || !all_zero(tl.data() + (2 * o + 1) * n2, n2))
          m2m_complex_ = true;
      swizzle_planes(tm, 8, m2m_sw);
      swizzle_planes(tl, 8, l2l_sw);
    }
    if (m2m_sw.empty()) m2m_sw.assign(sw, 0.0);
    if (l2l_sw.empty()) l2l_sw.assign(sw, 0.0);
    t_m2m_.upload(m2m_sw);
    t_l2l_.upload(l2l_sw);
    for (int l = 0; l + 1 < nlevels_; ++l) {
      // M2M into the parents of level l (boxes with children); narrow tasks (waves split
      // the rows) while wide ones would leave the GPU idle
      uint32_t nparents = 0;
      for (uint32_t b = 0; b < plan.nboxes(l); ++b) nparents += plan.first_child[l][b] != kNone;
      const int width = 2 * nparents < 64u * dev::COLS ? 16 : dev::COLS;
      ScheduleBuilder sb(width);
      std::vector<uint32_t> cols;
      std::vector<std::vector<std::pair<uint16_t, uint32_t>>> lists;
      for (uint32_t b = 0; b < plan.nboxes(l); ++b) {
        const uint32_t fc = plan.first_child[l][b];
        if (fc == kNone) continue;
        std::vector<std::pair<uint16_t, uint32_t>> re, im;
        for (int o = 0; o < 8; ++o) {
          const uint32_t c = gid(l + 1, fc + o);
          const uint16_t tr = uint16_t(16 * l + 2 * o), ti = uint16_t(tr + 1);
          re.push_back({tr, 3 * c});
          im.push_back({tr, 3 * c + 1});
          if (m2m_complex_) {
            re.push_back({ti, 3 * c + 2});
            im.push_back({ti, 3 * c});
          }
        }
        cols.push_back(3 * gid(l, b));
        lists.push_back(std::move(re));
        cols.push_back(3 * gid(l, b) + 1);
        lists.push_back(std::move(im));
        if (int(cols.size()) == width) {
          sb.add_task(cols, lists);
          cols.clear();
          lists.clear();
        }
      }
      if (!cols.empty()) sb.add_task(cols, lists);
      sb.finish(m2m_[l]);
    }
    for (int l = 1; l < nlevels_; ++l) {
      // L2L into the boxes of level l, by octant (a task shares its tables)
      const int width = 2 * plan.nboxes(l) < 64u * dev::COLS ? 16 : dev::COLS;
      ScheduleBuilder sb(width);
      for (int o = 0; o < 8; ++o) {
        std::vector<uint32_t> cols;
        std::vector<std::vector<std::pair<uint16_t, uint32_t>>> lists;
        const uint16_t tr = uint16_t(16 * (l - 1) + 2 * o), ti = uint16_t(tr + 1);
        for (uint32_t b = 0; b < plan.nboxes(l); ++b) {
          if (octant_of(plan.keys[l][b]) != o) continue;
          const uint32_t P = gid(l - 1, plan.parent[l][b]);
          std::vector<std::pair<uint16_t, uint32_t>> re{{tr, 3 * P}}, im{{tr, 3 * P + 1}};
          if (m2m_complex_) {
            re.push_back({ti, 3 * P + 2});
            im.push_back({ti, 3 * P});
          }
          cols.push_back(3 * gid(l, b));
          lists.push_back(std::move(re));
          cols.push_back(3 * gid(l, b) + 1);
          lists.push_back(std::move(im));
          if (int(cols.size()) == width) {
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
  // M2L levels (HelmholtzM2l, fmm-exec/src/helmholtz/operator.rs:448-479): under the dense
  // choice one complex table per offset on the levels whose tables fit the byte budget,
  // chosen from the finest level up as DenseM2l::select does
  // (fmm-tables/src/helmholtz/dense.rs:74-85: 316 x 2 (p+1)^4 f64 per level, never levels 0
  // and 1), and per-pair rotation on the others; under rotation, per-pair rotation on
  // every level.
  std::vector<char> dense_level(nlevels_, 0);
  if (opt_.m2l == kM2lDense) {
    const int64_t level_bytes = int64_t(316) * 2 * nc_ * nc_ * int64_t(sizeof(double));
    int64_t left = opt_.m2l_dense_budget > 0 ? opt_.m2l_dense_budget : kHelmholtzDenseBudget;
    for (int l = nlevels_ - 1; l >= 2; --l)
      if (opt_.m2l_dense_budget < 0 || level_bytes <= left) {
        left -= level_bytes;
        dense_level[l] = 1;
      }
  }
  // under the exponential M2L (HelmholtzM2l::PlaneWave, operator.rs:462-468) the levels
  // that the reference's rule gives a quadrature (merge.rs select without a tolerance:
  // iprec 2 up to degree 15, levels >= 2 whose kappa = 2 kh lies in FMM3D's tables), the
  // per-pair rotation on the others
  std::vector<char> pw_level(nlevels_, 0);
  if (opt_.m2l == kM2lPlaneWave)
    for (int l = 2; l < nlevels_; ++l) pw_level[l] = plan.v[l].size() > 0 && hpw_level_selected(l, p, kh[l]);
  {
    // Per pair: rotate the source multipole so the offset lies along +z, shift coaxially,
    // rotate back (the reference's RCR).  Pairs by level, then offset: blocks of at most
    // pb pairs of one (level, offset), in chunks whose scratch rows (two per pair) stay
    // within kRotScratchBytes; after each chunk a target adds that chunk's rows in the
    // order of its V list.
    std::vector<uint32_t> blk_first, blk_coax, pair_src, red_id, red_off{0}, red_ent;
    std::vector<uint16_t> blk_offset;
    std::vector<double> fwd, coax;
    uint32_t npairs = 0, ncoax = 0;
    int pb = 0;
    std::vector<std::vector<uint32_t>> level_pairs(nlevels_);  // pair index of each V entry
    for (int l = 2; l < nlevels_; ++l) {
      const Csr &V = plan.v[l];
      if (V.size() == 0 || dense_level[l] || pw_level[l]) continue;
      if (rot_rb_ == 0) {
        rot_rb_ = h_rcr_rotation_size(p);
        rot_cb_ = h_rcr_coaxial_size(p);
        h_rcr_rotations(p, fwd);
        pb = dev::h_rotation_pairs_per_block(p, rot_rb_, rot_cb_);
        hip_check(dev::h_rotation_allow_lds(p, rot_rb_, rot_cb_), "Helmholtz rotation M2L LDS");
      }
      m2l_pairs_ += V.size();
      // this level's distances |2 d| (by the exact integer |d|^2) and their coaxial blocks
      std::map<int, uint32_t> coax_of;
      for (uint32_t e = 0; e < V.size(); ++e) {
        int d[3];
        offset_of_index(V.aux[e], d);
        coax_of.emplace(d[0] * d[0] + d[1] * d[1] + d[2] * d[2], 0u);
      }
      std::vector<double> dist;
      for (auto &kv : coax_of) {
        kv.second = ncoax++;
        dist.push_back(2.0 * std::sqrt(double(kv.first)));
      }
      std::vector<double> c;
      h_rcr_coaxial(p, kh[l], dist, c);
      coax.insert(coax.end(), c.begin(), c.end());
      std::vector<std::vector<uint32_t>> by_t(316);
      for (uint32_t e = 0; e < V.size(); ++e) by_t[V.aux[e]].push_back(e);
      std::vector<uint32_t> &pair_of = level_pairs[l];
      pair_of.assign(V.size(), 0);
      for (int t = 0; t < 316; ++t) {
        if (by_t[t].empty()) continue;
        int d[3];
        offset_of_index(t, d);
        const uint32_t ci = coax_of.at(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        for (size_t k = 0; k < by_t[t].size(); ++k) {
          if (k % size_t(pb) == 0) {
            blk_first.push_back(npairs);
            blk_offset.push_back(static_cast<uint16_t>(t));
            blk_coax.push_back(ci);
          }
          const uint32_t e = by_t[t][k];
          pair_src.push_back(static_cast<uint32_t>(gid(l, V.ent[e])));
          pair_of[e] = npairs++;
        }
      }
    }
    if (npairs > 0) {
      const uint32_t nblocks = static_cast<uint32_t>(blk_offset.size());
      blk_first.push_back(npairs);
      // chunks of whole blocks, at most chunk_pairs pairs each (at least one block)
      const size_t chunk_pairs = std::max<size_t>(size_t(pb), dev::kRotScratchBytes / (2 * size_t(nc_) * sizeof(double)));
      rot_chunk_blk_.assign(1, 0);
      for (uint32_t bk = 0; bk < nblocks; ++bk)
        if (blk_first[bk + 1] - blk_first[rot_chunk_blk_.back()] > chunk_pairs) rot_chunk_blk_.push_back(bk);
      rot_chunk_blk_.push_back(nblocks);
      const size_t nchunks = rot_chunk_blk_.size() - 1;
      std::vector<uint32_t> chunk_q(nchunks + 1);
      for (size_t c = 0; c <= nchunks; ++c) chunk_q[c] = blk_first[rot_chunk_blk_[c]];
      rot_chunk_q_ = chunk_q;
      // per chunk, per target (level by level, box order) and plane: its rows of that chunk
      // in V order
      std::vector<std::vector<uint32_t>> cid(nchunks), coff(nchunks), cent(nchunks);
      for (size_t c = 0; c < nchunks; ++c) coff[c].push_back(0);
      for (int l = 2; l < nlevels_; ++l) {
        const Csr &V = plan.v[l];
        if (level_pairs[l].empty()) continue;
        for (uint32_t b = 0; b < plan.nboxes(l); ++b) {
          std::map<size_t, std::vector<uint32_t>> by_chunk;
          for (uint32_t e = V.off[b]; e < V.off[b + 1]; ++e) {
            const uint32_t q = level_pairs[l][e];
            const size_t c = size_t(std::upper_bound(chunk_q.begin(), chunk_q.end(), q) - chunk_q.begin()) - 1;
            by_chunk[c].push_back(q);
          }
          for (const auto &kv : by_chunk)
            for (int plane = 0; plane < 2; ++plane) {
              cid[kv.first].push_back(static_cast<uint32_t>(3 * gid(l, b) + plane));
              for (uint32_t q : kv.second) cent[kv.first].push_back(2 * (q - chunk_q[kv.first]) + plane);
              coff[kv.first].push_back(static_cast<uint32_t>(cent[kv.first].size()));
            }
        }
      }
      rot_chunk_row_.assign(1, 0);
      for (size_t c = 0; c < nchunks; ++c) {
        const uint32_t base = static_cast<uint32_t>(red_ent.size());
        red_id.insert(red_id.end(), cid[c].begin(), cid[c].end());
        for (size_t r = 1; r < coff[c].size(); ++r) red_off.push_back(base + coff[c][r]);
        red_ent.insert(red_ent.end(), cent[c].begin(), cent[c].end());
        rot_chunk_row_.push_back(static_cast<uint32_t>(red_id.size()));
      }
      size_t most = 0;
      for (size_t c = 0; c < nchunks; ++c) most = std::max<size_t>(most, chunk_q[c + 1] - chunk_q[c]);
      rot_fwd_.upload(fwd);
      rot_coax_.upload(coax);
      rot_blk_first_.upload(blk_first);
      rot_blk_offset_.upload(blk_offset);
      rot_blk_coax_.upload(blk_coax);
      rot_pair_src_.upload(pair_src);
      rot_slots_.alloc(2 * most * nc_);
      upload_rows(rot_red_, red_id, {}, red_off, red_ent);
    }
  }
  // M2L: per level the offsets present, ids base_l + 2 j + plane; tiles of 64 targets of
  // one octant class, classes interleaved, groups by (parent-level neighbour, octant)
  {
    std::vector<double> m2l_sw;
    std::vector<uint16_t> m2l16;
    const size_t sw16 = dev::swizzled_table_size_bf16(p);
    ScheduleBuilder all;
    std::vector<uint32_t> red_id, red_cell, red_off{0}, red_ent;
    uint32_t id_base = 0;
    for (int l = 2; l < nlevels_; ++l) {
      const Csr &V = plan.v[l];
      if (V.size() == 0 || !dense_level[l]) continue;
      m2l_pairs_ += V.size();
      std::vector<uint16_t> used(V.aux.begin(), V.aux.end());
      std::sort(used.begin(), used.end());
      used.erase(std::unique(used.begin(), used.end()), used.end());
      require(id_base + 2 * used.size() < 65536u, "Helmholtz: too many M2L tables", kTooLarge);
      std::vector<int> slot(316, -1);
      for (size_t j = 0; j < used.size(); ++j) slot[used[j]] = int(j);
      std::vector<double> t;
      h_m2l_tables(p, kh[l], used, t);
      if (opt_.precision == kMixed) {
        const size_t base = m2l16.size();
        m2l16.resize(base + 2 * used.size() * sw16);
        for (size_t i = 0; i < 2 * used.size(); ++i)
          dev::swizzle_table_bf16(p, t.data() + i * n2, m2l16.data() + base + i * sw16);
      } else {
        swizzle_planes(t, used.size(), m2l_sw);
      }
      // ranks of this level's table ids per target octant
      const uint32_t nid = id_base + 2 * uint32_t(used.size());
      std::vector<std::vector<uint16_t>> rank(8, std::vector<uint16_t>(nid, 0));
      for (int o = 0; o < 8; ++o)
        for (size_t j = 0; j < used.size(); ++j) {
          rank[o][id_base + 2 * j] = uint16_t(2 * m2l_rank(o, used[j]));
          rank[o][id_base + 2 * j + 1] = uint16_t(2 * m2l_rank(o, used[j]) + 1);
        }
      // a V entry's (re, im) column terms: Re L += Tr Re M + Ti (-Im M), Im L += Tr Im M + Ti Re M
      auto terms = [&](uint32_t e, std::vector<std::pair<uint16_t, uint32_t>> &re,
                       std::vector<std::pair<uint16_t, uint32_t>> &im) {
        const uint32_t s = gid(l, V.ent[e]);
        const uint16_t tr = uint16_t(id_base + 2 * slot[V.aux[e]]), ti = uint16_t(tr + 1);
        re.push_back({tr, 3 * s});
        re.push_back({ti, 3 * s + 2});
        im.push_back({tr, 3 * s + 1});
        im.push_back({ti, 3 * s});
      };
      std::vector<std::vector<uint32_t>> cls_cols(8);
      std::vector<std::vector<std::vector<std::pair<uint16_t, uint32_t>>>> cls_lists(8);
      for (uint32_t b = 0; b < plan.nboxes(l); ++b) {
        if (V.off[b + 1] == V.off[b]) continue;
        const int o = octant_of(plan.keys[l][b]);
        std::vector<std::pair<uint16_t, uint32_t>> re, im;
        for (uint32_t e = V.off[b]; e < V.off[b + 1]; ++e) terms(e, re, im);
        cls_cols[o].push_back(3 * gid(l, b));
        cls_lists[o].push_back(std::move(re));
        cls_cols[o].push_back(3 * gid(l, b) + 1);
        cls_lists[o].push_back(std::move(im));
      }
      const size_t C = dev::COLS;
      size_t most = 0;
      for (int o = 0; o < 8; ++o) most = std::max(most, cls_cols[o].size());
      ScheduleBuilder tile;
      for (size_t at = 0; at < most; at += C)
        for (int o = 0; o < 8; ++o) {
          if (at >= cls_cols[o].size()) continue;
          const size_t end = std::min(cls_cols[o].size(), at + C);
          std::vector<uint32_t> cols(cls_cols[o].begin() + at, cls_cols[o].begin() + end);
          std::vector<std::vector<std::pair<uint16_t, uint32_t>>> lists(cls_lists[o].begin() + at,
                                                                          cls_lists[o].begin() + end);
          tile.add_task(cols, lists, rank[o].data());
        }
      // Box mode needs both a dense level and enough tasks to fill the GPU: a task walks
      // all of its ~2 x 189 groups in sequence, so a level of a few dozen tasks (level 4
      // of a 1M-point tree holds 64) would be one long chain per CU.
      constexpr uint32_t kMinBoxTasks = 256;
      if (tile.pairs == 0 || (2 * tile.pairs >= tile.slots && tile.ntasks() >= kMinBoxTasks)) {
        all.append(tile);
      } else {
        // Sparse or small level: clustered mode as in the Laplace operator.  Each
        // (target, parent-level neighbour D) gets two scratch rows (re, im); a task holds
        // the rows of one (target octant, D) class, whose columns share their offsets
        // (at most 2 x 8 groups); a target plane's rows are added in ascending D after.
        struct Member {
          uint32_t row;
          std::vector<std::pair<uint16_t, uint32_t>> list;
        };
        std::vector<std::vector<Member>> cls(8 * 27);
        for (uint32_t b = 0; b < plan.nboxes(l); ++b) {
          if (V.off[b + 1] == V.off[b]) continue;
          const int o = octant_of(plan.keys[l][b]);
          std::map<int, std::vector<uint32_t>> by_D;
          for (uint32_t e = V.off[b]; e < V.off[b + 1]; ++e) {
            const int D = m2l_rank(o, V.aux[e]) / 8;
            require(D >= 0 && D < 27, "Helmholtz M2L schedule: V offset outside the parent's neighbours",
                    kUnsupportedLayout);
            by_D[D].push_back(e);
          }
          std::vector<uint32_t> rows_re, rows_im;
          for (const auto &kv : by_D) {
            std::vector<std::pair<uint16_t, uint32_t>> re, im;
            for (uint32_t e : kv.second) terms(e, re, im);
            const uint32_t r0 = static_cast<uint32_t>(3 * nbox_total_ + nslots_), r1 = r0 + 1;
            nslots_ += 2;
            rows_re.push_back(r0);
            rows_im.push_back(r1);
            cls[size_t(kv.first) * 8 + o].push_back({r0, std::move(re)});
            cls[size_t(kv.first) * 8 + o].push_back({r1, std::move(im)});
          }
          red_id.push_back(3 * gid(l, b));
          red_ent.insert(red_ent.end(), rows_re.begin(), rows_re.end());
          red_off.push_back(static_cast<uint32_t>(red_ent.size()));
          red_id.push_back(3 * gid(l, b) + 1);
          red_ent.insert(red_ent.end(), rows_im.begin(), rows_im.end());
          red_off.push_back(static_cast<uint32_t>(red_ent.size()));
        }
        for (int D = 0; D < 27; ++D) {
          size_t most_d = 0;
          for (int o = 0; o < 8; ++o) most_d = std::max(most_d, cls[size_t(D) * 8 + o].size());
          for (size_t at = 0; at < most_d; at += C)
            for (int o = 0; o < 8; ++o) {
              const auto &mem = cls[size_t(D) * 8 + o];
              if (at >= mem.size()) continue;
              std::vector<uint32_t> cols;
              std::vector<std::vector<std::pair<uint16_t, uint32_t>>> lists;
              for (size_t k = at; k < mem.size() && k < at + C; ++k) {
                cols.push_back(mem[k].row);
                lists.push_back(mem[k].list);
              }
              all.add_task(cols, lists, rank[o].data());
            }
        }
      }
      id_base = nid;
    }
    all.finish(m2l_);
    upload_rows(m2l_red_, red_id, red_cell, red_off, red_ent);
    if (opt_.precision == kMixed)
      slots32_.alloc(nslots_ * nc_);
    else
      slots_.alloc(nslots_ * nc_);
    if (opt_.precision == kMixed) {
      if (m2l16.empty()) m2l16.assign(sw16, 0);
      t_m2l16_.upload(m2l16);
      mult16_.alloc(3 * nbox_total_ * dev::split_expansion_size(p));
    } else {
      if (m2l_sw.empty()) m2l_sw.assign(sw, 0.0);
      t_m2l_.upload(m2l_sw);
    }
  }
  if (opt_.m2l == kM2lPlaneWave) build_plane_wave(plan, kh, pw_level);
  hip_check(hipDeviceSynchronize(), "Helmholtz operator build");
}

// The exponential M2L's levels (operator_plane_wave.rs; merge.rs MergeLevel): per level its
// quadrature and tables (hpw_level), its 104 source shifts and 48 child shifts, the merge's
// lists with wave blocks counted from the level's first box, and the level's waves and
// incoming waves (12 W f64 per box).
void HelmholtzOperator::build_plane_wave(const Plan &plan, const std::vector<cplx> &kh,
                                         const std::vector<char> &use) {
  const int p = opt_.p;
  const PwMergeGeometry &geo = plane_wave_merge_geometry();
  static_assert(kPwAccStride == dev::kPwAccStride, "one accumulator stride");
  std::vector<int> ints;
  std::vector<double> tab, shifts;
  struct Off {
    size_t start, len, ord, mode_start, trig, factor, wk, item, cosm, sinm, out_re, out_im, in_re, in_im, frames;
  };
  std::vector<Off> offs;
  size_t max_out = 0, max_local = 0, waves = 0;
  pw_wave_off_.assign(nlevels_ + 1, 0);
  pw_shift_off_.assign(nlevels_ + 1, 0);
  for (int l = 0; l < nlevels_; ++l) {
    pw_wave_off_[l] = waves;
    pw_shift_off_[l] = shifts.size();
    if (!use[l]) continue;
    const HpwLevel L = hpw_level(p, kh[l], hpw_quadrature(2.0 * kh[l]));
    Off o{};
    auto put_i = [&](const std::vector<int> &v) {
      const size_t at = ints.size();
      ints.insert(ints.end(), v.begin(), v.end());
      return at;
    };
    auto put_d = [&](const std::vector<double> &v) {
      const size_t at = tab.size();
      tab.insert(tab.end(), v.begin(), v.end());
      return at;
    };
    std::vector<int> wk(L.W), item;
    for (int k = 0; k < L.s; ++k) {
      for (int j = 0; j < L.len[k]; ++j) wk[L.start[k] + j] = k;
      for (int m = 0; m < L.ord[k]; ++m) item.push_back(k * (p + 1) + m);
    }
    std::stable_sort(item.begin(), item.end(),
                     [&](int a, int b) { return L.len[a / (p + 1)] > L.len[b / (p + 1)]; });
    o.start = put_i(L.start);
    o.len = put_i(L.len);
    o.ord = put_i(L.ord);
    o.mode_start = put_i(L.mode_start);
    o.trig = put_i(L.trig);
    o.factor = put_i(L.factor);
    o.wk = put_i(wk);
    o.item = put_i(item);
    o.cosm = put_d(L.cosm);
    o.sinm = put_d(L.sinm);
    o.out_re = put_d(L.out_re);
    o.out_im = put_d(L.out_im);
    o.in_re = put_d(L.in_re);
    o.in_im = put_d(L.in_im);
    o.frames = put_d(L.frame1);
    put_d(L.frame2);
    offs.push_back(o);
    pw_levels_.push_back(l);
    dev::HpwArgs a{};
    a.p = p;
    a.nodes = L.s;
    a.W = L.W;
    a.orders = L.orders;
    a.rb = static_cast<int>(L.frame1.size());
    a.nitem = static_cast<int>(item.size());
    pw_args_.push_back(a);
    max_out = std::max(max_out, dev::hpw_outgoing_lds(p, L.orders));
    max_local = std::max(max_local, dev::hpw_local_lds(p, L.W, L.orders));
    // the merge's shifts of this level: the source shifts, then the child shifts
    const size_t W2 = 2 * size_t(L.W), s0 = shifts.size();
    shifts.resize(s0 + (geo.shift_d.size() + size_t(kPwDirections) * kPwChildren) * W2);
    for (size_t i = 0; i < geo.shift_d.size(); ++i) {
      double dp[3];
      plane_wave_rotate(geo.shift_dir[i], geo.shift_d[i].data(), dp);
      hpw_shift(L, dp, (geo.shift_dir[i] & 1) == 0, &shifts[s0 + i * W2]);
    }
    for (int dir = 0; dir < kPwDirections; ++dir)
      for (int c = 0; c < kPwChildren; ++c) {
        const double d[3] = {((c >> 2) & 1) - 0.5, ((c >> 1) & 1) - 0.5, (c & 1) - 0.5};
        double dp[3];
        plane_wave_rotate(dir, d, dp);
        hpw_shift(L, dp, (dir & 1) == 0, &shifts[s0 + (geo.shift_d.size() + size_t(kPwChildren) * dir + c) * W2]);
      }
    waves += size_t(plan.nboxes(l)) * 12 * L.W;
    require(L.W * 2 <= (1 << 16) && waves < (size_t(1) << 40), "Helmholtz plane wave: too many waves", kTooLarge);
  }
  pw_wave_off_[nlevels_] = waves;
  pw_shift_off_[nlevels_] = shifts.size();
  if (pw_levels_.empty()) return;
  require(max_out <= 160 * 1024 && max_local <= 160 * 1024, "Helmholtz plane wave: the level's tables exceed LDS",
          kUnsupportedLayout);
  hip_check(dev::hpw_allow_lds(max_out, max_local), "Helmholtz plane wave LDS");
  pw_int_.upload(ints);
  pw_tab_.upload(tab);
  pw_shift_.upload(shifts);
  for (size_t i = 0; i < pw_args_.size(); ++i) {
    dev::HpwArgs &a = pw_args_[i];
    const Off &o = offs[i];
    a.start = pw_int_.ptr + o.start;
    a.len = pw_int_.ptr + o.len;
    a.ord = pw_int_.ptr + o.ord;
    a.mode_start = pw_int_.ptr + o.mode_start;
    a.trig = pw_int_.ptr + o.trig;
    a.factor = pw_int_.ptr + o.factor;
    a.wk = pw_int_.ptr + o.wk;
    a.item = pw_int_.ptr + o.item;
    a.cosm = pw_tab_.ptr + o.cosm;
    a.sinm = pw_tab_.ptr + o.sinm;
    a.out_re = pw_tab_.ptr + o.out_re;
    a.out_im = pw_tab_.ptr + o.out_im;
    a.in_re = pw_tab_.ptr + o.in_re;
    a.in_im = pw_tab_.ptr + o.in_im;
    a.frames = pw_tab_.ptr + o.frames;
  }
  pw_acc_first_.upload(geo.acc_first);
  pw_acc_mask_.upload(geo.acc_mask);
  const PwMergeLists lists = plane_wave_merge_lists(plan, box_base_, use, true);
  pw_par_level_ = lists.par_level;
  pw_tgt_level_ = lists.tgt_level;
  pw_pe_first_.upload(lists.pe_first);
  pw_pe_.upload(lists.pe);
  pw_children_.upload(lists.children);
  pw_targets_.upload(lists.targets);
  pw_waves_.alloc(waves);
  pw_inc_.alloc(waves);
  for (int l : pw_levels_) m2l_pairs_ += plan.v[l].size();
}

// Per level with a quadrature: the waves of its boxes, the merge of its parents into their
// children's incoming waves, the children's locals.
void HelmholtzOperator::launch_plane_wave() {
  const PwMergeGeometry &geo = plane_wave_merge_geometry();
  for (size_t i = 0; i < pw_levels_.size(); ++i) {
    const int l = pw_levels_[i];
    const dev::HpwArgs &a = pw_args_[i];
    const uint32_t first = box_base_[l], nb = box_base_[l + 1] - first;
    double *waves = pw_waves_.ptr + pw_wave_off_[l], *inc = pw_inc_.ptr + pw_wave_off_[l];
    dev::hpw_outgoing(stream_, a, nb, first, mult_.ptr, waves);
    dev::PwArgs pa{};
    pa.H = a.W;
    const double *sh = pw_shift_.ptr + pw_shift_off_[l];
    const dev::PwMerge g{pw_acc_first_.ptr, pw_acc_mask_.ptr, sh, sh + geo.shift_d.size() * 2 * size_t(a.W)};
    const uint32_t q0 = pw_par_level_[l], q1 = pw_par_level_[l + 1];
    timed(kKindM2lContract, [&] {
      dev::pw_merge(stream_, pa, g, q1 - q0, pw_pe_first_.ptr + size_t(q0) * 6 * dev::kPwAccStride, pw_pe_.ptr,
                    pw_children_.ptr + size_t(q0) * kPwChildren, first, waves, inc);
    });
    const uint32_t t0 = pw_tgt_level_[l], t1 = pw_tgt_level_[l + 1];
    dev::hpw_local(stream_, a, t1 - t0, pw_targets_.ptr + t0, first, inc, loc_.ptr);
  }
}

HelmholtzOperator::~HelmholtzOperator() {
  for (hipEvent_t e : events_) (void)hipEventDestroy(e);
  if (order_event_) (void)hipEventDestroy(order_event_);
  if (stream_) (void)hipStreamDestroy(stream_);
}

void HelmholtzOperator::wait_for(hipStream_t producer) {
  hip_check(hipEventRecord(order_event_, producer), "hipEventRecord");
  hip_check(hipStreamWaitEvent(stream_, order_event_, 0), "hipStreamWaitEvent");
}

size_t HelmholtzOperator::device_bytes() const {
  return mult_.bytes() + loc_.bytes() + t_m2m_.bytes() + t_l2l_.bytes() + t_m2l_.bytes() + t_m2l16_.bytes() +
         mult16_.bytes() + sx_.bytes() * 8 + slots_.bytes() + slots32_.bytes() + rot_fwd_.bytes() +
         rot_coax_.bytes() + rot_slots_.bytes() + pw_tab_.bytes() + pw_shift_.bytes() + pw_waves_.bytes() +
         pw_inc_.bytes();
}

void HelmholtzOperator::set_timing(bool enable) {
  sync();
  timing_ = enable;
  if (enable && events_.empty()) {
    events_.resize(512);
    for (auto &e : events_) hip_check(hipEventCreate(&e), "hipEventCreate");
  }
  marks_.clear();
  events_used_ = 0;
  for (int k = 0; k < 16; ++k) {
    kind_ms_[k] = 0.0;
    kind_launches_[k] = 0;
  }
}

template <class F>
void HelmholtzOperator::timed(int kind, F &&f) {
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

void HelmholtzOperator::collect_timings() {
  for (const Mark &m : marks_) {
    float ms = 0.0f;
    hip_check(hipEventElapsedTime(&ms, m.a, m.b), "hipEventElapsedTime");
    kind_ms_[m.kind] += ms;
    kind_launches_[m.kind] += 1;
  }
  marks_.clear();
  events_used_ = 0;
}

void HelmholtzOperator::kind_times(double *ms, int64_t *launches) {
  sync();
  for (int k = 0; k < kNumKinds; ++k) {
    if (ms) ms[k] = kind_ms_[k];
    if (launches) launches[k] = kind_launches_[k];
  }
}

void HelmholtzOperator::load_charges(const double *charges, bool on_device) {
  if (ns_ == 0) return;
  const double *src = charges;
  if (!on_device) {
    hip_check(hipMemcpyAsync(q_in_.ptr, charges, 2 * ns_ * sizeof(double), hipMemcpyHostToDevice, stream_),
              "charges");
    src = q_in_.ptr;
  }
  dev::h_gather(stream_, src, src_order_.ptr, ns_, qr_.ptr, qi_.ptr, src_rec_.count ? src_rec_.ptr : nullptr);
}

void HelmholtzOperator::reset() {
  dev::fill(stream_, mult_.ptr, mult_.count, 0.0);
  dev::fill(stream_, loc_.ptr, loc_.count, 0.0);
  dev::fill(stream_, phi_hat_.ptr, phi_hat_.count, 0.0);
  dev::fill(stream_, grad_hat_.ptr, grad_hat_.count, 0.0);
}

void HelmholtzOperator::launch_p2p() {
  const int damped = opt_.kim != 0.0;
  if (opt_.precision == kMixed)
    dev::h_p2p<float>(stream_, 0, nleaves_, opt_.gradients, damped, near_off_.ptr, near_ent_.ptr, leaf_cell_.ptr,
                      src_start_.ptr, tgt_start_.ptr, sx_.ptr, sy_.ptr, sz_.ptr, qr_.ptr, qi_.ptr, src_rec_.ptr,
                      tx_.ptr, ty_.ptr, tz_.ptr, hlev_.ptr, phi_hat_.ptr, grad_hat_.ptr);
  else
    dev::h_p2p<double>(stream_, 0, nleaves_, opt_.gradients, damped, near_off_.ptr, near_ent_.ptr, leaf_cell_.ptr,
                       src_start_.ptr, tgt_start_.ptr, sx_.ptr, sy_.ptr, sz_.ptr, qr_.ptr, qi_.ptr, nullptr, tx_.ptr,
                       ty_.ptr, tz_.ptr, hlev_.ptr, phi_hat_.ptr, grad_hat_.ptr);
}

void HelmholtzOperator::evaluate() {
  const int L = nlevels_ - 1, p = opt_.p;
  const uint32_t none = kNone;
  // scratch rows of the sparse M2L levels follow the 3 planes of every box
  const uint32_t first_slot = static_cast<uint32_t>(3 * nbox_total_);
  // upward: P2M of every leaf, then per level the negated planes and M2M into the parents
  timed(0, [&] {
    dev::h_expand(stream_, p, 0, p2m_rows_.rows, p2m_rows_.row_id.ptr, p2m_rows_.row_cell.ptr, p2m_rows_.row_off.ptr,
                  p2m_rows_.row_ent.ptr, leaf_cell_.ptr, src_start_.ptr, sx_.ptr, sy_.ptr, sz_.ptr, qr_.ptr, qi_.ptr,
                  hlev_.ptr, norm_.ptr, mult_.ptr);
  });
  for (int l = L - 1; l >= 0; --l)
    timed(1, [&] {
      dev::h_negate(stream_, p, box_base_[l + 1], box_base_[l + 2], mult_.ptr);
      const Schedule &s = m2m_[l];
      // levels near the root (fewer (task, M-tile) pairs than CUs): one workgroup per
      // task and M-tile (dev::m2m_split's per-tile form); else the narrow or wide
      // contraction with its LDS ring
      const uint32_t mtiles = uint32_t((nc_ + 15) / 16);
      if (s.ntasks && s.width == 16 && s.ntasks * mtiles < 256)
        dev::m2m_split(stream_, p, s.ntasks, s.task_group_off.ptr, s.group_table.ptr, s.group_src.ptr,
                       s.col_out.ptr, t_m2m_.ptr, mult_.ptr, mult_.ptr);
      else if (s.ntasks)
        dev::contract_f64(stream_, p, s.width, s.ntasks, s.task_group_off.ptr, s.group_table.ptr, s.group_src.ptr,
                          s.col_out.ptr, t_m2m_.ptr, mult_.ptr, mult_.ptr, none, nullptr, 0);
    });
  // M2L of every level in one launch
  if (!pw_levels_.empty()) timed(3, [&] { launch_plane_wave(); });
  // the per-pair rotation levels, chunk by chunk
  for (size_t c = 0; c + 1 < rot_chunk_blk_.size(); ++c)
    timed(3, [&] {
      const uint32_t b0 = rot_chunk_blk_[c], b1 = rot_chunk_blk_[c + 1];
      timed(kKindM2lContract, [&] {
        dev::h_m2l_rotation(stream_, p, b1 - b0, rot_chunk_q_[c], rot_blk_first_.ptr + b0, rot_blk_offset_.ptr + b0,
                            rot_blk_coax_.ptr + b0, rot_pair_src_.ptr, rot_fwd_.ptr, rot_coax_.ptr, rot_rb_, rot_cb_,
                            mult_.ptr, rot_slots_.ptr);
      });
      const uint32_t r0 = rot_chunk_row_[c], r1 = rot_chunk_row_[c + 1];
      dev::reduce_slots(stream_, p, r1 - r0, rot_red_.row_id.ptr + r0, rot_red_.row_off.ptr + r0,
                        rot_red_.row_ent.ptr, 0, rot_slots_.ptr, nullptr, loc_.ptr);
    });
  if (m2l_.ntasks)
    timed(3, [&] {
      if (opt_.precision == kMixed) {
        const size_t v0 = 3 * size_t(box_base_[2]), v1 = 3 * size_t(box_base_[L + 1]);
        const size_t sz = dev::split_expansion_size(p);
        dev::split_bf16(stream_, p, v1 - v0, mult_.ptr + v0 * nc_, mult16_.ptr + v0 * sz);
        timed(kKindM2lContract, [&] {
          dev::contract_bf16(stream_, p, m2l_.ntasks, m2l_.task_group_off.ptr, m2l_.group_table.ptr,
                             m2l_.group_src.ptr, m2l_.col_out.ptr, t_m2l16_.ptr, mult16_.ptr, loc_.ptr, first_slot,
                             slots32_.ptr, kM2lXcdRun);
        });
      } else {
        timed(kKindM2lContract, [&] {
          dev::contract_f64(stream_, p, m2l_.width, m2l_.ntasks, m2l_.task_group_off.ptr, m2l_.group_table.ptr,
                            m2l_.group_src.ptr, m2l_.col_out.ptr, t_m2l_.ptr, mult_.ptr, loc_.ptr, first_slot,
                            slots_.ptr, kM2lXcdRun);
        });
      }
      if (m2l_red_.rows)
        dev::reduce_slots(stream_, p, m2l_red_.rows, m2l_red_.row_id.ptr, m2l_red_.row_off.ptr, m2l_red_.row_ent.ptr,
                          first_slot, slots_.ptr, opt_.precision == kMixed ? slots32_.ptr : nullptr, loc_.ptr);
    });
  if (p2l_rows_.rows)
    timed(4, [&] {
      dev::h_expand(stream_, p, 1, p2l_rows_.rows, p2l_rows_.row_id.ptr, p2l_rows_.row_cell.ptr,
                    p2l_rows_.row_off.ptr, p2l_rows_.row_ent.ptr, leaf_cell_.ptr, src_start_.ptr, sx_.ptr, sy_.ptr,
                    sz_.ptr, qr_.ptr, qi_.ptr, hlev_.ptr, norm_.ptr, loc_.ptr);
    });
  // downward: L2L from the root
  for (int l = 1; l <= L; ++l)
    timed(2, [&] {
      if (m2m_complex_) dev::h_negate(stream_, p, box_base_[l - 1], box_base_[l], loc_.ptr);
      const Schedule &s = l2l_[l];
      if (s.ntasks)
        dev::contract_f64(stream_, p, s.width, s.ntasks, s.task_group_off.ptr, s.group_table.ptr, s.group_src.ptr,
                          s.col_out.ptr, t_l2l_.ptr, loc_.ptr, loc_.ptr, none, nullptr, 0);
    });
  // leaves: L2P, M2P, P2P
  timed(5, [&] {
    dev::h_evaluate(stream_, p, 0, opt_.gradients, l2p_rows_.rows, l2p_rows_.row_id.ptr, l2p_rows_.row_off.ptr,
                    l2p_rows_.row_ent.ptr, box_cell_.ptr, leaf_cell_.ptr, tgt_start_.ptr, tx_.ptr, ty_.ptr, tz_.ptr,
                    hlev_.ptr, norm_.ptr, loc_.ptr, phi_hat_.ptr, grad_hat_.ptr);
  });
  if (m2p_rows_.rows)
    timed(6, [&] {
      dev::h_evaluate(stream_, p, 1, opt_.gradients, m2p_rows_.rows, m2p_rows_.row_id.ptr, m2p_rows_.row_off.ptr,
                      m2p_rows_.row_ent.ptr, box_cell_.ptr, leaf_cell_.ptr, tgt_start_.ptr, tx_.ptr, ty_.ptr, tz_.ptr,
                      hlev_.ptr, norm_.ptr, mult_.ptr, phi_hat_.ptr, grad_hat_.ptr);
    });
  timed(7, [&] { launch_p2p(); });
}

void HelmholtzOperator::read_output(double *phi, double *grad, bool on_device) {
  const bool g = opt_.gradients && grad != nullptr;
  dev::h_output(stream_, nleaves_, tgt_start_.ptr, leaf_r_.ptr, tgt_order_.ptr, phi_hat_.ptr, grad_hat_.ptr,
                on_device ? phi : phi_.ptr, on_device ? grad : grad_.ptr, g);
  if (!on_device) {
    hip_check(hipMemcpyAsync(phi, phi_.ptr, 2 * nt_ * sizeof(double), hipMemcpyDeviceToHost, stream_), "phi");
    if (g) hip_check(hipMemcpyAsync(grad, grad_.ptr, 6 * nt_ * sizeof(double), hipMemcpyDeviceToHost, stream_), "grad");
  }
  sync();
}

void HelmholtzOperator::read_expansions(bool locals, int level, double *out) {
  require(level >= 0 && level < nlevels_, "read_expansions: level out of range");
  sync();
  const double *src = locals ? loc_.ptr : mult_.ptr;
  for (uint32_t b = box_base_[level]; b < box_base_[level + 1]; ++b)
    hip_check(hipMemcpy(out + size_t(b - box_base_[level]) * 2 * nc_, src + 3 * size_t(b) * nc_,
                        2 * nc_ * sizeof(double), hipMemcpyDeviceToHost),
              "read_expansions");
}

void HelmholtzOperator::sync() {
  hip_check(hipStreamSynchronize(stream_), "evaluation");
  hip_check(hipGetLastError(), "launch");
  if (!marks_.empty()) collect_timings();
}

void helmholtz_direct_sum(int device, double kre, double kim, const double *sources, const double *charges,
                          int64_t ns, const double *targets, int64_t nt, double *phi, double *grad, bool on_device) {
  require(ns >= 0 && nt >= 0, "helmholtz_direct_sum: negative size");
  if (nt == 0) return;
  hip_check(hipSetDevice(device), "hipSetDevice");
  hipStream_t s = nullptr;
  hip_check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
  const int nch = dev::h_direct_chunks(ns, nt);
  DevBuf<double> part, in, out;
  part.alloc(size_t(nch) * nt * 16);
  const double *src = sources, *q = charges, *tgt = targets;
  double *ph = phi, *gr = grad;
  if (!on_device) {
    in.alloc(5 * size_t(ns) + 3 * size_t(nt));
    out.alloc(8 * size_t(nt));
    hip_check(hipMemcpyAsync(in.ptr, sources, 3 * ns * sizeof(double), hipMemcpyHostToDevice, s), "sources");
    hip_check(hipMemcpyAsync(in.ptr + 3 * ns, charges, 2 * ns * sizeof(double), hipMemcpyHostToDevice, s),
              "charges");
    hip_check(hipMemcpyAsync(in.ptr + 5 * ns, targets, 3 * nt * sizeof(double), hipMemcpyHostToDevice, s),
              "targets");
    src = in.ptr;
    q = in.ptr + 3 * ns;
    tgt = in.ptr + 5 * ns;
    ph = out.ptr;
    gr = out.ptr + 2 * nt;
  }
  dev::h_direct_sum(s, kre, kim, src, q, ns, tgt, nt, grad != nullptr, part.ptr, nch, ph, grad ? gr : nullptr);
  if (!on_device) {
    hip_check(hipMemcpyAsync(phi, ph, 2 * nt * sizeof(double), hipMemcpyDeviceToHost, s), "phi");
    if (grad) hip_check(hipMemcpyAsync(grad, gr, 6 * nt * sizeof(double), hipMemcpyDeviceToHost, s), "grad");
  }
  const hipError_t e = hipStreamSynchronize(s);
  (void)hipStreamDestroy(s);
  hip_check(e, "helmholtz_direct_sum");
}

}  // namespace fmmh
