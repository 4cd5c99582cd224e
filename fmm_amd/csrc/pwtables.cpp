// pwtables.cpp -- plane-wave M2L: quadrature, frames, merge-and-shift geometry and the host
// restatements (docs/CONVENTIONS.md:1305-1434; fmm-tables/src/plane_wave/{mod,geometry,
// quadrature}.rs and fmm-kernels/src/plane_wave.rs:1-60 in the reference).
//
// Steps of one V pair (target cell - source cell = d, boxes), in the frame f of its
// direction (up: target on the positive side of the frame's axis, s = +1; down: s = -1):
//   multipole rotated into the frame, in box widths Mh_n^m = 2^-(n+1) M~_n^m;
//   A_k(m) = sum_{n>=m} lambda_k^n s^(n-m) Mh_n^m, m < F_k';
//   X(k, j) = (w_k / N_k) [A_k(0) + sum_{1<=m<F_k'} 2 (-i)^m Re(A_k(m) e^{i m a_j})], j < N_k / 2;
//   V = X e^{-s lambda_k d'_z + i lambda_k (d'_x cos a_j + d'_y sin a_j)}   (d' = Q_f d);
//   G_k(m) = 2 sum_j Re V e^{-i m a_j} (m even), 2 i sum_j Im V e^{-i m a_j} (m odd), m < F_k';
//   L~'_n^m += 2^-n lambda_k^n (-s)^(n-m) i^m G_k(m);   local rotated back out of the frame.
// Real charges make X(a + pi) = conj X(a), so half the a nodes carry the waves.  Merge and
// shift factors the shift of d into the slot's shift to the parent's centre (2.5 - a) and
// the parent's centre's shift to the child (b - 0.5): the accumulators of a parent sum the
// source-shifted waves once for every child they reach.
#include "pwtables.hpp"

#include <algorithm>
#include <cmath>
#include <complex>
#include <map>
#include <mutex>

#include "common.hpp"
#include "plan.hpp"

namespace fmmh {
namespace {

#include "pwquadrature.inc"

using cd = std::complex<double>;
constexpr double kPi = 3.141592653589793;
inline int idx(int n, int m) { return n * n + n + m; }

// The reference's lambda node count per degree p = 0..20 (fmm-tables/src/plane_wave/
// quadrature.rs:55-58): the fewest nodes whose M2L over a uniform level errs at most a tenth
// of degree p's truncation, made non-decreasing in p.
const int kDegreeNodes[21] = {3, 7, 9, 10, 12, 13, 15, 16, 17, 18, 20, 21, 21, 23, 26, 27, 27, 32, 32, 36, 36};

// first entry of table s in the flat arrays: sum_{t=2}^{s-1} t
inline int table_start(int s) { return (s - 1) * s / 2 - 1; }

const double kFrames[3][9] = {
    {1, 0, 0, 0, 1, 0, 0, 0, 1},
    {1, 0, 0, 0, 0, -1, 0, 1, 0},   // (x, y, z) -> (x, -z, y)
    {0, 0, -1, 0, 1, 0, 1, 0, 0},   // (x, y, z) -> (-z, y, x)
};

// x = F_f in: the multipole rotated into frame f
void rotate_into(const PlaneWave &pw, int f, const double *in, double *x) {
  const int p = pw.p, RB = rotation_block_size(p);
  for (int n = 0, bo = 0; n <= p; bo += (2 * n + 1) * (2 * n + 1), ++n) {
    const int k = 2 * n + 1;
    const double *blk = &pw.fwd[size_t(f) * RB + bo];
    for (int r = 0; r < k; ++r) {
      double s = 0.0;
      for (int c = 0; c < k; ++c) s += blk[r + c * k] * in[n * n + c];
      x[n * n + r] = s;
    }
  }
}

// X (re[H], im[H]) of the waves up (sgn = 1) or down (sgn = -1) of the frame multipole x
void waves_of(const PlaneWave &pw, const double *x, double sgn, double *X) {
  const int p = pw.p, P1 = p + 1, H = pw.H;
  const cd I(0.0, 1.0);
  for (int k = 0; k < pw.nodes; ++k) {
    const int F = pw.fourier[k];
    std::vector<cd> A(F, 0.0);
    for (int m = 0; m < F; ++m) {
      double s = 1.0;
      for (int n = m; n <= p; ++n, s *= sgn)
        A[m] += pw.opow[k * P1 + n] * s * cd(x[idx(n, m)], m > 0 ? x[idx(n, -m)] : 0.0);
    }
    const int Nk = pw.nalpha[k];
    for (int j = 0; j < Nk / 2; ++j) {
      const double a = 2.0 * kPi * j / Nk;
      cd v = A[0], mi = 1.0;  // (-i)^m
      for (int m = 1; m < F; ++m) {
        mi *= -I;
        v += 2.0 * mi * std::real(A[m] * std::exp(I * (m * a)));
      }
      v *= pw.weight[k] / Nk;
      X[pw.first[k] + j] = v.real();
      X[H + pw.first[k] + j] = v.imag();
    }
  }
}

// L ([n][m] complex, the frame's local) += the local of the incoming waves V, up or down
void local_add(const PlaneWave &pw, const double *V, double sgn, cd *L) {
  const int p = pw.p, P1 = p + 1, H = pw.H;
  const cd I(0.0, 1.0);
  for (int k = 0; k < pw.nodes; ++k) {
    const int F = pw.fourier[k], Nk = pw.nalpha[k];
    std::vector<cd> G(F, 0.0);
    for (int j = 0; j < Nk / 2; ++j) {
      const double a = 2.0 * kPi * j / Nk;
      const double vr = V[pw.first[k] + j], vi = V[H + pw.first[k] + j];
      for (int m = 0; m < F; ++m) {
        const cd e = std::exp(-I * (m * a));
        G[m] += (m % 2 == 0) ? 2.0 * vr * e : 2.0 * I * vi * e;
      }
    }
    cd im = 1.0;  // i^m
    for (int m = 0; m < F; ++m, im *= I) {
      double s = 1.0;
      for (int n = m; n <= p; ++n, s *= -sgn) L[n * P1 + m] += pw.ipow[k * P1 + n] * s * im * G[m];
    }
  }
}

// out += the frame local L rotated back out of frame f
void rotate_back_add(const PlaneWave &pw, int f, const cd *L, double *out) {
  const int p = pw.p, P1 = p + 1, N = ncoef(p), RB = rotation_block_size(p);
  std::vector<double> y(N, 0.0);
  for (int n = 0; n <= p; ++n)
    for (int m = 0; m <= n; ++m) {
      y[idx(n, m)] = L[n * P1 + m].real();
      if (m > 0) y[idx(n, -m)] = L[n * P1 + m].imag();
    }
  for (int n = 0, bo = 0; n <= p; bo += (2 * n + 1) * (2 * n + 1), ++n) {
    const int k = 2 * n + 1;
    const double *blk = &pw.bwd[size_t(f) * RB + bo];
    for (int r = 0; r < k; ++r) {
      double s = 0.0;
      for (int c = 0; c < k; ++c) s += blk[r + c * k] * y[n * n + c];
      out[n * n + r] += s;
    }
  }
}

void slot_coordinates(int slot, int a[3]) {
  a[0] = slot / 36;
  a[1] = (slot / 6) % 6;
  a[2] = slot % 6;
}

void child_bits(int child, int b[3]) {
  b[0] = (child >> 2) & 1;
  b[1] = (child >> 1) & 1;
  b[2] = child & 1;
}

// Direction::of (fmm-tables/src/plane_wave/geometry.rs): first of z, y, x with |d_i| >= 2
int direction_of(const int d[3]) {
  if (d[2] >= 2) return 0;
  if (d[2] <= -2) return 1;
  if (d[1] >= 2) return 2;
  if (d[1] <= -2) return 3;
  if (d[0] >= 2) return 4;
  if (d[0] <= -2) return 5;
  return -1;
}

}  // namespace

int plane_wave_nodes_for_degree(int p) {
  require(p >= 0 && p <= 20, "plane wave: the quadratures reach degree 20");
  return kDegreeNodes[p];
}

PlaneWave build_plane_wave(int p, int nodes) {
  require(p >= 0 && p <= 20, "plane wave: the quadratures reach degree 20");
  const int s = nodes == 0 ? kDegreeNodes[p] : nodes;
  require(s >= kPwMinNodes && s <= kPwMaxTableNodes, "plane wave: FMM3D's tables have 2 to 39 lambda nodes");
  PlaneWave pw;
  pw.p = p;
  pw.nodes = s;
  pw.max_error = kPwQuadMaxError[s - 2];
  const int t0 = table_start(s), P1 = p + 1;
  for (int k = 0; k < s; ++k) {
    pw.lambda.push_back(kPwQuadLambda[t0 + k]);
    pw.weight.push_back(kPwQuadWeight[t0 + k]);
    const int N = kPwQuadAlpha[t0 + k];
    pw.fmm3d_alpha.push_back(N);
    pw.nalpha.push_back(N + (N & 1));
    pw.fourier.push_back(std::min(kPwQuadFourier[t0 + k], P1));
  }
  for (int k = 0; k < s; ++k) {
    const int N = pw.nalpha[k];
    pw.first.push_back(pw.H);
    for (int j = 0; j < N / 2; ++j) {
      const double a = 2.0 * kPi * j / N;
      pw.hl.push_back(pw.lambda[k]);
      pw.hw.push_back(pw.weight[k] / N);
      pw.hc.push_back(std::cos(a));
      pw.hs.push_back(std::sin(a));
      pw.hk.push_back(k);
    }
    pw.H += N / 2;
    double ln = 1.0;  // lambda^n
    for (int n = 0; n <= p; ++n, ln *= pw.lambda[k]) {
      pw.opow.push_back(std::ldexp(ln, -(n + 1)));
      pw.ipow.push_back(std::ldexp(ln, -n));
    }
  }
  const int RB = rotation_block_size(p);
  pw.fwd.assign(size_t(3) * RB, 0.0);
  pw.bwd.assign(size_t(3) * RB, 0.0);
  for (int f = 0; f < 3; ++f) {
    double Qt[9];
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) Qt[3 * r + c] = kFrames[f][3 * c + r];
    rotation_blocks_of(p, kFrames[f], false, &pw.fwd[size_t(f) * RB]);
    rotation_blocks_of(p, Qt, true, &pw.bwd[size_t(f) * RB]);
  }
  return pw;
}

void plane_wave_rotate(int dir, const double d[3], double dp[3]) {
  const double *Q = kFrames[dir / 2];
  for (int r = 0; r < 3; ++r) dp[r] = Q[3 * r] * d[0] + Q[3 * r + 1] * d[1] + Q[3 * r + 2] * d[2];
}

void plane_wave_direction(const int d[3], int &frame, int &up, double dp[3]) {
  const int dir = direction_of(d);
  require(dir >= 0, "plane wave: not a V offset");
  frame = dir / 2;
  up = (dir & 1) ? -1 : 1;
  const double dd[3] = {double(d[0]), double(d[1]), double(d[2])};
  plane_wave_rotate(dir, dd, dp);
}

void plane_wave_shift(const PlaneWave &pw, int dir, const double d[3], double *out) {
  double dp[3];
  plane_wave_rotate(dir, d, dp);
  const double along = (dir & 1) ? dp[2] : -dp[2];
  for (int h = 0; h < pw.H; ++h) {
    const double lam = pw.hl[h], amp = std::exp(lam * along);
    const double ph = lam * (dp[0] * pw.hc[h] + dp[1] * pw.hs[h]);
    out[h] = amp * std::cos(ph);
    out[pw.H + h] = amp * std::sin(ph);
  }
}

int plane_wave_slot_direction(int slot, int child) {
  int a[3], b[3], d[3];
  slot_coordinates(slot, a);
  child_bits(child, b);
  for (int i = 0; i < 3; ++i) d[i] = b[i] + 2 - a[i];
  if (std::abs(d[0]) <= 1 && std::abs(d[1]) <= 1 && std::abs(d[2]) <= 1) return -1;
  const int axes[3] = {2, 1, 0};
  for (int q = 0; q < 3; ++q) {
    if (a[axes[q]] == 0) return 2 * q;
    if (a[axes[q]] == 5) return 2 * q + 1;
  }
  return direction_of(d);
}

const std::array<std::vector<PwAccumulator>, kPwDirections> &plane_wave_accumulators() {
  static std::array<std::vector<PwAccumulator>, kPwDirections> acc;
  static std::once_flag once;
  std::call_once(once, [] {
    for (int dir = 0; dir < kPwDirections; ++dir) {
      std::vector<std::pair<int, int>> masks;  // (mask, slot)
      for (int slot = 0; slot < kPwSlots; ++slot) {
        int mask = 0;
        for (int c = 0; c < kPwChildren; ++c)
          if (plane_wave_slot_direction(slot, c) == dir) mask |= 1 << c;
        if (mask) masks.push_back({mask, slot});
      }
      std::sort(masks.begin(), masks.end());
      for (const auto &ms : masks) {
        if (acc[dir].empty() || acc[dir].back().mask != ms.first) {
          acc[dir].push_back({static_cast<uint8_t>(ms.first), {}});
        }
        acc[dir].back().slots.push_back(ms.second);
      }
    }
  });
  return acc;
}

const PwMergeGeometry &plane_wave_merge_geometry() {
  static PwMergeGeometry g;
  static std::once_flag once;
  std::call_once(once, [] {
    const auto &accs = plane_wave_accumulators();
    g.acc_first.push_back(0);
    g.ent_first.push_back(0);
    std::map<std::array<double, 4>, int> shift_of;
    for (int dir = 0; dir < kPwDirections; ++dir) {
      require(accs[dir].size() < size_t(kPwAccStride), "plane wave: too many accumulators", kUnsupportedLayout);
      for (const PwAccumulator &a : accs[dir]) {
        g.acc_mask.push_back(a.mask);
        for (int slot : a.slots) {
          const double d[3] = {2.5 - slot / 36, 2.5 - (slot / 6) % 6, 2.5 - slot % 6};
          double dp[3];
          plane_wave_rotate(dir, d, dp);
          const std::array<double, 4> key{double(dir & 1), dp[0], dp[1], dp[2]};
          auto it = shift_of.find(key);
          if (it == shift_of.end()) {
            it = shift_of.emplace(key, int(shift_of.size())).first;
            g.shift_dir.push_back(dir);
            g.shift_d.push_back({d[0], d[1], d[2]});
          }
          g.ent_slot.push_back(slot);
          g.ent_shift.push_back(it->second);
        }
        g.ent_first.push_back(static_cast<int>(g.ent_slot.size()));
      }
      g.acc_first.push_back(static_cast<int>(g.acc_mask.size()));
    }
    require(g.shift_d.size() <= 128, "plane wave: too many source shifts", kUnsupportedLayout);
  });
  return g;
}

PwMergeLists plane_wave_merge_lists(const Plan &plan, const std::vector<uint32_t> &box_base,
                                    const std::vector<char> &use, bool per_level) {
  constexpr uint32_t kNoBox = 0xffffffffu;
  const PwMergeGeometry &g = plane_wave_merge_geometry();
  const int nlevels = plan.nlevels;
  const uint32_t box0 = nlevels > 2 ? box_base[2] : 0;
  require(nlevels <= 2 || uint64_t(box_base[nlevels] - box0) * 6 < (uint64_t(1) << 25),
          "plane wave: too many boxes for the merge's entry format", kTooLarge);
  PwMergeLists out;
  out.par_level.assign(nlevels + 1, 0);
  out.tgt_level.assign(nlevels + 1, 0);
  for (int l = 0; l < nlevels; ++l) {
    out.par_level[l] = static_cast<uint32_t>(out.children.size() / kPwChildren);
    out.tgt_level[l] = static_cast<uint32_t>(out.targets.size());
    if (l < 2 || !use[l]) continue;
    const Csr &V = plan.v[l];
    const uint32_t base = per_level ? box_base[l] : box0;
    for (uint32_t P = 0; P < plan.nboxes(l - 1); ++P) {
      const uint32_t c0 = plan.first_child[l - 1][P];
      if (c0 == kNone) continue;
      const Cell cp = decode(plan.keys[l - 1][P]);
      std::array<uint32_t, kPwSlots> sl;
      std::array<uint32_t, kPwChildren> ch;
      sl.fill(kNoBox);
      ch.fill(kNoBox);
      uint64_t entries = 0;
      for (uint32_t b = c0; b < c0 + kPwChildren; ++b) {
        const Cell cc = decode(plan.keys[l][b]);
        const int64_t bx = cc.i - 2 * cp.i, by = cc.j - 2 * cp.j, bz = cc.k - 2 * cp.k;
        require(bx >= 0 && bx <= 1 && by >= 0 && by <= 1 && bz >= 0 && bz <= 1 && plan.parent[l][b] == P,
                "plane wave: a parent's children are not its eight octants", kUnsupportedLayout);
        ch[4 * bx + 2 * by + bz] = box_base[l] + b;
        for (uint32_t e = V.off[b]; e < V.off[b + 1]; ++e) {
          const Cell cs = decode(plan.keys[l][V.ent[e]]);
          const int64_t ax = cs.i - 2 * cp.i + 2, ay = cs.j - 2 * cp.j + 2, az = cs.k - 2 * cp.k + 2;
          require(ax >= 0 && ax < 6 && ay >= 0 && ay < 6 && az >= 0 && az < 6,
                  "plane wave: a V entry outside the parent's slots", kUnsupportedLayout);
          const uint32_t src = box_base[l] + V.ent[e];
          uint32_t &slot = sl[36 * ax + 6 * ay + az];
          require(slot == kNoBox || slot == src, "plane wave: two boxes in one slot", kUnsupportedLayout);
          slot = src;
          ++entries;
        }
      }
      if (entries == 0) continue;
      uint64_t pairs = 0;
      for (int a = 0; a < kPwSlots; ++a)
        for (int c = 0; sl[a] != kNoBox && c < kPwChildren; ++c)
          if (ch[c] != kNoBox && plane_wave_slot_direction(a, c) >= 0) ++pairs;
      require(pairs == entries, "plane wave: the V lists are not the merge-and-shift pairs", kUnsupportedLayout);
      // the held entries of each direction's accumulators, in the accumulators' order
      for (int dir = 0; dir < kPwDirections; ++dir) {
        const size_t at = out.pe_first.size();
        out.pe_first.resize(at + kPwAccStride, 0);
        for (int a = g.acc_first[dir]; a < g.acc_first[dir + 1]; ++a) {
          out.pe_first[at + (a - g.acc_first[dir])] = static_cast<uint32_t>(out.pe.size());
          for (int e = g.ent_first[a]; e < g.ent_first[a + 1]; ++e) {
            const uint32_t src = sl[g.ent_slot[e]];
            if (src == kNoBox) continue;
            out.pe.push_back(((src - base) * 6 + dir) << 7 | static_cast<uint32_t>(g.ent_shift[e]));
          }
        }
        for (int a = g.acc_first[dir + 1] - g.acc_first[dir]; a < kPwAccStride; ++a)
          out.pe_first[at + a] = static_cast<uint32_t>(out.pe.size());
      }
      out.children.insert(out.children.end(), ch.begin(), ch.end());
      for (uint32_t b = c0; b < c0 + kPwChildren; ++b) out.targets.push_back(box_base[l] + b);
    }
  }
  out.par_level[nlevels] = static_cast<uint32_t>(out.children.size() / kPwChildren);
  out.tgt_level[nlevels] = static_cast<uint32_t>(out.targets.size());
  return out;
}

void plane_wave_m2l_host(const PlaneWave &pw, const int d[3], int dir, const double *in, double *out) {
  const int p = pw.p, P1 = p + 1, N = ncoef(p), H = pw.H;
  if (dir < 0) {
    dir = direction_of(d);
    require(dir >= 0, "plane wave: not a V offset");
  }
  const int f = dir / 2;
  const double sgn = (dir & 1) ? -1.0 : 1.0;
  std::vector<double> x(N), X(2 * size_t(H)), T(2 * size_t(H));
  rotate_into(pw, f, in, x.data());
  waves_of(pw, x.data(), sgn, X.data());
  const double dd[3] = {double(d[0]), double(d[1]), double(d[2])};
  plane_wave_shift(pw, dir, dd, T.data());
  for (int h = 0; h < H; ++h) {
    const double xr = X[h], xi = X[H + h], tr = T[h], ti = T[H + h];
    X[h] = xr * tr - xi * ti;
    X[H + h] = xr * ti + xi * tr;
  }
  std::vector<cd> L(size_t(P1) * P1, 0.0);
  local_add(pw, X.data(), sgn, L.data());
  rotate_back_add(pw, f, L.data(), out);
}

void plane_wave_merge_host(const PlaneWave &pw, const double *const *slots, double *out) {
  const int p = pw.p, P1 = p + 1, N = ncoef(p), H = pw.H, H2 = 2 * H;
  const auto &acc = plane_wave_accumulators();
  std::vector<double> x(N), X(H2), T(H2), sum(H2), vc(size_t(kPwChildren) * H2);
  std::vector<cd> L(size_t(kPwChildren) * P1 * P1);
  for (int f = 0; f < 3; ++f) {
    std::fill(L.begin(), L.end(), cd(0.0));
    for (int ud = 0; ud < 2; ++ud) {
      const int dir = 2 * f + ud;
      const double sgn = ud ? -1.0 : 1.0;
      std::fill(vc.begin(), vc.end(), 0.0);
      // merge: each accumulator sums its slots' waves shifted to the parent's centre
      for (const PwAccumulator &a : acc[dir]) {
        std::fill(sum.begin(), sum.end(), 0.0);
        for (int slot : a.slots) {
          if (!slots[slot]) continue;
          int c3[3];
          slot_coordinates(slot, c3);
          const double d[3] = {2.5 - c3[0], 2.5 - c3[1], 2.5 - c3[2]};
          rotate_into(pw, f, slots[slot], x.data());
          waves_of(pw, x.data(), sgn, X.data());
          plane_wave_shift(pw, dir, d, T.data());
          for (int h = 0; h < H; ++h) {
            sum[h] += X[h] * T[h] - X[H + h] * T[H + h];
            sum[H + h] += X[h] * T[H + h] + X[H + h] * T[h];
          }
        }
        for (int c = 0; c < kPwChildren; ++c)
          if (a.mask >> c & 1)
            for (int h = 0; h < H2; ++h) vc[size_t(c) * H2 + h] += sum[h];
      }
      // gather: each child's sum shifted from the parent's centre to the child, to its local
      for (int c = 0; c < kPwChildren; ++c) {
        int b[3];
        child_bits(c, b);
        const double d[3] = {b[0] - 0.5, b[1] - 0.5, b[2] - 0.5};
        plane_wave_shift(pw, dir, d, T.data());
        double *V = &vc[size_t(c) * H2];
        for (int h = 0; h < H; ++h) {
          const double vr = V[h], vi = V[H + h];
          V[h] = vr * T[h] - vi * T[H + h];
          V[H + h] = vr * T[H + h] + vi * T[h];
        }
        local_add(pw, V, sgn, &L[size_t(c) * P1 * P1]);
      }
    }
    for (int c = 0; c < kPwChildren; ++c) rotate_back_add(pw, f, &L[size_t(c) * P1 * P1], out + size_t(c) * N);
  }
}

}  // namespace fmmh
