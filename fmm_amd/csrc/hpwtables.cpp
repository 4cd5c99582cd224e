// hpwtables.cpp -- the Helmholtz exponential M2L of a level: FMM3D's quadrature mapped to
// kappa, its counts, the conversion tables and the host restatements (docs/CONVENTIONS.md
// §3.15 "Exponential M2L"; fmm-tables/src/helmholtz/plane_wave/{quadrature,pair,merge}.rs).
//
// Steps of one V pair along direction (frame f, up or down), stored forms of the level:
//   the multipole (re and im planes) rotated into the frame, per degree O^n(Q_f) x;
//   modes A_k(m) = sum_n P^out_k(n, m) M^{+m}_n, B_k(m) = the same with M^{-m}_n, m <= F_k' =
//   min(F_k, p), up the sum over n, down with (-1)^(n-m);
//   waves X(k, j) = sum_m A_k(m) cos m a_j + B_k(m) sin m a_j;
//   V = X e^{-+lambda_k d'_z + i t_k (d'_x cos a_j + d'_y sin a_j)} (d' = Q_f d; up: -);
//   L'^{+m}_n += P^in_k(n, m) sum_j V cos m a_j, L'^{-m}_n the same with sin, (-1)^(n-m) more
//   for down waves; the local rotated back by O^n(Q_f)^T.
// P^out = 2 pi e^beta (w_k / N_k) P_k, P^in = 4 pi e^beta (-1)^n P_k, P_k(n, m) = i^-n c_nm
// Pbar_n^m((s/kappa) i lambda_k, (s/kappa) t_k; s^2) (the solid normalised Legendre
// polynomials, a polynomial in their arguments, finite as kappa -> 0).
#include "hpwtables.hpp"

#include <algorithm>
#include <cmath>
#include <mutex>

#include "common.hpp"
#include "htables.hpp"
#include "pwtables.hpp"

namespace fmmh {
namespace {

#include "hpwquadrature.inc"

constexpr double kPi = 3.141592653589793;
constexpr double kSourceReach = 0.7071067811865476;  // sqrt 2 / 2: a point's lateral reach from its box centre
constexpr double kRhoMax = 5.656854249492381;        // 4 sqrt 2: the largest lateral offset of a V pair's points

// the domains' first nodes in kHpwNode / kHpwWeight
const std::vector<int> &domain_first() {
  static std::vector<int> first;
  static std::once_flag once;
  std::call_once(once, [] {
    first.assign(kHpwDomains + 1, 0);
    for (int d = 0; d < kHpwDomains; ++d) first[d + 1] = first[d] + kHpwCount[d];
  });
  return first;
}

// the breakpoints of Re kappa (hwts3dgetd): 0, 0.02, 0.04, 0.1, 0.2, 0.4, 1, pi/2, then pi to 16 pi
double re_break(int i) {
  static const double head[8] = {0.0, 0.02, 0.04, 0.10, 0.20, 0.40, 1.00, kPi / 2.0};
  return i < 8 ? head[i] : (i - 7) * kPi;
}

// the fewest n >= 1 with scale max_{m >= n} max_{rho <= reach} |J_m(t rho)| <= share, the
// maximum over rho at 64 equal steps
int tail_count(cplx t, double reach, double scale, double share) {
  const double x = std::abs(t) * reach;
  const int top = int(std::ceil(x)) * 2 + 80;
  std::vector<double> worst(top + 1, 0.0);
  for (int i = 1; i <= 64; ++i) {
    const std::vector<double> j = hpw_bessel_j_norms(t * (reach * i / 64.0), top);
    for (int m = 0; m <= top; ++m) worst[m] = std::max(worst[m], j[m]);
  }
  for (int m = top - 1; m >= 0; --m) worst[m] = std::max(worst[m], worst[m + 1]);
  for (int n = 1; n < top; ++n)
    if (scale * worst[n] <= share) return n;
  return top;
}

// c_nm Pbar_n^m(z, rho; r2) at n (n + 1) / 2 + m: the polar factors of the orthonormal solid
// harmonics at x = (rho cos a, rho sin a, z), r2 = x . x (no Condon-Shortley phase), by the
// recursion in n per order m.
void harmonic_factors(int p, cplx z, cplx rho, cplx r2, std::vector<cplx> &out) {
  out.assign(size_t(p + 1) * (p + 2) / 2, 0.0);
  cplx diag = 1.0;
  for (int m = 0; m <= p; ++m) {
    if (m > 0) diag *= rho * std::sqrt((2.0 * m - 1.0) / (2.0 * m));
    cplx prev = 0.0, cur = diag;
    for (int n = m; n <= p; ++n) {
      if (n == m + 1) {
        prev = cur;
        cur = z * cur * std::sqrt(2.0 * m + 1.0);
      } else if (n > m + 1) {
        const double cn = std::sqrt(double(n - m) * (n + m)), cn1 = std::sqrt(double(n - 1 - m) * (n - 1 + m));
        const cplx next = (z * cur * double(2 * n - 1) - r2 * prev * cn1) / cn;
        prev = cur;
        cur = next;
      }
      out[size_t(n) * (n + 1) / 2 + m] = cur * std::sqrt((m == 0 ? 1.0 : 2.0) * (2 * n + 1) / (4.0 * kPi));
    }
  }
}

cplx i_pow(int n) {
  switch (((n % 4) + 4) % 4) {
    case 0:
      return {1.0, 0.0};
    case 1:
      return {0.0, 1.0};
    case 2:
      return {-1.0, 0.0};
    default:
      return {0.0, -1.0};
  }
}

inline int idx(int n, int m) { return n * n + n + m; }

// Direction::of: the first of z, y, x with |d_i| >= 2
int direction_of(const int d[3]) {
  if (d[2] >= 2) return 0;
  if (d[2] <= -2) return 1;
  if (d[1] >= 2) return 2;
  if (d[1] <= -2) return 3;
  if (d[0] >= 2) return 4;
  if (d[0] <= -2) return 5;
  return -1;
}

const double kFrameQ[3][9] = {
    {1, 0, 0, 0, 1, 0, 0, 0, 1},
    {1, 0, 0, 0, 0, -1, 0, 1, 0},   // (x, y, z) -> (x, -z, y)
    {0, 0, -1, 0, 1, 0, 1, 0, 0},   // (x, y, z) -> (-z, y, x)
};

}  // namespace

int HpwQuadrature::waves() const {
  int w = 0;
  for (int n : alpha) w += n;
  return w;
}

std::vector<double> hpw_bessel_j_norms(cplx z, int top) {
  std::vector<double> out(top + 1, 0.0);
  const double r = std::abs(z);
  if (r < 1e-150) {
    out[0] = 1.0;
    return out;
  }
  if (z.imag() > 0.0) z = std::conj(z);  // |J_m(conj z)| = |J_m(z)|; Im z <= 0 keeps e^{iz} large
  const int start = std::max(top, int(std::ceil(r))) + 20 + int(std::ceil(std::sqrt(40.0 * std::max(r, 1.0))));
  std::vector<cplx> val(top + 1, 0.0);
  const cplx ipow[4] = {{1, 0}, {0, 1}, {-1, 0}, {0, -1}};
  cplx above = 0.0, cur = 1e-300, sum = 0.0;
  const cplx inv = 1.0 / z;
  for (int n = start; n >= 1; --n) {
    // cur = J_n, above = J_{n+1}, unnormalised
    const cplx below = inv * (2.0 * n) * cur - above;
    if (n <= top) val[n] = cur;
    sum += ipow[n % 4] * cur;
    above = cur;
    cur = below;
    if (std::abs(cur) > 1e250) {
      above *= 1e-250;
      cur *= 1e-250;
      sum *= 1e-250;
      for (cplx &v : val) v *= 1e-250;
    }
  }
  val[0] = cur;
  // e^{iz} = J_0 + 2 sum_{n >= 1} i^n J_n fixes the scale; divided by modulus and phase
  const cplx total = cur + 2.0 * sum;
  const double size = std::abs(total);
  const cplx norm = std::exp(cplx(0.0, 1.0) * z) * std::conj(total / size) / size;
  for (int m = 0; m <= top; ++m) out[m] = std::abs(val[m] * norm);
  return out;
}

int hpw_domain(cplx kappa) {
  const double re = kappa.real(), im = kappa.imag();
  if (!(re >= 0.0 && std::isfinite(re) && im >= 0.0 && std::isfinite(im))) return 0;
  int ix = 0, iy = 0;
  while (ix < 24 && re_break(ix) <= re) ++ix;
  while (iy < 20 && re_break(iy) <= im) ++iy;
  return (ix <= 23 && iy <= 19) ? ix + (iy - 1) * 23 : 0;
}

HpwQuadrature hpw_quadrature(cplx kappa) {
  HpwQuadrature q;
  q.kappa = kappa;
  q.domain = hpw_domain(kappa);
  if (q.domain == 0) return q;
  const int d = q.domain - 1, s = kHpwCount[d], f0 = domain_first()[d];
  q.stated_error = kHpwError[d];
  const cplx I(0.0, 1.0);
  for (int k = 0; k < s; ++k) {
    const double x = kHpwNode[f0 + k], w = kHpwWeight[f0 + k];
    cplx lam, wt;
    if (kHpwMap[d] == 1) {
      lam = std::sqrt(cplx(x * x, 0.0) - kappa * kappa);
      wt = w * x / lam;
    } else {
      const double u = 100.0 * x, a = std::fabs(kappa.real());
      lam = u - I * kappa + I * (a * u / (1.0 + u));
      wt = (1.0 + I * (a / ((1.0 + u) * (1.0 + u)))) * (100.0 * w);
    }
    q.lambda.push_back(lam);
    q.weight.push_back(wt);
    q.t.push_back(std::sqrt(lam * lam + kappa * kappa));
  }
  // counts: each node's share of the stated error bounds its Bessel tails
  const double share = q.stated_error / (4.0 * s);
  q.alpha.resize(s);
  q.fourier.resize(s);
#pragma omp parallel for schedule(dynamic, 1)
  for (int k = 0; k < s; ++k) {
    const double scale = 2.0 * std::abs(q.weight[k]) * std::exp(-q.lambda[k].real());
    const int f = tail_count(q.t[k], kSourceReach, scale, share) - 1;
    const int n = std::max(tail_count(q.t[k], kRhoMax, scale, share),
                           tail_count(q.t[k], kRhoMax - kSourceReach, scale, share) + f);
    q.alpha[k] = n;
    q.fourier[k] = f;
  }
  return q;
}

bool hpw_level_selected(int level, int p, cplx kh) {
  if (level < 2 || p > 15) return false;
  if (!(kh.real() > 0.0 && kh.imag() >= 0.0)) return false;
  return hpw_domain(2.0 * kh) != 0;
}

HpwLevel hpw_level(int p, cplx kh, const HpwQuadrature &q) {
  require(q.domain != 0, "helmholtz plane wave: kappa outside FMM3D's tables", kUnsupportedLayout);
  HpwLevel L;
  L.p = p;
  L.kh = kh;
  L.s = int(q.lambda.size());
  L.lambda = q.lambda;
  L.t = q.t;
  for (int k = 0; k < L.s; ++k) {
    require(q.alpha[k] >= 1, "helmholtz plane wave: an alpha count of 0", kUnsupportedLayout);
    L.start.push_back(L.W);
    L.len.push_back(q.alpha[k]);
    L.ord.push_back(std::min(q.fourier[k], p) + 1);
    L.mode_start.push_back(L.orders);
    L.trig.push_back(L.trig_size);
    L.factor.push_back(L.factor_size);
    L.W += q.alpha[k];
    L.orders += L.ord[k];
    L.trig_size += L.ord[k] * q.alpha[k];
    for (int m = 0; m < L.ord[k]; ++m) L.factor_size += p + 1 - m;
  }
  const double scale = h_scale(kh), beta = h_beta(kh), growth = std::exp(beta);
  const cplx kappa = 2.0 * kh, sk = scale / kappa, I(0.0, 1.0);
  const cplx r2(scale * scale, 0.0);
  std::vector<cplx> harm;
  for (int k = 0; k < L.s; ++k) {
    const int N = L.len[k];
    for (int j = 0; j < N; ++j) {
      L.acos.push_back(std::cos(2.0 * kPi * j / N));
      L.asin.push_back(std::sin(2.0 * kPi * j / N));
    }
    for (int m = 0; m < L.ord[k]; ++m)
      for (int j = 0; j < N; ++j) {
        const double a = 2.0 * kPi * double((m * j) % N) / N;  // m a_j reduced exactly
        L.cosm.push_back(std::cos(a));
        L.sinm.push_back(std::sin(a));
      }
    harmonic_factors(p, sk * I * L.lambda[k], sk * L.t[k], r2, harm);
    const cplx outgoing = q.weight[k] * (2.0 * kPi * growth / N);
    const double incoming = 4.0 * kPi * growth;
    for (int m = 0; m < L.ord[k]; ++m)
      for (int n = m; n <= p; ++n) {
        const cplx P = i_pow(-n) * harm[size_t(n) * (n + 1) / 2 + m];
        const cplx o = outgoing * P, in = P * (incoming * ((n & 1) ? -1.0 : 1.0));
        L.out_re.push_back(o.real());
        L.out_im.push_back(o.imag());
        L.in_re.push_back(in.real());
        L.in_im.push_back(in.imag());
      }
  }
  h_rotation_blocks(p, kFrameQ[1], L.frame1);
  h_rotation_blocks(p, kFrameQ[2], L.frame2);
  return L;
}

void hpw_shift(const HpwLevel &L, const double d[3], bool up, double *out) {
  const double sign = up ? -1.0 : 1.0;
  const cplx I(0.0, 1.0);
  for (int k = 0; k < L.s; ++k)
    for (int j = L.start[k]; j < L.start[k] + L.len[k]; ++j) {
      const double lateral = d[0] * L.acos[j] + d[1] * L.asin[j];
      const cplx e = std::exp(L.lambda[k] * (sign * d[2]) + I * L.t[k] * lateral);
      out[j] = e.real();
      out[L.W + j] = e.imag();
    }
}

void hpw_waves(const HpwLevel &L, int frame, const double *mult, double *up, double *down) {
  const int p = L.p, nc = (p + 1) * (p + 1), W = L.W;
  // the multipole in the frame: per plane and degree O^n x
  std::vector<double> x(2 * size_t(nc));
  const std::vector<double> *blk = frame == 1 ? &L.frame1 : frame == 2 ? &L.frame2 : nullptr;
  for (int pl = 0; pl < 2; ++pl) {
    const double *in = mult + size_t(pl) * nc;
    double *y = &x[size_t(pl) * nc];
    if (!blk) {
      std::copy(in, in + nc, y);
      continue;
    }
    for (int n = 0, at = 0; n <= p; at += (2 * n + 1) * (2 * n + 1), ++n) {
      const int w = 2 * n + 1;
      for (int i = 0; i < w; ++i) {
        double s = 0.0;
        for (int j = 0; j < w; ++j) s += (*blk)[at + i * w + j] * in[n * n + j];
        y[n * n + i] = s;
      }
    }
  }
  const double *re = x.data(), *im = x.data() + nc;
  for (int k = 0; k < L.s; ++k) {
    std::vector<cplx> Au(L.ord[k]), Bu(L.ord[k]), Ad(L.ord[k]), Bd(L.ord[k]);
    int f = L.factor[k];
    for (int m = 0; m < L.ord[k]; ++m) {
      cplx a[2] = {0.0, 0.0}, b[2] = {0.0, 0.0};  // even and odd n - m
      for (int n = m; n <= p; ++n, ++f) {
        const cplx P(L.out_re[f], L.out_im[f]);
        const int par = (n - m) & 1;
        a[par] += P * cplx(re[idx(n, m)], im[idx(n, m)]);
        if (m > 0) b[par] += P * cplx(re[idx(n, -m)], im[idx(n, -m)]);
      }
      Au[m] = a[0] + a[1];
      Ad[m] = a[0] - a[1];
      Bu[m] = b[0] + b[1];
      Bd[m] = b[0] - b[1];
    }
    for (int j = 0; j < L.len[k]; ++j) {
      cplx xu = 0.0, xd = 0.0;
      for (int m = 0; m < L.ord[k]; ++m) {
        const double c = L.cosm[L.trig[k] + m * L.len[k] + j], s = L.sinm[L.trig[k] + m * L.len[k] + j];
        xu += Au[m] * c + Bu[m] * s;
        xd += Ad[m] * c + Bd[m] * s;
      }
      const int v = L.start[k] + j;
      up[v] = xu.real();
      up[W + v] = xu.imag();
      down[v] = xd.real();
      down[W + v] = xd.imag();
    }
  }
}

void hpw_local_add(const HpwLevel &L, const double *V, bool up, double *local) {
  const int p = L.p, nc = (p + 1) * (p + 1), W = L.W;
  double *re = local, *im = local + nc;
  for (int k = 0; k < L.s; ++k) {
    int f = L.factor[k];
    for (int m = 0; m < L.ord[k]; ++m) {
      cplx gc = 0.0, gs = 0.0;
      for (int j = 0; j < L.len[k]; ++j) {
        const double c = L.cosm[L.trig[k] + m * L.len[k] + j], s = L.sinm[L.trig[k] + m * L.len[k] + j];
        const cplx v(V[L.start[k] + j], V[W + L.start[k] + j]);
        gc += v * c;
        gs += v * s;
      }
      for (int n = m; n <= p; ++n, ++f) {
        cplx Q(L.in_re[f], L.in_im[f]);
        if (!up && ((n - m) & 1)) Q = -Q;
        const cplx a = Q * gc;
        re[idx(n, m)] += a.real();
        im[idx(n, m)] += a.imag();
        if (m > 0) {
          const cplx b = Q * gs;
          re[idx(n, -m)] += b.real();
          im[idx(n, -m)] += b.imag();
        }
      }
    }
  }
}

void hpw_rotate_back_add(const HpwLevel &L, int frame, const double *local, double *out) {
  const int p = L.p, nc = (p + 1) * (p + 1);
  const std::vector<double> *blk = frame == 1 ? &L.frame1 : frame == 2 ? &L.frame2 : nullptr;
  for (int pl = 0; pl < 2; ++pl) {
    const double *x = local + size_t(pl) * nc;
    double *y = out + size_t(pl) * nc;
    if (!blk) {
      for (int i = 0; i < nc; ++i) y[i] += x[i];
      continue;
    }
    for (int n = 0, at = 0; n <= p; at += (2 * n + 1) * (2 * n + 1), ++n) {
      const int w = 2 * n + 1;
      for (int j = 0; j < w; ++j) {
        double s = 0.0;
        for (int i = 0; i < w; ++i) s += (*blk)[at + i * w + j] * x[n * n + i];
        y[n * n + j] += s;
      }
    }
  }
}

void hpw_m2l_host(const HpwLevel &L, const int d[3], int dir, const double *mult, double *out) {
  if (dir < 0) dir = direction_of(d);
  require(dir >= 0 && dir < 6, "helmholtz plane wave: not a V offset");
  const double dd[3] = {double(d[0]), double(d[1]), double(d[2])};
  double dp[3];
  plane_wave_rotate(dir, dd, dp);
  const bool up = (dir & 1) == 0;
  const double along = up ? dp[2] : -dp[2];
  require(along >= 2.0 && along <= 3.0, "helmholtz plane wave: the offset is not 2 or 3 boxes along the direction");
  const int nc = (L.p + 1) * (L.p + 1), W2 = 2 * L.W;
  std::vector<double> wu(W2), wd(W2), T(W2), V(W2), loc(2 * size_t(nc), 0.0);
  hpw_waves(L, dir / 2, mult, wu.data(), wd.data());
  hpw_shift(L, dp, up, T.data());
  const double *X = up ? wu.data() : wd.data();
  for (int j = 0; j < L.W; ++j) {
    const cplx v = cplx(X[j], X[L.W + j]) * cplx(T[j], T[L.W + j]);
    V[j] = v.real();
    V[L.W + j] = v.imag();
  }
  hpw_local_add(L, V.data(), up, loc.data());
  hpw_rotate_back_add(L, dir / 2, loc.data(), out);
}

void hpw_merge_host(const HpwLevel &L, const double *const *slots, double *out) {
  const int nc = (L.p + 1) * (L.p + 1), W = L.W, W2 = 2 * W;
  const auto &acc = plane_wave_accumulators();
  std::vector<double> wu(W2), wd(W2), T(W2), sum(W2), vc(size_t(kPwChildren) * W2);
  std::vector<double> loc(size_t(kPwChildren) * 2 * nc);
  for (int f = 0; f < 3; ++f) {
    std::fill(loc.begin(), loc.end(), 0.0);
    for (int ud = 0; ud < 2; ++ud) {
      const int dir = 2 * f + ud;
      const bool up = ud == 0;
      std::fill(vc.begin(), vc.end(), 0.0);
      for (const PwAccumulator &a : acc[dir]) {
        std::fill(sum.begin(), sum.end(), 0.0);
        for (int slot : a.slots) {
          if (!slots[slot]) continue;
          const double d[3] = {2.5 - slot / 36, 2.5 - (slot / 6) % 6, 2.5 - slot % 6};
          double dp[3];
          plane_wave_rotate(dir, d, dp);
          hpw_waves(L, f, slots[slot], wu.data(), wd.data());
          hpw_shift(L, dp, up, T.data());
          const double *X = up ? wu.data() : wd.data();
          for (int j = 0; j < W; ++j) {
            const cplx v = cplx(X[j], X[W + j]) * cplx(T[j], T[W + j]);
            sum[j] += v.real();
            sum[W + j] += v.imag();
          }
        }
        for (int c = 0; c < kPwChildren; ++c)
          if (a.mask >> c & 1)
            for (int j = 0; j < W2; ++j) vc[size_t(c) * W2 + j] += sum[j];
      }
      for (int c = 0; c < kPwChildren; ++c) {
        const double d[3] = {((c >> 2) & 1) - 0.5, ((c >> 1) & 1) - 0.5, (c & 1) - 0.5};
        double dp[3];
        plane_wave_rotate(dir, d, dp);
        hpw_shift(L, dp, up, T.data());
        double *V = &vc[size_t(c) * W2];
        for (int j = 0; j < W; ++j) {
          const cplx v = cplx(V[j], V[W + j]) * cplx(T[j], T[W + j]);
          V[j] = v.real();
          V[W + j] = v.imag();
        }
        hpw_local_add(L, V, up, &loc[size_t(c) * 2 * nc]);
      }
    }
    for (int c = 0; c < kPwChildren; ++c) hpw_rotate_back_add(L, f, &loc[size_t(c) * 2 * nc], out + size_t(c) * 2 * nc);
  }
}

}  // namespace fmmh
