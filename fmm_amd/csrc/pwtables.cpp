// pwtables.cpp -- plane-wave M2L: quadrature, frames and the host restatement of a pair
// (docs/CONVENTIONS.md:1305-1434; fmm-kernels/src/plane_wave.rs:779-1159 and
// fmm-tables/src/plane_wave/mod.rs:111 in the reference).
//
// Steps of one V pair (target cell - source cell = d, boxes), in the pair's frame f:
//   multipole rotated into the frame, in box widths Mh_n^m = 2^-(n+1) M~_n^m;
//   A_k(m) = sum_{n>=m} lambda_k^n s^(n-m) Mh_n^m  (s = +1 up, -1 down);
//   X(k, j) = (w_k / N_k) [A_k(0) + sum_{m>=1} 2 (-i)^m Re(A_k(m) e^{i m a_j})], j < N_k / 2;
//   V = X e^{-lambda_k |d'_z| + i lambda_k (d'_x cos a_j + d'_y sin a_j)}   (d' = Q_f d);
//   G_k(m) = 2 sum_j Re V e^{-i m a_j} (m even), 2 i sum_j Im V e^{-i m a_j} (m odd);
//   L~'_n^m += 2^-n lambda_k^n (-s)^(n-m) i^m G_k(m);   local rotated back out of the frame.
// Real charges make X(a + pi) = conj X(a), so half the a nodes carry the waves.
#include "pwtables.hpp"

#include <algorithm>
#include <cmath>
#include <complex>

#include "common.hpp"
#include "plan.hpp"

namespace fmmh {
namespace {

using cd = std::complex<double>;
constexpr double kPi = 3.141592653589793;
inline int idx(int n, int m) { return n * n + n + m; }

void gauss_legendre(int n, double a, double b, std::vector<double> &x, std::vector<double> &w) {
  x.resize(n);
  w.resize(n);
  for (int i = 0; i < n; ++i) {
    double t = std::cos(kPi * (i + 0.75) / (n + 0.5)), dp = 1.0;
    for (int it = 0; it < 100; ++it) {
      double p0 = 1.0, p1 = t;
      for (int k = 2; k <= n; ++k) {
        const double p2 = ((2.0 * k - 1.0) * t * p1 - (k - 1.0) * p0) / k;
        p0 = p1;
        p1 = p2;
      }
      dp = n * (t * p1 - p0) / (t * t - 1.0);
      const double dt = p1 / dp;
      t -= dt;
      if (std::fabs(dt) < 1e-16) break;
    }
    x[i] = a + 0.5 * (b - a) * (t + 1.0);
    w[i] = (b - a) / ((1.0 - t * t) * dp * dp);
  }
}

// Householder QR with column pivoting: the order in which the columns of A (m x n,
// column-major) are pivoted.
std::vector<int> pivot_order(std::vector<double> A, int m, int n) {
  std::vector<int> perm(n);
  for (int j = 0; j < n; ++j) perm[j] = j;
  std::vector<double> norms(n);
  for (int j = 0; j < n; ++j) {
    double s = 0.0;
    for (int i = 0; i < m; ++i) s += A[i + size_t(j) * m] * A[i + size_t(j) * m];
    norms[j] = s;
  }
  const int steps = std::min(m, n);
  for (int c = 0; c < steps; ++c) {
    int best = c;
    for (int j = c + 1; j < n; ++j)
      if (norms[j] > norms[best]) best = j;
    if (best != c) {
      for (int i = 0; i < m; ++i) std::swap(A[i + size_t(c) * m], A[i + size_t(best) * m]);
      std::swap(norms[c], norms[best]);
      std::swap(perm[c], perm[best]);
    }
    double s = 0.0;
    for (int i = c; i < m; ++i) s += A[i + size_t(c) * m] * A[i + size_t(c) * m];
    const double nrm = std::sqrt(s);
    if (nrm == 0.0) break;
    const double a0 = A[c + size_t(c) * m], alpha = a0 > 0 ? -nrm : nrm;
    std::vector<double> v(m, 0.0);
    for (int i = c; i < m; ++i) v[i] = A[i + size_t(c) * m];
    v[c] -= alpha;
    double vv = 0.0;
    for (int i = c; i < m; ++i) vv += v[i] * v[i];
    for (int j = c; j < n; ++j) {
      double t = 0.0;
      for (int i = c; i < m; ++i) t += v[i] * A[i + size_t(j) * m];
      t *= 2.0 / vv;
      for (int i = c; i < m; ++i) A[i + size_t(j) * m] -= t * v[i];
    }
    for (int j = c + 1; j < n; ++j) {
      const double r = A[c + size_t(j) * m];
      norms[j] = std::max(0.0, norms[j] - r * r);
    }
  }
  return perm;
}

// min |A x - b| for A m x k (column-major) by Householder QR.
std::vector<double> least_squares(std::vector<double> A, int m, int k, std::vector<double> b) {
  for (int c = 0; c < k; ++c) {
    double s = 0.0;
    for (int i = c; i < m; ++i) s += A[i + size_t(c) * m] * A[i + size_t(c) * m];
    const double nrm = std::sqrt(s);
    const double a0 = A[c + size_t(c) * m], alpha = a0 > 0 ? -nrm : nrm;
    std::vector<double> v(m, 0.0);
    for (int i = c; i < m; ++i) v[i] = A[i + size_t(c) * m];
    v[c] -= alpha;
    double vv = 0.0;
    for (int i = c; i < m; ++i) vv += v[i] * v[i];
    if (vv == 0.0) continue;
    for (int j = c; j < k; ++j) {
      double t = 0.0;
      for (int i = c; i < m; ++i) t += v[i] * A[i + size_t(j) * m];
      t *= 2.0 / vv;
      for (int i = c; i < m; ++i) A[i + size_t(j) * m] -= t * v[i];
    }
    double t = 0.0;
    for (int i = c; i < m; ++i) t += v[i] * b[i];
    t *= 2.0 / vv;
    for (int i = c; i < m; ++i) b[i] -= t * v[i];
  }
  std::vector<double> x(k, 0.0);
  for (int r = k - 1; r >= 0; --r) {
    double s = b[r];
    for (int q = r + 1; q < k; ++q) s -= A[r + size_t(q) * m] * x[q];
    x[r] = s / A[r + size_t(r) * m];
  }
  return x;
}

// sample points of the region 1 <= z <= 4, 0 <= rho <= sqrt(32) (Chebyshev in both)
void region(int nz, int nr, double shift, std::vector<double> &z, std::vector<double> &rho) {
  z.clear();
  rho.clear();
  const double rmax = std::sqrt(32.0);
  for (int a = 0; a < nz; ++a)
    for (int b = 0; b < nr; ++b) {
      const double tz = 0.5 * (1.0 - std::cos(kPi * (a + shift) / (nz - 1 + 2 * shift)));
      const double tr = 0.5 * (1.0 - std::cos(kPi * (b + shift) / (nr - 1 + 2 * shift)));
      z.push_back(1.0 + 3.0 * tz);
      rho.push_back(rmax * tr);
    }
}

double max_rel_error(const std::vector<double> &lam, const std::vector<double> &w, const std::vector<double> &z,
                     const std::vector<double> &rho) {
  double worst = 0.0;
  for (size_t i = 0; i < z.size(); ++i) {
    double s = 0.0;
    for (size_t k = 0; k < lam.size(); ++k) s += w[k] * std::exp(-lam[k] * z[i]) * std::cyl_bessel_j(0.0, lam[k] * rho[i]);
    const double r = std::sqrt(z[i] * z[i] + rho[i] * rho[i]);
    worst = std::max(worst, std::fabs(s * r - 1.0));
  }
  return worst;
}

const double kFrames[3][9] = {
    {1, 0, 0, 0, 1, 0, 0, 0, 1},
    {1, 0, 0, 0, 0, -1, 0, 1, 0},   // (x, y, z) -> (x, -z, y)
    {0, 0, -1, 0, 1, 0, 1, 0, 0},   // (x, y, z) -> (-z, y, x)
};

}  // namespace

PlaneWave build_plane_wave(int p, double eps) {
  require(p >= 0 && p <= 30, "plane wave: p must lie in 0..30");
  require(eps > 1e-14 && eps < 1e-1, "plane wave: tolerance out of range");
  PlaneWave pw;
  pw.p = p;
  pw.eps = eps;
  // fine rule on [0, Lambda], Lambda past the e^{-lambda z} tail (z >= 1, |u| <= sqrt 48)
  const double Lambda = std::log(20.0 / eps);
  std::vector<double> z, rho, zv, rv;
  region(28, 28, 0.0, z, rho);
  region(37, 41, 0.37, zv, rv);
  std::vector<double> fx, fw;
  int nf = 60;
  for (;; nf += 20) {
    gauss_legendre(nf, 0.0, Lambda, fx, fw);
    if (max_rel_error(fx, fw, z, rho) < 0.05 * eps || nf >= 400) break;
  }
  const int m = int(z.size());
  std::vector<double> A(size_t(m) * nf);
  for (int i = 0; i < m; ++i) {
    const double r = std::sqrt(z[i] * z[i] + rho[i] * rho[i]);
    for (int k = 0; k < nf; ++k) A[i + size_t(k) * m] = r * std::exp(-fx[k] * z[i]) * std::cyl_bessel_j(0.0, fx[k] * rho[i]);
  }
  const std::vector<int> order = pivot_order(A, m, nf);
  std::vector<double> lam, w;
  for (int K = 4; K <= nf; ++K) {
    std::vector<double> As(size_t(m) * K);
    std::vector<double> l(K);
    for (int c = 0; c < K; ++c) {
      l[c] = fx[order[c]];
      for (int i = 0; i < m; ++i) As[i + size_t(c) * m] = A[i + size_t(order[c]) * m];
    }
    const std::vector<double> x = least_squares(As, m, K, std::vector<double>(m, 1.0));
    if (max_rel_error(l, x, z, rho) < 0.25 * eps && max_rel_error(l, x, zv, rv) < eps) {
      lam = l;
      w = x;
      break;
    }
  }
  require(!lam.empty(), "plane wave: no quadrature reached the tolerance", kUnsupportedLayout);
  // nodes ascending
  std::vector<int> srt(lam.size());
  for (size_t i = 0; i < srt.size(); ++i) srt[i] = int(i);
  std::sort(srt.begin(), srt.end(), [&](int a, int b) { return lam[a] < lam[b]; });
  pw.nodes = int(lam.size());
  for (int i : srt) {
    pw.lambda.push_back(lam[i]);
    pw.weight.push_back(w[i]);
  }
  // alpha counts: J_q(lambda |d_perp|) for |d_perp| <= sqrt 18 resolved past the 2p of the
  // order pairs: N_k = 2p + 2 + n0, J_n(beta) < eps / 100 for n >= n0 > beta
  for (int k = 0; k < pw.nodes; ++k) {
    const double beta = pw.lambda[k] * std::sqrt(18.0);
    int n0 = int(std::ceil(beta));
    while (std::fabs(std::cyl_bessel_j(double(n0), beta)) > 0.01 * eps) ++n0;
    int N = 2 * p + 2 + n0;
    N += N & 1;
    pw.nalpha.push_back(N);
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

void plane_wave_direction(const int d[3], int &frame, int &up, double dp[3]) {
  const int axis_of_frame[3] = {2, 1, 0};
  frame = -1;
  for (int f = 0; f < 3; ++f)
    if (std::abs(d[axis_of_frame[f]]) >= 2) {
      frame = f;
      break;
    }
  require(frame >= 0, "plane wave: not a V offset");
  for (int r = 0; r < 3; ++r) dp[r] = kFrames[frame][3 * r] * d[0] + kFrames[frame][3 * r + 1] * d[1] + kFrames[frame][3 * r + 2] * d[2];
  up = dp[2] > 0 ? 1 : -1;
}

void plane_wave_m2l_host(const PlaneWave &pw, const int d[3], const double *in, double *out) {
  const int p = pw.p, N = ncoef(p), RB = rotation_block_size(p);
  int f, up;
  double dp[3];
  plane_wave_direction(d, f, up, dp);
  // rotate the multipole into the frame
  std::vector<double> x(N, 0.0), y(N, 0.0);
  for (int n = 0, bo = 0; n <= p; bo += (2 * n + 1) * (2 * n + 1), ++n) {
    const int k = 2 * n + 1;
    const double *blk = &pw.fwd[size_t(f) * RB + bo];
    for (int r = 0; r < k; ++r) {
      double s = 0.0;
      for (int c = 0; c < k; ++c) s += blk[r + c * k] * in[n * n + c];
      x[n * n + r] = s;
    }
  }
  // multipole in box widths, complex per order m >= 0
  std::vector<cd> Mh(size_t(p + 1) * (p + 1), 0.0);  // [n][m]
  for (int n = 0; n <= p; ++n) {
    const double sc = std::ldexp(1.0, -(n + 1));
    for (int m = 0; m <= n; ++m)
      Mh[n * (p + 1) + m] = sc * cd(x[idx(n, m)], m > 0 ? x[idx(n, -m)] : 0.0);
  }
  const double sgn = up > 0 ? 1.0 : -1.0;
  std::vector<cd> L(size_t(p + 1) * (p + 1), 0.0);  // [n][m] in the frame
  const cd I(0.0, 1.0);
  for (int k = 0; k < pw.nodes; ++k) {
    const double lam = pw.lambda[k];
    std::vector<cd> A(p + 1, 0.0), G(p + 1, 0.0);
    for (int m = 0; m <= p; ++m) {
      double ln = std::pow(lam, m), s = 1.0;
      for (int n = m; n <= p; ++n, ln *= lam, s *= sgn) A[m] += ln * s * Mh[n * (p + 1) + m];
    }
    const int Nk = pw.nalpha[k];
    for (int j = 0; j < Nk / 2; ++j) {
      const double a = 2.0 * kPi * j / Nk;
      cd X = A[0];
      cd mi = 1.0;  // (-i)^m
      for (int m = 1; m <= p; ++m) {
        mi *= -I;
        X += 2.0 * mi * std::real(A[m] * std::exp(I * (m * a)));
      }
      X *= pw.weight[k] / Nk;
      const cd T = std::exp(-lam * std::fabs(dp[2])) * std::exp(I * (lam * (dp[0] * std::cos(a) + dp[1] * std::sin(a))));
      const cd V = X * T;
      for (int m = 0; m <= p; ++m) {
        const cd e = std::exp(-I * (m * a));
        G[m] += (m % 2 == 0) ? 2.0 * V.real() * e : 2.0 * I * V.imag() * e;
      }
    }
    cd im = 1.0;  // i^m
    for (int m = 0; m <= p; ++m, im *= I) {
      double ln = std::ldexp(std::pow(lam, m), -m), s = 1.0;
      for (int n = m; n <= p; ++n, ln *= 0.5 * lam, s *= -sgn) L[n * (p + 1) + m] += ln * s * im * G[m];
    }
  }
  for (int n = 0; n <= p; ++n)
    for (int m = 0; m <= n; ++m) {
      y[idx(n, m)] = L[n * (p + 1) + m].real();
      if (m > 0) y[idx(n, -m)] = L[n * (p + 1) + m].imag();
    }
  // rotate the local back and add
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

}  // namespace fmmh
