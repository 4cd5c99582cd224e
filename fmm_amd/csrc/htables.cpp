// htables.cpp -- Helmholtz translation tables of a level: rotation, coaxial translation,
// rotation back, composed into dense complex matrices for the MFMA contraction
// (docs/CONVENTIONS.md:1622-1700, "Rotations" and "Coaxial translations"; the reference
// builds the same operators in fmm-tables/src/helmholtz/dense.rs and rcr.rs).
//
// A shift t = |t| e with e = Q^T e_z acts on coefficient vectors as T(t) = O(Q)^T A(|t|) O(Q):
// O(Q) the per-degree rotation blocks of the orthonormal real harmonics, Y_n(Q x) = O^n Y_n(x)
// (orthogonal, so the back rotation is the transpose), and A the coaxial translation, one
// (p+1-m)^2 block per order m shared by slots +m and -m.  The rotation blocks are solved in
// least squares from harmonics sampled on a spiral of points; the coaxial blocks come from
// the Gegenbauer start column
//   E^0_{n'0}(t) = (-1)^n' sqrt(2n'+1) f_n'(kh t)   (f = j: (S|S) = (R|R); f = h: (S|R))
// and the recursions in n and in m, run on the triangle n' >= n, the rest by the index
// symmetry E_{nn'} = (-1)^(n+n') E_{n'n}.  Scales of the stored form: M2M
// e^(beta_c - beta_P) s_c^n / s_P^n', L2L e^(beta_c - beta_P) s_c^n' / s_P^n times 1/2 (the
// change of units), M2L 4 pi i kh e^(2 beta) s^(n+n').
#include "htables.hpp"

#include <cmath>
#include <vector>

#include "common.hpp"

namespace fmmh {
namespace {

inline int idx(int n, int m) { return n * n + n + m; }
constexpr double kPi = 3.141592653589793;

// Racah regular harmonics (m >= 0) scaled to the orthonormal solid harmonics.
void solid(int p, const double x[3], const std::vector<double> &norm, double *out) {
  const int N = ncoef(p);
  std::vector<double> re(N, 0.0), im(N, 0.0);
  const double r2 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
  re[0] = 1.0;
  for (int m = 1; m <= p; ++m) {
    const double a = re[idx(m - 1, m - 1)], b = im[idx(m - 1, m - 1)];
    re[idx(m, m)] = (x[0] * a - x[1] * b) / (2.0 * m);
    im[idx(m, m)] = (x[0] * b + x[1] * a) / (2.0 * m);
  }
  for (int m = 0; m < p; ++m) {
    re[idx(m + 1, m)] = x[2] * re[idx(m, m)];
    im[idx(m + 1, m)] = x[2] * im[idx(m, m)];
  }
  for (int m = 0; m <= p; ++m)
    for (int n = m + 2; n <= p; ++n) {
      const double den = double((n + m) * (n - m));
      re[idx(n, m)] = ((2 * n - 1) * x[2] * re[idx(n - 1, m)] - r2 * re[idx(n - 2, m)]) / den;
      im[idx(n, m)] = ((2 * n - 1) * x[2] * im[idx(n - 1, m)] - r2 * im[idx(n - 2, m)]) / den;
    }
  for (int n = 0; n <= p; ++n) {
    out[idx(n, 0)] = norm[idx(n, 0)] * re[idx(n, 0)];
    for (int m = 1; m <= n; ++m) {
      out[idx(n, m)] = norm[idx(n, m)] * re[idx(n, m)];
      out[idx(n, -m)] = norm[idx(n, -m)] * im[idx(n, m)];
    }
  }
}

// j_n(z) (Miller's backward recurrence on g_n = j_n / z^n) and h_n(z) = h_n^(1)(z)
// (forward recurrence on e_n = h_n z^(n+1)) for n = 0 .. P.
void spherical_j(int P, cplx z, std::vector<cplx> &j) {
  const double az = std::abs(z);
  const int N = P + 30 + int(2.0 * az);
  const cplx z2 = z * z;
  std::vector<cplx> g(P + 1);
  cplx a = 0.0, b = 1.0;
  for (int n = N; n >= 1; --n) {
    const cplx c = (2.0 * n + 1.0) * b - z2 * a;
    a = b;
    b = c;
    if (n - 1 <= P) g[n - 1] = b;
    if (std::abs(b) > 1e200) {
      a *= 1e-200;
      b *= 1e-200;
      for (int k = n - 1; k <= P; ++k)
        if (k >= 0) g[k] *= 1e-200;
    }
  }
  cplx sc;
  if (az == 0.0) {
    sc = 1.0 / g[0];
  } else {
    const cplx g0 = std::sin(z) / z;
    if (az < 1.0 || P < 1) {
      sc = g0 / g[0];
    } else {
      const cplx g1 = (g0 - std::cos(z)) / z2;
      const double big = std::max(std::abs(g[0]), std::abs(g[1]));
      const cplx G0 = g[0] / big, G1 = g[1] / big;
      sc = (std::conj(G0) * g0 + std::conj(G1) * g1) / (std::norm(G0) + std::norm(G1)) / big;
    }
  }
  j.resize(P + 1);
  cplx zn = 1.0;
  for (int n = 0; n <= P; ++n, zn *= z) j[n] = g[n] * sc * zn;
}

void spherical_h(int P, cplx z, std::vector<cplx> &h) {
  const cplx I(0.0, 1.0), w = std::exp(I * z), z2 = z * z;
  std::vector<cplx> e(P + 2);
  e[0] = -I * w;
  e[1] = -(z + I) * w;
  for (int n = 1; n < P; ++n) e[n + 1] = (2.0 * n + 1.0) * e[n] - z2 * e[n - 1];
  h.resize(P + 1);
  cplx zn = z;
  for (int n = 0; n <= P; ++n, zn *= z) h[n] = e[n] / zn;
}

// Coaxial blocks E[m][n'][n] ((p+1)^3 values) from the start column f[0 .. 2p].
double ca(int n, int m) {
  return n < 0 ? 0.0 : std::sqrt(double(n + 1 + m) * (n + 1 - m) / ((2.0 * n + 1.0) * (2.0 * n + 3.0)));
}
double cb(int n, int m) {
  return n < 1 ? 0.0 : std::sqrt(double(n - m - 1) * (n - m) / ((2.0 * n - 1.0) * (2.0 * n + 1.0)));
}
double cg(int n, int m) {
  return n < 0 ? 0.0 : std::sqrt(double(n + m + 1) * (n + m + 2) / ((2.0 * n + 1.0) * (2.0 * n + 3.0)));
}

void coaxial(int p, const std::vector<cplx> &f, std::vector<cplx> &E) {
  const int R = 2 * p + 2, C = p + 1;
  std::vector<cplx> W(size_t(R) * C, 0.0), Wn(size_t(R) * C, 0.0);
  auto at = [C](std::vector<cplx> &A, int np, int n) -> cplx & { return A[size_t(np) * C + n]; };
  for (int np = 0; np <= 2 * p; ++np) at(W, np, 0) = ((np & 1) ? -1.0 : 1.0) * std::sqrt(2.0 * np + 1.0) * f[np];
  E.assign(size_t(C) * C * C, 0.0);
  for (int m = 0; m <= p; ++m) {
    if (m > 0) {
      std::fill(Wn.begin(), Wn.end(), cplx(0.0));
      for (int np = m; np <= 2 * p - m; ++np) {
        cplx v = cb(np + 1, m - 1) * at(W, np + 1, m - 1);
        if (np - 1 >= m - 1) v += cg(np - 1, m - 1) * at(W, np - 1, m - 1);
        at(Wn, np, m) = v / cg(m - 1, m - 1);
      }
      W.swap(Wn);
    }
    for (int n = m; n < p; ++n)
      for (int np = n + 1; np <= 2 * p - n - 1; ++np) {
        cplx v = -ca(np, m) * at(W, np + 1, n);
        if (n - 1 >= m) v += ca(n - 1, m) * at(W, np, n - 1);
        if (np - 1 >= m) v += ca(np - 1, m) * at(W, np - 1, n);
        at(W, np, n + 1) = v / ca(n, m);
      }
    for (int np = m; np <= p; ++np)
      for (int n = m; n <= p; ++n)
        E[(size_t(m) * C + np) * C + n] = np >= n ? at(W, np, n) : (((np + n) & 1) ? -1.0 : 1.0) * at(W, n, np);
  }
}

// Least-squares solve of X S = S' (k x k; S, S' k x K column-major) by the normal
// equations of the orthonormal harmonics (well conditioned on a spread of K >= 2k points).
void solve_block(int k, int K, const std::vector<double> &S, const std::vector<double> &Sp, double *X) {
  // G = S S^T (k x k), B = S' S^T; X = B G^-1
  std::vector<double> G(size_t(k) * k, 0.0), B(size_t(k) * k, 0.0);
  for (int s = 0; s < K; ++s)
    for (int a = 0; a < k; ++a)
      for (int b = 0; b < k; ++b) {
        G[a + size_t(b) * k] += S[a + size_t(s) * k] * S[b + size_t(s) * k];
        B[a + size_t(b) * k] += Sp[a + size_t(s) * k] * S[b + size_t(s) * k];
      }
  // Cholesky G = L L^T
  std::vector<double> L(size_t(k) * k, 0.0);
  for (int j = 0; j < k; ++j) {
    double d = G[j + size_t(j) * k];
    for (int q = 0; q < j; ++q) d -= L[j + size_t(q) * k] * L[j + size_t(q) * k];
    L[j + size_t(j) * k] = std::sqrt(d);
    for (int i = j + 1; i < k; ++i) {
      double v = G[i + size_t(j) * k];
      for (int q = 0; q < j; ++q) v -= L[i + size_t(q) * k] * L[j + size_t(q) * k];
      L[i + size_t(j) * k] = v / L[j + size_t(j) * k];
    }
  }
  // each row x of X: x G = b  <=>  G x^T = b^T (G symmetric)
  std::vector<double> y(k);
  for (int r = 0; r < k; ++r) {
    for (int i = 0; i < k; ++i) {
      double v = B[r + size_t(i) * k];
      for (int q = 0; q < i; ++q) v -= L[i + size_t(q) * k] * y[q];
      y[i] = v / L[i + size_t(i) * k];
    }
    for (int i = k - 1; i >= 0; --i) {
      double v = y[i];
      for (int q = i + 1; q < k; ++q) v -= L[q + size_t(i) * k] * X[r + size_t(q) * k];
      X[r + size_t(i) * k] = v / L[i + size_t(i) * k];
    }
  }
}

// Blocks O^n(Q) (row-major [m'+n][m+n]) of every degree, at offsets boff[n].
void rotation_blocks(int p, const double Q[9], const std::vector<double> &norm, std::vector<double> &O,
                     std::vector<int> &boff) {
  boff.assign(p + 2, 0);
  for (int n = 0; n <= p; ++n) boff[n + 1] = boff[n] + (2 * n + 1) * (2 * n + 1);
  O.assign(boff[p + 1], 0.0);
  const int K = 4 * p + 8, N = ncoef(p);
  std::vector<double> a(size_t(K) * N), b(size_t(K) * N);
  for (int s = 0; s < K; ++s) {
    const double z = 1.0 - (2.0 * s + 1.0) / K;
    const double rad = std::sqrt(1.0 - z * z), ph = 2.399963229728653 * s + 0.7;
    const double x[3] = {rad * std::cos(ph), rad * std::sin(ph), z};
    const double y[3] = {Q[0] * x[0] + Q[1] * x[1] + Q[2] * x[2], Q[3] * x[0] + Q[4] * x[1] + Q[5] * x[2],
                         Q[6] * x[0] + Q[7] * x[1] + Q[8] * x[2]};
    solid(p, x, norm, &b[size_t(s) * N]);
    solid(p, y, norm, &a[size_t(s) * N]);
  }
  for (int n = 0; n <= p; ++n) {
    const int k = 2 * n + 1;
    std::vector<double> S(size_t(k) * K), Sp(size_t(k) * K), X(size_t(k) * k);
    for (int s = 0; s < K; ++s)
      for (int c = 0; c < k; ++c) {
        S[c + size_t(s) * k] = b[size_t(s) * N + n * n + c];
        Sp[c + size_t(s) * k] = a[size_t(s) * N + n * n + c];
      }
    solve_block(k, K, S, Sp, X.data());
    for (int r = 0; r < k; ++r)
      for (int c = 0; c < k; ++c) O[boff[n] + r * k + c] = X[r + size_t(c) * k];
  }
}

// Q = R_y(-theta) R_z(-phi): the unit vector of t to e_z.
void rotation_to_z(const double t[3], double Q[9]) {
  const double r = std::sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
  const double th = std::acos(std::max(-1.0, std::min(1.0, t[2] / r))), ph = std::atan2(t[1], t[0]);
  const double ct = std::cos(th), st = std::sin(th), cp = std::cos(ph), sp = std::sin(ph);
  const double A[9] = {cp, sp, 0, -sp, cp, 0, 0, 0, 1};
  const double B[9] = {ct, 0, -st, 0, 1, 0, st, 0, ct};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0;
      for (int k = 0; k < 3; ++k) s += B[3 * i + k] * A[3 * k + j];
      Q[3 * i + j] = s;
    }
}

enum Kind { kM2M = 0, kL2L = 1, kM2L = 2 };

// One table of shift t (in the units whose wavenumber is kshift), appended to out.
void table(int p, Kind kind, cplx kshift, const double t[3], const std::vector<double> &rf,
           const std::vector<double> &cf, cplx common, const std::vector<double> &norm, double *out) {
  const int N = ncoef(p), C = p + 1;
  const double tn = std::sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
  std::vector<cplx> f;
  if (kind == kM2L)
    spherical_h(2 * p, kshift * tn, f);
  else
    spherical_j(2 * p, kshift * tn, f);
  std::vector<cplx> E;
  coaxial(p, f, E);
  double Q[9];
  rotation_to_z(t, Q);
  std::vector<double> O;
  std::vector<int> boff;
  rotation_blocks(p, Q, norm, O, boff);
  double *Tr = out, *Ti = out + size_t(N) * N;
  std::vector<cplx> y(N);
  for (int k = 0; k < N; ++k) {
    int nk = 0;
    while ((nk + 1) * (nk + 1) <= k) ++nk;
    const int ck = k - nk * nk, Wk = 2 * nk + 1;  // column of O^nk
    // y = A O e_k: degree nk of O e_k is O[:, ck]; the coaxial block maps degree nk to n'
    for (int np = 0; np <= p; ++np)
      for (int m = -np; m <= np; ++m) {
        const int am = m < 0 ? -m : m;
        cplx v = 0.0;
        if (am <= nk) {
          const double x = O[boff[nk] + (m + nk) * Wk + ck];
          v = E[(size_t(am) * C + np) * C + nk] * rf[np] * cf[nk] * x;
        }
        y[idx(np, m)] = common * v;
      }
    // O^T y
    for (int np = 0; np <= p; ++np) {
      const int W = 2 * np + 1;
      for (int c = 0; c < W; ++c) {
        cplx v = 0.0;
        for (int r = 0; r < W; ++r) v += O[boff[np] + r * W + c] * y[idx(np, r - np)];
        Tr[size_t(k) * N + idx(np, c - np)] = v.real();
        Ti[size_t(k) * N + idx(np, c - np)] = v.imag();
      }
    }
  }
}

void octant_offset(int o, double sign, double t[3]) {
  t[0] = sign * (((o >> 2) & 1) ? 0.5 : -0.5);
  t[1] = sign * (((o >> 1) & 1) ? 0.5 : -0.5);
  t[2] = sign * ((o & 1) ? 0.5 : -0.5);
}

}  // namespace

void h_norm_factors(int p, std::vector<double> &norm) {
  norm.assign(ncoef(p), 0.0);
  for (int n = 0; n <= p; ++n)
    for (int m = 0; m <= n; ++m) {
      double f = std::sqrt((2.0 * n + 1.0) / (4.0 * kPi));
      if (m == 0) {
        for (int q = 2; q <= n; ++q) f *= q;
      } else {
        double v = 2.0;
        for (int q = 2; q <= n + m; ++q) v *= q;
        for (int q = 2; q <= n - m; ++q) v *= q;
        f *= std::sqrt(v);
      }
      norm[idx(n, m)] = f;
      norm[idx(n, -m)] = f;
    }
}

void h_m2m_tables(int p, cplx kh, std::vector<double> &out) {
  // child (source, kh/2) -> parent (target, kh); t = parent - child in parent units
  std::vector<double> norm;
  h_norm_factors(p, norm);
  const size_t sz = 2 * size_t(ncoef(p)) * ncoef(p), base = out.size();
  out.resize(base + 8 * sz);
  const cplx kc = kh / 2.0;
  const double sp = h_scale(kh), sc = h_scale(kc);
  std::vector<double> rf(p + 1), cf(p + 1);
  for (int n = 0; n <= p; ++n) {
    rf[n] = std::pow(sp, -n);
    cf[n] = std::pow(sc, n);
  }
  const cplx common = std::exp(h_beta(kc) - h_beta(kh));
#pragma omp parallel for schedule(dynamic, 1)
  for (int o = 0; o < 8; ++o) {
    double t[3];
    octant_offset(o, -1.0, t);
    table(p, kM2M, kh, t, rf, cf, common, norm, out.data() + base + o * sz);
  }
}

void h_l2l_tables(int p, cplx kh, std::vector<double> &out) {
  // parent (source, kh) -> child (target, kh/2); t = child - parent in parent units
  std::vector<double> norm;
  h_norm_factors(p, norm);
  const size_t sz = 2 * size_t(ncoef(p)) * ncoef(p), base = out.size();
  out.resize(base + 8 * sz);
  const cplx kc = kh / 2.0;
  const double sp = h_scale(kh), sc = h_scale(kc);
  std::vector<double> rf(p + 1), cf(p + 1);
  for (int n = 0; n <= p; ++n) {
    rf[n] = std::pow(sc, n);
    cf[n] = std::pow(sp, -n);
  }
  const cplx common = 0.5 * std::exp(h_beta(kc) - h_beta(kh));
#pragma omp parallel for schedule(dynamic, 1)
  for (int o = 0; o < 8; ++o) {
    double t[3];
    octant_offset(o, 1.0, t);
    table(p, kL2L, kh, t, rf, cf, common, norm, out.data() + base + o * sz);
  }
}

void h_rotation_blocks(int p, const double Q[9], std::vector<double> &out) {
  std::vector<double> norm;
  h_norm_factors(p, norm);
  std::vector<int> boff;
  rotation_blocks(p, Q, norm, out, boff);
}

int h_rcr_rotation_size(int p) { return (p + 1) * (2 * p + 1) * (2 * p + 3) / 3; }

int h_rcr_coaxial_size(int p) {
  int s = 0;
  for (int m = 0; m <= p; ++m) s += (p + 1 - m) * (p + 1 - m);
  return s;
}

void h_rcr_rotations(int p, std::vector<double> &out) {
  std::vector<double> norm;
  h_norm_factors(p, norm);
  const int rb = h_rcr_rotation_size(p);
  out.assign(316 * size_t(rb), 0.0);
#pragma omp parallel for schedule(dynamic, 4)
  for (int i = 0; i < 316; ++i) {
    int d[3];
    offset_of_index(i, d);
    const double t[3] = {2.0 * d[0], 2.0 * d[1], 2.0 * d[2]};
    double Q[9];
    rotation_to_z(t, Q);
    std::vector<double> O;
    std::vector<int> boff;
    rotation_blocks(p, Q, norm, O, boff);
    for (int n = 0; n <= p; ++n) {
      const int w = 2 * n + 1;
      for (int r = 0; r < w; ++r)
        for (int c = 0; c < w; ++c) out[size_t(i) * rb + boff[n] + r + size_t(c) * w] = O[boff[n] + r * w + c];
    }
  }
}

void h_rcr_coaxial(int p, cplx kh, const std::vector<double> &dist, std::vector<double> &out) {
  const int cb = h_rcr_coaxial_size(p), C = p + 1;
  out.assign(2 * size_t(cb) * dist.size(), 0.0);
  const double s = h_scale(kh);
  std::vector<double> sp(p + 1);
  for (int n = 0; n <= p; ++n) sp[n] = std::pow(s, n);
  const cplx common = 4.0 * kPi * cplx(0.0, 1.0) * kh * std::exp(2.0 * h_beta(kh));
#pragma omp parallel for schedule(dynamic, 1)
  for (int i = 0; i < int(dist.size()); ++i) {
    std::vector<cplx> f, E;
    spherical_h(2 * p, kh * dist[i], f);
    coaxial(p, f, E);
    double *re = out.data() + 2 * size_t(cb) * i, *im = re + cb;
    int co = 0;
    for (int m = 0; m <= p; ++m) {
      const int w = p + 1 - m;
      for (int np = m; np <= p; ++np)
        for (int n = m; n <= p; ++n) {
          const cplx v = common * sp[np] * sp[n] * E[(size_t(m) * C + np) * C + n];
          re[co + (np - m) + (n - m) * w] = v.real();
          im[co + (np - m) + (n - m) * w] = v.imag();
        }
      co += w * w;
    }
  }
}

void h_m2l_tables(int p, cplx kh, const std::vector<uint16_t> &offsets, std::vector<double> &out) {
  std::vector<double> norm;
  h_norm_factors(p, norm);
  const size_t sz = 2 * size_t(ncoef(p)) * ncoef(p), base = out.size();
  out.resize(base + offsets.size() * sz);
  const double s = h_scale(kh);
  std::vector<double> rf(p + 1), cf(p + 1);
  for (int n = 0; n <= p; ++n) rf[n] = cf[n] = std::pow(s, n);
  const cplx common = 4.0 * kPi * cplx(0.0, 1.0) * kh * std::exp(2.0 * h_beta(kh));
#pragma omp parallel for schedule(dynamic, 1)
  for (int i = 0; i < int(offsets.size()); ++i) {
    int d[3];
    offset_of_index(offsets[i], d);
    const double t[3] = {2.0 * d[0], 2.0 * d[1], 2.0 * d[2]};
    table(p, kM2L, kh, t, rf, cf, common, norm, out.data() + base + size_t(i) * sz);
  }
}

}  // namespace fmmh
