// tables.cpp -- operator tables of CONVENTIONS §3.12 built entry by entry in f64.
//
// Each table is the translation of CONVENTIONS §3.11 at its canonical frame, written
// for real storage through the two order-sum identities of §3.11 "Real storage": for an
// input coefficient A^m = a + i b of degree n (slots +m and -m) and the shift harmonic B
// of the paired degree, read for any order by the (-1)^m rule,
//
//   M2M:  sum_l A^l B^{i-l} = A^0 B^i + sum_{l>0} (A^l B^{i-l} + (-1)^l conj(A^l) B^{i+l})
//   L2L, M2L: sum_m A^m B^{m-i} = A^0 B^{-i} + sum_{m>0} (A^m B^{m-i} + (-1)^m conj(A^m) B^{-m-i})
//
// so the coefficients of a and b in the real and imaginary outputs are closed forms in
// B: O(1) per entry, O(n^2) per table.  Layout: column-major (p+1)^2 x (p+1)^2, entry
// (i, k) at i + k n; M2M/L2L by octant o = 4x + 2y + z, M2L by offset index
// (fmm-tables/src/m2l.rs:67-160, octant.rs).
#include <cmath>
#include <vector>

#include "plan.hpp"

namespace fmmh {
namespace {

inline int idx(int n, int m) { return n * n + n + m; }

// Complex harmonics (re, im) of orders m >= 0, Cartesian recursions of CONVENTIONS §3.5.
struct Harm {
  int p;
  std::vector<double> re, im;
  explicit Harm(int p_) : p(p_), re(ncoef(p_), 0.0), im(ncoef(p_), 0.0) {}
  // order q of degree n by the (-1)^m rule; zero outside |q| <= n <= p
  void get(int n, int q, double &a, double &b) const {
    if (n < 0 || n > p || q > n || -q > n) {
      a = b = 0.0;
      return;
    }
    if (q >= 0) {
      a = re[idx(n, q)];
      b = im[idx(n, q)];
    } else {
      const double s = ((-q) & 1) ? -1.0 : 1.0;
      a = s * re[idx(n, -q)];
      b = -s * im[idx(n, -q)];
    }
  }
};

Harm regular(int p, const double x[3]) {
  Harm h(p);
  const double r2 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
  h.re[0] = 1.0;
  for (int m = 1; m <= p; ++m) {
    const double a = h.re[idx(m - 1, m - 1)], b = h.im[idx(m - 1, m - 1)];
    h.re[idx(m, m)] = (x[0] * a - x[1] * b) / (2.0 * m);
    h.im[idx(m, m)] = (x[0] * b + x[1] * a) / (2.0 * m);
  }
  for (int m = 0; m < p; ++m) {
    h.re[idx(m + 1, m)] = x[2] * h.re[idx(m, m)];
    h.im[idx(m + 1, m)] = x[2] * h.im[idx(m, m)];
  }
  for (int m = 0; m <= p; ++m)
    for (int n = m + 2; n <= p; ++n) {
      const double den = double((n + m) * (n - m));
      h.re[idx(n, m)] = ((2 * n - 1) * x[2] * h.re[idx(n - 1, m)] - r2 * h.re[idx(n - 2, m)]) / den;
      h.im[idx(n, m)] = ((2 * n - 1) * x[2] * h.im[idx(n - 1, m)] - r2 * h.im[idx(n - 2, m)]) / den;
    }
  return h;
}

Harm irregular(int p, const double x[3]) {
  Harm h(p);
  const double r2 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
  h.re[0] = 1.0 / std::sqrt(r2);
  for (int m = 1; m <= p; ++m) {
    const double a = h.re[idx(m - 1, m - 1)], b = h.im[idx(m - 1, m - 1)];
    const double f = (2.0 * m - 1.0) / r2;
    h.re[idx(m, m)] = f * (x[0] * a - x[1] * b);
    h.im[idx(m, m)] = f * (x[0] * b + x[1] * a);
  }
  for (int m = 0; m < p; ++m) {
    const double f = (2.0 * m + 1.0) * x[2] / r2;
    h.re[idx(m + 1, m)] = f * h.re[idx(m, m)];
    h.im[idx(m + 1, m)] = f * h.im[idx(m, m)];
  }
  for (int m = 0; m <= p; ++m)
    for (int n = m + 2; n <= p; ++n) {
      const double c = double((n - 1) * (n - 1) - m * m);
      h.re[idx(n, m)] = ((2 * n - 1) * x[2] * h.re[idx(n - 1, m)] - c * h.re[idx(n - 2, m)]) / r2;
      h.im[idx(n, m)] = ((2 * n - 1) * x[2] * h.im[idx(n - 1, m)] - c * h.im[idx(n - 2, m)]) / r2;
    }
  return h;
}

// Adds, for output (j, i >= 0) and input degree n, the entries from the order sum with
// first factor B(q1(m)) and second factor B(q2(m)) scaled by f.  `shift` reads the shift
// harmonic of the paired degree: B(q) = (br, bi).
template <class ShiftFn>
void add_block(double *T, int nc, int j, int i, int n, double f, int sign_first, ShiftFn B) {
  // sign_first = +1: q1 = m - i, q2 = -m - i   (L2L, M2L: B^{m-i}, B^{-m-i}, m = 0: B^{-i})
  // sign_first = -1: q1 = i - m, q2 = i + m    (M2M: B^{i-l}, B^{i+l}, l = 0: B^{i})
  const int rre = idx(j, i), rim = i > 0 ? idx(j, -i) : -1;
  {
    double br, bi;
    B(sign_first > 0 ? -i : i, br, bi);
    T[rre + size_t(idx(n, 0)) * nc] += f * br;
    if (rim >= 0) T[rim + size_t(idx(n, 0)) * nc] += f * bi;
  }
  for (int m = 1; m <= n; ++m) {
    const double s = (m & 1) ? -1.0 : 1.0;
    double b1r, b1i, b2r, b2i;
    if (sign_first > 0) {
      B(m - i, b1r, b1i);
      B(-m - i, b2r, b2i);
    } else {
      B(i - m, b1r, b1i);
      B(i + m, b2r, b2i);
    }
    const size_t ca = size_t(idx(n, m)) * nc, cb = size_t(idx(n, -m)) * nc;
    T[rre + ca] += f * (b1r + s * b2r);
    T[rre + cb] += f * (-b1i + s * b2i);
    if (rim >= 0) {
      T[rim + ca] += f * (b1i + s * b2i);
      T[rim + cb] += f * (b1r - s * b2r);
    }
  }
}

void m2m_table(int p, const double b[3], double rho, double *T) {
  const int nc = ncoef(p);
  const Harm h = regular(p, b);
  for (int j = 0; j <= p; ++j)
    for (int i = 0; i <= j; ++i)
      for (int k = 0; k <= j; ++k) {
        const double f = std::pow(rho, k);
        add_block(T, nc, j, i, k, f, -1, [&](int q, double &br, double &bi) {
          h.get(j - k, q, br, bi);
          bi = -bi;  // conj R_{j-k}
        });
      }
}

void l2l_table(int p, const double t[3], double sigma, double *T) {
  const int nc = ncoef(p);
  const Harm h = regular(p, t);
  for (int j = 0; j <= p; ++j)
    for (int i = 0; i <= j; ++i)
      for (int n = j; n <= p; ++n) {
        const double f = std::pow(sigma, j + 1);
        add_block(T, nc, j, i, n, f, +1, [&](int q, double &br, double &bi) { h.get(n - j, q, br, bi); });
      }
}

void m2l_table(int p, const double b[3], double sigma, double *T) {
  const int nc = ncoef(p);
  const Harm h = irregular(2 * p, b);
  for (int j = 0; j <= p; ++j)
    for (int i = 0; i <= j; ++i)
      for (int n = 0; n <= p; ++n) {
        const double f = (((j + i) & 1) ? -1.0 : 1.0) * std::pow(sigma, j + 1);
        add_block(T, nc, j, i, n, f, +1, [&](int q, double &br, double &bi) { h.get(n + j, q, br, bi); });
      }
}

}  // namespace

void build_tables(int p, std::vector<double> &m2m, std::vector<double> &l2l,
                  std::vector<double> &m2l) {
  require(p >= 0 && p <= 30, "build_tables: p must lie in 0..30");
  const size_t n2 = size_t(ncoef(p)) * ncoef(p);
  m2m.assign(8 * n2, 0.0);
  l2l.assign(8 * n2, 0.0);
  m2l.assign(316 * n2, 0.0);
  for (int o = 0; o < 8; ++o) {
    const double s[3] = {0.5 * (2 * ((o >> 2) & 1) - 1), 0.5 * (2 * ((o >> 1) & 1) - 1), 0.5 * (2 * (o & 1) - 1)};
    m2m_table(p, s, 0.5, m2m.data() + o * n2);
    l2l_table(p, s, 0.5, l2l.data() + o * n2);
  }
#pragma omp parallel for schedule(dynamic, 4)
  for (int t = 0; t < 316; ++t) {
    int d[3];
    offset_of_index(t, d);
    const double b[3] = {2.0 * d[0], 2.0 * d[1], 2.0 * d[2]};
    m2l_table(p, b, 1.0, m2l.data() + t * n2);
  }
}

}  // namespace fmmh
