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
#include <algorithm>
#include <cmath>
#include <cstdint>
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

// ---- rotation M2L (CONVENTIONS §3.11 "Rotation of coefficients", "Coaxial translations")
//
// T(d) = G(Q^T) A(|d|) F(Q), Q the rotation with Q d = |d| e_z.  F(Q) maps the multipole
// coefficients of a configuration to those of the configuration rotated by Q, G(R) the
// same for local coefficients; both act on each degree alone, a (2n+1)^2 block.  A(s) is
// the M2L table of the shift s e_z, which keeps the order: per order i a (p+1-i)^2 block
// from input degree n to output degree j, the same for slots +i and -i.
//
// The blocks are built from their definition: for sample points y, the coefficient vector
// of a unit charge at Q y is the block times the vector of the charge at y (multipoles:
// conj R_n(y), locals: conj I_n(y), in real storage); the block is the least-squares
// solution over 2 (2n+1) points, after scaling every coefficient to unit size so the
// factorials of the scaled harmonics do not spoil the solve.

// Real-storage coefficient vector of degree n of a unit charge at y: conj of the
// harmonic (slot +m: Re, slot -m: -Im), as P2M and P2L form them.
void unit_coefficients(const Harm &h, int n, double *out) {
  for (int m = -n; m <= n; ++m) {
    double a, b;
    h.get(n, m < 0 ? -m : m, a, b);
    out[m + n] = m < 0 ? -b : a;
  }
}

// Solves X S = S' in least squares for X (k x k), S and S' k x K column-major (K >= k),
// by Householder QR of S^T.
void least_squares_block(int k, int K, const std::vector<double> &S, const std::vector<double> &Sp,
                         std::vector<double> &X) {
  // A = S^T (K x k), B = Sp^T (K x k); minimise |A X^T - B|
  std::vector<double> A(size_t(K) * k), B(size_t(K) * k);
  for (int r = 0; r < K; ++r)
    for (int c = 0; c < k; ++c) {
      A[r + size_t(c) * K] = S[c + size_t(r) * k];
      B[r + size_t(c) * K] = Sp[c + size_t(r) * k];
    }
  for (int c = 0; c < k; ++c) {
    double norm = 0.0;
    for (int r = c; r < K; ++r) norm += A[r + size_t(c) * K] * A[r + size_t(c) * K];
    norm = std::sqrt(norm);
    const double a0 = A[c + size_t(c) * K];
    const double alpha = a0 > 0 ? -norm : norm;
    std::vector<double> v(K, 0.0);
    for (int r = c; r < K; ++r) v[r] = A[r + size_t(c) * K];
    v[c] -= alpha;
    double vv = 0.0;
    for (int r = c; r < K; ++r) vv += v[r] * v[r];
    if (vv == 0.0) continue;
    auto reflect = [&](std::vector<double> &M, int cols, int from) {
      for (int cc = from; cc < cols; ++cc) {
        double s = 0.0;
        for (int r = c; r < K; ++r) s += v[r] * M[r + size_t(cc) * K];
        s *= 2.0 / vv;
        for (int r = c; r < K; ++r) M[r + size_t(cc) * K] -= s * v[r];
      }
    };
    reflect(A, k, c);
    reflect(B, k, 0);
  }
  // back substitution: R X^T = (Q^T B)[0:k]
  X.assign(size_t(k) * k, 0.0);
  for (int cc = 0; cc < k; ++cc)
    for (int r = k - 1; r >= 0; --r) {
      double s = B[r + size_t(cc) * K];
      for (int q = r + 1; q < k; ++q) s -= A[r + size_t(q) * K] * X[cc + size_t(q) * k];
      X[cc + size_t(r) * k] = s / A[r + size_t(r) * K];
    }
}

// Rotation block of degree n (column-major, (2n+1)^2) for rotation matrix Rm (row-major
// 3 x 3): multipoles (local = false) or locals (local = true).
void rotation_block(int n, const double Rm[9], bool local, double *blk) {
  const int k = 2 * n + 1, K = 2 * k;
  std::vector<double> S(size_t(k) * K), Sp(size_t(k) * K);
  // sample points: a fixed spread on the unit sphere (golden-angle spiral, tilted)
  for (int s = 0; s < K; ++s) {
    const double z = 1.0 - (2.0 * s + 1.0) / K;
    const double rad = std::sqrt(1.0 - z * z), phi = 2.399963229728653 * s + 0.3;
    const double y[3] = {rad * std::cos(phi), rad * std::sin(phi), z};
    const double qy[3] = {Rm[0] * y[0] + Rm[1] * y[1] + Rm[2] * y[2], Rm[3] * y[0] + Rm[4] * y[1] + Rm[5] * y[2],
                          Rm[6] * y[0] + Rm[7] * y[1] + Rm[8] * y[2]};
    const Harm a = local ? irregular(n, y) : regular(n, y);
    const Harm b = local ? irregular(n, qy) : regular(n, qy);
    unit_coefficients(a, n, &S[size_t(s) * k]);
    unit_coefficients(b, n, &Sp[size_t(s) * k]);
  }
  // scale each coefficient to unit size over the samples (the same scale on both sides)
  std::vector<double> w(k);
  for (int c = 0; c < k; ++c) {
    double ss = 0.0;
    for (int s = 0; s < K; ++s) ss += S[c + size_t(s) * k] * S[c + size_t(s) * k];
    w[c] = ss > 0 ? 1.0 / std::sqrt(ss / K) : 1.0;
    for (int s = 0; s < K; ++s) {
      S[c + size_t(s) * k] *= w[c];
      Sp[c + size_t(s) * k] *= w[c];
    }
  }
  std::vector<double> X;
  least_squares_block(k, K, S, Sp, X);
  // unscale: block = W^-1 X W
  for (int c = 0; c < k; ++c)
    for (int r = 0; r < k; ++r) blk[r + size_t(c) * k] = X[r + size_t(c) * k] / w[r] * w[c];
}

}  // namespace

void rotation_blocks_of(int p, const double R[9], bool local, double *out) {
  size_t bo = 0;
  for (int n = 0; n <= p; ++n) {
    rotation_block(n, R, local, out + bo);
    bo += size_t(2 * n + 1) * (2 * n + 1);
  }
}

int rotation_block_size(int p) {
  int s = 0;
  for (int n = 0; n <= p; ++n) s += (2 * n + 1) * (2 * n + 1);
  return s;
}

int coaxial_block_size(int p) {
  int s = 0;
  for (int i = 0; i <= p; ++i) s += (p + 1 - i) * (p + 1 - i);
  return s;
}

void build_rotation_tables(int p, std::vector<double> &fwd, std::vector<double> &bwd, std::vector<double> &coax,
                           std::vector<uint16_t> &dist_of, std::vector<double> &dist2) {
  require(p >= 0 && p <= 30, "build_rotation_tables: p must lie in 0..30");
  const int RB = rotation_block_size(p), CB = coaxial_block_size(p), nc = ncoef(p);
  fwd.assign(size_t(316) * RB, 0.0);
  bwd.assign(size_t(316) * RB, 0.0);
  dist_of.assign(316, 0);
  dist2.clear();
  for (int t = 0; t < 316; ++t) {
    int d[3];
    offset_of_index(t, d);
    const double s2 = double(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    size_t at = 0;
    while (at < dist2.size() && dist2[at] != s2) ++at;
    if (at == dist2.size()) dist2.push_back(s2);
    dist_of[t] = static_cast<uint16_t>(at);
  }
#pragma omp parallel for schedule(dynamic, 4)
  for (int t = 0; t < 316; ++t) {
    int d[3];
    offset_of_index(t, d);
    const double len = std::sqrt(double(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]));
    const double ct = d[2] / len, st = std::sqrt(std::max(0.0, 1.0 - ct * ct));
    const double ph = std::atan2(double(d[1]), double(d[0]));
    const double cp = std::cos(ph), sp = std::sin(ph);
    // Q = R_y(-theta) R_z(-phi): Q d = |d| e_z
    const double Rz[9] = {cp, sp, 0, -sp, cp, 0, 0, 0, 1};
    const double Ry[9] = {ct, 0, -st, 0, 1, 0, st, 0, ct};
    double Q[9], Qt[9];
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) {
        double s = 0.0;
        for (int q = 0; q < 3; ++q) s += Ry[3 * r + q] * Rz[3 * q + c];
        Q[3 * r + c] = s;
      }
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) Qt[3 * r + c] = Q[3 * c + r];
    size_t bo = 0;
    for (int n = 0; n <= p; ++n) {
      rotation_block(n, Q, false, &fwd[size_t(t) * RB + bo]);
      rotation_block(n, Qt, true, &bwd[size_t(t) * RB + bo]);
      bo += size_t(2 * n + 1) * (2 * n + 1);
    }
  }
  // coaxial blocks: the M2L table of the shift (0, 0, 2 |d|) (box radius units), order by
  // order
  coax.assign(dist2.size() * CB, 0.0);
  std::vector<double> T(size_t(nc) * nc);
  for (size_t q = 0; q < dist2.size(); ++q) {
    std::fill(T.begin(), T.end(), 0.0);
    const double b[3] = {0.0, 0.0, 2.0 * std::sqrt(dist2[q])};
    m2l_table(p, b, 1.0, T.data());
    size_t bo = 0;
    for (int i = 0; i <= p; ++i) {
      const int w = p + 1 - i;
      for (int n = i; n <= p; ++n)
        for (int j = i; j <= p; ++j) coax[q * CB + bo + (j - i) + size_t(n - i) * w] = T[idx(j, i) + size_t(idx(n, i)) * nc];
      bo += size_t(w) * w;
    }
  }
}

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
