/*
 * harmonics.c -- oracle (test infrastructure): solid harmonics, leaf operators,
 * translations and operator tables, in f64, written from the reference's conventions.
 *
 * Real storage (CONVENTIONS §3.6): (p+1)^2 reals, idx(n, m) = n^2 + n + m; slot +m the
 * real part of the order-m quantity, slot -m its imaginary part; negative orders follow
 * X^{-m} = (-1)^m conj(X^m) (§3.3).  The translations are evaluated in full complex
 * storage straight from the formulas of CONVENTIONS §3.11 (the forms of
 * fmm-ref/src/direct.rs:75-210), and the tables by applying them to unit vectors, as
 * fmm-tables builds them from nd_fmm_ref::direct (fmm-tables/src/m2l.rs:114).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

static inline int idx(int n, int m) { return n * n + n + m; }
static inline int nc(int p) { return (p + 1) * (p + 1); }

/* Complex values (re, im) of orders m >= 0 per degree, recursions of CONVENTIONS §3.5
 * (fmm-math/src/harmonics.rs:118-260). */
static void regular_cplx(int p, const double x[3], double *re, double *im) {
  double r2 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
  for (int i = 0; i < nc(p); ++i) re[i] = im[i] = 0.0;
  re[idx(0, 0)] = 1.0;
  for (int m = 1; m <= p; ++m) {
    double a = re[idx(m - 1, m - 1)], b = im[idx(m - 1, m - 1)];
    re[idx(m, m)] = (x[0] * a - x[1] * b) / (2.0 * m);
    im[idx(m, m)] = (x[0] * b + x[1] * a) / (2.0 * m);
  }
  for (int m = 0; m < p; ++m) {
    re[idx(m + 1, m)] = x[2] * re[idx(m, m)];
    im[idx(m + 1, m)] = x[2] * im[idx(m, m)];
  }
  for (int m = 0; m <= p; ++m)
    for (int n = m + 2; n <= p; ++n) {
      double den = (double)((n + m) * (n - m));
      re[idx(n, m)] = ((2 * n - 1) * x[2] * re[idx(n - 1, m)] - r2 * re[idx(n - 2, m)]) / den;
      im[idx(n, m)] = ((2 * n - 1) * x[2] * im[idx(n - 1, m)] - r2 * im[idx(n - 2, m)]) / den;
    }
}

static void irregular_cplx(int p, const double x[3], double *re, double *im) {
  double r2 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
  for (int i = 0; i < nc(p); ++i) re[i] = im[i] = 0.0;
  re[idx(0, 0)] = 1.0 / sqrt(r2);
  for (int m = 1; m <= p; ++m) {
    double a = re[idx(m - 1, m - 1)], b = im[idx(m - 1, m - 1)];
    double f = (2.0 * m - 1.0) / r2;
    re[idx(m, m)] = f * (x[0] * a - x[1] * b);
    im[idx(m, m)] = f * (x[0] * b + x[1] * a);
  }
  for (int m = 0; m < p; ++m) {
    double f = (2.0 * m + 1.0) * x[2] / r2;
    re[idx(m + 1, m)] = f * re[idx(m, m)];
    im[idx(m + 1, m)] = f * im[idx(m, m)];
  }
  for (int m = 0; m <= p; ++m)
    for (int n = m + 2; n <= p; ++n) {
      double c = (double)((n - 1) * (n - 1) - m * m);
      re[idx(n, m)] = ((2 * n - 1) * x[2] * re[idx(n - 1, m)] - c * re[idx(n - 2, m)]) / r2;
      im[idx(n, m)] = ((2 * n - 1) * x[2] * im[idx(n - 1, m)] - c * im[idx(n - 2, m)]) / r2;
    }
}

/* Order q of degree n (any sign) from the m >= 0 complex arrays, by the (-1)^m rule;
 * zero when |q| > n or n < 0 or n > p. */
static inline void getc(int p, const double *re, const double *im, int n, int q, double *a,
                        double *b) {
  if (n < 0 || n > p || q > n || -q > n) {
    *a = *b = 0.0;
    return;
  }
  if (q >= 0) {
    *a = re[idx(n, q)];
    *b = im[idx(n, q)];
  } else {
    double s = ((-q) & 1) ? -1.0 : 1.0;
    *a = s * re[idx(n, -q)];
    *b = -s * im[idx(n, -q)];
  }
}

static void pack(int p, const double *re, const double *im, double *out) {
  for (int n = 0; n <= p; ++n) {
    out[idx(n, 0)] = re[idx(n, 0)];
    for (int m = 1; m <= n; ++m) {
      out[idx(n, m)] = re[idx(n, m)];
      out[idx(n, -m)] = im[idx(n, m)];
    }
  }
}

void orc_regular(int p, const double x[3], double *out) {
  double *re = malloc(sizeof(double) * nc(p)), *im = malloc(sizeof(double) * nc(p));
  regular_cplx(p, x, re, im);
  pack(p, re, im, out);
  free(re);
  free(im);
}

void orc_irregular(int p, const double x[3], double *out) {
  double *re = malloc(sizeof(double) * nc(p)), *im = malloc(sizeof(double) * nc(p));
  irregular_cplx(p, x, re, im);
  pack(p, re, im, out);
  free(re);
  free(im);
}

/* Gradient ladders (CONVENTIONS §3.11 "L2P and M2P"): with D the degree shift (-1 for
 * regular, +1 for irregular) and s_z (+1 regular, -1 irregular):
 *   dz X_n^m = s_z X_{n+D}^m,  dx X_n^m = (X_{n+D}^{m-1} - X_{n+D}^{m+1}) / 2,
 *   dy X_n^m = i (X_{n+D}^{m-1} + X_{n+D}^{m+1}) / 2. */
static void grad_from(int p, int pc, const double *re, const double *im, int shift, double sz,
                      double *gx, double *gy, double *gz) {
  for (int n = 0; n <= p; ++n)
    for (int m = 0; m <= n; ++m) {
      double am, bm, ap, bp, a0, b0;
      getc(pc, re, im, n + shift, m - 1, &am, &bm);
      getc(pc, re, im, n + shift, m + 1, &ap, &bp);
      getc(pc, re, im, n + shift, m, &a0, &b0);
      double xr = 0.5 * (am - ap), xi = 0.5 * (bm - bp);
      double yr = -0.5 * (bm + bp), yi = 0.5 * (am + ap);
      double zr = sz * a0, zi = sz * b0;
      gx[idx(n, m)] = xr;
      gy[idx(n, m)] = yr;
      gz[idx(n, m)] = zr;
      if (m > 0) {
        gx[idx(n, -m)] = xi;
        gy[idx(n, -m)] = yi;
        gz[idx(n, -m)] = zi;
      } else {
        (void)xi;
      }
    }
}

void orc_regular_grad(int p, const double x[3], double *gx, double *gy, double *gz) {
  double *re = malloc(sizeof(double) * nc(p)), *im = malloc(sizeof(double) * nc(p));
  regular_cplx(p, x, re, im);
  grad_from(p, p, re, im, -1, 1.0, gx, gy, gz);
  free(re);
  free(im);
}

void orc_irregular_grad(int p, const double x[3], double *gx, double *gy, double *gz) {
  double *re = malloc(sizeof(double) * nc(p + 1)), *im = malloc(sizeof(double) * nc(p + 1));
  irregular_cplx(p + 1, x, re, im);
  grad_from(p, p + 1, re, im, 1, -1.0, gx, gy, gz);
  free(re);
  free(im);
}

/* ---- leaf operators (fmm-ref/src/leaf.rs; CONVENTIONS §3.7, §3.11, §3.13) ------- */

/* P2M: M~ += sum q conj R(u), u = (y - c)/r. */
void orc_p2m(int p, const double c[3], double r, const double *src, const double *q,
             int64_t n, double *m) {
  double *h = malloc(sizeof(double) * nc(p));
  for (int64_t j = 0; j < n; ++j) {
    double u[3];
    for (int k = 0; k < 3; ++k) u[k] = (src[3 * j + k] - c[k]) / r;
    orc_regular(p, u, h);
    for (int d = 0; d <= p; ++d) {
      m[idx(d, 0)] += q[j] * h[idx(d, 0)];
      for (int o = 1; o <= d; ++o) {
        m[idx(d, o)] += q[j] * h[idx(d, o)];
        m[idx(d, -o)] -= q[j] * h[idx(d, -o)];
      }
    }
  }
  free(h);
}

/* P2L: L~ += sum q conj I(u), u = (y - c)/r. */
void orc_p2l(int p, const double c[3], double r, const double *src, const double *q,
             int64_t n, double *l) {
  double *h = malloc(sizeof(double) * nc(p));
  for (int64_t j = 0; j < n; ++j) {
    double u[3];
    for (int k = 0; k < 3; ++k) u[k] = (src[3 * j + k] - c[k]) / r;
    orc_irregular(p, u, h);
    for (int d = 0; d <= p; ++d) {
      l[idx(d, 0)] += q[j] * h[idx(d, 0)];
      for (int o = 1; o <= d; ++o) {
        l[idx(d, o)] += q[j] * h[idx(d, o)];
        l[idx(d, -o)] -= q[j] * h[idx(d, -o)];
      }
    }
  }
  free(h);
}

/* sum_m A_n^m X_n^m over all orders from real storage: A^0 X^0 + 2 sum (ReA ReX - ImA ImX). */
static double contract(int p, const double *a, const double *x) {
  double s = 0.0;
  for (int d = 0; d <= p; ++d) {
    s += a[idx(d, 0)] * x[idx(d, 0)];
    for (int o = 1; o <= d; ++o)
      s += 2.0 * (a[idx(d, o)] * x[idx(d, o)] - a[idx(d, -o)] * x[idx(d, -o)]);
  }
  return s;
}

/* L2P: phi += (1/r) sum L~ R(v), grad += (1/r^2) sum L~ grad R(v), v = (x - c)/r. */
void orc_l2p(int p, const double c[3], double r, const double *l, const double *tgt,
             int64_t n, double *phi, double *grad) {
  double *h = malloc(sizeof(double) * 4 * nc(p));
  for (int64_t i = 0; i < n; ++i) {
    double v[3];
    for (int k = 0; k < 3; ++k) v[k] = (tgt[3 * i + k] - c[k]) / r;
    orc_regular(p, v, h);
    phi[i] += contract(p, l, h) / r;
    if (grad) {
      orc_regular_grad(p, v, h + nc(p), h + 2 * nc(p), h + 3 * nc(p));
      for (int k = 0; k < 3; ++k) grad[3 * i + k] += contract(p, l, h + (k + 1) * nc(p)) / (r * r);
    }
  }
  free(h);
}

/* M2P: phi += (1/r) sum M~ I(v), grad += (1/r^2) sum M~ grad I(v), v = (x - c)/r. */
void orc_m2p(int p, const double c[3], double r, const double *m, const double *tgt,
             int64_t n, double *phi, double *grad) {
  double *h = malloc(sizeof(double) * 4 * nc(p));
  for (int64_t i = 0; i < n; ++i) {
    double v[3];
    for (int k = 0; k < 3; ++k) v[k] = (tgt[3 * i + k] - c[k]) / r;
    orc_irregular(p, v, h);
    phi[i] += contract(p, m, h) / r;
    if (grad) {
      orc_irregular_grad(p, v, h + nc(p), h + 2 * nc(p), h + 3 * nc(p));
      for (int k = 0; k < 3; ++k) grad[3 * i + k] += contract(p, m, h + (k + 1) * nc(p)) / (r * r);
    }
  }
  free(h);
}

/* ---- translations (CONVENTIONS §3.11), full complex storage ---------------------- */

typedef struct {
  int p;
  double *re, *im; /* full storage, index n^2 + n + m for -n <= m <= n */
} full;

static full full_from_real(int p, const double *x) {
  full f = {p, calloc(nc(p), sizeof(double)), calloc(nc(p), sizeof(double))};
  for (int n = 0; n <= p; ++n) {
    f.re[idx(n, 0)] = x[idx(n, 0)];
    for (int m = 1; m <= n; ++m) {
      double s = (m & 1) ? -1.0 : 1.0;
      f.re[idx(n, m)] = x[idx(n, m)];
      f.im[idx(n, m)] = x[idx(n, -m)];
      f.re[idx(n, -m)] = s * x[idx(n, m)];
      f.im[idx(n, -m)] = -s * x[idx(n, -m)];
    }
  }
  return f;
}

/* Harmonic of the shift in full storage, from the m >= 0 recursion. */
static full full_harmonic(int p, const double x[3], int irregular) {
  double *re = malloc(sizeof(double) * nc(p)), *im = malloc(sizeof(double) * nc(p));
  if (irregular)
    irregular_cplx(p, x, re, im);
  else
    regular_cplx(p, x, re, im);
  full f = {p, calloc(nc(p), sizeof(double)), calloc(nc(p), sizeof(double))};
  for (int n = 0; n <= p; ++n)
    for (int q = -n; q <= n; ++q) getc(p, re, im, n, q, &f.re[idx(n, q)], &f.im[idx(n, q)]);
  free(re);
  free(im);
  return f;
}

static void full_free(full *f) {
  free(f->re);
  free(f->im);
}

static inline void fget(const full *f, int n, int q, double *a, double *b) {
  if (n < 0 || n > f->p || q > n || -q > n) {
    *a = *b = 0.0;
    return;
  }
  *a = f->re[idx(n, q)];
  *b = f->im[idx(n, q)];
}

/* zero[n] = 1 when every coefficient of degree n is zero (skipped in the sums). */
static char *zero_degrees(int p, const double *x) {
  char *z = malloc(p + 1);
  for (int n = 0; n <= p; ++n) {
    z[n] = 1;
    for (int m = -n; m <= n; ++m)
      if (x[idx(n, m)] != 0.0) z[n] = 0;
  }
  return z;
}

/* M2M: M'_j^i = sum_{k<=j} sum_l rho^k M_k^l conj R_{j-k}^{i-l}(b). */
void orc_m2m(int p, const double b[3], double rho, const double *in, double *out) {
  full a = full_from_real(p, in), h = full_harmonic(p, b, 0);
  char *zero = zero_degrees(p, in);
  for (int j = 0; j <= p; ++j)
    for (int i = 0; i <= j; ++i) {
      double sr = 0, si = 0;
      for (int k = 0; k <= j; ++k) {
        if (zero[k]) continue;
        double pk = pow(rho, k);
        for (int l = -k; l <= k; ++l) {
          double ar, ai, hr, hi;
          fget(&a, k, l, &ar, &ai);
          fget(&h, j - k, i - l, &hr, &hi);
          hi = -hi; /* conj */
          sr += pk * (ar * hr - ai * hi);
          si += pk * (ar * hi + ai * hr);
        }
      }
      out[idx(j, i)] += sr;
      if (i > 0) out[idx(j, -i)] += si;
    }
  full_free(&a);
  full_free(&h);
  free(zero);
}

/* L2L: L'_j^i = sigma^{j+1} sum_{n>=j} sum_m L_n^m R_{n-j}^{m-i}(t). */
void orc_l2l(int p, const double t[3], double sigma, const double *in, double *out) {
  full a = full_from_real(p, in), h = full_harmonic(p, t, 0);
  char *zero = zero_degrees(p, in);
  for (int j = 0; j <= p; ++j)
    for (int i = 0; i <= j; ++i) {
      double sr = 0, si = 0;
      for (int n = j; n <= p; ++n) {
        if (zero[n]) continue;
        for (int m = -n; m <= n; ++m) {
          double ar, ai, hr, hi;
          fget(&a, n, m, &ar, &ai);
          fget(&h, n - j, m - i, &hr, &hi);
          sr += ar * hr - ai * hi;
          si += ar * hi + ai * hr;
        }
      }
      double f = pow(sigma, j + 1);
      out[idx(j, i)] += f * sr;
      if (i > 0) out[idx(j, -i)] += f * si;
    }
  full_free(&a);
  full_free(&h);
  free(zero);
}

/* M2L: L'_j^i = (-1)^{j+i} sigma^{j+1} sum_n sum_m M_n^m I_{n+j}^{m-i}(b). */
void orc_m2l(int p, const double b[3], double sigma, const double *in, double *out) {
  full a = full_from_real(p, in), h = full_harmonic(2 * p, b, 1);
  char *zero = zero_degrees(p, in);
  for (int j = 0; j <= p; ++j)
    for (int i = 0; i <= j; ++i) {
      double sr = 0, si = 0;
      for (int n = 0; n <= p; ++n) {
        if (zero[n]) continue;
        for (int m = -n; m <= n; ++m) {
          double ar, ai, hr, hi;
          fget(&a, n, m, &ar, &ai);
          fget(&h, n + j, m - i, &hr, &hi);
          sr += ar * hr - ai * hi;
          si += ar * hi + ai * hr;
        }
      }
      double f = (((j + i) & 1) ? -1.0 : 1.0) * pow(sigma, j + 1);
      out[idx(j, i)] += f * sr;
      if (i > 0) out[idx(j, -i)] += f * si;
    }
  full_free(&a);
  full_free(&h);
  free(zero);
}

/* ---- offsets and tables (CONVENTIONS §3.12) --------------------------------------- */

static int kappa(int t) {
  int v = t + 1;
  return v < 0 ? 0 : (v > 3 ? 3 : v);
}

/* index(d), CONVENTIONS §3.12 "V-list offsets"; -1 outside the 316 offsets. */
int orc_offset_index(int dx, int dy, int dz) {
  if (dx < -3 || dx > 3 || dy < -3 || dy > 3 || dz < -3 || dz > 3) return -1;
  if (abs(dx) <= 1 && abs(dy) <= 1 && abs(dz) <= 1) return -1;
  int v = 49 * (dx + 3) + 7 * (dy + 3) + (dz + 3) - 9 * kappa(dx);
  if (abs(dx) <= 1) v -= 3 * kappa(dy) + (abs(dy) <= 1 ? kappa(dz) : 0);
  return v;
}

/* The 316 offsets in lexicographic (dx, dy, dz) order (V_LIST_DIRECTIONS,
 * fmm-plan/src/interaction_manager.rs:86). */
void orc_offset_of_index(int index, int d[3]) {
  int count = 0;
  for (int x = -3; x <= 3; ++x)
    for (int y = -3; y <= 3; ++y)
      for (int z = -3; z <= 3; ++z) {
        if (abs(x) <= 1 && abs(y) <= 1 && abs(z) <= 1) continue;
        if (count == index) {
          d[0] = x;
          d[1] = y;
          d[2] = z;
          return;
        }
        ++count;
      }
  d[0] = d[1] = d[2] = 0;
}

typedef void (*translation)(int, const double *, double, const double *, double *);

/* Column k of the matrix = the operator applied to unit vector k; column-major, entry
 * (i, k) at i + k n. */
static void build_matrix(int p, translation op, const double s[3], double ratio, double *out) {
  int n = nc(p);
  double *e = calloc(n, sizeof(double));
  for (int k = 0; k < n; ++k) {
    e[k] = 1.0;
    op(p, s, ratio, e, out + (size_t)k * n);
    e[k] = 0.0;
  }
  free(e);
}

void orc_table_m2m(int p, double *out) {
  size_t n2 = (size_t)nc(p) * nc(p);
  memset(out, 0, sizeof(double) * 8 * n2);
  for (int o = 0; o < 8; ++o) {
    double b[3] = {0.5 * (2 * ((o >> 2) & 1) - 1), 0.5 * (2 * ((o >> 1) & 1) - 1),
                   0.5 * (2 * (o & 1) - 1)};
    build_matrix(p, orc_m2m, b, 0.5, out + o * n2);
  }
}

void orc_table_l2l(int p, double *out) {
  size_t n2 = (size_t)nc(p) * nc(p);
  memset(out, 0, sizeof(double) * 8 * n2);
  for (int o = 0; o < 8; ++o) {
    double t[3] = {0.5 * (2 * ((o >> 2) & 1) - 1), 0.5 * (2 * ((o >> 1) & 1) - 1),
                   0.5 * (2 * (o & 1) - 1)};
    build_matrix(p, orc_l2l, t, 0.5, out + o * n2);
  }
}

void orc_table_m2l(int p, double *out) {
  size_t n2 = (size_t)nc(p) * nc(p);
  memset(out, 0, sizeof(double) * 316 * n2);
  for (int t = 0; t < 316; ++t) {
    int d[3];
    orc_offset_of_index(t, d);
    double b[3] = {2.0 * d[0], 2.0 * d[1], 2.0 * d[2]};
    build_matrix(p, orc_m2l, b, 1.0, out + t * n2);
  }
}
