/*
 * hexpansion.c -- oracle (test infrastructure): Helmholtz expansions, leaf operators and
 * translations in f64, written from the reference's conventions for the Helmholtz family
 * (docs/CONVENTIONS.md:1435-1700, "The Helmholtz family"; nd_fmm_ref::helmholtz,
 * fmm-ref/src/helmholtz/leaf.rs and rcr.rs).
 *
 * Basis: real orthonormal harmonics Y_n^m (cos part in slot +m, sin part in slot -m, no
 * Condon-Shortley phase), as solid harmonics S_n^m(x) = |x|^n Y_n^m(x^), obtained from the
 * Racah harmonics R_n^m of harmonics.c by S = sqrt((2n+1)/4pi) nu R with nu = n! for m = 0
 * and sqrt(2 (n+m)! (n-m)!) on Re / Im R_n^m otherwise (CONVENTIONS:1496-1506).
 *
 * Everything works in a level's units (half-width 1, circumradius a = sqrt(3), wavenumber
 * kh = k r_l; CONVENTIONS:1592-1620): s = min(|kh| a, 1), beta = Im(kh) a.  Stored
 * coefficients Mh = e^-beta M / s^n and Lh = e^beta L s^n (CONVENTIONS:1528-1545), with
 *   M_n^m = sum q j_n(kh rho) Y_n^m,   L_n^m = 4 pi i kh sum q h_n(kh rho) Y_n^m.
 * The radial functions enter through g_n(z) = j_n(z) / z^n (even, entire) and
 * e_n(z) = h_n(z) z^(n+1), so that j_n Y = (kh)^n g_n S and h_n Y = e_n S / (kh^(n+1) |x|^(2n+1)).
 *
 * A complex expansion is two real planes (real part, then imaginary part) of (p+1)^2 values.
 * Translations are dense complex matrices (column-major, entry (i, k) at i + k (p+1)^2, real
 * plane then imaginary plane) built as rotation, coaxial shift, rotation back: rotation
 * blocks by exact quadrature of Y_n(Q x) Y_n(x)^T over the sphere, coaxial blocks by the
 * recursions of CONVENTIONS:1656-1690 (start column from Gegenbauer's addition theorem).
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

typedef double complex cplx;

static const double kPi = 3.141592653589793;
static inline int idx(int n, int m) { return n * n + n + m; }
static inline int nc(int p) { return (p + 1) * (p + 1); }

/* ---- basis ------------------------------------------------------------------------ */

static double norm_factor(int n, int m) {
  double f = sqrt((2.0 * n + 1.0) / (4.0 * kPi));
  if (m == 0) {
    for (int k = 2; k <= n; ++k) f *= k;
    return f;
  }
  double v = 2.0;
  for (int k = 2; k <= n + m; ++k) v *= k;
  for (int k = 2; k <= n - m; ++k) v *= k;
  return f * sqrt(v);
}

void orc_h_solid(int p, const double x[3], double *out) {
  orc_regular(p, x, out);
  for (int n = 0; n <= p; ++n)
    for (int m = -n; m <= n; ++m) out[idx(n, m)] *= norm_factor(n, m < 0 ? -m : m);
}

void orc_h_solid_grad(int p, const double x[3], double *gx, double *gy, double *gz) {
  orc_regular_grad(p, x, gx, gy, gz);
  for (int n = 0; n <= p; ++n)
    for (int m = -n; m <= n; ++m) {
      const double f = norm_factor(n, m < 0 ? -m : m);
      gx[idx(n, m)] *= f;
      gy[idx(n, m)] *= f;
      gz[idx(n, m)] *= f;
    }
}

/* ---- radial functions ------------------------------------------------------------- */

/* g_n(z) = j_n(z) / z^n for n = 0..P: Miller's backward recurrence
 * g_{n-1} = (2n+1) g_n - z^2 g_{n+1}, normalised by g_0 = sin z / z (and, where |z| >= 1,
 * together with g_1 = (sin z / z - cos z) / z^2 by least squares, which keeps it away
 * from the zeros of either). */
static void radial_j(int P, cplx z, cplx *g) {
  const double az = cabs(z);
  const int N = P + 30 + (int)(2.0 * az);
  const cplx z2 = z * z;
  cplx a = 0.0, b = 1.0;
  for (int n = N; n >= 1; --n) {
    const cplx c = (2.0 * n + 1.0) * b - z2 * a;
    a = b;
    b = c;
    if (n - 1 <= P) g[n - 1] = b;
    if (cabs(b) > 1e200) {
      a *= 1e-200;
      b *= 1e-200;
      for (int k = n - 1; k <= P; ++k) g[k] *= 1e-200;
    }
  }
  const cplx g0 = az == 0.0 ? 1.0 : csin(z) / z;
  cplx sc;
  if (az < 1.0) {
    sc = g0 / g[0];
  } else {
    const cplx g1 = (csin(z) / z - ccos(z)) / z2;
    const double big = fmax(cabs(g[0]), cabs(g[1]));
    const cplx G0 = g[0] / big, G1 = g[1] / big;
    sc = (conj(G0) * g0 + conj(G1) * g1) / (creal(G0 * conj(G0)) + creal(G1 * conj(G1))) / big;
  }
  for (int k = 0; k <= P; ++k) g[k] *= sc;
}

/* e_n(z) = h_n(z) z^(n+1), n = 0..P, h_n = h_n^(1): e_0 = -i e^{iz}, e_1 = -(z + i) e^{iz},
 * e_{n+1} = (2n+1) e_n - z^2 e_{n-1} (forward, the dominant direction). */
static void radial_h(int P, cplx z, cplx *e) {
  const cplx w = cexp(I * z), z2 = z * z;
  e[0] = -I * w;
  if (P >= 1) e[1] = -(z + I) * w;
  for (int n = 1; n < P; ++n) e[n + 1] = (2.0 * n + 1.0) * e[n] - z2 * e[n - 1];
}

void orc_h_radial(int P, double zre, double zim, double *j, double *h) {
  cplx *g = malloc(sizeof(cplx) * (P + 1)), *e = malloc(sizeof(cplx) * (P + 1));
  const cplx z = zre + I * zim;
  radial_j(P, z, g);
  radial_h(P, z, e);
  cplx zn = 1.0;
  for (int n = 0; n <= P; ++n) {
    const cplx jn = g[n] * zn;
    zn *= z;
    const cplx hn = e[n] / zn;
    j[2 * n] = creal(jn);
    j[2 * n + 1] = cimag(jn);
    h[2 * n] = creal(hn);
    h[2 * n + 1] = cimag(hn);
  }
  free(g);
  free(e);
}

typedef struct {
  cplx k;
  double s, beta;
} level_t;

static level_t level_of(cplx kh) {
  level_t L;
  L.k = kh;
  L.s = fmin(cabs(kh) * sqrt(3.0), 1.0);
  L.beta = cimag(kh) * sqrt(3.0);
  return L;
}

/* ---- leaf operators (stored form, level units) -------------------------------------- */

/* P2M: Mh += e^-beta (kh/s)^n q g_n(kh |u|) S_n^m(u), u = (y - c) / r. */
void orc_h_p2m(int p, double kre, double kim, const double c[3], double r, const double *src, const double *q,
               int64_t n, double *out) {
  const level_t L = level_of(kre + I * kim);
  const int N = nc(p);
  double *S = malloc(sizeof(double) * N);
  cplx *g = malloc(sizeof(cplx) * (p + 2)), *pw = malloc(sizeof(cplx) * (p + 1));
  pw[0] = exp(-L.beta);
  for (int d = 1; d <= p; ++d) pw[d] = pw[d - 1] * L.k / L.s;
  for (int64_t j = 0; j < n; ++j) {
    double u[3];
    for (int k = 0; k < 3; ++k) u[k] = (src[3 * j + k] - c[k]) / r;
    const double rho = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    radial_j(p, L.k * rho, g);
    orc_h_solid(p, u, S);
    const cplx qj = q[2 * j] + I * q[2 * j + 1];
    for (int d = 0; d <= p; ++d) {
      const cplx f = pw[d] * g[d] * qj;
      for (int m = -d; m <= d; ++m) {
        out[idx(d, m)] += creal(f) * S[idx(d, m)];
        out[N + idx(d, m)] += cimag(f) * S[idx(d, m)];
      }
    }
  }
  free(S);
  free(g);
  free(pw);
}

/* L2P: phi += (1/r) F(v), grad += (1/r^2) grad F(v), v = (x - c) / r, with
 * F = e^-beta sum Lh (kh/s)^n g_n(kh |v|) S_n^m(v) and grad g_n(kh|v|) = -kh^2 g_{n+1} v. */
void orc_h_l2p(int p, double kre, double kim, const double c[3], double r, const double *lc, const double *tgt,
               int64_t n, double *phi, double *grad) {
  const level_t L = level_of(kre + I * kim);
  const int N = nc(p);
  double *S = malloc(sizeof(double) * 4 * N);
  cplx *g = malloc(sizeof(cplx) * (p + 2)), *pw = malloc(sizeof(cplx) * (p + 1));
  pw[0] = exp(-L.beta);
  for (int d = 1; d <= p; ++d) pw[d] = pw[d - 1] * L.k / L.s;
  for (int64_t i = 0; i < n; ++i) {
    double v[3];
    for (int k = 0; k < 3; ++k) v[k] = (tgt[3 * i + k] - c[k]) / r;
    const double rho = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    radial_j(p + 1, L.k * rho, g);
    orc_h_solid(p, v, S);
    orc_h_solid_grad(p, v, S + N, S + 2 * N, S + 3 * N);
    cplx f = 0, gr[3] = {0, 0, 0}, rad = 0;
    for (int d = 0; d <= p; ++d) {
      const cplx cn = pw[d] * g[d], dn = -L.k * L.k * pw[d] * g[d + 1];
      for (int m = -d; m <= d; ++m) {
        const int t = idx(d, m);
        const cplx l = lc[t] + I * lc[N + t];
        f += cn * l * S[t];
        rad += dn * l * S[t];
        for (int k = 0; k < 3; ++k) gr[k] += cn * l * S[(k + 1) * N + t];
      }
    }
    phi[2 * i] += creal(f) / r;
    phi[2 * i + 1] += cimag(f) / r;
    if (grad)
      for (int k = 0; k < 3; ++k) {
        const cplx gk = (gr[k] + rad * v[k]) / (r * r);
        grad[6 * i + 2 * k] += creal(gk);
        grad[6 * i + 2 * k + 1] += cimag(gk);
      }
  }
  free(S);
  free(g);
  free(pw);
}

/* P2L: Lh += 4 pi i e^beta (s/kh)^n q e_n(kh |u|) S_n^m(u) / |u|^(2n+1), u = (y - c) / r
 * (the charges' positions in the units of the target box). */
void orc_h_p2l(int p, double kre, double kim, const double c[3], double r, const double *src, const double *q,
               int64_t n, double *out) {
  const level_t L = level_of(kre + I * kim);
  const int N = nc(p);
  double *S = malloc(sizeof(double) * N);
  cplx *e = malloc(sizeof(cplx) * (p + 2));
  for (int64_t j = 0; j < n; ++j) {
    double u[3];
    for (int k = 0; k < 3; ++k) u[k] = (src[3 * j + k] - c[k]) / r;
    const double rho2 = u[0] * u[0] + u[1] * u[1] + u[2] * u[2], rho = sqrt(rho2);
    radial_h(p, L.k * rho, e);
    orc_h_solid(p, u, S);
    const cplx qj = q[2 * j] + I * q[2 * j + 1];
    cplx f = 4.0 * kPi * I * exp(L.beta) * qj / rho;
    for (int d = 0; d <= p; ++d) {
      const cplx fd = f * e[d];
      for (int m = -d; m <= d; ++m) {
        out[idx(d, m)] += creal(fd) * S[idx(d, m)];
        out[N + idx(d, m)] += cimag(fd) * S[idx(d, m)];
      }
      f *= L.s / (L.k * rho2);
    }
  }
  free(S);
  free(e);
}

/* M2P: phi += (1/r) F(v), grad += (1/r^2) grad F(v), v = (x - c) / r, with
 * F = 4 pi i e^beta sum (s/kh)^n Mh e_n(kh|v|) S_n^m(v) / |v|^(2n+1); e_n'(z) = z e_{n-1}(z)
 * (e_0' = e^{iz}). */
void orc_h_m2p(int p, double kre, double kim, const double c[3], double r, const double *mc, const double *tgt,
               int64_t n, double *phi, double *grad) {
  const level_t L = level_of(kre + I * kim);
  const int N = nc(p);
  double *S = malloc(sizeof(double) * 4 * N);
  cplx *e = malloc(sizeof(cplx) * (p + 2));
  for (int64_t i = 0; i < n; ++i) {
    double v[3];
    for (int k = 0; k < 3; ++k) v[k] = (tgt[3 * i + k] - c[k]) / r;
    const double rho2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2], rho = sqrt(rho2);
    const cplx z = L.k * rho;
    radial_h(p, z, e);
    orc_h_solid(p, v, S);
    orc_h_solid_grad(p, v, S + N, S + 2 * N, S + 3 * N);
    cplx f = 0, gr[3] = {0, 0, 0}, rad = 0;
    cplx w = 4.0 * kPi * I * exp(L.beta) / rho; /* 4 pi i e^beta (s/kh)^n / rho^(2n+1) */
    for (int d = 0; d <= p; ++d) {
      const cplx ed = e[d], dd = d == 0 ? cexp(I * z) : z * e[d - 1];
      /* radial factor h(rho) = w e_d; dh/drho / rho = w (kh e_d' - (2d+1) e_d / rho) / rho */
      const cplx hv = w * ed, hr = w * (L.k * dd - (2.0 * d + 1.0) * ed / rho) / rho;
      for (int m = -d; m <= d; ++m) {
        const int t = idx(d, m);
        const cplx mm = mc[t] + I * mc[N + t];
        f += hv * mm * S[t];
        rad += hr * mm * S[t];
        for (int k = 0; k < 3; ++k) gr[k] += hv * mm * S[(k + 1) * N + t];
      }
      w *= L.s / (L.k * rho2);
    }
    phi[2 * i] += creal(f) / r;
    phi[2 * i + 1] += cimag(f) / r;
    if (grad)
      for (int k = 0; k < 3; ++k) {
        const cplx gk = (gr[k] + rad * v[k]) / (r * r);
        grad[6 * i + 2 * k] += creal(gk);
        grad[6 * i + 2 * k + 1] += cimag(gk);
      }
  }
  free(S);
  free(e);
}

/* ---- rotations -------------------------------------------------------------------- */

/* Gauss-Legendre nodes and weights on [-1, 1] (Newton on P_n from the Chebyshev guess). */
static void gauss_legendre(int n, double *x, double *w) {
  for (int i = 0; i < n; ++i) {
    double t = cos(kPi * (i + 0.75) / (n + 0.5)), dp = 1.0;
    for (int it = 0; it < 100; ++it) {
      double p0 = 1.0, p1 = t;
      for (int k = 2; k <= n; ++k) {
        const double p2 = ((2.0 * k - 1.0) * t * p1 - (k - 1.0) * p0) / k;
        p0 = p1;
        p1 = p2;
      }
      if (n == 1) p0 = 1.0;
      dp = n * (t * p1 - p0) / (t * t - 1.0);
      const double dt = p1 / dp;
      t -= dt;
      if (fabs(dt) < 1e-16) break;
    }
    double p0 = 1.0, p1 = t;
    for (int k = 2; k <= n; ++k) {
      const double p2 = ((2.0 * k - 1.0) * t * p1 - (k - 1.0) * p0) / k;
      p0 = p1;
      p1 = p2;
    }
    if (n == 1) p0 = 1.0;
    dp = n * (t * p1 - p0) / (t * t - 1.0);
    x[i] = t;
    w[i] = 2.0 / ((1.0 - t * t) * dp * dp);
  }
}

/* Blocks O^n(Q), Y_n(Q x) = O^n Y_n(x), degree n at offset sum_{n'<n} (2n'+1)^2, row-major
 * [m' + n][m + n]: O = integral of Y_n(Q x) Y_n(x)^T over the sphere, by a product rule
 * exact for degree 2p (Gauss-Legendre in cos theta, trapezoid in phi). */
void orc_h_rotation_blocks(int p, const double Q[9], double *out) {
  const int NG = p + 1, NP = 2 * p + 2, N = nc(p);
  double *xg = malloc(sizeof(double) * NG), *wg = malloc(sizeof(double) * NG);
  double *a = malloc(sizeof(double) * N), *b = malloc(sizeof(double) * N);
  gauss_legendre(NG, xg, wg);
  int total = 0;
  for (int n = 0; n <= p; ++n) total += (2 * n + 1) * (2 * n + 1);
  memset(out, 0, sizeof(double) * total);
  for (int i = 0; i < NG; ++i)
    for (int j = 0; j < NP; ++j) {
      const double ph = 2.0 * kPi * j / NP, st = sqrt(1.0 - xg[i] * xg[i]);
      const double x[3] = {st * cos(ph), st * sin(ph), xg[i]};
      const double y[3] = {Q[0] * x[0] + Q[1] * x[1] + Q[2] * x[2], Q[3] * x[0] + Q[4] * x[1] + Q[5] * x[2],
                           Q[6] * x[0] + Q[7] * x[1] + Q[8] * x[2]};
      orc_h_solid(p, y, a);
      orc_h_solid(p, x, b);
      const double w = wg[i] * 2.0 * kPi / NP;
      int off = 0;
      for (int n = 0; n <= p; ++n) {
        const int W = 2 * n + 1;
        for (int r = 0; r < W; ++r)
          for (int c = 0; c < W; ++c) out[off + r * W + c] += w * a[idx(n, r - n)] * b[idx(n, c - n)];
        off += W * W;
      }
    }
  free(xg);
  free(wg);
  free(a);
  free(b);
}

/* Q = R_y(-theta) R_z(-phi): takes the unit vector of (theta, phi) to +z. */
static void rotation_to_z(const double t[3], double Q[9]) {
  const double r = sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
  const double th = acos(fmax(-1.0, fmin(1.0, t[2] / r))), ph = atan2(t[1], t[0]);
  const double ct = cos(th), st = sin(th), cp = cos(ph), sp = sin(ph);
  /* R_z(-phi) = [[cp, sp, 0], [-sp, cp, 0], [0, 0, 1]]; R_y(-theta) = [[ct, 0, -st], [0, 1, 0], [st, 0, ct]] */
  const double A[9] = {cp, sp, 0, -sp, cp, 0, 0, 0, 1};
  const double B[9] = {ct, 0, -st, 0, 1, 0, st, 0, ct};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0;
      for (int k = 0; k < 3; ++k) s += B[3 * i + k] * A[3 * k + j];
      Q[3 * i + j] = s;
    }
}

/* ---- coaxial translations ---------------------------------------------------------- */

/* E^m_{n'n}(t) for 0 <= m <= p, m <= n', n <= p, out[(m (p+1) + n') (p+1) + n], from the
 * start column f[n'] = f_{n'}(k t), n' = 0..2p (CONVENTIONS:1656-1690):
 *   E^0_{n'0} = (-1)^n' sqrt(2n'+1) f_n',
 *   a_n E_{n',n+1} = a_{n-1} E_{n',n-1} - a_n' E_{n'+1,n} + a_{n'-1} E_{n'-1,n},
 *   g_m^m E^{m+1}_{n',m+1} = b^m_{n'+1} E^m_{n'+1,m} + g^m_{n'-1} E^m_{n'-1,m},
 * on the triangle n' >= n, the rest by the index symmetry (CONVENTIONS:1646-1650). */
static double ca(int n, int m) {
  if (n < 0) return 0.0;
  return sqrt((double)(n + 1 + m) * (n + 1 - m) / ((2.0 * n + 1.0) * (2.0 * n + 3.0)));
}
static double cb(int n, int m) {
  if (n < 1) return 0.0;
  return sqrt((double)(n - m - 1) * (n - m) / ((2.0 * n - 1.0) * (2.0 * n + 1.0)));
}
static double cg(int n, int m) {
  if (n < 0) return 0.0;
  return sqrt((double)(n + m + 1) * (n + m + 2) / ((2.0 * n + 1.0) * (2.0 * n + 3.0)));
}

static void coaxial(int p, const cplx *f, cplx *out) {
  const int R = 2 * p + 2, C = p + 1;
  cplx *W = calloc((size_t)R * C, sizeof(cplx)), *Wn = calloc((size_t)R * C, sizeof(cplx));
#define AT(A, np, n) A[(size_t)(np) * C + (n)]
  for (int np = 0; np <= 2 * p; ++np) AT(W, np, 0) = ((np & 1) ? -1.0 : 1.0) * sqrt(2.0 * np + 1.0) * f[np];
  for (int m = 0; m <= p; ++m) {
    if (m > 0) {
      /* column n = m of order m from column m-1 of order m-1 (held in W) */
      memset(Wn, 0, sizeof(cplx) * (size_t)R * C);
      for (int np = m; np <= 2 * p - m; ++np) {
        cplx v = cb(np + 1, m - 1) * AT(W, np + 1, m - 1);
        if (np - 1 >= m - 1) v += cg(np - 1, m - 1) * AT(W, np - 1, m - 1);
        AT(Wn, np, m) = v / cg(m - 1, m - 1);
      }
      cplx *t = W;
      W = Wn;
      Wn = t;
    }
    /* columns n = m+1 .. p of order m, on the triangle n' >= n (which the recursion
     * closes over; below it the entries are small differences of large ones) */
    for (int n = m; n < p; ++n)
      for (int np = n + 1; np <= 2 * p - n - 1; ++np) {
        cplx v = -ca(np, m) * AT(W, np + 1, n);
        if (n - 1 >= m) v += ca(n - 1, m) * AT(W, np, n - 1);
        if (np - 1 >= m) v += ca(np - 1, m) * AT(W, np - 1, n);
        AT(W, np, n + 1) = v / ca(n, m);
      }
    /* the rest by the index symmetry E_{n'n} = (-1)^{n+n'} E_{nn'} */
    for (int np = 0; np <= p; ++np)
      for (int n = 0; n <= p; ++n) {
        cplx v = 0.0;
        if (np >= m && n >= m) v = np >= n ? AT(W, np, n) : (((np + n) & 1) ? -1.0 : 1.0) * AT(W, n, np);
        out[((size_t)m * C + np) * C + n] = v;
      }
  }
#undef AT
  free(W);
  free(Wn);
}

void orc_h_coaxial(int p, int singular_start, double kre, double kim, double t, double *out) {
  const int P2 = 2 * p;
  cplx *f = malloc(sizeof(cplx) * (P2 + 1)), *g = malloc(sizeof(cplx) * (P2 + 1));
  const cplx z = (kre + I * kim) * t;
  if (singular_start) {
    radial_h(P2, z, g);
    cplx zn = z;
    for (int n = 0; n <= P2; ++n, zn *= z) f[n] = g[n] / zn;
  } else {
    radial_j(P2, z, g);
    cplx zn = 1.0;
    for (int n = 0; n <= P2; ++n, zn *= z) f[n] = g[n] * zn;
  }
  const size_t M = (size_t)(p + 1) * (p + 1) * (p + 1);
  cplx *E = malloc(sizeof(cplx) * M);
  coaxial(p, f, E);
  for (size_t i = 0; i < M; ++i) {
    out[2 * i] = creal(E[i]);
    out[2 * i + 1] = cimag(E[i]);
  }
  free(f);
  free(g);
  free(E);
}

/* ---- dense translation matrices ---------------------------------------------------- */

/* T = O(Q)^T A O(Q), A the coaxial blocks scaled per entry by scale(n', n):
 * out (complex, column-major, real plane then imaginary plane) += T. */
static void compose(int p, const double t[3], const cplx *E, const cplx *row_f, const cplx *col_f, cplx common,
                    double *out) {
  const int N = nc(p);
  double Q[9];
  rotation_to_z(t, Q);
  int total = 0;
  for (int n = 0; n <= p; ++n) total += (2 * n + 1) * (2 * n + 1);
  double *O = malloc(sizeof(double) * total);
  int *boff = malloc(sizeof(int) * (p + 1));
  orc_h_rotation_blocks(p, Q, O);
  for (int n = 0, o = 0; n <= p; ++n) {
    boff[n] = o;
    o += (2 * n + 1) * (2 * n + 1);
  }
  cplx *x = malloc(sizeof(cplx) * N), *y = malloc(sizeof(cplx) * N);
  const int C = p + 1;
  for (int k = 0; k < N; ++k) {
    /* column k: rotate the unit vector e_k (degree nk) */
    int nk = 0;
    while ((nk + 1) * (nk + 1) <= k) ++nk;
    const int mk = k - nk * nk - nk, Wk = 2 * nk + 1;
    for (int i = 0; i < N; ++i) x[i] = 0;
    for (int r = 0; r < Wk; ++r) x[idx(nk, r - nk)] = O[boff[nk] + r * Wk + (mk + nk)];
    /* coaxial: y_{n', m} = sum_n A^{|m|}_{n'n} x_{n, m} */
    for (int i = 0; i < N; ++i) y[i] = 0;
    for (int np = 0; np <= p; ++np)
      for (int m = -np; m <= np; ++m) {
        const int am = m < 0 ? -m : m;
        cplx v = 0;
        for (int n = am; n <= p; ++n)
          v += E[((size_t)am * C + np) * C + n] * row_f[np] * col_f[n] * x[idx(n, m)];
        y[idx(np, m)] = common * v;
      }
    /* rotate back: x = O^T y */
    for (int np = 0; np <= p; ++np) {
      const int W = 2 * np + 1;
      for (int c = 0; c < W; ++c) {
        cplx v = 0;
        for (int r = 0; r < W; ++r) v += O[boff[np] + r * W + c] * y[idx(np, r - np)];
        out[(size_t)k * N + idx(np, c - np)] += creal(v);
        out[(size_t)N * N + (size_t)k * N + idx(np, c - np)] += cimag(v);
      }
    }
  }
  free(O);
  free(boff);
  free(x);
  free(y);
}

static void translation(int p, int kind, cplx kt, cplx ks, const double t[3], double factor, double *out) {
  /* kind 0: M2M (S|S) child (source, ks) -> parent (target, kt); 1: L2L (R|R) parent -> child;
   * 2: M2L (S|R), same level (ks = kt) */
  const level_t Lt = level_of(kt), Ls = level_of(ks);
  const int C = p + 1;
  const double tn = sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
  /* the wavenumber of the shift is that of the level whose units t is in */
  const cplx kshift = kind == 1 ? ks : kt;
  double *Er = malloc(sizeof(double) * 2 * (size_t)C * C * C);
  orc_h_coaxial(p, kind == 2, creal(kshift), cimag(kshift), tn, Er);
  cplx *E = malloc(sizeof(cplx) * (size_t)C * C * C);
  for (size_t i = 0; i < (size_t)C * C * C; ++i) E[i] = Er[2 * i] + I * Er[2 * i + 1];
  cplx *rf = malloc(sizeof(cplx) * C), *cf = malloc(sizeof(cplx) * C);
  cplx common;
  /* stored source Mh/Lh -> unscaled: times e^{+-beta_s} s_s^{-+n}; unscaled -> stored target */
  for (int n = 0; n <= p; ++n) {
    if (kind == 0) {
      cf[n] = pow(Ls.s, n);      /* M = e^beta_s s_s^n Mh */
      rf[n] = pow(Lt.s, -n);     /* Mh = e^-beta_t M / s_t^n */
    } else if (kind == 1) {
      cf[n] = pow(Ls.s, -n);     /* L = e^-beta_s Lh / s_s^n */
      rf[n] = pow(Lt.s, n);      /* Lh = e^beta_t s_t^n L */
    } else {
      cf[n] = pow(Ls.s, n);
      rf[n] = pow(Lt.s, n);
    }
  }
  if (kind == 0)
    common = exp(Ls.beta - Lt.beta);
  else if (kind == 1)
    common = exp(Lt.beta - Ls.beta);
  else
    common = 4.0 * kPi * I * kt * exp(Ls.beta + Lt.beta);
  compose(p, t, E, rf, cf, common * factor, out);
  free(Er);
  free(E);
  free(rf);
  free(cf);
}

/* M2M from the child of octant o (o = 4x + 2y + z bits) into its parent; kh is the parent's
 * level wavenumber (the child's is kh / 2).  t = parent - child centre in parent units. */
void orc_h_m2m(int p, double kre, double kim, int o, double *out) {
  const double t[3] = {((o >> 2) & 1) ? -0.5 : 0.5, ((o >> 1) & 1) ? -0.5 : 0.5, (o & 1) ? -0.5 : 0.5};
  const cplx kp = kre + I * kim;
  memset(out, 0, sizeof(double) * 2 * (size_t)nc(p) * nc(p));
  translation(p, 0, kp, kp / 2.0, t, 1.0, out);
}

/* L2L from the parent (level wavenumber kh) into its child of octant o, in the child's
 * units (the factor 1/2 of the change of units).  t = child - parent centre. */
void orc_h_l2l(int p, double kre, double kim, int o, double *out) {
  const double t[3] = {((o >> 2) & 1) ? 0.5 : -0.5, ((o >> 1) & 1) ? 0.5 : -0.5, (o & 1) ? 0.5 : -0.5};
  const cplx kp = kre + I * kim;
  memset(out, 0, sizeof(double) * 2 * (size_t)nc(p) * nc(p));
  translation(p, 1, kp / 2.0, kp, t, 0.5, out);
}

/* M2L between boxes of one level whose cells differ by d = target - source (level units:
 * t = 2 d). */
void orc_h_m2l(int p, double kre, double kim, const int d[3], double *out) {
  const double t[3] = {2.0 * d[0], 2.0 * d[1], 2.0 * d[2]};
  const cplx k = kre + I * kim;
  memset(out, 0, sizeof(double) * 2 * (size_t)nc(p) * nc(p));
  translation(p, 2, k, k, t, 1.0, out);
}

/* y += T x for a complex table T (planes) and complex vectors (planes). */
void orc_h_apply(int p, const double *T, const double *x, double *y) {
  const int N = nc(p);
  const double *Tr = T, *Ti = T + (size_t)N * N;
  for (int k = 0; k < N; ++k) {
    const double xr = x[k], xi = x[N + k];
    if (xr == 0.0 && xi == 0.0) continue;
    for (int i = 0; i < N; ++i) {
      const double a = Tr[(size_t)k * N + i], b = Ti[(size_t)k * N + i];
      y[i] += a * xr - b * xi;
      y[N + i] += b * xr + a * xi;
    }
  }
}
