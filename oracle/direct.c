/*
 * direct.c -- oracle (test infrastructure): the reference P2P terms, the compensated
 * f64 direct sum and the error norms.
 *
 * Term: fmm-ref/src/p2p.rs:370-403 -- d = x - y, r2 = (d0^2 + d1^2) + d2^2, r = sqrt(r2),
 * potential q / r, gradient -(q dk) / (r2 r); a pair with x == y exactly is skipped.
 * Neumaier: fmm-ref/src/p2p.rs:407-431.  Norms: fmm-validate/src/metrics.rs:51-170.
 */
#include <math.h>
#include "oracle.h"

typedef struct { double sum, comp; } neumaier;

static inline void nm_add(neumaier *s, double term) {
  double t = s->sum + term;
  if (fabs(s->sum) >= fabs(term))
    s->comp += (s->sum - t) + term;
  else
    s->comp += (term - t) + s->sum;
  s->sum = t;
}

static inline double nm_value(const neumaier *s) { return s->sum + s->comp; }

double orc_neumaier_sum(const double *terms, int64_t n) {
  neumaier s = {0.0, 0.0};
  for (int64_t i = 0; i < n; ++i) nm_add(&s, terms[i]);
  return nm_value(&s);
}

/* Returns 0 when the pair coincides exactly (x == y in all components). */
static inline int term(const double *x, const double *y, double q, double *phi, double g[3]) {
  if (x[0] == y[0] && x[1] == y[1] && x[2] == y[2]) return 0;
  double d[3] = {x[0] - y[0], x[1] - y[1], x[2] - y[2]};
  double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
  double r = sqrt(r2);
  *phi = q / r;
  double r3 = r2 * r;
  for (int k = 0; k < 3; ++k) g[k] = -(q * d[k] / r3);
  return 1;
}

/* nd_fmm_ref::p2p::p2p (fmm-ref/src/p2p.rs:90-121): plain accumulation, input order. */
void orc_p2p(const double *src, const double *q, int64_t ns, const double *tgt, int64_t nt,
             double *phi, double *grad) {
  for (int64_t i = 0; i < nt; ++i) {
    for (int64_t j = 0; j < ns; ++j) {
      double tp, tg[3];
      if (!term(tgt + 3 * i, src + 3 * j, q[j], &tp, tg)) continue;
      phi[i] = phi[i] + tp;
      if (grad)
        for (int k = 0; k < 3; ++k) grad[3 * i + k] = grad[3 * i + k] + tg[k];
    }
  }
}

/* nd_fmm_ref::p2p::direct_sum (fmm-ref/src/p2p.rs:147-205): compensated per target,
 * then one final addition into the output. */
void orc_direct_sum(const double *src, const double *q, int64_t ns, const double *tgt,
                    int64_t nt, double *phi, double *grad) {
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t i = 0; i < nt; ++i) {
    neumaier s = {0, 0}, g[3] = {{0, 0}, {0, 0}, {0, 0}};
    for (int64_t j = 0; j < ns; ++j) {
      double tp, tg[3];
      if (!term(tgt + 3 * i, src + 3 * j, q[j], &tp, tg)) continue;
      nm_add(&s, tp);
      if (grad)
        for (int k = 0; k < 3; ++k) nm_add(&g[k], tg[k]);
    }
    phi[i] += nm_value(&s);
    if (grad)
      for (int k = 0; k < 3; ++k) grad[3 * i + k] += nm_value(&g[k]);
  }
}

/* ErrorNorms (fmm-validate/src/metrics.rs:51-170): out[0] relative L2, out[1] relative
 * max; for vec3 each entry is a 3-vector measured by its Euclidean norm. */
void orc_error_norms(const double *approx, const double *exact, int64_t n, int vec3,
                     double out[2]) {
  double esq = 0, rsq = 0, emax = 0, rmax = 0;
  for (int64_t i = 0; i < n; ++i) {
    double e, r;
    if (vec3) {
      double a = approx[3 * i] - exact[3 * i], b = approx[3 * i + 1] - exact[3 * i + 1],
             c = approx[3 * i + 2] - exact[3 * i + 2];
      e = sqrt(a * a + b * b + c * c);
      r = sqrt(exact[3 * i] * exact[3 * i] + exact[3 * i + 1] * exact[3 * i + 1] +
               exact[3 * i + 2] * exact[3 * i + 2]);
    } else {
      e = fabs(approx[i] - exact[i]);
      r = fabs(exact[i]);
    }
    esq += e * e;
    rsq += r * r;
    if (e > emax) emax = e;
    if (r > rmax) rmax = r;
  }
  double l2n = sqrt(esq), l2d = sqrt(rsq);
  out[0] = l2d == 0.0 ? (l2n == 0.0 ? 0.0 : INFINITY) : l2n / l2d;
  out[1] = rmax == 0.0 ? (emax == 0.0 ? 0.0 : INFINITY) : emax / rmax;
}
