/*
 * helmholtz.c -- oracle (test infrastructure): the Helmholtz P2P term and the f64
 * compensated direct sum of nd_fmm_ref::helmholtz::p2p (fmm-ref/src/helmholtz/p2p.rs).
 *
 * Term (p2p.rs module docs, "The term"): d = x - y, r2 = (d0^2 + d1^2) + d2^2, r = sqrt(r2),
 *   e = exp(-Im k r) (cos(Re k r) + i sin(Re k r)),   phi += (q e) / r,
 *   g = ((-Im k r - 1) + i Re k r) e / (r2 r),          grad_k += (q g) d_k,
 * no 1/(4 pi); a pair with x == y exactly is skipped.  direct_sum: each real component
 * summed with Neumaier's compensation (fmm-ref/src/p2p.rs:407-431), added once.
 * Complex values are interleaved (re, im).
 */
#include <math.h>
#include "oracle.h"

typedef struct { double sum, comp; } nm;

static inline void nm_add(nm *s, double t) {
  double u = s->sum + t;
  if (fabs(s->sum) >= fabs(t))
    s->comp += (s->sum - u) + t;
  else
    s->comp += (t - u) + s->sum;
  s->sum = u;
}

void orc_helmholtz_direct_sum(double kre, double kim, const double *src, const double *q, int64_t ns,
                              const double *tgt, int64_t nt, double *pot, double *grad) {
  for (int64_t i = 0; i < nt; ++i) {
    nm p[2] = {{0, 0}, {0, 0}}, g[6] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}};
    const double *x = tgt + 3 * i;
    for (int64_t j = 0; j < ns; ++j) {
      const double *y = src + 3 * j;
      if (x[0] == y[0] && x[1] == y[1] && x[2] == y[2]) continue;
      const double d[3] = {x[0] - y[0], x[1] - y[1], x[2] - y[2]};
      const double r2 = (d[0] * d[0] + d[1] * d[1]) + d[2] * d[2];
      const double r = sqrt(r2);
      const double damp = exp(-kim * r), ph = kre * r;
      const double er = damp * cos(ph), ei = damp * sin(ph);
      const double qr = q[2 * j], qi = q[2 * j + 1];
      /* q e */
      const double qer = qr * er - qi * ei, qei = qr * ei + qi * er;
      nm_add(&p[0], qer / r);
      nm_add(&p[1], qei / r);
      if (grad) {
        const double ar = -kim * r - 1.0, ai = kre * r, r3 = r2 * r;
        /* g = (a e) / r3, then q g */
        const double gr = (ar * er - ai * ei) / r3, gi = (ar * ei + ai * er) / r3;
        const double qgr = qr * gr - qi * gi, qgi = qr * gi + qi * gr;
        for (int k = 0; k < 3; ++k) {
          nm_add(&g[2 * k], qgr * d[k]);
          nm_add(&g[2 * k + 1], qgi * d[k]);
        }
      }
    }
    pot[2 * i] += p[0].sum + p[0].comp;
    pot[2 * i + 1] += p[1].sum + p[1].comp;
    if (grad)
      for (int c = 0; c < 6; ++c) grad[6 * i + c] += g[c].sum + g[c].comp;
  }
}
