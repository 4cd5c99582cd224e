/*
 * cpu_path.c -- the CPU baseline (test and benchmark infrastructure, not the product): a
 * C restatement of the reference's multi-threaded host evaluation over the oracle's tree
 * and lists, timed beside the GPU path.
 *
 * What it restates (tbetcke/fmm, host path of nd_fmm_exec):
 *   - threads: every level call hands targets (boxes or leaves) to the threads of a pool,
 *     each target's body serial and in its row's order (fmm-exec/src/operator.rs:240-260);
 *     here OpenMP, dynamic scheduling over targets or blocks of targets;
 *   - P2P: targets in SIMD lanes, each source broadcast to every lane in input order, with
 *     the sources of another leaf first mapped into the target leaf's frame
 *     (fmm-simd/src/lib.rs:58-75, fmm-exec/src/operator.rs:1388-1490).  On AVX-512F: 16
 *     targets per block (two vectors of 8), 1/r from vrsqrt14pd refined by two Newton
 *     steps, r^2 = fma(dz, dz, fma(dy, dy, dx dx)), rho cleared where r^2 = 0, phi += q rho,
 *     grad -= (q rho) rho^2 d; a partial block runs on a copy padded with its last target;
 *   - M2L under the Batched strategy (fmm-exec/src/operator.rs:128-140, 416-420): the
 *     level's boxes in blocks of 64 consecutive targets; per block, for each offset index
 *     d = 0..315 in turn, the block's pairs of offset d in one call of the column kernel
 *     (fmm-simd/src/matrix.rs:166): C[:, c_j] += A_d B[:, b_j], four source columns
 *     packed together, A read once per four columns, each element's chain over k
 *     ascending with one fma per step (MulAdd::Fused, the default from p = 6,
 *     fmm-exec/src/fmm.rs:501-518);
 *   - M2M and L2L as products over columns with the octant tables (the Batched strategy's
 *     families, fmm-exec/src/tables.rs:13): per block of parents, octant by octant;
 *   - P2M, L2P, P2L, M2P: the scalar leaf operators of the reference crate (here the
 *     oracle's, harmonics.c), one target per thread at a time.
 * Without AVX-512F the P2P and column kernels run their scalar loops (same order).
 */
#include <immintrin.h>
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"
#include "tree_internal.h"

#define BLOCK_TARGETS 64

static inline int ncoef(int p) { return (p + 1) * (p + 1); }

int orc_cpu_isa(void) {
  __builtin_cpu_init();
  return __builtin_cpu_supports("avx512f") ? 2 : 0;
}

int orc_cpu_threads(void) { return omp_get_max_threads(); }

/* ---- P2P: targets in lanes ------------------------------------------------------------ */

static void p2p_scalar(const double *y, const double *q, int64_t ns, const double *x, int64_t nt, double *phi,
                       double *grad) {
  for (int64_t i = 0; i < nt; ++i) {
    double p = phi[i], g0 = 0, g1 = 0, g2 = 0;
    if (grad) {
      g0 = grad[3 * i];
      g1 = grad[3 * i + 1];
      g2 = grad[3 * i + 2];
    }
    for (int64_t j = 0; j < ns; ++j) {
      const double d0 = x[3 * i] - y[3 * j], d1 = x[3 * i + 1] - y[3 * j + 1], d2 = x[3 * i + 2] - y[3 * j + 2];
      const double r2 = fma(d2, d2, fma(d1, d1, d0 * d0));
      const double rho = r2 == 0.0 ? 0.0 : 1.0 / sqrt(r2);
      const double qr = q[j] * rho;
      p += qr;
      if (grad) {
        const double t = qr * (rho * rho);
        g0 -= t * d0;
        g1 -= t * d1;
        g2 -= t * d2;
      }
    }
    phi[i] = p;
    if (grad) {
      grad[3 * i] = g0;
      grad[3 * i + 1] = g1;
      grad[3 * i + 2] = g2;
    }
  }
}

__attribute__((target("avx512f"))) static inline __m512d rsqrt_pd(__m512d r2) {
  const __m512d half = _mm512_set1_pd(0.5), three_halves = _mm512_set1_pd(1.5);
  __m512d y = _mm512_rsqrt14_pd(r2);
  const __m512d h = _mm512_mul_pd(half, r2);
  for (int it = 0; it < 2; ++it) {  // Newton: y (3/2 - r2/2 y^2)
    const __m512d yy = _mm512_mul_pd(y, y);
    y = _mm512_mul_pd(y, _mm512_fnmadd_pd(h, yy, three_halves));
  }
  const __mmask8 nz = _mm512_cmp_pd_mask(r2, _mm512_setzero_pd(), _CMP_NEQ_OQ);
  return _mm512_maskz_mov_pd(nz, y);
}

/* one block of 16 targets (two vectors), SoA copies tx/ty/tz, outputs accumulated in place */
__attribute__((target("avx512f"))) static void p2p_block16(const double *y, const double *q, int64_t ns,
                                                           const double *tx, const double *ty, const double *tz,
                                                           double *ph, double *gx, double *gy, double *gz,
                                                           int grad) {
  __m512d x0[2], x1[2], x2[2], p[2], a0[2], a1[2], a2[2];
  for (int v = 0; v < 2; ++v) {
    x0[v] = _mm512_loadu_pd(tx + 8 * v);
    x1[v] = _mm512_loadu_pd(ty + 8 * v);
    x2[v] = _mm512_loadu_pd(tz + 8 * v);
    p[v] = _mm512_loadu_pd(ph + 8 * v);
    a0[v] = _mm512_loadu_pd(gx + 8 * v);
    a1[v] = _mm512_loadu_pd(gy + 8 * v);
    a2[v] = _mm512_loadu_pd(gz + 8 * v);
  }
  for (int64_t j = 0; j < ns; ++j) {
    const __m512d y0 = _mm512_set1_pd(y[3 * j]), y1 = _mm512_set1_pd(y[3 * j + 1]),
                  y2 = _mm512_set1_pd(y[3 * j + 2]), qj = _mm512_set1_pd(q[j]);
    for (int v = 0; v < 2; ++v) {
      const __m512d d0 = _mm512_sub_pd(x0[v], y0), d1 = _mm512_sub_pd(x1[v], y1), d2 = _mm512_sub_pd(x2[v], y2);
      const __m512d r2 = _mm512_fmadd_pd(d2, d2, _mm512_fmadd_pd(d1, d1, _mm512_mul_pd(d0, d0)));
      const __m512d rho = rsqrt_pd(r2);
      const __m512d qr = _mm512_mul_pd(qj, rho);
      p[v] = _mm512_add_pd(p[v], qr);
      if (grad) {
        const __m512d t = _mm512_mul_pd(qr, _mm512_mul_pd(rho, rho));
        a0[v] = _mm512_fnmadd_pd(t, d0, a0[v]);
        a1[v] = _mm512_fnmadd_pd(t, d1, a1[v]);
        a2[v] = _mm512_fnmadd_pd(t, d2, a2[v]);
      }
    }
  }
  for (int v = 0; v < 2; ++v) {
    _mm512_storeu_pd(ph + 8 * v, p[v]);
    _mm512_storeu_pd(gx + 8 * v, a0[v]);
    _mm512_storeu_pd(gy + 8 * v, a1[v]);
    _mm512_storeu_pd(gz + 8 * v, a2[v]);
  }
}

static void p2p_avx512(const double *y, const double *q, int64_t ns, const double *x, int64_t nt, double *phi,
                       double *grad) {
  double tx[16], ty[16], tz[16], ph[16], gx[16], gy[16], gz[16];
  for (int64_t i0 = 0; i0 < nt; i0 += 16) {
    const int cnt = (int)(nt - i0 < 16 ? nt - i0 : 16);
    for (int l = 0; l < 16; ++l) {
      const int64_t i = i0 + (l < cnt ? l : cnt - 1);  // padded with the last target
      tx[l] = x[3 * i];
      ty[l] = x[3 * i + 1];
      tz[l] = x[3 * i + 2];
      ph[l] = phi[i];
      gx[l] = grad ? grad[3 * i] : 0.0;
      gy[l] = grad ? grad[3 * i + 1] : 0.0;
      gz[l] = grad ? grad[3 * i + 2] : 0.0;
    }
    p2p_block16(y, q, ns, tx, ty, tz, ph, gx, gy, gz, grad != NULL);
    for (int l = 0; l < cnt; ++l) {
      phi[i0 + l] = ph[l];
      if (grad) {
        grad[3 * (i0 + l)] = gx[l];
        grad[3 * (i0 + l) + 1] = gy[l];
        grad[3 * (i0 + l) + 2] = gz[l];
      }
    }
  }
}

/* ---- the column kernel: C[:, c_j] += A B[:, b_j] ------------------------------------ */

static void columns_scalar(int n, const double *a, const double *const *b, double *const *c, int ncols) {
  for (int j = 0; j < ncols; ++j)
    for (int i = 0; i < n; ++i) {
      double s = c[j][i];
      for (int k = 0; k < n; ++k) s = fma(a[(size_t)k * n + i], b[j][k], s);
      c[j][i] = s;
    }
}

/* four columns at a time, packed k-major (pack[4 k + c]); rows in runs of 32 (four
 * vectors, the last run masked); each element's chain over k ascending, one fma a step */
__attribute__((target("avx512f"))) static void columns_avx512(int n, const double *a, const double *const *b,
                                                             double *const *c, int ncols, double *pack) {
  for (int j0 = 0; j0 < ncols; j0 += 4) {
    const int cnt = ncols - j0 < 4 ? ncols - j0 : 4;
    for (int k = 0; k < n; ++k)
      for (int cc = 0; cc < 4; ++cc) pack[4 * k + cc] = cc < cnt ? b[j0 + cc][k] : 0.0;
    for (int i0 = 0; i0 < n; i0 += 32) {
      __mmask8 m[4];
      for (int r = 0; r < 4; ++r) {
        const int left = n - (i0 + 8 * r);
        m[r] = left >= 8 ? (__mmask8)0xFF : left > 0 ? (__mmask8)((1u << left) - 1) : (__mmask8)0;
      }
      __m512d acc[4][4];
      for (int cc = 0; cc < 4; ++cc)
        for (int r = 0; r < 4; ++r)
          acc[cc][r] = cc < cnt ? _mm512_maskz_loadu_pd(m[r], c[j0 + cc] + i0 + 8 * r) : _mm512_setzero_pd();
      for (int k = 0; k < n; ++k) {
        const double *ak = a + (size_t)k * n + i0;
        __m512d av[4];
        for (int r = 0; r < 4; ++r) av[r] = _mm512_maskz_loadu_pd(m[r], ak + 8 * r);
        for (int cc = 0; cc < 4; ++cc) {
          const __m512d bk = _mm512_set1_pd(pack[4 * k + cc]);
          for (int r = 0; r < 4; ++r) acc[cc][r] = _mm512_fmadd_pd(av[r], bk, acc[cc][r]);
        }
      }
      for (int cc = 0; cc < cnt; ++cc)
        for (int r = 0; r < 4; ++r) _mm512_mask_storeu_pd(c[j0 + cc] + i0 + 8 * r, m[r], acc[cc][r]);
    }
  }
}

static void columns(int isa, int n, const double *a, const double *const *b, double *const *c, int ncols,
                    double *pack) {
  if (ncols <= 0) return;
  if (isa == 2)
    columns_avx512(n, a, b, c, ncols, pack);
  else
    columns_scalar(n, a, b, c, ncols);
}

/* ---- the evaluation ------------------------------------------------------------------- */

void orc_fmm_evaluate_cpu(orc_fmm *f, const double *q_caller, double *phi, double *grad, int threads) {
  const int p = f->p, n = ncoef(p), L = f->nlevels - 1, isa = orc_cpu_isa();
  const size_t n2 = (size_t)n * n;
  if (threads > 0) omp_set_num_threads(threads);
  if (!f->mult) {
    f->mult = calloc(f->nlevels, sizeof(double *));
    f->loc = calloc(f->nlevels, sizeof(double *));
    for (int l = 0; l < f->nlevels; ++l) {
      f->mult[l] = malloc(((size_t)f->nbox[l] * n + 1) * sizeof(double));
      f->loc[l] = malloc(((size_t)f->nbox[l] * n + 1) * sizeof(double));
    }
  }
  double **mult = f->mult, **loc = f->loc;
  for (int l = 0; l < f->nlevels; ++l) {
    memset(mult[l], 0, ((size_t)f->nbox[l] * n + 1) * sizeof(double));
    memset(loc[l], 0, ((size_t)f->nbox[l] * n + 1) * sizeof(double));
  }
  double *q = malloc(sizeof(double) * (f->ns + 1));
  double *ph = calloc(f->nt + 1, sizeof(double)), *gr = calloc(3 * f->nt + 3, sizeof(double));
  double *gp = f->gradients ? gr : NULL;
  const double zero3[3] = {0, 0, 0};
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < f->ns; ++i) q[i] = q_caller[f->src_order[i]];

  /* upward: P2M per leaf, then M2M level by level, octant by octant per block of parents */
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t lf = 0; lf < f->nleaves; ++lf) {
    const int l = orc_key_level(f->leaf_key[lf]);
    const int64_t b = f->leaf_box[lf], s0 = f->src_start[lf], s1 = f->src_start[lf + 1];
    orc_p2m(p, zero3, 1.0, f->us + 3 * s0, q + s0, s1 - s0, mult[l] + (size_t)b * n);
  }
  for (int l = L; l >= 1; --l) {
    const int64_t np = f->nbox[l - 1];
#pragma omp parallel
    {
      const double *bcol[BLOCK_TARGETS];
      double *ccol[BLOCK_TARGETS];
      double *pack = malloc(sizeof(double) * 4 * n);
#pragma omp for schedule(dynamic, 1)
      for (int64_t b0 = 0; b0 < np; b0 += BLOCK_TARGETS) {
        const int64_t b1 = b0 + BLOCK_TARGETS < np ? b0 + BLOCK_TARGETS : np;
        for (int o = 0; o < 8; ++o) {
          int cnt = 0;
          for (int64_t b = b0; b < b1; ++b) {
            const int64_t c0 = f->first_child[l - 1][b];
            if (c0 < 0) continue;
            bcol[cnt] = mult[l] + (size_t)(c0 + o) * n;
            ccol[cnt++] = mult[l - 1] + (size_t)b * n;
          }
          columns(isa, n, f->m2m + o * n2, bcol, ccol, cnt, pack);
        }
      }
      free(pack);
    }
  }

  /* downward, per level: L2L (each box from its parent), M2L batched by offset, P2L */
  for (int l = 1; l <= L; ++l) {
    const int64_t nb = f->nbox[l];
#pragma omp parallel
    {
      const double *bcol[BLOCK_TARGETS];
      double *ccol[BLOCK_TARGETS];
      double *pack = malloc(sizeof(double) * 4 * n);
      int64_t *cursor = malloc(sizeof(int64_t) * BLOCK_TARGETS);
#pragma omp for schedule(dynamic, 1)
      for (int64_t b0 = 0; b0 < nb; b0 += BLOCK_TARGETS) {
        const int64_t b1 = b0 + BLOCK_TARGETS < nb ? b0 + BLOCK_TARGETS : nb;
        for (int o = 0; o < 8; ++o) {
          int cnt = 0;
          for (int64_t b = b0; b < b1; ++b) {
            int64_t d[4];
            orc_decode(f->keys[l][b], d);
            if ((int)(((d[1] & 1) << 2) | ((d[2] & 1) << 1) | (d[3] & 1)) != o) continue;
            const int64_t pb = orc_find_key(f->keys[l - 1], f->nbox[l - 1], orc_parent(f->keys[l][b]));
            bcol[cnt] = loc[l - 1] + (size_t)pb * n;
            ccol[cnt++] = loc[l] + (size_t)b * n;
          }
          columns(isa, n, f->l2l + o * n2, bcol, ccol, cnt, pack);
        }
        /* M2L: offsets in index order; a row holds each offset at most once, ascending */
        for (int64_t b = b0; b < b1; ++b) cursor[b - b0] = f->v[l].off[b];
        for (int t = 0; t < 316; ++t) {
          int cnt = 0;
          for (int64_t b = b0; b < b1; ++b) {
            int64_t *e = &cursor[b - b0];
            if (*e < f->v[l].off[b + 1] && f->v[l].aux[*e] == t) {
              bcol[cnt] = mult[l] + (size_t)f->v[l].ent[*e] * n;
              ccol[cnt++] = loc[l] + (size_t)b * n;
              ++*e;
            }
          }
          columns(isa, n, f->m2l + (size_t)t * n2, bcol, ccol, cnt, pack);
        }
        for (int64_t b = b0; b < b1; ++b)
          for (int64_t e = f->x[l].off[b]; e < f->x[l].off[b + 1]; ++e) {
            const int64_t sl = f->x[l].ent[e], s0 = f->src_start[sl], s1 = f->src_start[sl + 1];
            double c[3], r;
            orc_relative_frame(f->keys[l][b], f->leaf_key[sl], c, &r);
            orc_p2l(p, c, r, f->us + 3 * s0, q + s0, s1 - s0, loc[l] + (size_t)b * n);
          }
      }
      free(cursor);
      free(pack);
    }
  }

  /* leaves: L2P, M2P (W list), P2P (near row, sources mapped into the target frame) */
  int64_t maxs = 1;
  for (int64_t lf = 0; lf < f->nleaves; ++lf)
    if (f->src_start[lf + 1] - f->src_start[lf] > maxs) maxs = f->src_start[lf + 1] - f->src_start[lf];
#pragma omp parallel
  {
    double *y = malloc(sizeof(double) * 3 * (size_t)maxs);
#pragma omp for schedule(dynamic, 4)
    for (int64_t lf = 0; lf < f->nleaves; ++lf) {
      const int l = orc_key_level(f->leaf_key[lf]);
      const int64_t r = lf - f->leaf_off[l], t0 = f->tgt_start[lf], t1 = f->tgt_start[lf + 1], b = f->leaf_box[lf];
      double *pt = ph + t0, *gt = gp ? gp + 3 * t0 : NULL;
      orc_l2p(p, zero3, 1.0, loc[l] + (size_t)b * n, f->ut + 3 * t0, t1 - t0, pt, gt);
      for (int64_t e = f->w[l].off[r]; e < f->w[l].off[r + 1]; ++e) {
        const int64_t sb = f->w[l].ent[e];
        const int sl = (int)f->w[l].aux[e];
        double c[3], rr;
        orc_relative_frame(f->keys[sl][sb], f->leaf_key[lf], c, &rr);
        orc_m2p(p, c, rr, mult[sl] + (size_t)sb * n, f->ut + 3 * t0, t1 - t0, pt, gt);
      }
      for (int64_t e = f->u[l].off[r]; e < f->u[l].off[r + 1]; ++e) {
        const int64_t sl = f->u[l].ent[e], s0 = f->src_start[sl], s1 = f->src_start[sl + 1];
        const double *ys = f->us + 3 * s0;
        if (sl != lf) {
          double c[3], rr;
          orc_relative_frame(f->leaf_key[sl], f->leaf_key[lf], c, &rr);
          for (int64_t j = 0; j < s1 - s0; ++j)
            for (int k = 0; k < 3; ++k) y[3 * j + k] = c[k] + rr * ys[3 * j + k];
          ys = y;
        }
        if (isa == 2)
          p2p_avx512(ys, q + s0, s1 - s0, f->ut + 3 * t0, t1 - t0, pt, gt);
        else
          p2p_scalar(ys, q + s0, s1 - s0, f->ut + 3 * t0, t1 - t0, pt, gt);
      }
    }
    free(y);
  }

  /* output pass: phi = phi^ / (4 pi r_t), grad = g^ / (4 pi r_t^2) */
#pragma omp parallel for schedule(static)
  for (int64_t lf = 0; lf < f->nleaves; ++lf) {
    const int lvl = orc_key_level(f->leaf_key[lf]);
    const double rt = f->side / ldexp(1.0, lvl + 1);
    const double s1 = 4.0 * 3.141592653589793 * rt, s2 = 4.0 * 3.141592653589793 * rt * rt;
    for (int64_t i = f->tgt_start[lf]; i < f->tgt_start[lf + 1]; ++i) {
      const int64_t c = f->tgt_order[i];
      phi[c] = ph[i] / s1;
      if (grad && f->gradients)
        for (int k = 0; k < 3; ++k) grad[3 * c + k] = gr[3 * i + k] / s2;
    }
  }
  free(q);
  free(ph);
  free(gr);
}
