// helmholtz.hip -- device kernels of the Helmholtz FMM (kernel e^{ikr}/r, complex charges):
// the leaf expansions (P2M, P2L), their evaluations (L2P, M2P), the near field (P2P), the
// output pass and a compensated direct sum.  The translations (M2M, L2L, M2L) are dense
// complex tables run on the MFMA contraction of contract.hip.
//
// Conventions (docs/CONVENTIONS.md:1435-1700; nd_fmm_ref::helmholtz, fmm-ref/src/helmholtz/
// leaf.rs and p2p.rs): real orthonormal harmonics Y_n^m (cos in slot +m, sin in slot -m, no
// Condon-Shortley phase), computed as S_n^m(x) = |x|^n Y_n^m(x^) = norm[idx] x (Re or Im of
// the Racah harmonic R_n^m) by the Cartesian recursions; every level in its own units
// (half-width 1, wavenumber kh, s = min(|kh| sqrt 3, 1), beta = Im(kh) sqrt 3); stored
// coefficients Mh = e^-beta M / s^n, Lh = e^beta L s^n with
//   M_n^m = sum q j_n(kh rho) Y_n^m,   L_n^m = 4 pi i kh sum q h_n(kh rho) Y_n^m.
// Radial functions enter as g_n(z) = j_n(z) / z^n (Miller's backward recurrence) and
// e_n(z) = h_n(z) z^(n+1) (forward recurrence), so that
//   j_n Y = kh^n g_n S,   h_n Y = e_n S / (kh^(n+1) |x|^(2n+1)),
// both regular in the coordinates; grad g_n(kh |x|) = -kh^2 g_{n+1} x and e_n' = z e_{n-1}.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "hkernels.hpp"
#include "kernels.hpp"

namespace fmmh {
namespace dev {
namespace {

constexpr double kPi = 3.141592653589793;
constexpr int HT = 16;  // points per tile of the expansion kernel (16 x 16 harmonic columns)

struct cx {
  double re, im;
};
__device__ __forceinline__ cx cadd(cx a, cx b) { return {a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ cx csub(cx a, cx b) { return {a.re - b.re, a.im - b.im}; }
__device__ __forceinline__ cx cmul(cx a, cx b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ __forceinline__ cx cscale(cx a, double s) { return {a.re * s, a.im * s}; }
__device__ __forceinline__ cx cdiv(cx a, cx b) {
  const double d = 1.0 / (b.re * b.re + b.im * b.im);
  return {(a.re * b.re + a.im * b.im) * d, (a.im * b.re - a.re * b.im) * d};
}
__device__ __forceinline__ cx cconj(cx a) { return {a.re, -a.im}; }
__device__ __forceinline__ double cabs1(cx a) { return fabs(a.re) + fabs(a.im); }
// sin, cos and exp(i z) of a complex argument
__device__ __forceinline__ void csincos(cx z, cx &sn, cx &cs) {
  double sa, ca;
  sincos(z.re, &sa, &ca);
  const double eb = exp(z.im), emb = exp(-z.im);
  const double ch = 0.5 * (eb + emb), sh = 0.5 * (eb - emb);
  sn = {sa * ch, ca * sh};
  cs = {ca * ch, -sa * sh};
}
__device__ __forceinline__ cx cexpi(cx z) {  // e^{iz} = e^{-Im z} (cos Re z + i sin Re z)
  double sa, ca;
  sincos(z.re, &sa, &ca);
  const double d = exp(-z.im);
  return {d * ca, d * sa};
}

// Frame of box/leaf cell cs seen from cell ct (exact from the integer cells): c = (C(s) -
// C(t)) 2^(l_t - L) in t's units, r = 2^(l_t - l_s) (docs/CONVENTIONS.md:787-816).
__device__ __forceinline__ void cell_frame(const uint32_t *cs, const uint32_t *ct, double c[3], double &r) {
  const int ls = int(cs[0]), lt = int(ct[0]);
  const int L = ls > lt ? ls : lt;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const int64_t a = int64_t(2 * cs[k + 1] + 1) << (L - ls);
    const int64_t b = int64_t(2 * ct[k + 1] + 1) << (L - lt);
    c[k] = ldexp(double(a - b), lt - L);
  }
  r = ldexp(1.0, lt - ls);
}

constexpr int kRad = kMaxDegree + 2;  // radial values per point: degrees 0 .. p + 1

// g_n(z) = j_n(z) / z^n for n = 0 .. P (P <= kMaxDegree + 1) by Miller's backward
// recurrence g_{n-1} = (2n+1) g_n - z^2 g_{n+1} from index N = P + 30 + 2|z|, normalised by
// g_0 = sin z / z, or, for |z| >= 1, by g_0 and g_1 = (sin z / z - cos z) / z^2 together
// in least squares (away from the zeros of either).
__device__ __forceinline__ void radial_j(int P, cx z, cx (&g)[kRad]) {
  const double az = sqrt(z.re * z.re + z.im * z.im);
  const int N = P + 30 + int(2.0 * az);
  const cx z2 = cmul(z, z);
  cx a = {0.0, 0.0}, b = {1.0, 0.0};
  for (int n = N; n > P; --n) {  // ends with b = g_P, a = g_{P+1} (unnormalised)
    const cx c = csub(cscale(b, 2.0 * n + 1.0), cmul(z2, a));
    a = b;
    b = c;
    if (cabs1(b) > 1e150) {
      a = cscale(a, 1e-150);
      b = cscale(b, 1e-150);
    }
  }
#pragma unroll
  for (int n = kRad - 1; n >= 1; --n) {
    if (n > P) continue;
    if (n == P) g[n] = b;
    const cx c = csub(cscale(b, 2.0 * n + 1.0), cmul(z2, a));
    a = b;
    b = c;
    g[n - 1] = b;
  }
  cx sc;
  if (az == 0.0) {
    sc = cdiv({1.0, 0.0}, g[0]);
  } else {
    cx sn, cs;
    csincos(z, sn, cs);
    const cx g0 = cdiv(sn, z);
    if (az < 1.0) {
      sc = cdiv(g0, g[0]);
    } else {
      const cx g1 = cdiv(csub(g0, cs), z2);
      const double big = fmax(cabs1(g[0]), cabs1(g[1]));
      const cx G0 = cscale(g[0], 1.0 / big), G1 = cscale(g[1], 1.0 / big);
      const cx num = cadd(cmul(cconj(G0), g0), cmul(cconj(G1), g1));
      const double den = G0.re * G0.re + G0.im * G0.im + G1.re * G1.re + G1.im * G1.im;
      sc = cscale(num, 1.0 / (den * big));
    }
  }
#pragma unroll
  for (int n = 0; n < kRad; ++n)
    if (n <= P) g[n] = cmul(g[n], sc);
}

// e_n(z) = h_n(z) z^(n+1), n = 0 .. P: e_0 = -i e^{iz}, e_1 = -(z + i) e^{iz},
// e_{n+1} = (2n+1) e_n - z^2 e_{n-1}; w = e^{iz} is returned for e_0' = e^{iz}.
__device__ __forceinline__ void radial_h(int P, cx z, cx (&e)[kRad], cx &w) {
  w = cexpi(z);
  const cx z2 = cmul(z, z);
  e[0] = {w.im, -w.re};
  e[1] = cmul({-z.re, -z.im - 1.0}, w);
#pragma unroll
  for (int n = 1; n + 1 < kRad; ++n)
    if (n < P) e[n + 1] = csub(cscale(e[n], 2.0 * n + 1.0), cmul(z2, e[n - 1]));
}

struct Level {
  cx k;
  double s, beta;
};
__device__ __forceinline__ Level level_of(const double *hlev, int l) {
  return {{hlev[4 * l], hlev[4 * l + 1]}, hlev[4 * l + 2], hlev[4 * l + 3]};
}

// ---- P2M / P2L ---------------------------------------------------------------------------
// One block of 256 threads per output box; thread k owns slot k.  Points go through LDS in
// tiles of 16: their radial weights W_j(n) (complex) and their harmonics S_j(k) (16 points x
// 16 orders, one recursion column per thread); then slot k adds W_j(n) S_j(k) over the tile.
template <bool SING>
__global__ __launch_bounds__(256) void k_h_expand(int p, uint32_t rows, const uint32_t *__restrict__ row_box,
                                                  const uint32_t *__restrict__ row_cell,
                                                  const uint32_t *__restrict__ row_off,
                                                  const uint32_t *__restrict__ row_ent,
                                                  const uint32_t *__restrict__ leaf_cell,
                                                  const uint32_t *__restrict__ src_start, const double *sx,
                                                  const double *sy, const double *sz, const double *qr,
                                                  const double *qi, const double *hlev, const double *norm,
                                                  double *coeffs) {
  __shared__ double H[HT][kMaxCoef];
  __shared__ double Wr[HT][kMaxDegree + 1], Wi[HT][kMaxDegree + 1];
  __shared__ double Pt[HT][4];
  __shared__ double rinv[kMaxDegree + 1][kMaxDegree + 1];  // [d][m]: 1/((d+m)(d-m)), d > m
  __shared__ double dinv[kMaxDegree + 1];                   // [q]: 1/(2q), q > 0
  const uint32_t row = xcd_index(rows);
  if (row >= rows) return;
  const int P1 = p + 1, nc = P1 * P1;
  const int k = threadIdx.x;
  {
    const int d = k / (kMaxDegree + 1), m = k % (kMaxDegree + 1);
    rinv[d][m] = d > m ? 1.0 / double((d + m) * (d - m)) : 0.0;
    if (k <= kMaxDegree) dinv[k] = k > 0 ? 1.0 / (2.0 * k) : 0.0;
  }
  int n = 0;
  while ((n + 1) * (n + 1) <= k) ++n;
  const uint32_t *cell = row_cell + 4 * size_t(row);
  const Level L = level_of(hlev, int(cell[0]));
  double ar = 0.0, ai = 0.0;
  for (uint32_t e = row_off[row]; e < row_off[row + 1]; ++e) {
    const uint32_t sl = row_ent[e];
    const uint32_t s0 = src_start[sl], s1 = src_start[sl + 1];
    double c[3] = {0.0, 0.0, 0.0}, r = 1.0;
    if (SING) cell_frame(cell, leaf_cell + 4 * size_t(sl), c, r);  // c^(t|s), r^(t|s)
    const double inv_r = 1.0 / r;
    for (uint32_t sb = s0; sb < s1; sb += HT) {
      const int cnt = int((s1 - sb) < uint32_t(HT) ? (s1 - sb) : uint32_t(HT));
      __syncthreads();
      if (k < cnt) {
        const uint32_t sidx = sb + k;
        double x = sx[sidx], y = sy[sidx], z = sz[sidx];
        if (SING) {
          x = (x - c[0]) * inv_r;
          y = (y - c[1]) * inv_r;
          z = (z - c[2]) * inv_r;
        }
        const double r2 = x * x + y * y + z * z, rho = sqrt(r2);
        Pt[k][0] = x;
        Pt[k][1] = y;
        Pt[k][2] = z;
        Pt[k][3] = r2;
        const cx q = {qr[sidx], qi[sidx]};
        const cx zz = cscale(L.k, rho);
        cx v[kRad];
        if (!SING) {
          // W(n) = e^-beta (kh/s)^n g_n(kh rho) q
          radial_j(p, zz, v);
          cx f = cscale(q, exp(-L.beta));
          const cx step = cscale(L.k, 1.0 / L.s);
#pragma unroll
          for (int d = 0; d <= kMaxDegree; ++d)
            if (d <= p) {
              const cx w = cmul(f, v[d]);
              Wr[k][d] = w.re;
              Wi[k][d] = w.im;
              f = cmul(f, step);
            }
        } else {
          // W(n) = 4 pi i e^beta (s/kh)^n e_n(kh rho) q / rho^(2n+1)
          cx w0;
          radial_h(p, zz, v, w0);
          const double a = 4.0 * kPi * exp(L.beta) / rho;
          cx f = cmul({0.0, a}, q);
          const cx step = cdiv({L.s / r2, 0.0}, L.k);
#pragma unroll
          for (int d = 0; d <= kMaxDegree; ++d)
            if (d <= p) {
              const cx w = cmul(f, v[d]);
              Wr[k][d] = w.re;
              Wi[k][d] = w.im;
              f = cmul(f, step);
            }
        }
      }
      __syncthreads();
      {
        // harmonics: thread (j, m) runs column m of the Racah recursion for point j
        const int j = k >> 4, m = k & 15;
        if (j < cnt && m <= p) {
          const double x = Pt[j][0], y = Pt[j][1], z = Pt[j][2], r2 = Pt[j][3];
          double rr = 1.0, ri = 0.0;
          for (int q = 1; q <= m; ++q) {
            const double inv = dinv[q];
            const double nr = (x * rr - y * ri) * inv, ni = (x * ri + y * rr) * inv;
            rr = nr;
            ri = ni;
          }
          double pr = 0.0, pi = 0.0;
          for (int d = m; d <= p; ++d) {
            if (d == m + 1) {
              pr = rr;
              pi = ri;
              rr *= z;
              ri *= z;
            } else if (d >= m + 2) {
              const double inv = rinv[d][m];
              const double nr = ((2.0 * d - 1.0) * z * rr - r2 * pr) * inv;
              const double ni = ((2.0 * d - 1.0) * z * ri - r2 * pi) * inv;
              pr = rr;
              pi = ri;
              rr = nr;
              ri = ni;
            }
            // unnormalised: slot k's normalisation multiplies its sum at the end
            const int t = d * d + d;
            H[j][t + m] = rr;
            if (m > 0) H[j][t - m] = ri;
          }
        }
      }
      __syncthreads();
      if (k < nc)
        for (int j = 0; j < cnt; ++j) {
          const double h = H[j][k];
          ar += Wr[j][n] * h;
          ai += Wi[j][n] * h;
        }
    }
  }
  if (k < nc) {
    const size_t box = row_box[row];
    const double nk = norm[k];
    coeffs[(3 * box) * nc + k] += nk * ar;
    coeffs[(3 * box + 1) * nc + k] += nk * ai;
  }
}

// ---- L2P / M2P ---------------------------------------------------------------------------
// One wave per target leaf, one target per lane (chunks of 64).  With C the entry box's
// coefficients (two planes in LDS) and v the target in the box's units,
//   F = sum_n c_n sum_m C_n^m S_n^m(v),   grad F = sum_n c_n grad(...) + (sum_n d_n ...) v,
// L2P: c_n = e^-beta (kh/s)^n g_n, d_n = -kh^2 e^-beta (kh/s)^n g_{n+1};
// M2P: w_n = 4 pi i e^beta (s/kh)^n / rho^(2n+1), c_n = w_n e_n,
//      d_n = w_n (kh e_n' - (2n+1) e_n / rho) / rho.
// The gradient of sum_m C_n^m S_n^m is a degree n-1 sum sum_q D_{n-1}^q S_{n-1}^q with
// derivative coefficients D built once per box from the Racah ladder (dz R_n^m = R_{n-1}^m,
// dx R_n^m = (R_{n-1}^{m-1} - R_{n-1}^{m+1}) / 2, dy R_n^m = i (R_{n-1}^{m-1} + R_{n-1}^{m+1}) / 2,
// R^{-1} = -conj R^1): with alpha_m = N_nm (a_{n,m} - i a_{n,-m}) (alpha_0 = N_n0 a_{n,0}) for
// one real plane a and F_n = alpha_0 R^0 + sum_{m>0} Re(alpha_m R^m),
//   z: beta_q = alpha_q;
//   x: beta_0 = Re(alpha_1) / 2, beta_1 = alpha_2 / 2 - alpha_0, beta_q = (alpha_{q+1} - alpha_{q-1}) / 2;
//   y: beta_0 = -Im(alpha_1) / 2, beta_1 = i (alpha_2 / 2 + alpha_0), beta_q = i (alpha_{q+1} + alpha_{q-1}) / 2;
// and D_{n-1,+q} = Re(beta_q) / N, D_{n-1,-q} = -Im(beta_q) / N.  Each target then runs the
// harmonic recursion column by column (orders m, degrees n = m .. p) with scalar state.
template <bool SING, bool GRAD>
__global__ __launch_bounds__(64) void k_h_evaluate(int p, uint32_t rows, const uint32_t *__restrict__ row_leaf,
                                                   const uint32_t *__restrict__ row_off,
                                                   const uint32_t *__restrict__ row_box,
                                                   const uint32_t *__restrict__ box_cell,
                                                   const uint32_t *__restrict__ leaf_cell,
                                                   const uint32_t *__restrict__ tgt_start, const double *tx,
                                                   const double *ty, const double *tz, const double *hlev,
                                                   const double *norm, const double *coeffs, double *phi_hat,
                                                   double *grad_hat) {
  // coefficients scaled by their normalisation, (re, im) planes interleaved: C[t] =
  // N_t (c_re[t], c_im[t]); derivative coefficients likewise, D[q][t] = N_t (D_re, D_im)
  __shared__ double2 C[kMaxCoef];
  __shared__ double2 D[3][kMaxCoef];
  __shared__ double Nm[kMaxCoef];
  __shared__ double rinv[kMaxDegree + 1][kMaxDegree + 1];  // [n][m]: 1/((n+m)(n-m)), n >= m + 2
  const uint32_t row = xcd_index(rows);
  if (row >= rows) return;
  const int P1 = p + 1, nc = P1 * P1;
  const int lane = threadIdx.x;
  for (int t = lane; t < nc; t += 64) Nm[t] = norm[t];
  for (int t = lane; t < (kMaxDegree + 1) * (kMaxDegree + 1); t += 64) {
    const int n = t / (kMaxDegree + 1), m = t % (kMaxDegree + 1);
    rinv[n][m] = n > m ? 1.0 / double((n + m) * (n - m)) : 0.0;
  }
  const uint32_t lf = row_leaf[row];
  const uint32_t t0 = tgt_start[lf], t1 = tgt_start[lf + 1];
  const uint32_t *lcell = leaf_cell + 4 * size_t(lf);
  for (uint32_t tb = t0; tb < t1; tb += 64) {
    const uint32_t t = tb + lane;
    const bool active = t < t1;
    double u[3] = {0.0, 0.0, 0.0};
    if (active) {
      u[0] = tx[t];
      u[1] = ty[t];
      u[2] = tz[t];
    }
    cx phi = {0.0, 0.0}, g[3] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
    for (uint32_t e = row_off[row]; e < row_off[row + 1]; ++e) {
      const uint32_t box = row_box[e];
      __syncthreads();
      for (int i = lane; i < nc; i += 64)
        C[i] = make_double2(Nm[i] * coeffs[(3 * size_t(box)) * nc + i],
                            Nm[i] * coeffs[(3 * size_t(box) + 1) * nc + i]);
      __syncthreads();
      if (GRAD)
        for (int i = lane; i < p * p; i += 64) {
          // derivative coefficient slot i = (n', m') of degree n' = n - 1
          int np = 0;
          while ((np + 1) * (np + 1) <= i) ++np;
          const int mp = i - np * np - np, q = mp < 0 ? -mp : mp, n = np + 1;
          const int base = n * n + n;
          // alpha_m of degree n for both planes (the coefficients carry N already)
          auto alpha = [&](int pl, int m, double &xr, double &xi) {
            if (m > n) {
              xr = xi = 0.0;
            } else if (m == 0) {
              xr = pl == 0 ? C[base].x : C[base].y;
              xi = 0.0;
            } else {
              xr = pl == 0 ? C[base + m].x : C[base + m].y;
              xi = -(pl == 0 ? C[base - m].x : C[base - m].y);
            }
          };
          const double inv = Nm[i] / Nm[np * np + np + q];
          double out[2][3];
#pragma unroll
          for (int pl = 0; pl < 2; ++pl) {
            double ur, ui, vr, vi, wr, wi;  // alpha_{q+1}, alpha_{q-1}, alpha_q
            alpha(pl, q + 1, ur, ui);
            if (q >= 1)
              alpha(pl, q - 1, vr, vi);
            else
              vr = vi = 0.0;
            alpha(pl, q, wr, wi);
            double bx_r, bx_i, by_r, by_i;
            if (q == 0) {
              bx_r = 0.5 * ur;
              bx_i = 0.0;
              by_r = -0.5 * ui;
              by_i = 0.0;
            } else if (q == 1) {
              bx_r = 0.5 * ur - vr;
              bx_i = 0.5 * ui - vi;
              by_r = -(0.5 * ui + vi);
              by_i = 0.5 * ur + vr;
            } else {
              bx_r = 0.5 * (ur - vr);
              bx_i = 0.5 * (ui - vi);
              by_r = -0.5 * (ui + vi);
              by_i = 0.5 * (ur + vr);
            }
            // slot +q takes Re(beta) / N, slot -q takes -Im(beta) / N (times N_i)
            const bool neg = mp < 0;
            out[pl][0] = (neg ? -bx_i : bx_r) * inv;
            out[pl][1] = (neg ? -by_i : by_r) * inv;
            out[pl][2] = (neg ? -wi : wr) * inv;
          }
#pragma unroll
          for (int d = 0; d < 3; ++d) D[d][i] = make_double2(out[0][d], out[1][d]);
        }
      double c[3] = {0.0, 0.0, 0.0}, r = 1.0;
      if (SING) cell_frame(box_cell + 4 * size_t(box), lcell, c, r);  // c^(s|t), r^(s|t)
      const Level L = level_of(hlev, SING ? int(box_cell[4 * size_t(box)]) : int(lcell[0]));
      const double inv_r = 1.0 / r;
      const double v0 = (u[0] - c[0]) * inv_r, v1 = (u[1] - c[1]) * inv_r, v2 = (u[2] - c[2]) * inv_r;
      const double r2 = v0 * v0 + v1 * v1 + v2 * v2, rho = sqrt(r2);
      // radial factors c_n (and for M2P d_n) of this lane's target, in registers (degree
      // loops unrolled).  L2P: d_n = -k^2 e^-beta (k/s)^n g_{n+1} = -k s c_{n+1}, so the
      // radial sum runs over c_{n+1} and takes the factor -k s once at the end.
      cx cn[kMaxDegree + 2], dn[SING ? kMaxDegree + 1 : 1];
      {
        const cx zz = cscale(L.k, rho);
        cx v[kRad];
        if (!SING) {
          radial_j(p + 1, zz, v);
          cx f = {exp(-L.beta), 0.0};
          const cx step = cscale(L.k, 1.0 / L.s);
#pragma unroll
          for (int d = 0; d <= kMaxDegree; ++d) {
            cn[d] = cmul(f, v[d]);
            f = cmul(f, step);
          }
        } else {
          cx w0;
          radial_h(p, zz, v, w0);
          cx w = {0.0, 4.0 * kPi * exp(L.beta) / rho};
          const cx step = cdiv({L.s / r2, 0.0}, L.k);
          const double inv_rho = 1.0 / rho;
#pragma unroll
          for (int d = 0; d <= kMaxDegree; ++d) {
            const cx ed = v[d], dd = d == 0 ? w0 : cmul(zz, v[d > 0 ? d - 1 : 0]);
            cn[d] = cmul(w, ed);
            const cx hr = cscale(csub(cmul(L.k, dd), cscale(ed, (2.0 * d + 1.0) * inv_rho)), inv_rho);
            if (SING) dn[SING ? d : 0] = cmul(w, hr);
            w = cmul(w, step);
          }
        }
        cn[kMaxDegree + 1] = {0.0, 0.0};
      }
      __syncthreads();
      cx f = {0.0, 0.0}, rad = {0.0, 0.0}, gr[3] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
      double dr = 1.0, di = 0.0;  // R_m^m
      for (int m = 0; m <= p; ++m) {
        if (m > 0) {
          const double inv = 1.0 / (2.0 * m);
          const double nr = (v0 * dr - v1 * di) * inv, ni = (v0 * di + v1 * dr) * inv;
          dr = nr;
          di = ni;
        }
        double rr = dr, ri = di, pr = 0.0, pi = 0.0;
#pragma unroll
        for (int n = 0; n <= kMaxDegree; ++n) {
          if (n < m || n > p) continue;
          if (n == m + 1) {
            pr = rr;
            pi = ri;
            rr *= v2;
            ri *= v2;
          } else if (n >= m + 2) {
            const double inv = rinv[n][m];
            const double nr = ((2.0 * n - 1.0) * v2 * rr - r2 * pr) * inv;
            const double ni = ((2.0 * n - 1.0) * v2 * ri - r2 * pi) * inv;
            pr = rr;
            pi = ri;
            rr = nr;
            ri = ni;
          }
          const int tp = n * n + n + m, tm = n * n + n - m;
          const double sm = m > 0 ? ri : 0.0;
          const double2 a = C[tp], b = C[tm];
          const cx tsum = {a.x * rr + b.x * sm, a.y * rr + b.y * sm};
          f = cadd(f, cmul(cn[n], tsum));
          if (GRAD) {
            rad = cadd(rad, cmul(SING ? dn[SING ? n : 0] : cn[n + 1], tsum));
            if (n < p) {
#pragma unroll
              for (int q = 0; q < 3; ++q) {
                const double2 da = D[q][tp], db = D[q][tm];
                const cx ds = {da.x * rr + db.x * sm, da.y * rr + db.y * sm};
                gr[q] = cadd(gr[q], cmul(cn[n + 1], ds));
              }
            }
          }
        }
      }
      phi = cadd(phi, cscale(f, inv_r));
      if (GRAD) {
        if (!SING) rad = cmul(cscale(L.k, -L.s), rad);
        const double s2 = inv_r * inv_r, vv[3] = {v0, v1, v2};
#pragma unroll
        for (int q = 0; q < 3; ++q) g[q] = cadd(g[q], cscale(cadd(gr[q], cscale(rad, vv[q])), s2));
      }
    }
    if (active) {
      phi_hat[2 * size_t(t)] += phi.re;
      phi_hat[2 * size_t(t) + 1] += phi.im;
      if (GRAD)
#pragma unroll
        for (int q = 0; q < 3; ++q) {
          grad_hat[6 * size_t(t) + 2 * q] += g[q].re;
          grad_hat[6 * size_t(t) + 2 * q + 1] += g[q].im;
        }
    }
  }
}

// ---- P2P ---------------------------------------------------------------------------------
// One wave per target leaf.  Lanes are (phase, target): with n_t <= 32 targets the wave
// runs 64 / n_t phases that split each source tile; phases are added in order at the end.
// Every near leaf's sources go through LDS in tiles of 64 with y^ = c^ + r^ u_s formed in
// f64.  A pair with r^2 == 0 adds nothing (selects, no branch).  T = float: separations
// from double-float coordinates (hi + lo), the pair arithmetic in f32, each source leaf's
// f32 sums added to f64 accumulators.
// Pair kernel pieces per precision.  f64: sqrt, division and sincos of the math library.
// f32: 1/r from v_rsq_f32 (1 ulp) and r = r^2 / r; sine and cosine from v_sin_f32 and
// v_cos_f32 of the fraction of x / (2 pi).
template <typename T>
struct PairMath;

template <>
struct PairMath<double> {
  __device__ static __forceinline__ void rinv(double r2, double &r, double &ir) {
    r = sqrt(r2);
    ir = r2 == 0.0 ? 0.0 : 1.0 / r;
  }
  __device__ static __forceinline__ void sincos(double x, double *s, double *c) { ::sincos(x, s, c); }
};

template <>
struct PairMath<float> {
  __device__ static __forceinline__ void rinv(float r2, float &r, float &ir) {
    const float y = __builtin_amdgcn_rsqf(r2);
    ir = r2 == 0.0f ? 0.0f : y;
    r = r2 * ir;
  }
  __device__ static __forceinline__ void sincos(float x, float *s, float *c) {
    // in revolutions: t = x / (2 pi), reduced to [0, 1) by v_fract_f32; v_sin_f32 and
    // v_cos_f32 take revolutions
    const float t = __builtin_amdgcn_fractf(x * 0.159154943091895336f);
    *s = __builtin_amdgcn_sinf(t);
    *c = __builtin_amdgcn_cosf(t);
  }
};

// Leaves of more than 32 targets in the f32 path: each lane holds one target (TWO: two,
// the second 64 further on) and all lanes take the same source at each step, read by
// uniform loads from its 32-byte record (u_hi x, y, z, u_lo x | u_lo y, z, q re, q im: the
// source's own leaf frame, f32); no LDS staging, no barriers.  Per near entry the targets
// move into the source leaf's frame, x' = c'' + rho u_t (rho = r_t / r_s), and the pair
// term is formed in the source leaf's units with that leaf's k' = k r_s: in target units
// e^{i k_t r_t} / r_t = rho e^{i k' r'} / r' and its gradient is rho^2 (i k' r' - 1)
// e^{i k' r'} d' / r'^3, so an entry's f32 sums (at most 128 sources) go into the f64
// sums scaled by rho and rho^2 (powers of two).
template <bool GRAD, bool DAMP, bool TWO>
__device__ __forceinline__ void h_p2p_uniform(uint32_t lf, uint32_t tb, uint32_t n, uint32_t e0, uint32_t e1,
                                              const uint32_t *__restrict__ near_ent,
                                              const uint32_t *__restrict__ leaf_cell,
                                              const uint32_t *__restrict__ src_start,
                                              const float4 *__restrict__ rec, const double *__restrict__ hlev,
                                              const double *__restrict__ tx, const double *__restrict__ ty,
                                              const double *__restrict__ tz, double *__restrict__ phi_hat,
                                              double *__restrict__ grad_hat) {
  constexpr int A = TWO ? 2 : 1;
  const int lane = threadIdx.x;
  double u[A][3];
  bool ok[A];
#pragma unroll
  for (int a = 0; a < A; ++a) {
    const uint32_t i = tb + uint32_t(lane + 64 * a);
    ok[a] = uint32_t(lane + 64 * a) < n;
    u[a][0] = ok[a] ? tx[i] : 0.0;
    u[a][1] = ok[a] ? ty[i] : 0.0;
    u[a][2] = ok[a] ? tz[i] : 0.0;
  }
  double acc[A][8];
#pragma unroll
  for (int a = 0; a < A; ++a)
#pragma unroll
    for (int q = 0; q < 8; ++q) acc[a][q] = 0.0;
  const uint32_t *ct = leaf_cell + 4 * size_t(lf);
  for (uint32_t e = e0; e < e1; ++e) {
    const uint32_t sl = near_ent[e];
    const uint32_t s0 = src_start[sl], s1 = src_start[sl + 1];
    if (s0 == s1) continue;
    const uint32_t *cs = leaf_cell + 4 * size_t(sl);
    double c[3], rho;
    cell_frame(ct, cs, c, rho);  // the target leaf in the source leaf's frame
    const int ls = int(cs[0]);
    const float kre = float(hlev[4 * ls]), kim = float(hlev[4 * ls + 1]);
    float xh[A][3], xl[A][3];
#pragma unroll
    for (int a = 0; a < A; ++a)
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const double x = fma(rho, u[a][k], c[k]);
        xh[a][k] = float(x);
        xl[a][k] = float(x - double(xh[a][k]));
      }
    for (uint32_t jb = s0; jb < s1; jb += 128) {
      const uint32_t je = (s1 - jb) < 128u ? s1 : jb + 128;
      float loc[A][8];
#pragma unroll
      for (int a = 0; a < A; ++a)
#pragma unroll
        for (int q = 0; q < 8; ++q) loc[a][q] = 0.0f;
      // off the own leaf the smallest normal number enters the first product: it keeps a
      // separation rounded to zero finite (no r^2 == 0 test) and changes no r^2 above
      // 2^-102; the own leaf keeps the test (a pair with r^2 == 0 adds nothing)
      auto run = [&](auto check) {
#pragma unroll 2
        for (uint32_t j = jb; j < je; ++j) {
          const float4 h = rec[2 * size_t(j)], l = rec[2 * size_t(j) + 1];
#pragma unroll
          for (int a = 0; a < A; ++a) {
            const float dx = (xh[a][0] - h.x) + (xl[a][0] - h.w);
            const float dy = (xh[a][1] - h.y) + (xl[a][1] - l.x);
            const float dz = (xh[a][2] - h.z) + (xl[a][2] - l.y);
            float rr, ir;
            if (decltype(check)::value) {
              PairMath<float>::rinv(dx * dx + dy * dy + dz * dz, rr, ir);
            } else {
              const float r2 = fmaf(dz, dz, fmaf(dy, dy, fmaf(dx, dx, 1.17549435e-38f)));
              ir = __builtin_amdgcn_rsqf(r2);
              rr = r2 * ir;
            }
            float sn, cs2;
            PairMath<float>::sincos(kre * rr, &sn, &cs2);
            const float dmp = DAMP ? expf(-kim * rr) : 1.0f;
            const float er = dmp * cs2 * ir, ei = dmp * sn * ir;
            const float q0 = l.z, q1 = l.w;
            loc[a][0] += q0 * er - q1 * ei;
            loc[a][1] += q0 * ei + q1 * er;
            if (GRAD) {
              const float ar = DAMP ? -kim * rr - 1.0f : -1.0f, ai = kre * rr, ir2 = ir * ir;
              const float gr = (ar * er - ai * ei) * ir2, gi = (ar * ei + ai * er) * ir2;
              const float qgr = q0 * gr - q1 * gi, qgi = q0 * gi + q1 * gr;
              loc[a][2] += qgr * dx;
              loc[a][3] += qgi * dx;
              loc[a][4] += qgr * dy;
              loc[a][5] += qgi * dy;
              loc[a][6] += qgr * dz;
              loc[a][7] += qgi * dz;
            }
          }
        }
      };
      if (sl == lf)
        run(std::true_type{});
      else
        run(std::false_type{});
      const double s2 = rho * rho;
#pragma unroll
      for (int a = 0; a < A; ++a) {
        acc[a][0] += double(loc[a][0]) * rho;
        acc[a][1] += double(loc[a][1]) * rho;
        if (GRAD)
#pragma unroll
          for (int q = 2; q < 8; ++q) acc[a][q] += double(loc[a][q]) * s2;
      }
    }
  }
#pragma unroll
  for (int a = 0; a < A; ++a) {
    if (!ok[a]) continue;
    const size_t t = tb + uint32_t(lane + 64 * a);
    phi_hat[2 * t] += acc[a][0];
    phi_hat[2 * t + 1] += acc[a][1];
    if (GRAD)
#pragma unroll
      for (int q = 0; q < 6; ++q) grad_hat[6 * t + q] += acc[a][2 + q];
  }
}

template <typename T, bool GRAD, bool DAMP>
__global__ __launch_bounds__(64) void k_h_p2p(uint32_t leaf_first, uint32_t count,
                                              const uint32_t *__restrict__ near_off,
                                              const uint32_t *__restrict__ near_ent,
                                              const uint32_t *__restrict__ leaf_cell,
                                              const uint32_t *__restrict__ src_start,
                                              const uint32_t *__restrict__ tgt_start, const double *__restrict__ sx,
                                              const double *__restrict__ sy, const double *__restrict__ sz,
                                              const double *__restrict__ qr, const double *__restrict__ qi,
                                              const float4 *__restrict__ rec, const double *__restrict__ tx,
                                              const double *__restrict__ ty, const double *__restrict__ tz,
                                              const double *__restrict__ hlev, double *__restrict__ phi_hat,
                                              double *__restrict__ grad_hat) {
  constexpr bool F32 = sizeof(T) == 4;
  __shared__ T ys[64][8];  // y hi (3), y lo (3, f32 only), q re, q im
  __shared__ double red[64][8];
  const uint32_t item = xcd_index(count);
  if (item >= count) return;
  const uint32_t lf = leaf_first + item;
  const uint32_t t0 = tgt_start[lf], t1 = tgt_start[lf + 1];
  if (t1 == t0) return;
  const int lane = threadIdx.x;
  const uint32_t *ct = leaf_cell + 4 * size_t(lf);
  const int lev = int(ct[0]);
  const T kre = T(hlev[4 * lev]), kim = T(hlev[4 * lev + 1]);
  const uint32_t ntl = t1 - t0;
  if (F32 && ntl > 32) {
    const uint32_t e0 = near_off[lf], e1 = near_off[lf + 1];
    for (uint32_t tb = t0; tb < t1; tb += 128) {
      const uint32_t n = (t1 - tb) < 128u ? (t1 - tb) : 128u;
      if (n > 64)
        h_p2p_uniform<GRAD, DAMP, true>(lf, tb, n, e0, e1, near_ent, leaf_cell, src_start, rec, hlev, tx, ty, tz,
                                        phi_hat, grad_hat);
      else
        h_p2p_uniform<GRAD, DAMP, false>(lf, tb, n, e0, e1, near_ent, leaf_cell, src_start, rec, hlev, tx, ty, tz,
                                         phi_hat, grad_hat);
    }
    return;
  }
  const int phases = ntl >= 64 ? 1 : (ntl > 32 ? 1 : int(64 / ntl));
  const int span = ntl >= 64 ? 64 : int(ntl);  // targets per chunk
  for (uint32_t tb = t0; tb < t1; tb += uint32_t(span)) {
    const int j = lane % span, ph = lane / span;
    const uint32_t t = tb + uint32_t(j);
    const bool active = ph < phases && t < t1;
    double ux = 0.0, uy = 0.0, uz = 0.0;
    if (active) {
      ux = tx[t];
      uy = ty[t];
      uz = tz[t];
    }
    const T xh = T(ux), yh = T(uy), zh = T(uz);
    const T xl = F32 ? T(ux - double(xh)) : T(0), yl = F32 ? T(uy - double(yh)) : T(0),
            zl = F32 ? T(uz - double(zh)) : T(0);
    double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (uint32_t e = near_off[lf]; e < near_off[lf + 1]; ++e) {
      const uint32_t sl = near_ent[e];
      const uint32_t s0 = src_start[sl], s1 = src_start[sl + 1];
      double c[3] = {0.0, 0.0, 0.0}, r = 1.0;
      if (sl != lf) cell_frame(leaf_cell + 4 * size_t(sl), ct, c, r);  // c^(s|t), r^(s|t)
      T loc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
      for (uint32_t sb = s0; sb < s1; sb += 64) {
        const int cnt = int((s1 - sb) < 64u ? (s1 - sb) : 64u);
        __syncthreads();
        if (lane < cnt) {
          const uint32_t s = sb + lane;
          const double y0 = fma(r, sx[s], c[0]), y1 = fma(r, sy[s], c[1]), y2 = fma(r, sz[s], c[2]);
          const T h0 = T(y0), h1 = T(y1), h2 = T(y2);
          ys[lane][0] = h0;
          ys[lane][1] = h1;
          ys[lane][2] = h2;
          ys[lane][3] = F32 ? T(y0 - double(h0)) : T(0);
          ys[lane][4] = F32 ? T(y1 - double(h1)) : T(0);
          ys[lane][5] = F32 ? T(y2 - double(h2)) : T(0);
          ys[lane][6] = T(qr[s]);
          ys[lane][7] = T(qi[s]);
        }
        __syncthreads();
        if (active)
          for (int i = ph; i < cnt; i += phases) {
            T dx = xh - ys[i][0], dy = yh - ys[i][1], dz = zh - ys[i][2];
            if (F32) {
              dx += xl - ys[i][3];
              dy += yl - ys[i][4];
              dz += zl - ys[i][5];
            }
            const T r2 = dx * dx + dy * dy + dz * dz;
            T rr, ir;
            PairMath<T>::rinv(r2, rr, ir);
            T sn, cs;
            PairMath<T>::sincos(kre * rr, &sn, &cs);
            const T dmp = DAMP ? T(exp(-kim * rr)) : T(1);
            const T er = dmp * cs * ir, ei = dmp * sn * ir;
            const T q0 = ys[i][6], q1 = ys[i][7];
            loc[0] += q0 * er - q1 * ei;
            loc[1] += q0 * ei + q1 * er;
            if (GRAD) {
              const T ar = DAMP ? -kim * rr - T(1) : T(-1), ai = kre * rr, ir2 = ir * ir;
              const T gr = (ar * er - ai * ei) * ir2, gi = (ar * ei + ai * er) * ir2;
              const T qgr = q0 * gr - q1 * gi, qgi = q0 * gi + q1 * gr;
              loc[2] += qgr * dx;
              loc[3] += qgi * dx;
              loc[4] += qgr * dy;
              loc[5] += qgi * dy;
              loc[6] += qgr * dz;
              loc[7] += qgi * dz;
            }
          }
      }
#pragma unroll
      for (int q = 0; q < 8; ++q) acc[q] += double(loc[q]);
    }
    // phases in order
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 8; ++q) red[lane][q] = acc[q];
    __syncthreads();
    if (ph == 0 && active) {
      for (int h = 1; h < phases; ++h)
#pragma unroll
        for (int q = 0; q < 8; ++q) acc[q] += red[h * span + j][q];
      phi_hat[2 * size_t(t)] += acc[0];
      phi_hat[2 * size_t(t) + 1] += acc[1];
      if (GRAD)
#pragma unroll
        for (int q = 0; q < 6; ++q) grad_hat[6 * size_t(t) + q] += acc[2 + q];
    }
  }
}

__global__ void k_h_negate(int nc, size_t b0, size_t b1, double *coeffs) {
  const size_t n = (b1 - b0) * size_t(nc);
  for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) {
    const size_t b = b0 + i / nc, k = i % nc;
    coeffs[(3 * b + 2) * nc + k] = -coeffs[(3 * b + 1) * nc + k];
  }
}

__global__ void k_h_gather(const double *q, const uint32_t *order, size_t n, double *qr, double *qi, float *rec) {
  for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) {
    const size_t c = order[i];
    qr[i] = q[2 * c];
    qi[i] = q[2 * c + 1];
    if (rec) {
      rec[8 * i + 6] = float(q[2 * c]);
      rec[8 * i + 7] = float(q[2 * c + 1]);
    }
  }
}

__global__ void k_h_output(uint32_t nleaves, const uint32_t *tgt_start, const double *leaf_r, const uint32_t *order,
                           const double *phi_hat, const double *grad_hat, double *phi, double *grad, int gradients) {
  const uint32_t lf = blockIdx.x;
  if (lf >= nleaves) return;
  const double r = leaf_r[lf];
  const double s1 = 4.0 * kPi * r, s2 = 4.0 * kPi * r * r;
  for (uint32_t i = tgt_start[lf] + threadIdx.x; i < tgt_start[lf + 1]; i += blockDim.x) {
    const size_t c = order[i];
    phi[2 * c] = phi_hat[2 * size_t(i)] / s1;
    phi[2 * c + 1] = phi_hat[2 * size_t(i) + 1] / s1;
    if (gradients)
      for (int q = 0; q < 6; ++q) grad[6 * c + q] = grad_hat[6 * size_t(i) + q] / s2;
  }
}

// Compensated direct sum: the term of nd_fmm_ref::helmholtz::p2p::direct_sum
// (fmm-ref/src/helmholtz/p2p.rs:155): d = x - y, r2 = (d0^2 + d1^2) + d2^2, e = e^{-Im k r}
// (cos Re k r + i sin Re k r), phi += q e / r, grad += q ((ik r - 1) e / (r2 r)) d; a source
// equal to the target is skipped; eight real sums with Neumaier's compensation per chunk of
// sources, chunks added in order.
struct Neumaier {
  double s = 0.0, c = 0.0;
  __device__ void add(double t) {
    const double u = s + t;
    c += fabs(s) >= fabs(t) ? (s - u) + t : (t - u) + s;
    s = u;
  }
};

__global__ __launch_bounds__(256) void k_h_direct_partial(double kre, double kim, const double *src,
                                                          const double *q, int64_t ns, const double *tgt,
                                                          int64_t nt, int64_t chunk, int gradients, double *part) {
  __shared__ double ys[256][5];
  const int64_t t = int64_t(blockIdx.x) * 256 + threadIdx.x;
  const int64_t s0 = int64_t(blockIdx.y) * chunk, s1 = s0 + chunk < ns ? s0 + chunk : ns;
  double x[3] = {0.0, 0.0, 0.0};
  if (t < nt)
    for (int k = 0; k < 3; ++k) x[k] = tgt[3 * t + k];
  Neumaier acc[8];
  for (int64_t sb = s0; sb < s1; sb += 256) {
    const int cnt = int((s1 - sb) < 256 ? (s1 - sb) : 256);
    __syncthreads();
    if (int(threadIdx.x) < cnt) {
      const int64_t s = sb + threadIdx.x;
      ys[threadIdx.x][0] = src[3 * s];
      ys[threadIdx.x][1] = src[3 * s + 1];
      ys[threadIdx.x][2] = src[3 * s + 2];
      ys[threadIdx.x][3] = q[2 * s];
      ys[threadIdx.x][4] = q[2 * s + 1];
    }
    __syncthreads();
    if (t < nt)
      for (int j = 0; j < cnt; ++j) {
        const double y0 = ys[j][0], y1 = ys[j][1], y2 = ys[j][2];
        if (x[0] == y0 && x[1] == y1 && x[2] == y2) continue;
        const double d[3] = {x[0] - y0, x[1] - y1, x[2] - y2};
        const double r2 = (d[0] * d[0] + d[1] * d[1]) + d[2] * d[2];
        const double r = sqrt(r2);
        double sn, cs;
        sincos(kre * r, &sn, &cs);
        const double dmp = exp(-kim * r);
        const double er = dmp * cs, ei = dmp * sn;
        const double q0 = ys[j][3], q1 = ys[j][4];
        const double qer = q0 * er - q1 * ei, qei = q0 * ei + q1 * er;
        acc[0].add(qer / r);
        acc[1].add(qei / r);
        if (gradients) {
          const double ar = -kim * r - 1.0, ai = kre * r, r3 = r2 * r;
          const double gr = (ar * er - ai * ei) / r3, gi = (ar * ei + ai * er) / r3;
          const double qgr = q0 * gr - q1 * gi, qgi = q0 * gi + q1 * gr;
          for (int k = 0; k < 3; ++k) {
            acc[2 + 2 * k].add(qgr * d[k]);
            acc[3 + 2 * k].add(qgi * d[k]);
          }
        }
      }
  }
  if (t < nt) {
    double *o = part + (size_t(blockIdx.y) * nt + t) * 16;
    for (int k = 0; k < 8; ++k) {
      o[2 * k] = acc[k].s;
      o[2 * k + 1] = acc[k].c;
    }
  }
}

__global__ void k_h_direct_combine(int64_t nt, int nchunks, int gradients, const double *part, double *phi,
                                   double *grad) {
  const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  Neumaier acc[8];
  for (int c = 0; c < nchunks; ++c) {
    const double *o = part + (size_t(c) * nt + t) * 16;
    for (int k = 0; k < 8; ++k) {
      acc[k].add(o[2 * k]);
      acc[k].add(o[2 * k + 1]);
    }
  }
  phi[2 * t] = acc[0].s + acc[0].c;
  phi[2 * t + 1] = acc[1].s + acc[1].c;
  if (gradients)
    for (int k = 0; k < 6; ++k) grad[6 * t + k] = acc[2 + k].s + acc[2 + k].c;
}

}  // namespace

void h_expand(hipStream_t s, int p, int singular, uint32_t rows, const uint32_t *row_box, const uint32_t *row_cell,
              const uint32_t *row_off, const uint32_t *row_ent, const uint32_t *leaf_cell, const uint32_t *src_start,
              const double *sx, const double *sy, const double *sz, const double *qr, const double *qi,
              const double *hlev, const double *norm, double *coeffs) {
  if (rows == 0) return;
  if (singular)
    hipLaunchKernelGGL(k_h_expand<true>, dim3(xcd_grid(rows)), dim3(256), 0, s, p, rows, row_box, row_cell, row_off,
                       row_ent, leaf_cell, src_start, sx, sy, sz, qr, qi, hlev, norm, coeffs);
  else
    hipLaunchKernelGGL(k_h_expand<false>, dim3(xcd_grid(rows)), dim3(256), 0, s, p, rows, row_box, row_cell,
                       row_off, row_ent, leaf_cell, src_start, sx, sy, sz, qr, qi, hlev, norm, coeffs);
}

void h_evaluate(hipStream_t s, int p, int singular, int gradients, uint32_t rows, const uint32_t *row_leaf,
                const uint32_t *row_off, const uint32_t *row_box, const uint32_t *box_cell,
                const uint32_t *leaf_cell, const uint32_t *tgt_start, const double *tx, const double *ty,
                const double *tz, const double *hlev, const double *norm, const double *coeffs, double *phi_hat,
                double *grad_hat) {
  if (rows == 0) return;
#define FMMH_HEVAL(S, G)                                                                                        \
  hipLaunchKernelGGL((k_h_evaluate<S, G>), dim3(xcd_grid(rows)), dim3(64), 0, s, p, rows, row_leaf, row_off,   \
                     row_box, box_cell, leaf_cell, tgt_start, tx, ty, tz, hlev, norm, coeffs, phi_hat, grad_hat)
  if (singular) {
    if (gradients)
      FMMH_HEVAL(true, true);
    else
      FMMH_HEVAL(true, false);
  } else {
    if (gradients)
      FMMH_HEVAL(false, true);
    else
      FMMH_HEVAL(false, false);
  }
#undef FMMH_HEVAL
}

template <typename T>
void h_p2p(hipStream_t s, uint32_t leaf_first, uint32_t count, int gradients, int damped, const uint32_t *near_off,
           const uint32_t *near_ent, const uint32_t *leaf_cell, const uint32_t *src_start, const uint32_t *tgt_start,
           const double *sx, const double *sy, const double *sz, const double *qr, const double *qi,
           const float *rec, const double *tx, const double *ty, const double *tz, const double *hlev,
           double *phi_hat, double *grad_hat) {
  if (count == 0) return;
  const float4 *r4 = reinterpret_cast<const float4 *>(rec);
#define FMMH_HP2P(G, D)                                                                                         \
  hipLaunchKernelGGL((k_h_p2p<T, G, D>), dim3(xcd_grid(count)), dim3(64), 0, s, leaf_first, count, near_off,   \
                     near_ent, leaf_cell, src_start, tgt_start, sx, sy, sz, qr, qi, r4, tx, ty, tz, hlev,       \
                     phi_hat, grad_hat)
  if (gradients) {
    if (damped)
      FMMH_HP2P(true, true);
    else
      FMMH_HP2P(true, false);
  } else {
    if (damped)
      FMMH_HP2P(false, true);
    else
      FMMH_HP2P(false, false);
  }
#undef FMMH_HP2P
}

template void h_p2p<double>(hipStream_t, uint32_t, uint32_t, int, int, const uint32_t *, const uint32_t *,
                            const uint32_t *, const uint32_t *, const uint32_t *, const double *, const double *,
                            const double *, const double *, const double *, const float *, const double *,
                            const double *, const double *, const double *, double *, double *);
template void h_p2p<float>(hipStream_t, uint32_t, uint32_t, int, int, const uint32_t *, const uint32_t *,
                           const uint32_t *, const uint32_t *, const uint32_t *, const double *, const double *,
                           const double *, const double *, const double *, const float *, const double *,
                           const double *, const double *, const double *, double *, double *);

void h_negate(hipStream_t s, int p, size_t b0, size_t b1, double *coeffs) {
  if (b1 <= b0) return;
  const int nc = (p + 1) * (p + 1);
  const size_t n = (b1 - b0) * size_t(nc);
  const unsigned blocks = unsigned(n / 256 + 1 < 4096 ? n / 256 + 1 : 4096);
  hipLaunchKernelGGL(k_h_negate, dim3(blocks), dim3(256), 0, s, nc, b0, b1, coeffs);
}

void h_gather(hipStream_t s, const double *q, const uint32_t *order, size_t n, double *qr, double *qi, float *rec) {
  if (n == 0) return;
  const unsigned blocks = unsigned(n / 256 + 1 < 4096 ? n / 256 + 1 : 4096);
  hipLaunchKernelGGL(k_h_gather, dim3(blocks), dim3(256), 0, s, q, order, n, qr, qi, rec);
}

void h_source_records(const double *sx, const double *sy, const double *sz, size_t n, float *host_rec) {
  for (size_t j = 0; j < n; ++j) {
    const double u[3] = {sx[j], sy[j], sz[j]};
    float hi[3], lo[3];
    for (int k = 0; k < 3; ++k) {
      hi[k] = float(u[k]);
      lo[k] = float(u[k] - double(hi[k]));
    }
    float *r = host_rec + 8 * j;
    r[0] = hi[0];
    r[1] = hi[1];
    r[2] = hi[2];
    r[3] = lo[0];
    r[4] = lo[1];
    r[5] = lo[2];
    r[6] = 0.0f;
    r[7] = 0.0f;
  }
}

void h_output(hipStream_t s, uint32_t nleaves, const uint32_t *tgt_start, const double *leaf_r,
              const uint32_t *order, const double *phi_hat, const double *grad_hat, double *phi, double *grad,
              int gradients) {
  if (nleaves == 0) return;
  hipLaunchKernelGGL(k_h_output, dim3(nleaves), dim3(64), 0, s, nleaves, tgt_start, leaf_r, order, phi_hat, grad_hat,
                     phi, grad, gradients);
}

int h_direct_chunks(int64_t ns, int64_t nt) { return direct_chunks(ns, nt); }

void h_direct_sum(hipStream_t s, double kre, double kim, const double *src, const double *q, int64_t ns,
                  const double *tgt, int64_t nt, int gradients, double *part, int nchunks, double *phi,
                  double *grad) {
  if (nt == 0) return;
  const int64_t chunk = (ns + nchunks - 1) / nchunks;
  hipLaunchKernelGGL(k_h_direct_partial, dim3(unsigned((nt + 255) / 256), unsigned(nchunks)), dim3(256), 0, s, kre,
                     kim, src, q, ns, tgt, nt, chunk, gradients, part);
  hipLaunchKernelGGL(k_h_direct_combine, dim3(unsigned((nt + 255) / 256)), dim3(256), 0, s, nt, nchunks, gradients,
                     part, phi, grad);
}

}  // namespace dev
}  // namespace fmmh
