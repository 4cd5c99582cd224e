// hpwave.hip -- the Helmholtz exponential (plane-wave) M2L on the device by merge and shift
// (the steps of hpwtables.cpp; reference fmm-exec/src/helmholtz/operator_plane_wave.rs and
// fmm-tables/src/helmholtz/plane_wave/{pair,merge}.rs).
//
// One level per launch (each level has its own quadrature, W complex wave values per
// direction):
//   k_hpw_outgoing: one 512-thread block per box of the level.  The stored multipole (real
//     and imaginary planes) rotated into the three frames (thread = (frame, plane, slot)),
//     the modes A_k(m), B_k(m) up and down (thread = (frame, node, order)), the waves of the
//     six directions (thread = (frame, wave value), cos and sin m a_j by a recurrence in m):
//     waves[box][dir][re, im][W].
//   the merge: k_pw_merge of pwave.hip (the accumulators and child shifts are complex in
//     both families; W in place of H).
//   k_hpw_local: one 512-thread block per target box, frame by frame: its up, then its
//     down incoming waves staged in LDS and summed over the wave values per (node, order;
//     cos and sin m a_j by a recurrence in j), the frame's local per (n, m) from both, the rotation back per (plane, slot), summed over the frames and added to the
//     box's local (planes 0 and 1).
// f64 throughout (the reference's exponential M2L is f64 only).
#include <hip/hip_runtime.h>

#include <cstdint>

#include "hpwkernels.hpp"
#include "kernels.hpp"

namespace fmmh {
namespace dev {
namespace {

constexpr int kHpwThreads = 512;

__device__ __forceinline__ int degree_of(int k) {
  int n = 0;
  while ((n + 1) * (n + 1) <= k) ++n;
  return n;
}

// sum_j row[j * stride] x[j], j < w: four independent partial sums (their table loads in
// flight together), added in a fixed order
__device__ __forceinline__ double strided_dot(const double *__restrict__ row, int stride, int w, const double *x) {
  double v0 = 0.0, v1 = 0.0, v2 = 0.0, v3 = 0.0;
  int j = 0;
  for (; j + 3 < w; j += 4) {
    v0 = fma(row[j * stride], x[j], v0);
    v1 = fma(row[(j + 1) * stride], x[j + 1], v1);
    v2 = fma(row[(j + 2) * stride], x[j + 2], v2);
    v3 = fma(row[(j + 3) * stride], x[j + 3], v3);
  }
  for (; j < w; ++j) v0 = fma(row[j * stride], x[j], v0);
  return (v0 + v1) + (v2 + v3);
}

// first factor of order m of a node whose factors start at f0: per order m' < m the
// degrees m'..p
__device__ __forceinline__ int factor_of(int f0, int p, int m) { return f0 + m * (p + 1) - m * (m - 1) / 2; }

__global__ __launch_bounds__(kHpwThreads) void k_hpw_outgoing(HpwArgs a, uint32_t nboxes, uint32_t first,
                                                              const double *__restrict__ M,
                                                              double *__restrict__ waves) {
  extern __shared__ double sm[];
  const int p = a.p, nc = (p + 1) * (p + 1), W = a.W, O = a.orders;
  double *x = sm;             // [frame][plane][nc]
  double *md = x + 6 * nc;    // [frame][order][up A, up B, down A, down B] complex
  const uint32_t item = xcd_index(nboxes);
  if (item >= nboxes) return;
  const size_t box = size_t(first) + item;
  const int t = threadIdx.x;
  for (int q = t; q < 2 * nc; q += kHpwThreads) x[q] = M[(3 * box + q / nc) * nc + q % nc];
  __syncthreads();
  // frames 1 and 2: per plane and degree O^n x (row-major blocks)
  for (int it = t; it < 4 * nc; it += kHpwThreads) {
    const int f = 1 + it / (2 * nc), r = it % (2 * nc), pl = r / nc, k = r % nc;
    const int n = degree_of(k), w = 2 * n + 1, i = k - n * n;
    const double *blk = a.frames + size_t(f - 1) * a.rb + n * (2 * n - 1) * (2 * n + 1) / 3 + i * w;
    x[(2 * f + pl) * nc + k] = strided_dot(blk, 1, w, x + pl * nc + n * n);
  }
  __syncthreads();
  // modes: per order the sums over n of P^out_k(n, m) M^{+-m}_n, even and odd n - m apart
  for (int it = t; it < 3 * a.nitem; it += kHpwThreads) {
    const int f = it % 3, q = a.item[it / 3], k = q / (p + 1), m = q % (p + 1);
    const int f0 = factor_of(a.factor[k], p, m);
    const double *xr = x + (2 * f) * nc, *xi = x + (2 * f + 1) * nc;
    double acc[2][4] = {{0, 0, 0, 0}, {0, 0, 0, 0}};  // [parity][A re, A im, B re, B im]
    for (int n = m; n <= p; ++n) {
      const double fr = a.out_re[f0 + n - m], fi = a.out_im[f0 + n - m];
      const int par = (n - m) & 1, ip = n * n + n + m;
      acc[par][0] = fma(fr, xr[ip], fma(-fi, xi[ip], acc[par][0]));
      acc[par][1] = fma(fr, xi[ip], fma(fi, xr[ip], acc[par][1]));
      if (m > 0) {
        const int im = n * n + n - m;
        acc[par][2] = fma(fr, xr[im], fma(-fi, xi[im], acc[par][2]));
        acc[par][3] = fma(fr, xi[im], fma(fi, xr[im], acc[par][3]));
      }
    }
    double *o = md + (size_t(f) * O + a.mode_start[k] + m) * 8;
    for (int c = 0; c < 4; ++c) {
      o[c] = acc[0][c] + acc[1][c];      // up
      o[4 + c] = acc[0][c] - acc[1][c];  // down
    }
  }
  __syncthreads();
  // waves: X(k, j) = sum_m A(m) cos m a_j + B(m) sin m a_j, up and down
  double *wv = waves + size_t(item) * 12 * W;
  for (int it = t; it < 3 * W; it += kHpwThreads) {
    const int f = it / W, j = it % W, k = a.wk[j], jj = j - a.start[k], N = a.len[k], om = a.ord[k];
    const double *mk = md + (size_t(f) * O + a.mode_start[k]) * 8;
    // cos and sin of m a_j by the recurrence in m from row 1 of the node's table (a_j)
    const double c1 = om > 1 ? a.cosm[a.trig[k] + N + jj] : 1.0, s1 = om > 1 ? a.sinm[a.trig[k] + N + jj] : 0.0;
    double c = 1.0, s = 0.0;
    double ur = 0.0, ui = 0.0, dr = 0.0, di = 0.0;
    for (int m = 0; m < om; ++m) {
      if (m > 0) {
        const double nc_ = c * c1 - s * s1;
        s = s * c1 + c * s1;
        c = nc_;
      }
      const double *mm = mk + 8 * m;
      ur = fma(mm[0], c, fma(mm[2], s, ur));
      ui = fma(mm[1], c, fma(mm[3], s, ui));
      dr = fma(mm[4], c, fma(mm[6], s, dr));
      di = fma(mm[5], c, fma(mm[7], s, di));
    }
    wv[(2 * f) * 2 * W + j] = ur;
    wv[(2 * f) * 2 * W + W + j] = ui;
    wv[(2 * f + 1) * 2 * W + j] = dr;
    wv[(2 * f + 1) * 2 * W + W + j] = di;
  }
}

__global__ __launch_bounds__(kHpwThreads) void k_hpw_local(HpwArgs a, uint32_t ntargets,
                                                           const uint32_t *__restrict__ target, uint32_t box0,
                                                           const double *__restrict__ inc, double *__restrict__ L) {
  extern __shared__ double sm[];
  const int p = a.p, P1 = p + 1, nc = P1 * P1, W = a.W, O = a.orders, s = a.nodes;
  double *V = sm;              // one direction's waves [re, im][W]
  double *g = V + 2 * W;       // [up/down][order][cos re, cos im, sin re, sin im]
  double *ys = g + 8 * O;      // the frame's local [plane][nc]
  const uint32_t row = xcd_index(ntargets);
  if (row >= ntargets) return;
  const int t = threadIdx.x;
  const uint32_t box = target[row];
  const double *in = inc + size_t(box - box0) * 12 * W;
  double acc = 0.0;  // plane t / nc, slot t % nc
  for (int f = 0; f < 3; ++f)
  for (int ud = 0; ud < 2; ++ud) {
    // one direction's waves at a time (half the LDS of both: two blocks per CU)
    for (int q = t; q < 2 * W; q += kHpwThreads) V[q] = in[size_t(2 * f + ud) * 2 * W + q];
    __syncthreads();
    // the sums over the wave values of V cos m a_j and V sin m a_j
    for (int it = t; it < a.nitem; it += kHpwThreads) {
      const int q = a.item[it], k = q / P1, m = q % P1;
      const int N = a.len[k], st = a.start[k];
      const double *vr = V + st, *vi = vr + W;
      // cos and sin of m a_j by the recurrence in j from the step m a_1 (the table's row m)
      const double cd = N > 1 ? a.cosm[a.trig[k] + m * N + 1] : 1.0, sd = N > 1 ? a.sinm[a.trig[k] + m * N + 1] : 0.0;
      double c = 1.0, sv = 0.0;
      double cr = 0.0, ci = 0.0, sr = 0.0, si = 0.0;
      for (int j = 0; j < N; ++j) {
        cr = fma(vr[j], c, cr);
        ci = fma(vi[j], c, ci);
        sr = fma(vr[j], sv, sr);
        si = fma(vi[j], sv, si);
        const double nc_ = c * cd - sv * sd;
        sv = sv * cd + c * sd;
        c = nc_;
      }
      double *o = g + (size_t(ud) * O + a.mode_start[k] + m) * 4;
      o[0] = cr;
      o[1] = ci;
      o[2] = sr;
      o[3] = si;
    }
    __syncthreads();
    if (ud == 0) continue;
    // the frame's local: slot +m from the cos sums, -m from the sin sums, up and down
    const int tri = P1 * (P1 + 1) / 2;
    for (int it = t; it < tri; it += kHpwThreads) {
      int r = it, n = 0;
      while (r > n) r -= ++n;
      const int m = r;
      const double sg = ((n - m) & 1) ? -1.0 : 1.0;  // the down waves' (-1)^(n-m)
      double pr = 0.0, pi = 0.0, mr = 0.0, mi = 0.0;
      for (int k = 0; k < s; ++k) {
        if (m >= a.ord[k]) continue;
        const int fi = factor_of(a.factor[k], p, m) + n - m;
        const double qr = a.in_re[fi], qi = a.in_im[fi];
        const double *gu = g + (size_t(a.mode_start[k]) + m) * 4, *gd = g + (size_t(O) + a.mode_start[k] + m) * 4;
        // cos: Q (gu) + sg Q (gd)
        const double c0 = gu[0] + sg * gd[0], c1 = gu[1] + sg * gd[1];
        pr = fma(qr, c0, fma(-qi, c1, pr));
        pi = fma(qr, c1, fma(qi, c0, pi));
        const double s0 = gu[2] + sg * gd[2], s1 = gu[3] + sg * gd[3];
        mr = fma(qr, s0, fma(-qi, s1, mr));
        mi = fma(qr, s1, fma(qi, s0, mi));
      }
      ys[n * n + n + m] = pr;
      ys[nc + n * n + n + m] = pi;
      if (m > 0) {
        ys[n * n + n - m] = mr;
        ys[nc + n * n + n - m] = mi;
      }
    }
    __syncthreads();
    // back out of the frame: per plane and degree O^n(Q)^T L'
    if (t < 2 * nc) {
      const int pl = t / nc, j = t % nc, n = degree_of(j), w = 2 * n + 1, jj = j - n * n;
      const double *y = ys + pl * nc + n * n;
      if (f == 0) {
        acc += y[jj];
      } else {
        const double *blk = a.frames + size_t(f - 1) * a.rb + n * (2 * n - 1) * (2 * n + 1) / 3 + jj;
        acc += strided_dot(blk, w, w, y);
      }
    }
    __syncthreads();
  }
  if (t < 2 * nc) L[(3 * size_t(box) + t / nc) * nc + t % nc] += acc;
}

}  // namespace

size_t hpw_outgoing_lds(int p, int orders) {
  const size_t nc = size_t(p + 1) * (p + 1);
  return sizeof(double) * (6 * nc + 24 * size_t(orders));
}

size_t hpw_local_lds(int p, int W, int orders) {
  const size_t nc = size_t(p + 1) * (p + 1);
  return sizeof(double) * (2 * size_t(W) + 8 * size_t(orders) + 2 * nc);
}

hipError_t hpw_allow_lds(size_t outgoing, size_t local) {
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(&k_hpw_outgoing),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, int(outgoing));
  if (e != hipSuccess) return e;
  return hipFuncSetAttribute(reinterpret_cast<const void *>(&k_hpw_local), hipFuncAttributeMaxDynamicSharedMemorySize,
                             int(local));
}

void hpw_outgoing(hipStream_t s, const HpwArgs &a, uint32_t nboxes, uint32_t first, const double *M, double *waves) {
  if (nboxes == 0) return;
  hipLaunchKernelGGL(k_hpw_outgoing, dim3(xcd_grid(nboxes)), dim3(kHpwThreads), hpw_outgoing_lds(a.p, a.orders), s, a,
                     nboxes, first, M, waves);
}

void hpw_local(hipStream_t s, const HpwArgs &a, uint32_t ntargets, const uint32_t *target, uint32_t box0,
               const double *inc, double *L) {
  if (ntargets == 0) return;
  hipLaunchKernelGGL(k_hpw_local, dim3(xcd_grid(ntargets)), dim3(kHpwThreads), hpw_local_lds(a.p, a.W, a.orders), s,
                     a, ntargets, target, box0, inc, L);
}

}  // namespace dev
}  // namespace fmmh
