// pwave.hip -- plane-wave M2L on the device by merge and shift (the steps of pwtables.cpp;
// reference fmm-kernels/src/plane_wave.rs:1-60: outgoing, merge, gather, local).
//
// Three kernels, all levels of a call in one launch each:
//   k_pw_outgoing: one 512-thread block per source box, the three frames at once: the
//     multipole rotated into each frame (thread = (frame, slot)), A_k(m) formed for up and
//     down (thread = (frame, node, order), orders m < F_k'), and the waves X of the H half
//     nodes written for the six directions (thread = (frame, half node)):
//     waves[box][2 f + down][re, im][H].
//   k_pw_merge: one block per (parent, direction), a thread per half node.  Each of the
//     direction's accumulators sums its slots' waves times their shifts to the parent's
//     centre; each child adds the accumulators that serve it and multiplies by its shift
//     from the parent's centre: inc[child][dir][re, im][H].  The merge and the gather of
//     the reference in one pass, the accumulators in registers.
//   k_pw_local: one 512-thread block per target box, the six directions at once: G_k(m)
//     from the incoming waves (thread = (direction, node, order), the e^{-i m a_j} by a
//     recurrence in j), every frame's local L'_n^m from its up and down sums (thread =
//     (frame, n, m)), the three rotations back summed per slot and added to the box's local.
// f64 throughout (the plane-wave M2L is an f64 operator in the reference).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "kernels.hpp"
#include "pwkernels.hpp"

namespace fmmh {
namespace dev {
namespace {

constexpr int kPwLocalThreads = 512;
constexpr uint32_t kNoBox = 0xffffffffu;

// LDS words of k_pw_local's waves region (the six directions' waves, later the three
// frames' locals and their rotations back)
__host__ __device__ inline size_t pw_local_wave_words(int p, int H) {
  const size_t a = 12 * size_t(H), b = 6 * size_t(p + 1) * (p + 1);  // later ys[3][nc], part[3][nc]
  return a > b ? a : b;
}

// sum_c blk[c * w] x[c], c < w: four independent partial sums (their table loads in
// flight together), added in a fixed order
__device__ __forceinline__ double column_dot(const double *__restrict__ blk, int w, const double *x) {
  double v0 = 0.0, v1 = 0.0, v2 = 0.0, v3 = 0.0;
  int c = 0;
  for (; c + 3 < w; c += 4) {
    v0 = fma(blk[c * w], x[c], v0);
    v1 = fma(blk[(c + 1) * w], x[c + 1], v1);
    v2 = fma(blk[(c + 2) * w], x[c + 2], v2);
    v3 = fma(blk[(c + 3) * w], x[c + 3], v3);
  }
  for (; c < w; ++c) v0 = fma(blk[c * w], x[c], v0);
  return (v0 + v1) + (v2 + v3);
}

// slot k -> degree
__device__ __forceinline__ int degree_of(int k) {
  int n = 0;
  while ((n + 1) * (n + 1) <= k) ++n;
  return n;
}

// One block per source box, the three frames at once: the rotations into the frames
// (thread = (frame, slot)), the modes A_k(m) up and down (thread = (frame, node, order)),
// the waves of the six directions (thread = (frame, half node)).
__global__ __launch_bounds__(kPwLocalThreads) void k_pw_outgoing(PwArgs a, uint32_t nboxes, uint32_t box0,
                                                                 const double *__restrict__ M,
                                                                 double *__restrict__ waves) {
  extern __shared__ double sm[];
  const int p = a.p, P1 = p + 1, nc = P1 * P1, H = a.H, s = a.nodes;
  double *Ms = sm, *xs = sm + nc;          // the multipole; its rotations into the frames [3][nc]
  double *A = xs + 3 * nc;                 // [3][s][P1][up re, up im, down re, down im]
  const uint32_t item = xcd_index(nboxes);
  if (item >= nboxes) return;
  const size_t box = size_t(box0) + item;
  const int t = threadIdx.x;
  for (int k = t; k < nc; k += kPwLocalThreads) Ms[k] = M[box * nc + k];
  __syncthreads();
  // rotate into the frames: x_n = F_n M_n (column-major blocks)
  for (int it = t; it < 3 * nc; it += kPwLocalThreads) {
    const int f = it / nc, k = it % nc;
    const int n = degree_of(k), w = 2 * n + 1, r = k - n * n;
    const double *blk = a.fwd + size_t(f) * a.rb + n * (2 * n - 1) * (2 * n + 1) / 3;
    xs[f * nc + k] = column_dot(blk + r, w, Ms + n * n);
  }
  __syncthreads();
  // A_k(m) = sum_{n >= m} lambda^n s^(n-m) 2^-(n+1) x_n^m, up (s = 1) and down (s = -1)
  for (int it = t; it < 3 * a.nitem; it += kPwLocalThreads) {
    const int f = it % 3, q = a.gitem[it / 3], k = q / P1, m = q % P1;
    const double *pw = a.opow + k * P1, *x = xs + f * nc;
    double ur = 0.0, ui = 0.0, dr = 0.0, di = 0.0, sg = 1.0;
    for (int n = m; n <= p; ++n) {
      const double ln = pw[n], re = x[n * n + n + m], im = m > 0 ? x[n * n + n - m] : 0.0;
      ur = fma(ln, re, ur);
      ui = fma(ln, im, ui);
      dr = fma(sg * ln, re, dr);
      di = fma(sg * ln, im, di);
      sg = -sg;
    }
    double *o = A + ((f * s + k) * P1 + m) * 4;
    o[0] = ur;
    o[1] = ui;
    o[2] = dr;
    o[3] = di;
  }
  __syncthreads();
  // X(k, j) = (w_k / N_k) [A(0) + sum_{1 <= m < F_k'} 2 (-i)^m Re(A(m) e^{i m a_j})]
  double *wv = waves + size_t(item) * 12 * H;
  for (int it = t; it < 3 * H; it += kPwLocalThreads) {
    const int f = it / H, h = it % H;
    const int k = a.hk[h], F = a.fourier[k];
    const double *Ak = A + (f * s + k) * P1 * 4;
    const double c1 = a.hc[h], s1 = a.hs[h];
    double cr = 1.0, ci = 0.0;  // e^{i m a}
    double xur = Ak[0], xui = 0.0, xdr = Ak[2], xdi = 0.0;
    for (int m = 1; m < F; ++m) {
      const double nr = cr * c1 - ci * s1, ni = cr * s1 + ci * c1;
      cr = nr;
      ci = ni;
      // Re(A e^{i m a}) for up and down
      const double ru = 2.0 * (Ak[4 * m] * cr - Ak[4 * m + 1] * ci);
      const double rd = 2.0 * (Ak[4 * m + 2] * cr - Ak[4 * m + 3] * ci);
      // (-i)^m: 1, -i, -1, i
      switch (m & 3) {
        case 0:
          xur += ru;
          xdr += rd;
          break;
        case 1:
          xui -= ru;
          xdi -= rd;
          break;
        case 2:
          xur -= ru;
          xdr -= rd;
          break;
        default:
          xui += ru;
          xdi += rd;
          break;
      }
    }
    const double wn = a.hw[h];
    wv[(2 * f) * 2 * H + h] = wn * xur;
    wv[(2 * f) * 2 * H + H + h] = wn * xui;
    wv[(2 * f + 1) * 2 * H + h] = wn * xdr;
    wv[(2 * f + 1) * 2 * H + H + h] = wn * xdi;
  }
}

// One block per (parent, direction); threads stride over the half nodes.  The held
// entries of the parent's accumulators of the direction are listed ahead (pe_first per
// (parent, direction) and accumulator, pe = wave block << 7 | source shift), so the loads
// of successive entries do not wait on one another.
__global__ __launch_bounds__(1024) void k_pw_merge(PwArgs a, PwMerge g, uint32_t nparents,
                                                   const uint32_t *__restrict__ pe_first,
                                                   const uint32_t *__restrict__ pe,
                                                   const uint32_t *__restrict__ children, uint32_t box0,
                                                   const double *__restrict__ waves, double *__restrict__ inc) {
  const uint32_t item = xcd_index(nparents * 6);
  if (item >= nparents * 6) return;
  const uint32_t parent = item / 6;
  const int dir = int(item % 6), H = a.H;
  const uint32_t *fo = pe_first + size_t(item) * kPwAccStride;
  const uint32_t *ch = children + size_t(parent) * 8;
  const int a0 = g.acc_first[dir], na = g.acc_first[dir + 1] - a0;
  for (int h = threadIdx.x; h < H; h += blockDim.x) {
    double vr[8], vi[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) vr[c] = vi[c] = 0.0;
    for (int ac = 0; ac < na; ++ac) {
      double sr = 0.0, si = 0.0;
      const uint32_t e1 = fo[ac + 1];
#pragma unroll 4
      for (uint32_t e = fo[ac]; e < e1; ++e) {
        const uint32_t w = pe[e];
        const double *X = waves + size_t(w >> 7) * 2 * H;
        const double *T = g.src_shift + size_t(w & 127u) * 2 * H;
        const double xr = X[h], xi = X[H + h], tr = T[h], ti = T[H + h];
        sr = fma(xr, tr, fma(-xi, ti, sr));
        si = fma(xr, ti, fma(xi, tr, si));
      }
      const int mask = g.acc_mask[a0 + ac];
#pragma unroll
      for (int c = 0; c < 8; ++c)
        if ((mask >> c) & 1) {
          vr[c] += sr;
          vi[c] += si;
        }
    }
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      const uint32_t cb = ch[c];
      if (cb == kNoBox) continue;
      const double *T = g.child_shift + size_t(8 * dir + c) * 2 * H;
      const double tr = T[h], ti = T[H + h];
      double *V = inc + (size_t(cb - box0) * 6 + dir) * 2 * H;
      V[h] = vr[c] * tr - vi[c] * ti;
      V[H + h] = vr[c] * ti + vi[c] * tr;
    }
  }
}

// One block per target box.  All six directions at once: the incoming waves staged in LDS,
// then every (direction, node, order) sum G (items by decreasing alpha count, dealt to the
// threads in turn), then every frame's local (thread = (frame, n, m)), stored over the
// waves, then the rotations back (thread = slot).
__global__ __launch_bounds__(kPwLocalThreads) void k_pw_local(PwArgs a, uint32_t ntargets,
                                                         const uint32_t *__restrict__ target, uint32_t box0,
                                                         const double *__restrict__ inc, double *__restrict__ L) {
  extern __shared__ double sm[];
  const int p = a.p, P1 = p + 1, nc = P1 * P1, H = a.H, s = a.nodes;
  double *V = sm;                      // [6][re, im][H], later the frames' locals ys[3][nc]
  double *Gr = sm + pw_local_wave_words(p, H);  // [6][s][P1]
  double *Gi = Gr + 6 * s * P1;
  double *ip = Gi + 6 * s * P1;        // the incoming powers [s][P1]
  int *kf = reinterpret_cast<int *>(ip + s * P1);  // per node: first half node, N_k, F_k'
  int *gq = kf + 3 * kPwMaxNodes;      // the (node, order) items
  const uint32_t row = xcd_index(ntargets);
  if (row >= ntargets) return;
  const int t = threadIdx.x;
  const uint32_t box = target[row];
  const double *in = inc + size_t(box - box0) * 12 * H;
  for (int h = t; h < 12 * H; h += kPwLocalThreads) V[h] = in[h];
  for (int q = t; q < s * P1; q += kPwLocalThreads) ip[q] = a.ipow[q];
  for (int k = t; k < s; k += kPwLocalThreads) {
    kf[3 * k] = a.first[k];
    kf[3 * k + 1] = a.nalpha[k];
    kf[3 * k + 2] = a.fourier[k];
  }
  for (int q = t; q < a.nitem; q += kPwLocalThreads) gq[q] = a.gitem[q];
  __syncthreads();
  // G_k(m) = 2 sum_j Re V e^{-i m a_j} (m even), 2 i sum_j Im V e^{-i m a_j} (m odd)
  for (int it = t; it < 6 * a.nitem; it += kPwLocalThreads) {
    const int dir = it % 6, q = gq[it / 6], k = q / P1, m = q % P1;
    const int h0 = kf[3 * k], nh = kf[3 * k + 1] / 2;
    double sr, si;
    sincos(-2.0 * 3.141592653589793 * m / kf[3 * k + 1], &si, &sr);  // e^{-i m 2 pi / N_k}
    double er = 1.0, ei = 0.0, gr = 0.0, gi = 0.0;
    const bool even = (m & 1) == 0;
    const double *v = V + dir * 2 * H + (even ? 0 : H) + h0;
    for (int j = 0; j < nh; ++j) {
      gr = fma(v[j], er, gr);
      gi = fma(v[j], ei, gi);
      const double nr = er * sr - ei * si, ni = er * si + ei * sr;
      er = nr;
      ei = ni;
    }
    // even: 2 (gr + i gi); odd: 2 i (gr + i gi) = 2 (-gi + i gr)
    Gr[(dir * s + k) * P1 + m] = even ? 2.0 * gr : -2.0 * gi;
    Gi[(dir * s + k) * P1 + m] = even ? 2.0 * gi : 2.0 * gr;
  }
  __syncthreads();
  // the frame's local: L'_n^m = sum_{up, down} sum_{k: m < F_k'} 2^-n lambda_k^n (-s)^(n-m) i^m
  // G_k(m) (s = 1 up, -1 down), stored as the real coefficients of the frame
  double *ys = V;
  const int tri = P1 * (P1 + 1) / 2;
  for (int it = t; it < 3 * tri; it += kPwLocalThreads) {
    const int f = it / tri;
    int r = it % tri, n = 0;
    while (r > n) r -= ++n;
    const int m = r;
    double ur = 0.0, ui = 0.0, dr = 0.0, di = 0.0;
    for (int k = 0; k < s; ++k) {
      if (m >= kf[3 * k + 2]) continue;
      const double c = ip[k * P1 + n];
      const int gu = ((2 * f) * s + k) * P1 + m, gd = ((2 * f + 1) * s + k) * P1 + m;
      ur = fma(c, Gr[gu], ur);
      ui = fma(c, Gi[gu], ui);
      dr = fma(c, Gr[gd], dr);
      di = fma(c, Gi[gd], di);
    }
    // up carries (-1)^(n-m)
    const double sr = ((n - m) & 1) ? dr - ur : dr + ur, si = ((n - m) & 1) ? di - ui : di + ui;
    // times i^m
    double rr = sr, ri = si;
    switch (m & 3) {
      case 1:
        rr = -si;
        ri = sr;
        break;
      case 2:
        rr = -sr;
        ri = -si;
        break;
      case 3:
        rr = si;
        ri = -sr;
        break;
      default:
        break;
    }
    ys[f * nc + n * n + n + m] = rr;
    if (m > 0) ys[f * nc + n * n + n - m] = ri;
  }
  __syncthreads();
  // each frame's rotation back in its own thread, then the three added per slot in frame order
  double *part = ys + 3 * nc;
  for (int it = t; it < 3 * nc; it += kPwLocalThreads) {
    const int f = it / nc, k = it % nc, n = degree_of(k), w = 2 * n + 1, r = k - n * n;
    const double *blk = a.bwd + size_t(f) * a.rb + n * (2 * n - 1) * (2 * n + 1) / 3;
    part[it] = column_dot(blk + r, w, ys + f * nc + n * n);
  }
  __syncthreads();
  if (t < nc) L[size_t(box) * nc + t] += (part[t] + part[nc + t]) + part[2 * nc + t];
}

}  // namespace

size_t pw_local_lds(int p, int H, int nodes) {
  const size_t P1 = p + 1;
  return sizeof(double) * (pw_local_wave_words(p, H) + 13 * size_t(nodes) * P1) +
         sizeof(int) * (3 * size_t(kPwMaxNodes) + size_t(kPwMaxNodes) * P1);
}

void pw_outgoing(hipStream_t s, const PwArgs &a, uint32_t nboxes, uint32_t box0, const double *M, double *waves) {
  if (nboxes == 0) return;
  const size_t P1 = a.p + 1, lds = sizeof(double) * (4 * P1 * P1 + 12 * size_t(a.nodes) * P1);
  hipLaunchKernelGGL(k_pw_outgoing, dim3(xcd_grid(nboxes)), dim3(kPwLocalThreads), lds, s, a, nboxes, box0, M, waves);
}

void pw_merge(hipStream_t s, const PwArgs &a, const PwMerge &g, uint32_t nparents, const uint32_t *pe_first,
              const uint32_t *pe, const uint32_t *children, uint32_t box0, const double *waves, double *inc) {
  if (nparents == 0) return;
  const int threads = a.H >= 1024 ? 1024 : (a.H + 63) / 64 * 64;
  hipLaunchKernelGGL(k_pw_merge, dim3(xcd_grid(nparents * 6)), dim3(threads), 0, s, a, g, nparents, pe_first, pe,
                     children, box0, waves, inc);
}

hipError_t pw_allow_lds(size_t bytes) {
  return hipFuncSetAttribute(reinterpret_cast<const void *>(&k_pw_local), hipFuncAttributeMaxDynamicSharedMemorySize,
                             int(bytes));
}

void pw_local(hipStream_t s, const PwArgs &a, uint32_t ntargets, const uint32_t *target, uint32_t box0,
              const double *inc, double *L) {
  if (ntargets == 0) return;
  hipLaunchKernelGGL(k_pw_local, dim3(xcd_grid(ntargets)), dim3(kPwLocalThreads), pw_local_lds(a.p, a.H, a.nodes), s, a, ntargets,
                     target, box0, inc, L);
}

}  // namespace dev
}  // namespace fmmh
