// pwave.hip -- plane-wave M2L on the device by merge and shift (the steps of pwtables.cpp;
// reference fmm-kernels/src/plane_wave.rs:1-60: outgoing, merge, gather, local).
//
// Three kernels, all levels of a call in one launch each:
//   k_pw_outgoing: one 256-thread block per source box.  For each frame f the multipole is
//     rotated into the frame (thread = slot), A_k(m) formed for up and down (thread =
//     (node, order), orders m < F_k'), and the waves X of the H half nodes written for the
//     six directions (thread = half node): waves[box][2 f + down][re, im][H].
//   k_pw_merge: one block per (parent, direction), a thread per half node.  Each of the
//     direction's accumulators sums its slots' waves times their shifts to the parent's
//     centre; each child adds the accumulators that serve it and multiplies by its shift
//     from the parent's centre: inc[child][dir][re, im][H].  The merge and the gather of
//     the reference in one pass, the accumulators in registers.
//   k_pw_local: one 256-thread block per target box.  Per direction G_k(m) from the
//     incoming waves (thread = (node, order), the e^{-i m a_j} by a recurrence in j), the
//     frame's local L'_n^m (thread = slot); after up and down the frame's local is rotated
//     back and added to the slot's sum, which is added to the box's local at the end.
// f64 throughout (the plane-wave M2L is an f64 operator in the reference).
#include <hip/hip_runtime.h>

#include <cstdint>

#include "kernels.hpp"
#include "pwkernels.hpp"

namespace fmmh {
namespace dev {
namespace {

constexpr int kPwThreads = 256;
constexpr uint32_t kNoBox = 0xffffffffu;

// slot k -> degree
__device__ __forceinline__ int degree_of(int k) {
  int n = 0;
  while ((n + 1) * (n + 1) <= k) ++n;
  return n;
}

__global__ __launch_bounds__(kPwThreads) void k_pw_outgoing(PwArgs a, uint32_t nboxes, uint32_t box0,
                                                            const double *__restrict__ M, double *__restrict__ waves) {
  __shared__ double Ms[kMaxCoef], xs[kMaxCoef];
  __shared__ double Au[kPwMaxNodes][kMaxDegree + 1][2], Ad[kPwMaxNodes][kMaxDegree + 1][2];
  const uint32_t item = xcd_index(nboxes);
  if (item >= nboxes) return;
  const size_t box = size_t(box0) + item;
  const int p = a.p, P1 = p + 1, nc = P1 * P1, H = a.H, s = a.nodes;
  const int t = threadIdx.x;
  for (int k = t; k < nc; k += kPwThreads) Ms[k] = M[box * nc + k];
  double *wv = waves + size_t(item) * 12 * H;
  for (int f = 0; f < 3; ++f) {
    __syncthreads();
    // rotate into the frame: x_n = F_n M_n (column-major blocks)
    for (int k = t; k < nc; k += kPwThreads) {
      const int n = degree_of(k), w = 2 * n + 1, r = k - n * n;
      const double *blk = a.fwd + size_t(f) * a.rb + n * (2 * n - 1) * (2 * n + 1) / 3;
      double v = 0.0;
      for (int c = 0; c < w; ++c) v = fma(blk[r + c * w], Ms[n * n + c], v);
      xs[k] = v;
    }
    __syncthreads();
    // A_k(m) = sum_{n >= m} lambda^n s^(n-m) 2^-(n+1) x_n^m, up (s = 1) and down (s = -1)
    for (int q = t; q < s * P1; q += kPwThreads) {
      const int k = q / P1, m = q % P1;
      if (m >= a.fourier[k]) continue;
      const double *pw = a.opow + k * P1;
      double ur = 0.0, ui = 0.0, dr = 0.0, di = 0.0, sg = 1.0;
      for (int n = m; n <= p; ++n) {
        const double ln = pw[n], re = xs[n * n + n + m], im = m > 0 ? xs[n * n + n - m] : 0.0;
        ur = fma(ln, re, ur);
        ui = fma(ln, im, ui);
        dr = fma(sg * ln, re, dr);
        di = fma(sg * ln, im, di);
        sg = -sg;
      }
      Au[k][m][0] = ur;
      Au[k][m][1] = ui;
      Ad[k][m][0] = dr;
      Ad[k][m][1] = di;
    }
    __syncthreads();
    // X(k, j) = (w_k / N_k) [A(0) + sum_{1 <= m < F_k'} 2 (-i)^m Re(A(m) e^{i m a_j})]
    for (int h = t; h < H; h += kPwThreads) {
      const int k = a.hk[h], F = a.fourier[k];
      const double c1 = a.hc[h], s1 = a.hs[h];
      double cr = 1.0, ci = 0.0;  // e^{i m a}
      double xur = Au[k][0][0], xui = 0.0, xdr = Ad[k][0][0], xdi = 0.0;
      for (int m = 1; m < F; ++m) {
        const double nr = cr * c1 - ci * s1, ni = cr * s1 + ci * c1;
        cr = nr;
        ci = ni;
        // Re(A e^{i m a}) for up and down
        const double ru = 2.0 * (Au[k][m][0] * cr - Au[k][m][1] * ci);
        const double rd = 2.0 * (Ad[k][m][0] * cr - Ad[k][m][1] * ci);
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
}

// One block per (parent, direction); threads stride over the half nodes.
__global__ __launch_bounds__(1024) void k_pw_merge(PwArgs a, PwMerge g, uint32_t nparents,
                                                   const uint32_t *__restrict__ slots,
                                                   const uint32_t *__restrict__ children, uint32_t box0,
                                                   const double *__restrict__ waves, double *__restrict__ inc) {
  const uint32_t item = xcd_index(nparents * 6);
  if (item >= nparents * 6) return;
  const uint32_t parent = item / 6;
  const int dir = int(item % 6), H = a.H;
  const uint32_t *sl = slots + size_t(parent) * 216;
  const uint32_t *ch = children + size_t(parent) * 8;
  const int a0 = g.acc_first[dir], a1 = g.acc_first[dir + 1];
  for (int h = threadIdx.x; h < H; h += blockDim.x) {
    double vr[8], vi[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) vr[c] = vi[c] = 0.0;
    for (int ac = a0; ac < a1; ++ac) {
      double sr = 0.0, si = 0.0;
      for (int e = g.ent_first[ac]; e < g.ent_first[ac + 1]; ++e) {
        const uint32_t src = sl[g.ent_slot[e]];
        if (src == kNoBox) continue;
        const double *X = waves + (size_t(src - box0) * 6 + dir) * 2 * H;
        const double *T = g.src_shift + size_t(g.ent_shift[e]) * 2 * H;
        const double xr = X[h], xi = X[H + h], tr = T[h], ti = T[H + h];
        sr = fma(xr, tr, fma(-xi, ti, sr));
        si = fma(xr, ti, fma(xi, tr, si));
      }
      const int mask = g.acc_mask[ac];
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

__global__ __launch_bounds__(kPwThreads) void k_pw_local(PwArgs a, uint32_t ntargets,
                                                         const uint32_t *__restrict__ target, uint32_t box0,
                                                         const double *__restrict__ inc, double *__restrict__ L) {
  extern __shared__ double sm[];
  const int p = a.p, P1 = p + 1, nc = P1 * P1, H = a.H, s = a.nodes;
  double *Vr = sm, *Vi = sm + H;                           // incoming waves of one direction
  double *Gr = Vi + H, *Gi = Gr + kPwMaxNodes * P1;       // G_k(m)
  double *Lr = Gi + kPwMaxNodes * P1, *Li = Lr + P1 * P1;  // the frame's local, [n][m]
  double *ys = Li + P1 * P1;                              // real storage of the frame's local
  const uint32_t row = xcd_index(ntargets);
  if (row >= ntargets) return;
  const int t = threadIdx.x;
  const uint32_t box = target[row];
  const double *in = inc + size_t(box - box0) * 12 * H;
  double acc = 0.0;  // slot t's sum over the frames
  for (int f = 0; f < 3; ++f) {
    for (int q = t; q < P1 * P1; q += kPwThreads) Lr[q] = Li[q] = 0.0;
    for (int ud = 0; ud < 2; ++ud) {
      const int dir = 2 * f + ud;
      __syncthreads();
      for (int h = t; h < 2 * H; h += kPwThreads) Vr[h] = in[size_t(dir) * 2 * H + h];
      __syncthreads();
      // G_k(m) = 2 sum_j Re V e^{-i m a_j} (m even), 2 i sum_j Im V e^{-i m a_j} (m odd)
      for (int q = t; q < s * P1; q += kPwThreads) {
        const int k = q / P1, m = q % P1;
        if (m >= a.fourier[k]) continue;
        const int h0 = a.first[k], nh = a.nalpha[k] / 2;
        double sr, si;
        sincos(-2.0 * 3.141592653589793 * m / a.nalpha[k], &si, &sr);  // e^{-i m 2 pi / N_k}
        double er = 1.0, ei = 0.0, gr = 0.0, gi = 0.0;
        const bool even = (m & 1) == 0;
        const double *v = even ? Vr + h0 : Vi + h0;
        for (int j = 0; j < nh; ++j) {
          gr = fma(v[j], er, gr);
          gi = fma(v[j], ei, gi);
          const double nr = er * sr - ei * si, ni = er * si + ei * sr;
          er = nr;
          ei = ni;
        }
        // even: 2 (gr + i gi); odd: 2 i (gr + i gi) = 2 (-gi + i gr)
        Gr[k * P1 + m] = even ? 2.0 * gr : -2.0 * gi;
        Gi[k * P1 + m] = even ? 2.0 * gi : 2.0 * gr;
      }
      __syncthreads();
      // L'_n^m += sum_{k: m < F_k'} 2^-n lambda_k^n (-s)^(n-m) i^m G_k(m)   (s = 1 up, -1 down)
      for (int q = t; q < P1 * P1; q += kPwThreads) {
        const int n = q / P1, m = q % P1;
        if (m > n) continue;
        double sr = 0.0, si = 0.0;
        for (int k = 0; k < s; ++k) {
          if (m >= a.fourier[k]) continue;
          const double c = a.ipow[k * P1 + n];
          sr = fma(c, Gr[k * P1 + m], sr);
          si = fma(c, Gi[k * P1 + m], si);
        }
        if (ud == 0 && ((n - m) & 1)) {
          sr = -sr;
          si = -si;
        }
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
        Lr[q] += rr;
        Li[q] += ri;
      }
    }
    __syncthreads();
    for (int k = t; k < nc; k += kPwThreads) {
      const int n = degree_of(k), m = k - n * n - n;
      ys[k] = m >= 0 ? Lr[n * P1 + m] : Li[n * P1 - m];
    }
    __syncthreads();
    if (t < nc) {
      const int n = degree_of(t), w = 2 * n + 1, r = t - n * n;
      const double *blk = a.bwd + size_t(f) * a.rb + n * (2 * n - 1) * (2 * n + 1) / 3;
      double v = 0.0;
      for (int c = 0; c < w; ++c) v = fma(blk[r + c * w], ys[n * n + c], v);
      acc += v;
    }
  }
  if (t < nc) L[size_t(box) * nc + t] += acc;
}

}  // namespace

size_t pw_local_lds(int p, int H) {
  const int P1 = p + 1;
  return sizeof(double) * (2 * size_t(H) + 2 * size_t(kPwMaxNodes) * P1 + 2 * size_t(P1) * P1 + size_t(P1) * P1);
}

void pw_outgoing(hipStream_t s, const PwArgs &a, uint32_t nboxes, uint32_t box0, const double *M, double *waves) {
  if (nboxes == 0) return;
  hipLaunchKernelGGL(k_pw_outgoing, dim3(xcd_grid(nboxes)), dim3(kPwThreads), 0, s, a, nboxes, box0, M, waves);
}

void pw_merge(hipStream_t s, const PwArgs &a, const PwMerge &g, uint32_t nparents, const uint32_t *slots,
              const uint32_t *children, uint32_t box0, const double *waves, double *inc) {
  if (nparents == 0) return;
  const int threads = a.H >= 1024 ? 1024 : (a.H + 63) / 64 * 64;
  hipLaunchKernelGGL(k_pw_merge, dim3(xcd_grid(nparents * 6)), dim3(threads), 0, s, a, g, nparents, slots, children,
                     box0, waves, inc);
}

hipError_t pw_allow_lds(size_t bytes) {
  return hipFuncSetAttribute(reinterpret_cast<const void *>(&k_pw_local), hipFuncAttributeMaxDynamicSharedMemorySize,
                             int(bytes));
}

void pw_local(hipStream_t s, const PwArgs &a, uint32_t ntargets, const uint32_t *target, uint32_t box0,
              const double *inc, double *L) {
  if (ntargets == 0) return;
  hipLaunchKernelGGL(k_pw_local, dim3(xcd_grid(ntargets)), dim3(kPwThreads), pw_local_lds(a.p, a.H), s, a, ntargets,
                     target, box0, inc, L);
}

}  // namespace dev
}  // namespace fmmh
