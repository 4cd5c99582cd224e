// rotation.hip -- rotation-based M2L (fmm-kernels/src/rotation.rs:441,729; CONVENTIONS
// §3.11 "Rotation of coefficients", "Coaxial translations").
//
// For a V pair (target t, source s, offset d) the local of t gains G(Q^T) A(|d|) F(Q)
// M_s: F rotates the multipole so that d lies along +z (one (2n+1)^2 block per degree),
// A is the coaxial M2L of the distance (one (p+1-i)^2 block per order, the same for
// slots +i and -i), G rotates the local back.  7,319 multiply-adds per pair at p = 12
// against 28,561 for the dense table, and 23 KB of tables per offset instead of 228 KB.
//
// Pairs are grouped by offset: a workgroup takes up to PB pairs of one offset, loads that
// offset's F and G blocks and its distance's A blocks into LDS once, stages the PB source
// multipoles, and applies the three steps to all of them, each thread forming one output
// slot k = n^2 + n + m for 8 pairs at a time (every table entry read from LDS serves 8
// pairs; the pairs' coefficients of one degree are read by all of that degree's threads
// at once).  Each pair's result goes to its own scratch row; a second pass adds a
// target's rows into its f64 local in the order of its V list.  T = double (plain path)
// or float (mixed: multipoles rounded to f32 once per evaluation, f32 steps and rows).  Tables are
// column-major per block (entry (row r, column c) of a degree-n block at r + c (2n+1)).
#include <hip/hip_runtime.h>

#include <cstdint>

#include "kernels.hpp"

namespace fmmh {
namespace dev {

constexpr int kRotThreads = 256;
constexpr int kRotJ = 8;  // pairs per thread item

// LDS of one workgroup: F and G blocks (rb each), A blocks (cb), two staging buffers of
// PB rows of nc coefficients (row stride nc + 1)
template <typename T>
static size_t rot_lds(int p, int rb, int cb, int pb) {
  const int nc = (p + 1) * (p + 1);
  return sizeof(T) * (size_t(2) * rb + cb + size_t(2) * pb * (nc + 1));
}

template <typename T>
__global__ __launch_bounds__(kRotThreads) void k_m2l_rot_pairs(int p, int pb, uint32_t nblocks, uint32_t q_base,
                                                               const uint32_t *__restrict__ blk_first,
                                                               const uint16_t *__restrict__ blk_offset,
                                                               const uint32_t *__restrict__ pair_src,
                                                               const uint16_t *__restrict__ dist_of,
                                                               const T *__restrict__ fwd, const T *__restrict__ bwd,
                                                               const T *__restrict__ coax, int rb, int cb,
                                                               const T *__restrict__ M, T *__restrict__ S) {
  extern __shared__ __align__(16) unsigned char smem[];
  const uint32_t blk = xcd_index(nblocks);
  if (blk >= nblocks) return;  // uniform over the workgroup
  const int nc = (p + 1) * (p + 1), ld = nc + 1;
  T *tF = reinterpret_cast<T *>(smem);
  T *tG = tF + rb;
  T *tA = tG + rb;
  T *X = tA + cb;
  T *Y = X + size_t(pb) * ld;
  const int tid = threadIdx.x;
  const uint32_t t = blk_offset[blk];
  const uint32_t q0 = blk_first[blk], np = blk_first[blk + 1] - q0;  // np <= pb
  const uint32_t dist = dist_of[t];
  // the block's source boxes first, so the multipole loads below carry no dependent
  // index load and can all be in flight together
  __shared__ uint32_t src[64];
  if (tid < pb) src[tid] = uint32_t(tid) < np ? pair_src[q0 + tid] : 0u;
#pragma unroll 4
  for (int i = tid; i < rb; i += kRotThreads) {
    tF[i] = fwd[size_t(t) * rb + i];
    tG[i] = bwd[size_t(t) * rb + i];
  }
#pragma unroll 4
  for (int i = tid; i < cb; i += kRotThreads) tA[i] = coax[size_t(dist) * cb + i];
  __syncthreads();
#pragma unroll 4
  for (int i = tid; i < pb * nc; i += kRotThreads) {
    const int j = i / nc, k = i - j * nc;
    X[j * ld + k] = uint32_t(j) < np ? M[size_t(src[j]) * nc + k] : T(0);
  }
  __syncthreads();
  const int items = nc * (pb / kRotJ);
  // 1. rotate: Y_j = F X_j, per degree block
  for (int it = tid; it < items; it += kRotThreads) {
    const int jg = it / nc, k = it - jg * nc;
    int n = 0;
    while ((n + 1) * (n + 1) <= k) ++n;
    const int r = k - n * n, w = 2 * n + 1, bo = n * (2 * n - 1) * (2 * n + 1) / 3;
    const T *x = X + size_t(jg) * kRotJ * ld + n * n;
    T y[kRotJ];
#pragma unroll
    for (int jj = 0; jj < kRotJ; ++jj) y[jj] = T(0);
    for (int c = 0; c < w; ++c) {
      const T f = tF[bo + r + c * w];
#pragma unroll
      for (int jj = 0; jj < kRotJ; ++jj) y[jj] = fma(f, x[jj * ld + c], y[jj]);
    }
    T *yo = Y + size_t(jg) * kRotJ * ld + k;
#pragma unroll
    for (int jj = 0; jj < kRotJ; ++jj) yo[jj * ld] = y[jj];
  }
  __syncthreads();
  // 2. coaxial M2L, order i = |m|: X_j(n, m) = sum_q A_i[n - i][q - i] Y_j(q, m)
  for (int it = tid; it < items; it += kRotThreads) {
    const int jg = it / nc, k = it - jg * nc;
    int n = 0;
    while ((n + 1) * (n + 1) <= k) ++n;
    const int m = k - n * n - n, i = m < 0 ? -m : m, cw = p + 1 - i;
    int co = 0;  // sum of (p+1-i')^2 over i' < i
    for (int q = 0; q < i; ++q) co += (p + 1 - q) * (p + 1 - q);
    const T *a = tA + co + (n - i);
    const T *yv = Y + size_t(jg) * kRotJ * ld;
    T z[kRotJ];
#pragma unroll
    for (int jj = 0; jj < kRotJ; ++jj) z[jj] = T(0);
    for (int q = i; q <= p; ++q) {
      const T av = a[(q - i) * cw];
      const int kq = q * q + q + m;
#pragma unroll
      for (int jj = 0; jj < kRotJ; ++jj) z[jj] = fma(av, yv[jj * ld + kq], z[jj]);
    }
    T *xo = X + size_t(jg) * kRotJ * ld + k;
#pragma unroll
    for (int jj = 0; jj < kRotJ; ++jj) xo[jj * ld] = z[jj];
  }
  __syncthreads();
  // 3. rotate back: row_j = G X_j, into the pairs' scratch rows
  for (int it = tid; it < items; it += kRotThreads) {
    const int jg = it / nc, k = it - jg * nc;
    int n = 0;
    while ((n + 1) * (n + 1) <= k) ++n;
    const int r = k - n * n, w = 2 * n + 1, bo = n * (2 * n - 1) * (2 * n + 1) / 3;
    const T *x = X + size_t(jg) * kRotJ * ld + n * n;
    T y[kRotJ];
#pragma unroll
    for (int jj = 0; jj < kRotJ; ++jj) y[jj] = T(0);
    for (int c = 0; c < w; ++c) {
      const T g = tG[bo + r + c * w];
#pragma unroll
      for (int jj = 0; jj < kRotJ; ++jj) y[jj] = fma(g, x[jj * ld + c], y[jj]);
    }
#pragma unroll
    for (int jj = 0; jj < kRotJ; ++jj) {
      const uint32_t j = uint32_t(jg * kRotJ + jj);
      if (j < np) S[size_t(q0 - q_base + j) * nc + k] = y[jj];
    }
  }
}

// 32 pairs per block, or the most that fit (the dynamic part; the 256 bytes of source
// indices come on top)
constexpr size_t kRotLdsMax = 160 * 1024 - 1024;
template <typename T>
int rotation_pairs_per_block(int p, int rb, int cb) {
  int pb = 32;
  while (pb > kRotJ && rot_lds<T>(p, rb, cb, pb) > kRotLdsMax) pb /= 2;
  return pb;
}

template <typename T>
hipError_t rotation_allow_lds(int p, int rb, int cb) {
  const int pb = rotation_pairs_per_block<T>(p, rb, cb);
  const size_t bytes = rot_lds<T>(p, rb, cb, pb);
  if (bytes > kRotLdsMax) return hipErrorInvalidValue;
  return hipFuncSetAttribute(reinterpret_cast<const void *>(&k_m2l_rot_pairs<T>),
                             hipFuncAttributeMaxDynamicSharedMemorySize, int(bytes));
}

__global__ void k_to_f32(size_t n, const double *__restrict__ a, float *__restrict__ b) {
  for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
    b[i] = float(a[i]);
}

void to_f32(hipStream_t s, size_t n, const double *a, float *b) {
  if (n == 0) return;
  const size_t blocks = (n + 255) / 256;
  hipLaunchKernelGGL(k_to_f32, dim3(blocks < 65536 ? blocks : 65536), dim3(256), 0, s, n, a, b);
}

template <typename T>
void m2l_rotation(hipStream_t s, int p, uint32_t nblocks, uint32_t q_base, const uint32_t *blk_first,
                  const uint16_t *blk_offset, const uint32_t *pair_src, const uint16_t *dist_of, const T *fwd,
                  const T *bwd, const T *coax, int rb, int cb, const T *M, T *S) {
  if (nblocks == 0) return;
  const int pb = rotation_pairs_per_block<T>(p, rb, cb);
  hipLaunchKernelGGL((k_m2l_rot_pairs<T>), dim3(xcd_grid(nblocks)), dim3(kRotThreads), rot_lds<T>(p, rb, cb, pb), s,
                     p, pb, nblocks, q_base, blk_first, blk_offset, pair_src, dist_of, fwd, bwd, coax, rb, cb, M, S);
}

// ---- Helmholtz: the same three steps on complex expansions ------------------------------------
//
// A pair's multipole is two real planes (re, im); the rotation blocks are real, so steps
// 1 and 3 treat the two planes as two rows, and the back rotation is the transpose of the
// forward one (orthogonal blocks: only F is staged).  The coaxial blocks are complex (real
// plane, then imaginary plane) and per (level, distance): step 2 forms
// Z = (Ar Yr - Ai Yi) + i (Ar Yi + Ai Yr).  Row 2 j of the staging buffers is pair j's
// real plane, row 2 j + 1 its imaginary plane; the pair's result goes to scratch rows
// 2 q (real) and 2 q + 1 (imaginary) for the ordered reduction.
constexpr int kHRotJ = 4;  // complex pairs per thread item of step 2

static size_t h_rot_lds(int p, int rb, int cb, int pb) {
  const int nc = (p + 1) * (p + 1);
  return sizeof(double) * (size_t(rb) + 2 * size_t(cb) + size_t(2) * 2 * pb * (nc + 1));
}

__global__ __launch_bounds__(kRotThreads) void k_h_m2l_rot_pairs(int p, int pb, uint32_t nblocks, uint32_t q_base,
                                                                 const uint32_t *__restrict__ blk_first,
                                                                 const uint16_t *__restrict__ blk_offset,
                                                                 const uint32_t *__restrict__ blk_coax,
                                                                 const uint32_t *__restrict__ pair_src,
                                                                 const double *__restrict__ fwd,
                                                                 const double *__restrict__ coax, int rb, int cb,
                                                                 const double *__restrict__ M,
                                                                 double *__restrict__ S) {
  extern __shared__ __align__(16) unsigned char smem[];
  const uint32_t blk = xcd_index(nblocks);
  if (blk >= nblocks) return;  // uniform over the workgroup
  const int nc = (p + 1) * (p + 1), ld = nc + 1, rows = 2 * pb;
  double *tF = reinterpret_cast<double *>(smem);
  double *tAr = tF + rb;
  double *tAi = tAr + cb;
  double *X = tAi + cb;
  double *Y = X + size_t(rows) * ld;
  const int tid = threadIdx.x;
  const uint32_t t = blk_offset[blk], a = blk_coax[blk];
  const uint32_t q0 = blk_first[blk], np = blk_first[blk + 1] - q0;  // np <= pb
  __shared__ uint32_t src[64];
  if (tid < pb) src[tid] = uint32_t(tid) < np ? pair_src[q0 + tid] : 0u;
#pragma unroll 4
  for (int i = tid; i < rb; i += kRotThreads) tF[i] = fwd[size_t(t) * rb + i];
#pragma unroll 4
  for (int i = tid; i < cb; i += kRotThreads) {
    tAr[i] = coax[2 * size_t(a) * cb + i];
    tAi[i] = coax[(2 * size_t(a) + 1) * cb + i];
  }
  __syncthreads();
  // the pairs' source planes: box s holds planes 3 s (re) and 3 s + 1 (im)
#pragma unroll 4
  for (int i = tid; i < rows * nc; i += kRotThreads) {
    const int r = i / nc, k = i - r * nc, j = r >> 1;
    X[r * ld + k] = uint32_t(j) < np ? M[(3 * size_t(src[j]) + (r & 1)) * nc + k] : 0.0;
  }
  __syncthreads();
  const int items = nc * (rows / kRotJ);
  // 1. rotate every row: Y_r = F X_r, per degree block
  for (int it = tid; it < items; it += kRotThreads) {
    const int jg = it / nc, k = it - jg * nc;
    int n = 0;
    while ((n + 1) * (n + 1) <= k) ++n;
    const int r = k - n * n, w = 2 * n + 1, bo = n * (2 * n - 1) * (2 * n + 1) / 3;
    const double *x = X + size_t(jg) * kRotJ * ld + n * n;
    double y[kRotJ];
#pragma unroll
    for (int jj = 0; jj < kRotJ; ++jj) y[jj] = 0.0;
    for (int c = 0; c < w; ++c) {
      const double f = tF[bo + r + c * w];
#pragma unroll
      for (int jj = 0; jj < kRotJ; ++jj) y[jj] = fma(f, x[jj * ld + c], y[jj]);
    }
    double *yo = Y + size_t(jg) * kRotJ * ld + k;
#pragma unroll
    for (int jj = 0; jj < kRotJ; ++jj) yo[jj * ld] = y[jj];
  }
  __syncthreads();
  // 2. complex coaxial M2L, order i = |m|: Z(n, m) = sum_q A_i[n - i][q - i] Y(q, m)
  const int citems = nc * (pb / kHRotJ);
  for (int it = tid; it < citems; it += kRotThreads) {
    const int jg = it / nc, k = it - jg * nc;
    int n = 0;
    while ((n + 1) * (n + 1) <= k) ++n;
    const int m = k - n * n - n, i = m < 0 ? -m : m, cw = p + 1 - i;
    int co = 0;  // sum of (p+1-i')^2 over i' < i
    for (int q = 0; q < i; ++q) co += (p + 1 - q) * (p + 1 - q);
    const double *ar = tAr + co + (n - i), *ai = tAi + co + (n - i);
    const double *yv = Y + size_t(jg) * 2 * kHRotJ * ld;
    double zr[kHRotJ], zi[kHRotJ];
#pragma unroll
    for (int jj = 0; jj < kHRotJ; ++jj) zr[jj] = zi[jj] = 0.0;
    for (int q = i; q <= p; ++q) {
      const double vr = ar[(q - i) * cw], vi = ai[(q - i) * cw];
      const int kq = q * q + q + m;
#pragma unroll
      for (int jj = 0; jj < kHRotJ; ++jj) {
        const double yr = yv[(2 * jj) * ld + kq], yi = yv[(2 * jj + 1) * ld + kq];
        zr[jj] = fma(vr, yr, fma(-vi, yi, zr[jj]));
        zi[jj] = fma(vr, yi, fma(vi, yr, zi[jj]));
      }
    }
    double *xo = X + size_t(jg) * 2 * kHRotJ * ld + k;
#pragma unroll
    for (int jj = 0; jj < kHRotJ; ++jj) {
      xo[(2 * jj) * ld] = zr[jj];
      xo[(2 * jj + 1) * ld] = zi[jj];
    }
  }
  __syncthreads();
  // 3. rotate back with the transpose: row_r = F^T X_r, into the pairs' scratch rows
  for (int it = tid; it < items; it += kRotThreads) {
    const int jg = it / nc, k = it - jg * nc;
    int n = 0;
    while ((n + 1) * (n + 1) <= k) ++n;
    const int r = k - n * n, w = 2 * n + 1, bo = n * (2 * n - 1) * (2 * n + 1) / 3;
    const double *x = X + size_t(jg) * kRotJ * ld + n * n;
    double y[kRotJ];
#pragma unroll
    for (int jj = 0; jj < kRotJ; ++jj) y[jj] = 0.0;
    for (int c = 0; c < w; ++c) {
      const double f = tF[bo + c + r * w];
#pragma unroll
      for (int jj = 0; jj < kRotJ; ++jj) y[jj] = fma(f, x[jj * ld + c], y[jj]);
    }
#pragma unroll
    for (int jj = 0; jj < kRotJ; ++jj) {
      const uint32_t row = uint32_t(jg * kRotJ + jj);
      if ((row >> 1) < np) S[(2 * size_t(q0 - q_base) + row) * nc + k] = y[jj];
    }
  }
}

// pairs per block: 32, or the most (a multiple of kHRotJ) whose tables and staging fit
int h_rotation_pairs_per_block(int p, int rb, int cb) {
  int pb = 32;
  while (pb > kHRotJ && h_rot_lds(p, rb, cb, pb) > kRotLdsMax) pb /= 2;
  return pb;
}

hipError_t h_rotation_allow_lds(int p, int rb, int cb) {
  const size_t bytes = h_rot_lds(p, rb, cb, h_rotation_pairs_per_block(p, rb, cb));
  if (bytes > kRotLdsMax) return hipErrorInvalidValue;
  return hipFuncSetAttribute(reinterpret_cast<const void *>(&k_h_m2l_rot_pairs),
                             hipFuncAttributeMaxDynamicSharedMemorySize, int(bytes));
}

void h_m2l_rotation(hipStream_t s, int p, uint32_t nblocks, uint32_t q_base, const uint32_t *blk_first,
                    const uint16_t *blk_offset, const uint32_t *blk_coax, const uint32_t *pair_src, const double *fwd,
                    const double *coax, int rb, int cb, const double *M, double *S) {
  if (nblocks == 0) return;
  const int pb = h_rotation_pairs_per_block(p, rb, cb);
  hipLaunchKernelGGL(k_h_m2l_rot_pairs, dim3(xcd_grid(nblocks)), dim3(kRotThreads), h_rot_lds(p, rb, cb, pb), s, p,
                     pb, nblocks, q_base, blk_first, blk_offset, blk_coax, pair_src, fwd, coax, rb, cb, M, S);
}

template int rotation_pairs_per_block<double>(int, int, int);
template int rotation_pairs_per_block<float>(int, int, int);
template hipError_t rotation_allow_lds<double>(int, int, int);
template hipError_t rotation_allow_lds<float>(int, int, int);
template void m2l_rotation<double>(hipStream_t, int, uint32_t, uint32_t, const uint32_t *, const uint16_t *,
                                   const uint32_t *, const uint16_t *, const double *, const double *, const double *,
                                   int, int, const double *, double *);
template void m2l_rotation<float>(hipStream_t, int, uint32_t, uint32_t, const uint32_t *, const uint16_t *,
                                  const uint32_t *, const uint16_t *, const float *, const float *, const float *, int,
                                  int, const float *, float *);

}  // namespace dev
}  // namespace fmmh
