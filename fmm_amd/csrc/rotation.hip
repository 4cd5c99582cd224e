// rotation.hip -- rotation-based M2L (fmm-kernels/src/rotation.rs:441,729; CONVENTIONS
// §3.11 "Rotation of coefficients", "Coaxial translations").
//
// For a V pair (target t, source s, offset d) the local of t gains G(Q^T) A(|d|) F(Q)
// M_s: F rotates the multipole so that d lies along +z (one (2n+1)^2 block per degree),
// A is the coaxial M2L of the distance (one (p+1-i)^2 block per order, the same for
// slots +i and -i), G rotates the local back.  7,319 multiply-adds per pair at p = 12
// against 28,561 for the dense table, and 23 KB of tables per offset instead of 228 KB.
//
// One workgroup per target box with V entries (all levels in one launch); thread k owns
// coefficient slot k = n^2 + n + m through the three steps of every entry, in the row's
// order (offsets ascending), and keeps the target's sum in a register; the step vectors
// live in LDS between barriers.  T = double (plain path) or float (mixed path: the
// multipole is rounded to f32, the steps and the row sum are f32, the row sum is added
// into the f64 local).  Tables are column-major so that the threads of one degree read
// consecutive entries.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "kernels.hpp"

namespace fmmh {
namespace dev {

template <typename T>
__global__ __launch_bounds__(256) void k_m2l_rotation(int p, uint32_t rows, const uint32_t *__restrict__ row_box,
                                                      const uint32_t *__restrict__ row_off,
                                                      const uint32_t *__restrict__ ent_src,
                                                      const uint16_t *__restrict__ ent_offset,
                                                      const uint16_t *__restrict__ dist_of, const T *__restrict__ fwd,
                                                      const T *__restrict__ bwd, const T *__restrict__ coax, int rb,
                                                      int cb, const double *__restrict__ M, double *__restrict__ L) {
  __shared__ T x[kMaxCoef], y[kMaxCoef];
  const uint32_t row = xcd_index(rows);
  if (row >= rows) return;
  const int nc = (p + 1) * (p + 1);
  const int k = threadIdx.x;
  const bool active = k < nc;
  // slot k = n^2 + n + m
  int n = 0;
  while ((n + 1) * (n + 1) <= k) ++n;
  const int r = k - n * n, m = r - n, w = 2 * n + 1;
  const int bo = n * (2 * n - 1) * (2 * n + 1) / 3;  // sum of (2n'+1)^2 over n' < n
  const int i = m < 0 ? -m : m;
  int co = 0;  // sum of (p+1-i')^2 over i' < i
  for (int q = 0; q < i; ++q) co += (p + 1 - q) * (p + 1 - q);
  const int cw = p + 1 - i;
  const uint32_t box = row_box[row];
  T acc = T(0);
  for (uint32_t e = row_off[row]; e < row_off[row + 1]; ++e) {
    const uint32_t s = ent_src[e];
    const uint32_t t = ent_offset[e];
    __syncthreads();
    if (active) x[k] = T(M[size_t(s) * nc + k]);
    __syncthreads();
    // 1. rotate the multipole: y_n = F_n x_n
    if (active) {
      const T *f = fwd + size_t(t) * rb + bo + r;
      T v = T(0);
      for (int c = 0; c < w; ++c) v = fma(f[c * w], x[n * n + c], v);
      y[k] = v;
    }
    __syncthreads();
    // 2. coaxial M2L, order i: x_(j,m) = sum_n A_i[j][n] y_(n,m)
    if (active) {
      const T *a = coax + size_t(dist_of[t]) * cb + co + (n - i);
      T v = T(0);
      for (int q = i; q <= p; ++q) v = fma(a[(q - i) * cw], y[q * q + q + m], v);
      x[k] = v;
    }
    __syncthreads();
    // 3. rotate the local back: acc += G_n x_n
    if (active) {
      const T *g = bwd + size_t(t) * rb + bo + r;
      T v = T(0);
      for (int c = 0; c < w; ++c) v = fma(g[c * w], x[n * n + c], v);
      acc += v;
    }
  }
  if (active) L[size_t(box) * nc + k] += double(acc);
}

template <typename T>
void m2l_rotation(hipStream_t s, int p, uint32_t rows, const uint32_t *row_box, const uint32_t *row_off,
                  const uint32_t *ent_src, const uint16_t *ent_offset, const uint16_t *dist_of, const T *fwd,
                  const T *bwd, const T *coax, int rb, int cb, const double *M, double *L) {
  if (rows == 0) return;
  hipLaunchKernelGGL((k_m2l_rotation<T>), dim3(xcd_grid(rows)), dim3(256), 0, s, p, rows, row_box, row_off, ent_src,
                     ent_offset, dist_of, fwd, bwd, coax, rb, cb, M, L);
}

template void m2l_rotation<double>(hipStream_t, int, uint32_t, const uint32_t *, const uint32_t *, const uint32_t *,
                                   const uint16_t *, const uint16_t *, const double *, const double *,
                                   const double *, int, int, const double *, double *);
template void m2l_rotation<float>(hipStream_t, int, uint32_t, const uint32_t *, const uint32_t *, const uint32_t *,
                                  const uint16_t *, const uint16_t *, const float *, const float *, const float *,
                                  int, int, const double *, double *);

}  // namespace dev
}  // namespace fmmh
