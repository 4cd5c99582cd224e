// hkernels.hpp -- launch wrappers of the Helmholtz kernels (helmholtz.hip).
//
// Complex expansions are planes of (p+1)^2 doubles; box b of the Helmholtz operator holds
// three planes at vector ids 3b (real part), 3b + 1 (imaginary part) and 3b + 2 (the
// negated imaginary part, written by `negate`, so that a complex table product is four
// real contractions with no sign per column).  Point values (potentials, gradients,
// charges) are interleaved complex: (re, im) per value.  Level constants hlev[4 l + 0..3] =
// (Re kh_l, Im kh_l, s_l, beta_l), kh_l = k r_l the level's wavenumber in its own units.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace fmmh {
namespace dev {

// P2M (singular = 0; row_ent the row's own leaf, unit frame) or P2L (singular = 1; row_ent
// the X-list leaves, frames from row_cell and leaf_cell) into planes 3 row_box[r] + {0, 1}.
void h_expand(hipStream_t s, int p, int singular, uint32_t rows, const uint32_t *row_box, const uint32_t *row_cell,
              const uint32_t *row_off, const uint32_t *row_ent, const uint32_t *leaf_cell, const uint32_t *src_start,
              const double *sx, const double *sy, const double *sz, const double *qr, const double *qi,
              const double *hlev, const double *norm, double *coeffs);
// L2P (singular = 0; one entry per row, the leaf's own box) or M2P (singular = 1; the W-list
// boxes) at the targets of each row's leaf; phi_hat 2 and grad_hat 6 doubles per target.
void h_evaluate(hipStream_t s, int p, int singular, int gradients, uint32_t rows, const uint32_t *row_leaf,
                const uint32_t *row_off, const uint32_t *row_box, const uint32_t *box_cell,
                const uint32_t *leaf_cell, const uint32_t *tgt_start, const double *tx, const double *ty,
                const double *tz, const double *hlev, const double *norm, const double *coeffs, double *phi_hat,
                double *grad_hat);
// P2P over the leaves leaf_first .. leaf_first + count: e^{i kh r} / r in the target leaf's
// units, computed in T (double, or float with double-float separations and per-leaf f32
// sums added in f64; T = float also reads the source records of h_source_records for the
// leaves of more than 32 targets, rec may be null for T = double).
template <typename T>
void h_p2p(hipStream_t s, uint32_t leaf_first, uint32_t count, int gradients, int damped, const uint32_t *near_off,
           const uint32_t *near_ent, const uint32_t *leaf_cell, const uint32_t *src_start, const uint32_t *tgt_start,
           const double *sx, const double *sy, const double *sz, const double *qr, const double *qi,
           const float *rec, const double *tx, const double *ty, const double *tz, const double *hlev,
           double *phi_hat, double *grad_hat);
// f32 P2P source records, 8 floats per source in leaf order: (u_hi x, y, z, u_lo x, y, z,
// q re, q im) of the own-leaf coordinates u = hi + lo; the charges are written by h_gather
void h_source_records(const double *sx, const double *sy, const double *sz, size_t n, float *host_rec);
// planes 3b + 2 = -(planes 3b + 1) for boxes b0 .. b1
void h_negate(hipStream_t s, int p, size_t b0, size_t b1, double *coeffs);
// caller-order interleaved complex charges to leaf-order planes
void h_gather(hipStream_t s, const double *q, const uint32_t *order, size_t n, double *qr, double *qi,
              float *rec = nullptr);
// phi = phi_hat / (4 pi r_t), grad = grad_hat / (4 pi r_t^2), in caller order
void h_output(hipStream_t s, uint32_t nleaves, const uint32_t *tgt_start, const double *leaf_r,
              const uint32_t *order, const double *phi_hat, const double *grad_hat, double *phi, double *grad,
              int gradients);
// Compensated f64 direct sum of q e^{ikr}/r (no 1/(4 pi)), complex interleaved charges and
// outputs; part holds nchunks * nt * 16 doubles.
int h_direct_chunks(int64_t ns, int64_t nt);
void h_direct_sum(hipStream_t s, double kre, double kim, const double *src, const double *q, int64_t ns,
                  const double *tgt, int64_t nt, int gradients, double *part, int nchunks, double *phi,
                  double *grad);

}  // namespace dev
}  // namespace fmmh
