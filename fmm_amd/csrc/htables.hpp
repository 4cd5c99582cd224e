// htables.hpp -- Helmholtz translation tables of one level (htables.cpp).
//
// A table is a dense complex (p+1)^2 x (p+1)^2 matrix from stored source coefficients to
// stored target coefficients, as two column-major real planes (real part, then imaginary
// part; entry (i, k) at i + k (p+1)^2).  kh is a level's wavenumber in its own units
// (k r_l); the stored form and the level units are those of the Helmholtz kernels.
#pragma once

#include <complex>
#include <cstdint>
#include <vector>

namespace fmmh {

using cplx = std::complex<double>;

// Orthonormal real solid harmonics S_n^m(x) = |x|^n Y_n^m(x^) in slots n^2 + n + m, and the
// factor norm(n, m) that takes the Racah harmonic's real (m >= 0) or imaginary (m < 0)
// part to them.
void h_norm_factors(int p, std::vector<double> &norm);

// M2M from the child of octant o into its parent (kh: the parent's level), L2L from the
// parent into that child, 8 tables each (octant order); M2L of the listed offset indices
// (d = target cell - source cell).  Each table 2 (p+1)^4 values, appended to out.
void h_m2m_tables(int p, cplx kh, std::vector<double> &out);
void h_l2l_tables(int p, cplx kh, std::vector<double> &out);
void h_m2l_tables(int p, cplx kh, const std::vector<uint16_t> &offsets, std::vector<double> &out);

// The M2L as rotation, coaxial translation and rotation back, in pieces for the device's
// per-pair path (the reference's RCR, fmm-exec/src/helmholtz/operator.rs:448-461):
//  - h_rcr_rotations: per offset index 0..315 the degree blocks of O(Q) (Q takes the
//    offset's direction to +z), block n column-major ((r, c) at r + c (2n+1)) from
//    n (2n-1)(2n+1)/3 on, h_rcr_rotation_size(p) values per offset; the rotation back is
//    the transpose (O is orthogonal);
//  - h_rcr_coaxial: per distance |t| (level units, t = 2 d) the coaxial blocks of order
//    m = 0..p, block m ((p+1-m)^2, entry (n' - m) + (n - m)(p+1-m): target degree n',
//    source degree n) from sum_{m'<m} (p+1-m')^2 on, scaled as the stored forms
//    (4 pi i kh e^{2 beta} s^(n+n')), real plane then imaginary plane:
//    2 h_rcr_coaxial_size(p) values per distance.
// O^T C O equals the h_m2l_tables table of the offset to rounding.
int h_rcr_rotation_size(int p);
int h_rcr_coaxial_size(int p);
void h_rcr_rotations(int p, std::vector<double> &out);
void h_rcr_coaxial(int p, cplx kh, const std::vector<double> &dist, std::vector<double> &out);

// The degree blocks O^n(Q) of a rotation Q (Y_n(Q x) = O^n Y_n(x)) of the orthonormal real
// harmonics, row-major ((r, c) at r (2n+1) + c) from n (2n-1)(2n+1)/3 on: a coefficient
// vector goes into the frame x' = Q x by O^n and back by its transpose.
void h_rotation_blocks(int p, const double Q[9], std::vector<double> &out);

// Level constants in a level's units: s = min(|kh| sqrt 3, 1), beta = Im(kh) sqrt 3.
inline double h_scale(cplx kh) {
  const double v = std::abs(kh) * 1.7320508075688772;
  return v < 1.0 ? v : 1.0;
}
inline double h_beta(cplx kh) { return kh.imag() * 1.7320508075688772; }

}  // namespace fmmh
