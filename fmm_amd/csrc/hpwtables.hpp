// hpwtables.hpp -- the exponential (plane-wave) M2L of the Helmholtz FMM on one level
// (hpwtables.cpp): FMM3D's quadrature of kappa = k w at precision iprec 2 with its alpha and
// Fourier counts, the level's conversion tables, and host restatements of one V pair's
// translation and of one parent's merge and shift (docs/CONVENTIONS.md §3.15 "Exponential
// M2L"; fmm-tables/src/helmholtz/plane_wave/{quadrature,pair,merge}.rs in the reference).
//
// For a target above a source by u (box widths), 1 <= u_z <= 4, |u_x|, |u_y| <= 4, and
// kappa = k w (Re, Im >= 0),
//   e^{i kappa |u|} / |u| ~ sum_k (w_k / N_k) sum_{j < N_k} e^{-lambda_k u_z + i t_k (u_x cos a_j + u_y sin a_j)},
// a_j = 2 pi j / N_k, t_k = sqrt(lambda_k^2 + kappa^2), complex nodes lambda_k and weights
// w_k mapped from FMM3D's real table of kappa's domain; a source keeps its orders |m| <= F_k.
// Every wave value is complex (no half-node symmetry): W = sum_k N_k values per direction,
// stored as W real parts, then W imaginary parts.
#pragma once

#include <complex>
#include <cstdint>
#include <vector>

namespace fmmh {

using cplx = std::complex<double>;

// One level's quadrature (quadrature.rs Quadrature): kappa's domain (1-based, 0 outside the
// tables), FMM3D's stated error, and per node lambda_k, t_k, the weight (of d lambda), the
// alpha count N_k and the Fourier count F_k.
struct HpwQuadrature {
  cplx kappa;
  int domain = 0;
  double stated_error = 0.0;
  std::vector<cplx> lambda, t, weight;
  std::vector<int> alpha, fourier;
  int waves() const;
};

// The domain of kappa at iprec 2 (hwts3dgetd; 0 outside: Re kappa >= 16 pi, Im kappa >= 12 pi
// or a negative part).
int hpw_domain(cplx kappa);
// The quadrature of kappa (domain 0 and no nodes outside the tables).
HpwQuadrature hpw_quadrature(cplx kappa);
// |J_0(z)|, .., |J_top(z)| of complex z (quadrature.rs bessel_j_norms).
std::vector<double> hpw_bessel_j_norms(cplx z, int top);

// The reference's rule for a level without a tolerance (merge.rs precision_for_degree and
// select): the exponential M2L at iprec 2 on levels >= 2 of degree <= 15 whose
// kappa = 2 kh lies in the tables.
bool hpw_level_selected(int level, int p, cplx kh);

// One level's tables (pair.rs PlaneWaveLevel), kh = k r_l the level's wavenumber in its
// units (kappa = 2 kh): per node its first wave value (start), its values N_k (len), its
// orders min(F_k, p) + 1 (orders), its first order among all (mode_start), its first cos/sin
// row (trig: orders rows of len) and its first factor (factor: per order m the degrees
// n = m..p); cos and sin of m a_j; the outgoing factors 2 pi e^beta (w_k / N_k) P_k(n, m) and
// the incoming 4 pi e^beta (-1)^n P_k(n, m), P_k(n, m) = i^-n c_nm Pbar_n^m((s/kappa) i
// lambda_k, (s/kappa) t_k; s^2); the frame rotations O^n(Q) of frames 1 and 2 (row-major per
// degree).
struct HpwLevel {
  int p = 0;
  cplx kh;
  int s = 0, W = 0, orders = 0, trig_size = 0, factor_size = 0;
  std::vector<cplx> lambda, t;
  std::vector<int> start, len, ord, mode_start, trig, factor;
  std::vector<double> acos, asin;            // [W]: cos a_j, sin a_j
  std::vector<double> cosm, sinm;            // [trig_size]
  std::vector<double> out_re, out_im, in_re, in_im;  // [factor_size]
  std::vector<double> frame1, frame2;        // rotation blocks, sum_n (2n+1)^2 each
};
HpwLevel hpw_level(int p, cplx kh, const HpwQuadrature &q);

// The shift factors of the up (up = true) or down waves by d (box widths, in the
// direction's frame): e^{-+lambda_k d_z + i t_k (d_x cos a_j + d_y sin a_j)}, 2W reals.
void hpw_shift(const HpwLevel &L, const double d[3], bool up, double *out);
// The up and down waves (2W reals each) of a stored multipole (two planes) in frame f.
void hpw_waves(const HpwLevel &L, int frame, const double *mult, double *up, double *down);
// local (two planes, in the frame) += the local of incoming waves V (2W reals), up or down.
void hpw_local_add(const HpwLevel &L, const double *V, bool up, double *local);
// out (two planes) += the frame local rotated back out of frame f.
void hpw_rotate_back_add(const HpwLevel &L, int frame, const double *local, double *out);

// Host restatement of one V pair's exponential M2L (pair.rs m2l) of offset d (target cell
// - source cell) along direction dir (-1: the first axis of z, y, x with |d_i| >= 2):
// out (two planes) += M2L(mult), stored forms of the level.
void hpw_m2l_host(const HpwLevel &L, const int d[3], int dir, const double *mult, double *out);
// Host restatement of one parent's merge and shift: slots[a] the stored multipole (two
// planes) of slot a or nullptr; out[c] (two planes each) += the M2L of every pair slot ->
// child c along the merge's directions.
void hpw_merge_host(const HpwLevel &L, const double *const *slots, double *out);

}  // namespace fmmh
