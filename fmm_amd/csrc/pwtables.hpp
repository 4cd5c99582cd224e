// pwtables.hpp -- the plane-wave M2L of the Laplace FMM (pwtables.cpp): the quadrature,
// the three frames and a host restatement of one V pair's translation.
//
// In box widths u (half the scaled coordinates of a level), for 1 <= z <= 4 and
// x^2 + y^2 <= 32,
//   1/|u| ~ sum_k (w_k / N_k) sum_{j < N_k} e^{-lambda_k z + i lambda_k (x cos a_j + y sin a_j)},
// a_j = 2 pi j / N_k (docs/CONVENTIONS.md:1325-1345).  The nodes lambda_k and weights w_k are
// built here (not FMM3D's tables): a Gauss-Legendre rule on [0, Lambda] compressed by
// column-pivoted QR to the fewest nodes whose least-squares weights reproduce 1/|u| on the
// region to the tolerance; N_k (even) is the count at which the trapezoid rule in a resolves
// J_q(lambda_k rho) for every order pair of degree p and every lateral centre offset of a V
// pair.
#pragma once

#include <cstdint>
#include <vector>

namespace fmmh {

struct PlaneWave {
  int p = 0;
  double eps = 0.0;
  int nodes = 0;  // s
  int H = 0;      // half the alpha nodes: sum_k N_k / 2 (the waves of real charges)
  std::vector<double> lambda, weight;
  std::vector<int> nalpha, first;  // N_k (even); node k's first half node
  // per half node (k, j < N_k / 2): lambda_k, w_k / N_k, cos a_j, sin a_j, node k
  std::vector<double> hl, hw, hc, hs;
  std::vector<int> hk;
  // frames 0 (z), 1 (R_x(pi/2): (x, y, z) -> (x, -z, y)), 2 (R_y(-pi/2): (x, y, z) -> (-z, y, x)):
  // per frame the multipole rotation into it and the local rotation back out of it,
  // rotation_block_size(p) values each (column-major blocks per degree)
  std::vector<double> fwd, bwd;
};

PlaneWave build_plane_wave(int p, double eps);

// The direction of a V pair with offset d (target cell - source cell): the first axis of
// z, y, x along which |d_i| >= 2, its frame f (0, 1, 2), up (+1: target on the positive side)
// or down (-1), and the offset in the frame's coordinates (box widths).
void plane_wave_direction(const int d[3], int &frame, int &up, double dp[3]);

// Host restatement of one V pair's plane-wave M2L (scaled storage of the level):
// out += M2L_pw(d) in.  The device kernels follow the same steps.
void plane_wave_m2l_host(const PlaneWave &pw, const int d[3], const double *in, double *out);

}  // namespace fmmh
