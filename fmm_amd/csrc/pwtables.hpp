// pwtables.hpp -- the plane-wave M2L of the Laplace FMM (pwtables.cpp): the quadrature,
// the three frames, the merge-and-shift geometry with its shifts, and host restatements of
// one V pair's translation and of one parent's merge and shift.
//
// In box widths u (half the scaled coordinates of a level), for 1 <= z <= 4 and
// x^2 + y^2 <= 32,
//   1/|u| ~ sum_k (w_k / N_k) sum_{j < N_k} e^{-lambda_k z + i lambda_k (x cos a_j + y sin a_j)},
// a_j = 2 pi j / N_k (docs/CONVENTIONS.md:1325-1345).  The nodes lambda_k, weights w_k, alpha
// counts N_k and Fourier mode counts F_k are FMM3D's tables for separation 1, the ones the
// reference uses (fmm-tables/src/plane_wave/quadrature.rs:1-20); N_k is rounded up to even
// so that half the alpha nodes carry the waves of real charges, and orders m >= F_k' =
// min(F_k, p + 1) are dropped for lambda_k.
#pragma once

#include <array>
#include <cstdint>
#include <vector>

namespace fmmh {

constexpr int kPwSlots = 216;    // the children of a parent's 27 colleagues: 6 x 6 x 6 boxes
constexpr int kPwChildren = 8;
constexpr int kPwDirections = 6; // up, down, north, south, east, west (2 frame + down)
constexpr int kPwMinNodes = 2, kPwMaxTableNodes = 39;

// The number of lambda nodes of degree p's quadrature, p <= 20: the reference's measured
// rule (fmm-tables/src/plane_wave/quadrature.rs:48-58).
int plane_wave_nodes_for_degree(int p);

struct PlaneWave {
  int p = 0;
  int nodes = 0;           // s
  int H = 0;               // half the alpha nodes: sum_k N_k / 2 (the waves of real charges)
  double max_error = 0.0;  // FMM3D's stated maximum error of the quadrature (box width 1)
  std::vector<double> lambda, weight;
  std::vector<int> nalpha, fourier, first;  // N_k (even); F_k' = min(F_k, p + 1); first half node
  std::vector<int> fmm3d_alpha;             // N_k as FMM3D states it (before rounding to even)
  // per half node (k, j < N_k / 2): lambda_k, w_k / N_k, cos a_j, sin a_j, node k
  std::vector<double> hl, hw, hc, hs;
  std::vector<int> hk;
  // powers: outgoing opow[k (p+1) + n] = lambda_k^n 2^-(n+1), incoming ipow[k (p+1) + n] =
  // lambda_k^n 2^-n
  std::vector<double> opow, ipow;
  // frames 0 (z), 1 (R_x(pi/2): (x, y, z) -> (x, -z, y)), 2 (R_y(-pi/2): (x, y, z) -> (-z, y, x)):
  // per frame the multipole rotation into it and the local rotation back out of it,
  // rotation_block_size(p) values each (column-major blocks per degree)
  std::vector<double> fwd, bwd;
};

// Degree p's quadrature with `nodes` lambda nodes (0: plane_wave_nodes_for_degree(p)).
PlaneWave build_plane_wave(int p, int nodes = 0);

// The direction of a V pair with offset d (target cell - source cell): the first axis of
// z, y, x along which |d_i| >= 2, its frame f (0, 1, 2), up (+1: target on the positive side)
// or down (-1), and the offset in the frame's coordinates (box widths).
void plane_wave_direction(const int d[3], int &frame, int &up, double dp[3]);
// d (box widths) in the frame of direction dir: (lateral, lateral, along).
void plane_wave_rotate(int dir, const double d[3], double dp[3]);
// The shift of every half node of direction dir by d (box widths, unrotated), re then im:
// e^{-+lambda dp_z + i lambda (dp_x cos a + dp_y sin a)} (up: -, down: +), dp = Q_f d.
void plane_wave_shift(const PlaneWave &pw, int dir, const double d[3], double *out);

// Merge and shift (docs/CONVENTIONS.md "Merge and shift"; fmm-tables/src/plane_wave/
// geometry.rs): slot a in {0..5}^3 of a parent P is the box 2P - 2 + a one level down,
// numbered 36 a_x + 6 a_y + a_z; child b in {0, 1}^3 is numbered 4 b_x + 2 b_y + b_z.  The
// direction of the pair slot -> child, or -1 if they are adjacent: the axis z, y, x along
// which the slot lies at least two boxes from every child (coordinate 0: the positive
// direction, 5: the negative), else the direction of the pair's offset.
int plane_wave_slot_direction(int slot, int child);

// The accumulators of one direction: the slots grouped by the set of children (mask) they
// reach in that direction, by ascending mask, slots ascending.
struct PwAccumulator {
  uint8_t mask = 0;
  std::vector<int> slots;
};
const std::array<std::vector<PwAccumulator>, kPwDirections> &plane_wave_accumulators();

// The merge's tables (the same for every parent and kernel): the accumulators of
// direction d are acc_first[d] .. acc_first[d + 1], accumulator a serves the children of
// acc_mask[a] and sums the entries ent_first[a] .. ent_first[a + 1], entry e the slot
// ent_slot[e] shifted by the source shift ent_shift[e]; source shift i takes the waves of
// direction shift_dir[i] by shift_d[i] (box widths, unrotated: a slot's offset to the
// parent's centre); at most kPwAccStride - 1 accumulators per direction.
constexpr int kPwAccStride = 11;
struct PwMergeGeometry {
  std::vector<int> acc_first, acc_mask, ent_first, ent_slot, ent_shift, shift_dir;
  std::vector<std::array<double, 3>> shift_d;
};
const PwMergeGeometry &plane_wave_merge_geometry();

struct Plan;
// The merge's lists of the levels l with use[l] (l >= 2): per level (of the children) its
// parents' held entries (per parent and direction kPwAccStride offsets into pe, pe = wave
// block (box - box0) 6 + dir shifted left by 7, or'd with the source shift), the parents'
// children (8 box ids per parent, child 4 b_x + 2 b_y + b_z) and the target boxes (the
// children, in parent order); par_level[l] .. par_level[l + 1] the parents of level l's
// children, tgt_level likewise.  Box ids are box_base[level] + index; the wave blocks count
// from box0 = box_base[2] (one buffer for all levels) or, with per_level, from each level's
// first box box_base[l] (one buffer per level).
// Checks that every V entry of a child is a slot of its parent and every held, non-adjacent
// (slot, child) pair a V entry.
struct PwMergeLists {
  std::vector<uint32_t> pe_first, pe, children, targets, par_level, tgt_level;
};
PwMergeLists plane_wave_merge_lists(const Plan &plan, const std::vector<uint32_t> &box_base,
                                    const std::vector<char> &use, bool per_level = false);

// Host restatement of one V pair's plane-wave M2L (scaled storage of the level): out +=
// M2L_pw(d) in, in direction dir (-1: plane_wave_direction's).  The device kernels follow
// the same steps.
void plane_wave_m2l_host(const PlaneWave &pw, const int d[3], int dir, const double *in, double *out);
// Host restatement of one parent's merge and shift: slots[a] the multipole of slot a or
// nullptr; out[c] += the plane-wave M2L of every pair slot -> child c (8 x (p+1)^2).
void plane_wave_merge_host(const PlaneWave &pw, const double *const *slots, double *out);

}  // namespace fmmh
