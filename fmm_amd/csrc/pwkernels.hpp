// pwkernels.hpp -- launch wrappers of the plane-wave M2L kernels (pwave.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace fmmh {
namespace dev {

constexpr int kPwMaxNodes = 39;  // FMM3D's largest quadrature; the kernels hold per-node data in LDS

// Device tables of one plane-wave quadrature (see pwtables.hpp), passed by value.
struct PwArgs {
  int p, nodes, H, rb;
  const int *first, *nalpha, *fourier;  // [nodes]: first half node, N_k (even), F_k'
  const double *opow, *ipow;            // [nodes][p + 1]: lambda^n 2^-(n+1), lambda^n 2^-n
  const double *hw, *hc, *hs;           // [H]: w_k / N_k, cos a_j, sin a_j
  const int *hk;                        // [H]: node of the half node
  const double *fwd, *bwd;              // [3][rb]: multipole rotation into frame f, local rotation back
  int nitem;                            // the (node, order) pairs with order < F_k'
  const int *gitem;                     // [nitem]: k (p + 1) + m, by decreasing N_k
};

// The merge-and-shift geometry (the same for every parent): the accumulators of direction
// d are acc_first[d] .. acc_first[d + 1], accumulator a serves the children of acc_mask[a];
// the source shifts src_shift[2][H] (re then im) are indexed by the entries; child c of
// direction d is shifted by child_shift[(8 d + c)][2][H].
struct PwMerge {
  const int *acc_first, *acc_mask;
  const double *src_shift, *child_shift;
};

// waves[(box - box0) * 6 + dir][2][H] of the boxes box0 .. box0 + nboxes (multipoles M)
void pw_outgoing(hipStream_t s, const PwArgs &a, uint32_t nboxes, uint32_t box0, const double *M, double *waves);
constexpr int kPwAccStride = 11;  // accumulator offsets per (parent, direction): at most 10 + 1
// Per parent and direction: the accumulators of its held slots' waves (the entries of
// accumulator a of item (parent, dir) are pe[pe_first[(6 parent + dir) kPwAccStride + a]
// .. + a + 1], each wave block (box - box0) 6 + dir shifted left by 7 or'd with its
// source shift), and each child's (children[8] per parent, box ids or 0xffffffff) incoming
// waves inc[(child - box0) * 6 + dir][2][H] = (sum of its accumulators) x its child shift.
void pw_merge(hipStream_t s, const PwArgs &a, const PwMerge &g, uint32_t nparents, const uint32_t *pe_first,
              const uint32_t *pe, const uint32_t *children, uint32_t box0, const double *waves, double *inc);
// L[box] += the local of each target box's incoming waves in the six directions
void pw_local(hipStream_t s, const PwArgs &a, uint32_t ntargets, const uint32_t *target, uint32_t box0,
              const double *inc, double *L);
// LDS bytes of k_pw_local
size_t pw_local_lds(int p, int H, int nodes);
// raise k_pw_local's dynamic LDS limit to `bytes` (once per process)
hipError_t pw_allow_lds(size_t bytes);

}  // namespace dev
}  // namespace fmmh
