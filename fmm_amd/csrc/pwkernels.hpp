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
};

// The merge-and-shift geometry (the same for every parent): the accumulators of direction
// d are acc_first[d] .. acc_first[d + 1], accumulator a serves the children of acc_mask[a]
// and sums the entries ent_first[a] .. ent_first[a + 1], entry e the slot ent_slot[e]
// shifted by src_shift[ent_shift[e]] ([2][H] each, re then im); child c of direction d
// is shifted by child_shift[(8 d + c)][2][H].
struct PwMerge {
  const int *acc_first, *acc_mask, *ent_first, *ent_slot, *ent_shift;
  const double *src_shift, *child_shift;
};

// waves[(box - box0) * 6 + dir][2][H] of the boxes box0 .. box0 + nboxes (multipoles M)
void pw_outgoing(hipStream_t s, const PwArgs &a, uint32_t nboxes, uint32_t box0, const double *M, double *waves);
// Per parent (slots[216] and children[8] per parent, box ids or 0xffffffff) and direction:
// the accumulators of its slots' waves, and each child's incoming waves
// inc[(child - box0) * 6 + dir][2][H] = (sum of its accumulators) x its child shift.
void pw_merge(hipStream_t s, const PwArgs &a, const PwMerge &g, uint32_t nparents, const uint32_t *slots,
              const uint32_t *children, uint32_t box0, const double *waves, double *inc);
// L[box] += the local of each target box's incoming waves in the six directions
void pw_local(hipStream_t s, const PwArgs &a, uint32_t ntargets, const uint32_t *target, uint32_t box0,
              const double *inc, double *L);
// LDS bytes of k_pw_local
size_t pw_local_lds(int p, int H);
// raise k_pw_local's dynamic LDS limit to `bytes` (once per process)
hipError_t pw_allow_lds(size_t bytes);

}  // namespace dev
}  // namespace fmmh
