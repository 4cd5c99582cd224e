// pwkernels.hpp -- launch wrappers of the plane-wave M2L kernels (pwave.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace fmmh {
namespace dev {

constexpr int kPwMaxNodes = 80;  // quadrature nodes the kernels hold in LDS

// Device tables of one plane-wave quadrature (see pwtables.hpp), passed by value.
struct PwArgs {
  int p, nodes, H, rb;
  const double *lambda;       // [nodes]
  const int *first, *nalpha;  // [nodes]: first half node, N_k
  const double *hw, *hc, *hs; // [H]: w_k / N_k, cos a_j, sin a_j
  const int *hk;              // [H]: node of the half node
  const double *fwd, *bwd;    // [3][rb]: multipole rotation into frame f, local rotation back
  const double *shift;        // [316][2][H]: T(d') of each offset (re, im)
  const int *dir_of;          // [316]: 2 frame + (down ? 1 : 0)
};

// LDS bytes of k_pw_incoming
size_t pw_incoming_lds(int p, int H);
// waves[(box - box0) * 6 + dir][2][H] of the boxes box0 .. box0 + nboxes (multipoles M)
void pw_outgoing(hipStream_t s, const PwArgs &a, uint32_t nboxes, uint32_t box0, const double *M, double *waves);
// L[row_box[r]] += plane-wave M2L of the row's V entries (source gid, offset index)
void pw_incoming(hipStream_t s, const PwArgs &a, uint32_t rows, const uint32_t *row_box, const uint32_t *row_off,
                 const uint32_t *ent_src, const uint16_t *ent_offset, uint32_t box0, const double *waves, double *L);
// raise the kernel's dynamic LDS limit to `bytes` (once per process)
hipError_t pw_allow_lds(size_t bytes);

}  // namespace dev
}  // namespace fmmh
