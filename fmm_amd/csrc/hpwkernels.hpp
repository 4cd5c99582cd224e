// hpwkernels.hpp -- launch wrappers of the Helmholtz exponential M2L kernels (hpwave.hip);
// the merge is pwave.hip's pw_merge with W complex wave values per direction.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace fmmh {
namespace dev {

// One level's tables (hpwtables.hpp HpwLevel) on the device, passed by value.
struct HpwArgs {
  int p, nodes, W, orders, rb;                            // rb: sum_n (2n+1)^2
  const int *start, *len, *ord, *mode_start, *trig, *factor;  // [nodes]
  const int *wk;                                          // [W]: node of each wave value
  int nitem;
  const int *item;                                        // (node, order) pairs: k (p + 1) + m, by decreasing N_k
  const double *cosm, *sinm;                              // cos m a_j, sin m a_j per node, order, value
  const double *out_re, *out_im, *in_re, *in_im;          // the outgoing and incoming factors
  const double *frames;                                   // [2][rb]: O^n of frames 1 and 2, row-major
};

// waves[(box - first) * 6 + dir][re, im][W] of the boxes first .. first + nboxes
// (stored multipoles M, three planes per box)
void hpw_outgoing(hipStream_t s, const HpwArgs &a, uint32_t nboxes, uint32_t first, const double *M, double *waves);
// L[box] (planes 0 and 1) += the local of each target box's incoming waves
// inc[(box - box0) * 6 + dir][re, im][W]
void hpw_local(hipStream_t s, const HpwArgs &a, uint32_t ntargets, const uint32_t *target, uint32_t box0,
               const double *inc, double *L);
size_t hpw_outgoing_lds(int p, int orders);
size_t hpw_local_lds(int p, int W, int orders);
// raise both kernels' dynamic LDS limits (once per process)
hipError_t hpw_allow_lds(size_t outgoing, size_t local);

}  // namespace dev
}  // namespace fmmh
