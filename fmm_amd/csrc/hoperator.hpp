// hoperator.hpp -- the Helmholtz device operator: an FMM of e^{ikr}/(4 pi r) with complex
// charges on one MI355X, over the same plan (tree, lists, leaf-scaled points) as the
// Laplace operator.  The reference evaluates its Helmholtz FMM on the host only
// (nd_fmm_exec::helmholtz, fmm-exec/src/helmholtz/fmm.rs:71; operator
// fmm-exec/src/helmholtz/operator.rs:448-480); this operator keeps its evaluator's order
// (P2M, M2M, M2L, P2L, L2L, L2P, M2P, P2P) and its stored forms, and runs every kind on the
// GPU: leaf kernels in f64, the translations as dense complex tables per level on the MFMA
// contraction, P2P in f64 (plain) or f32 with f64 sums (mixed, where M2L also runs in
// bf16x3 on MFMA).
#pragma once

#include <hip/hip_runtime.h>

#include <complex>
#include <cstdint>
#include <vector>

#include "hpwkernels.hpp"
#include "operator.hpp"
#include "plan.hpp"
#include "pwkernels.hpp"

namespace fmmh {

struct HelmholtzOptions {
  int p = 8;
  double kre = 1.0, kim = 0.0;  // wavenumber k in the caller's units: k != 0, Re k >= 0, Im k >= 0
  int precision = kPlainF64;
  int gradients = 1;
  int device = 0;
  int m2l = 0;  // kM2lDense: dense complex tables per level on MFMA where they fit the
                // budget, per-pair rotation elsewhere; kM2lRotation: per-pair rotation,
                // coaxial shift, rotation back (f64) on every level, tables O(p^3) per distance;
                // kM2lPlaneWave: the exponential M2L by merge and shift (f64) on the levels
                // with a quadrature of the reference's rule, per-pair rotation elsewhere
  int64_t m2l_dense_budget = 0;  // bytes of dense tables (316 x 2 (p+1)^4 f64 per level): 0 the
                                 // reference's default (1 GiB), negative no limit
};

// DEFAULT_DENSE_BUDGET of fmm-exec/src/helmholtz/operator.rs:783, 1 GiB
constexpr int64_t kHelmholtzDenseBudget = int64_t(1) << 30;

class HelmholtzOperator {
 public:
  HelmholtzOperator(const Plan &plan, const HelmholtzOptions &options);
  ~HelmholtzOperator();

  // complex charges (re, im) per source in caller order, host or device pointer
  void load_charges(const double *charges, bool on_device);
  void reset();
  void evaluate();
  // phi[2 nt], grad[6 nt] (x re, x im, y re, ...) in caller order; the one sync
  void read_output(double *phi, double *grad, bool on_device);
  void sync();
  // order this operator's stream after the work queued on `producer` so far
  void wait_for(hipStream_t producer);
  // the stored coefficients of one level, box-major, real plane then imaginary plane
  void read_expansions(bool locals, int level, double *out);
  void set_timing(bool enable);
  void kind_times(double *ms, int64_t *launches);

  int nlevels() const { return nlevels_; }
  int64_t nsources() const { return ns_; }
  int64_t ntargets() const { return nt_; }
  uint64_t p2p_pairs() const { return p2p_pairs_; }
  uint64_t m2l_pairs() const { return m2l_pairs_; }
  size_t device_bytes() const;

 private:
  template <class F>
  void timed(int kind, F &&f);
  void collect_timings();
  void launch_p2p();

  HelmholtzOptions opt_;
  int nlevels_ = 0, nc_ = 0;
  bool m2m_complex_ = true;  // M2M / L2L tables have an imaginary part (not for real k)
  hipStream_t stream_ = nullptr;
  hipEvent_t order_event_ = nullptr;  // wait_for: a caller's stream before this one
  int64_t ns_ = 0, nt_ = 0;
  uint32_t nleaves_ = 0;
  size_t nbox_total_ = 0;
  std::vector<uint32_t> box_base_;
  uint64_t p2p_pairs_ = 0, m2l_pairs_ = 0;

  DevBuf<uint32_t> leaf_cell_, box_cell_, src_start_, tgt_start_, src_order_, tgt_order_;
  DevBuf<uint32_t> near_off_, near_ent_;
  DevBuf<double> sx_, sy_, sz_, tx_, ty_, tz_, qr_, qi_, q_in_;
  DevBuf<float> src_rec_;  // mixed: P2P source records (dev::h_source_records), charges by h_gather
  DevBuf<double> leaf_r_, hlev_, norm_;
  DevBuf<double> mult_, loc_, phi_hat_, grad_hat_, phi_, grad_;
  DevBuf<double> t_m2m_, t_l2l_, t_m2l_;
  DevBuf<uint16_t> t_m2l16_, mult16_;

  Rows p2m_rows_, l2p_rows_, m2p_rows_, p2l_rows_;
  std::vector<Schedule> m2m_, l2l_;  // per level (M2M into the parents of level l, L2L into level l)
  Schedule m2l_;
  // sparse M2L levels: scratch rows (ids from 3 nbox_total_ on) of one (target plane,
  // parent-level neighbour) each, and the ordered reduction of a target plane's rows
  size_t nslots_ = 0;
  DevBuf<double> slots_;
  DevBuf<float> slots32_;
  Rows m2l_red_;
  // rotation M2L: the offsets' rotation blocks, the (level, distance) coaxial blocks, blocks
  // of pairs of one (level, offset), two scratch rows per pair and their ordered reduction
  int rot_rb_ = 0, rot_cb_ = 0;
  std::vector<uint32_t> rot_chunk_blk_, rot_chunk_q_, rot_chunk_row_;  // per chunk: first block, pair, row
  DevBuf<double> rot_fwd_, rot_coax_, rot_slots_;
  DevBuf<uint32_t> rot_blk_first_, rot_blk_coax_, rot_pair_src_;
  DevBuf<uint16_t> rot_blk_offset_;
  Rows rot_red_;
  // exponential M2L: per level with a quadrature its tables (one buffer per kind, the levels
  // after one another), its merge shifts, its waves and incoming waves, and the merge's
  // parents, children and targets by level
  std::vector<int> pw_levels_;
  std::vector<dev::HpwArgs> pw_args_;
  std::vector<size_t> pw_wave_off_, pw_shift_off_;
  DevBuf<int> pw_int_;
  DevBuf<double> pw_tab_, pw_shift_, pw_waves_, pw_inc_;
  DevBuf<int> pw_acc_first_, pw_acc_mask_;
  DevBuf<uint32_t> pw_pe_first_, pw_pe_, pw_children_, pw_targets_;
  std::vector<uint32_t> pw_par_level_, pw_tgt_level_;
  void build_plane_wave(const Plan &plan, const std::vector<std::complex<double>> &kh, const std::vector<char> &use);
  void launch_plane_wave();

  bool timing_ = false;
  struct Mark {
    int kind;
    hipEvent_t a, b;
  };
  std::vector<hipEvent_t> events_;
  std::vector<Mark> marks_;
  size_t events_used_ = 0;
  double kind_ms_[16] = {0};
  int64_t kind_launches_[16] = {0};
};

// Compensated f64 direct sum of q e^{ikr}/r and its gradient (no 1/(4 pi)); complex
// charges and outputs interleaved, host or device pointers; synchronous.
void helmholtz_direct_sum(int device, double kre, double kim, const double *sources, const double *charges,
                          int64_t ns, const double *targets, int64_t nt, double *phi, double *grad, bool on_device);

}  // namespace fmmh
