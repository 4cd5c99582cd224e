// operator.hpp -- the device operator: everything an evaluation needs resident on one
// MI355X, and the level calls of the reference's operator interface
// (FmmOperator, fmm-plan/src/operator.rs:434-497; DeviceOperator,
// fmm-exec/src/device.rs:5212-5446).
//
// Build once (allocate and upload plan, points and tables), then per evaluation: load
// charges, reset, the level calls in the evaluator's order (all queued on one stream),
// and one read of the output (the output pass and the single synchronisation).
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <cstdint>
#include <string>
#include <vector>

#include "plan.hpp"
#include "pwkernels.hpp"

namespace fmmh {

constexpr int kNumKinds = 10;        // timing kinds: the operator kinds, other, M2L contraction
constexpr int kKindM2lContract = 9;  // the M2L contraction kernel alone
// M2L tasks go to the XCDs in runs of 8: the eight octant classes of one block of parents
constexpr uint32_t kM2lXcdRun = 8;

enum Precision : int {
  kPlainF64 = 0,  // every operator in f64
  kMixed = 1,     // P2P in f32 (double-float separations), M2L in bf16x3 on MFMA
};

void hip_check(hipError_t e, const char *what);

template <typename T>
struct DevBuf {
  T *ptr = nullptr;
  size_t count = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  DevBuf(DevBuf &&o) noexcept : ptr(o.ptr), count(o.count) {
    o.ptr = nullptr;
    o.count = 0;
  }
  DevBuf &operator=(DevBuf &&o) noexcept {
    if (this != &o) {
      release();
      ptr = o.ptr;
      count = o.count;
      o.ptr = nullptr;
      o.count = 0;
    }
    return *this;
  }
  ~DevBuf() { release(); }
  void release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    count = 0;
  }
  void alloc(size_t n) {
    release();
    count = n;
    if (n) hip_check(hipMalloc(&ptr, n * sizeof(T)), "hipMalloc");
  }
  void upload(const std::vector<T> &v) {
    alloc(v.size());
    if (!v.empty()) hip_check(hipMemcpy(ptr, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice), "upload");
  }
  size_t bytes() const { return count * sizeof(T); }
};

// One contraction launch: tasks of COLS output columns, groups of (table, sources).
struct Schedule {
  int width = 128;  // columns per task
  uint32_t ntasks = 0;
  uint64_t pairs = 0;  // real (column, group) pairs with a source
  uint64_t slots = 0;  // columns x groups issued
  DevBuf<uint32_t> task_group_off, group_src, col_out;
  DevBuf<uint16_t> group_table;
};

// A row list for the leaf kernels: rows with an output (box or leaf) and CSR entries.
struct Rows {
  uint32_t rows = 0;
  DevBuf<uint32_t> row_id, row_cell, row_off, row_ent;
};

enum M2lMethod : int {
  kM2lDense = 0,     // dense tables, grouped contraction on MFMA
  kM2lRotation = 1,  // rotation, coaxial shift, rotation back (f32 in the mixed precision)
  kM2lPlaneWave = 2, // plane waves in six directions by merge and shift (f64 in both precisions;
                     // approximate to the quadrature's tolerance)
};

constexpr int kTuneLevels = 17;  // levels 0..16

// Launch-shape decisions an autotuner may take (tune.hip); 0 everywhere is the static rule.
struct Tuning {
  int p2p_uniform_min = 0;                  // mixed P2P: chunks of at least this many targets read
                                            // sources by uniform loads (0: 33)
  int m2l_stage = 0;                        // bf16x3 M2L: 0/1 one k-step per pipeline stage,
                                            // 2 two k-steps in a ring of two buffers
  std::array<int, kTuneLevels> m2l_mode{};  // per level: 0 static, 1 box mode, 2 clustered mode
  std::array<int, kTuneLevels> l2l_width{}; // per level: 0 static, else 16, 64 or 128 columns per task
  std::array<int, kTuneLevels> m2m_form{};  // per parent level: 0 static, 1 per-tile, 2 split
};

struct OperatorOptions {
  int p = 6;
  int precision = kPlainF64;
  int gradients = 1;
  int device = 0;
  int m2l = kM2lDense;
  Tuning tuning;
};

class DeviceOperator {
 public:
  DeviceOperator(const Plan &plan, const OperatorOptions &options);
  ~DeviceOperator();

  // begin_evaluation: charges in caller order (host or device pointer).
  void load_charges(const double *charges, bool on_device);
  void reset();
  // level calls (FmmOperator, fmm-plan/src/operator.rs:434-497)
  void p2m(int level);
  void m2m(int level);
  void m2l(int level);
  void p2l(int level);
  void l2l(int level);
  void l2p(int level);
  void m2p(int level);
  void p2p(int level);
  // the evaluator's fixed order (fmm-plan/src/operator.rs:408-425), all levels
  void evaluate();
  // reset() + evaluate() as one hipGraph launch: captured on the first call (timing must
  // be off), replayed afterwards; the charges are loaded and the output read as usual.
  void evaluate_graph();
  // output pass + download in caller order (phi[nt], grad[3 nt]); the one sync
  void read_output(double *phi, double *grad, bool on_device);
  // Vector-Jacobian product of the potentials (sources = targets): given the charges q,
  // the cotangent g = dL/dphi and the forward gradient field (grad phi, 3 per point), all
  // in caller order, writes grad_q = dL/dq and, when grad_pos is not null, grad_pos =
  // dL/dpositions.  Runs one more evaluation with charges g (the adjoint field), so it
  // overwrites the operator's charges and expansions.  Buffers for host arrays are
  // allocated on the first call.
  void vjp(const double *q, const double *g, const double *field, bool on_device, double *grad_q,
           double *grad_pos);
  void sync();
  // order this operator's stream after the work queued on `producer` so far (a caller's
  // stream that writes the charges or cotangents this operator reads from device memory)
  void wait_for(hipStream_t producer);
  // copy the multipoles (locals = false) or locals of one level to host, box-major
  void read_expansions(bool locals, int level, double *out);
  // per-kind device times (HIP events on this operator's stream)
  void set_timing(bool enable);
  void kind_times(double *ms, int64_t *launches);
  // the mixed P2P's uniform-source threshold (targets per chunk), applied from the next launch
  void set_p2p_uniform_min(int n) { opt_.tuning.p2p_uniform_min = n; }
  // the bf16x3 M2L's pipeline stage shape (Tuning::m2l_stage), from the next launch
  void set_m2l_stage(int stage) { opt_.tuning.m2l_stage = stage; }
  int p2p_uniform_min() const { return opt_.tuning.p2p_uniform_min > 0 ? opt_.tuning.p2p_uniform_min : 33; }
  // the M2L mode each level's schedule took: 0 none, 1 box mode, 2 clustered mode
  int m2l_mode_of(int level) const { return level >= 0 && level < kTuneLevels ? m2l_mode_taken_[level] : 0; }
  int l2l_width_of(int level) const { return level >= 0 && level < int(l2l_.size()) ? l2l_[level].width : 0; }

  int nlevels() const { return nlevels_; }
  int64_t nsources() const { return ns_; }
  int64_t ntargets() const { return nt_; }
  hipStream_t stream() const { return stream_; }
  const OperatorOptions &options() const { return opt_; }
  // statistics for reports and the roofline
  uint64_t p2p_pairs() const { return p2p_pairs_; }
  uint64_t m2l_pairs() const { return m2l_.pairs; }
  uint64_t m2l_slots() const { return m2l_.slots; }
  size_t device_bytes() const;
  double *device_locals() const { return loc_.ptr; }
  double *device_multipoles() const { return mult_.ptr; }
  size_t nboxes_total() const { return nbox_total_; }

 private:
  void build_schedules(const Plan &plan);
  template <class F>
  void timed(int kind, F &&f);
  void collect_timings();
  void launch_p2p(uint32_t leaf_first, uint32_t count);
  void launch_contract(const Schedule &s, uint32_t task_first, uint32_t ntasks, const double *X, double *Y,
                       const DevBuf<double> &tables, uint32_t xchunk = 0);
  void launch_m2l(int level_first, int level_last);
  void launch_p2l(int level_first, int level_last);

  OperatorOptions opt_;
  int nlevels_ = 0, nc_ = 0;
  hipStream_t stream_ = nullptr;
  hipEvent_t order_event_ = nullptr;  // wait_for: a caller's stream before this one
  int64_t ns_ = 0, nt_ = 0;
  uint32_t nleaves_ = 0;
  size_t nbox_total_ = 0;
  std::vector<uint32_t> box_base_, leaf_off_;
  uint64_t p2p_pairs_ = 0;

  DevBuf<uint32_t> leaf_cell_, box_cell_, src_start_, tgt_start_, src_order_, tgt_order_;
  DevBuf<uint32_t> near_off_, near_ent_;
  DevBuf<double> sx_, sy_, sz_, tx_, ty_, tz_, q_, q_in_;
  DevBuf<double> leaf_r_;  // leaf radius in domain units (output scaling)
  hipGraph_t graph_ = nullptr;
  hipGraphExec_t graph_exec_ = nullptr;
  DevBuf<double> mult_, loc_, phi_hat_, grad_hat_, phi_, grad_;
  DevBuf<double> vjp_in_, vjp_out_;  // host-array staging of vjp (5 n and 4 n doubles)
  bool same_points_ = false;
  DevBuf<double> t_m2m_, t_l2l_, t_m2l_;
  DevBuf<uint16_t> t_m2l16_, mult16_;  // mixed precision: split tables and multipoles
  std::array<std::array<uint16_t, 316>, 8> m2l_rank_;  // M2L group order per target octant
  // rotation M2L: rows of target boxes with V entries (by level), entry (source, offset)
  DevBuf<uint32_t> rot_box_, rot_off_, rot_src_;
  DevBuf<uint16_t> rot_offset_, rot_dist_;
  DevBuf<double> rot_fwd_, rot_bwd_, rot_coax_;
  DevBuf<float> rot_fwd32_, rot_bwd32_, rot_coax32_;
  std::vector<uint32_t> rot_level_;
  int rot_rb_ = 0, rot_cb_ = 0;
  // rotation M2L: the pairs grouped by (level, offset) in blocks of one offset; pair_of
  // maps each V entry (rot rows' order) to its pair's scratch row
  DevBuf<uint32_t> rot_blk_first_, rot_pair_src_;
  // rotation M2L in chunks of whole blocks of one level (scratch rows within
  // kRotScratchBytes): per chunk its first block, first pair and first reduction row; the
  // chunks of level l from rot_chunk_level_[l]
  std::vector<uint32_t> rot_chunk_blk_, rot_chunk_q_, rot_chunk_row_, rot_chunk_level_;
  Rows rot_red_;
  DevBuf<uint16_t> rot_blk_offset_;
  std::vector<uint32_t> rot_blk_level_;
  DevBuf<double> rot_slots_;
  DevBuf<float> rot_slots32_, mult32_;
  void build_rotation(const Plan &plan);
  void build_plane_wave(const Plan &plan);
  // plane-wave M2L: quadrature tables, the merge-and-shift geometry, per level (of the
  // children) the parents' held accumulator entries and children and the target boxes,
  // the waves and the incoming waves of the boxes of levels >= 2
  dev::PwArgs pw_{};
  dev::PwMerge pw_merge_{};
  DevBuf<double> pw_opow_, pw_ipow_, pw_hw_, pw_hc_, pw_hs_, pw_fwd_, pw_bwd_, pw_src_shift_, pw_child_shift_;
  DevBuf<double> pw_waves_, pw_inc_;
  DevBuf<int> pw_first_, pw_nalpha_, pw_fourier_, pw_hk_, pw_gitem_;
  DevBuf<int> pw_acc_first_, pw_acc_mask_;
  DevBuf<uint32_t> pw_pe_first_, pw_pe_, pw_children_, pw_targets_;
  std::vector<uint32_t> pw_par_level_, pw_tgt_level_;
  uint32_t pw_box0_ = 0, pw_nbox_ = 0;
  size_t nslots_ = 0;                  // clustered-mode scratch rows (slot ids follow the box ids)
  DevBuf<double> slots_;               // plain f64: scratch rows
  DevBuf<float> slots32_;              // mixed: scratch rows (one group's f32 sum each)
  DevBuf<float> src_rec_;              // mixed: P2P source records, 4 words per source
  DevBuf<uint8_t> near_close_;         // mixed: near entries with a close pair (p2p_close_flags)

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

  Rows p2m_rows_, l2p_rows_, m2p_rows_;
  std::vector<uint32_t> p2m_level_, l2p_level_, m2p_level_;  // row ranges per level
  Rows p2l_rows_;                                            // every level, by level
  std::vector<uint32_t> p2l_level_;
  std::vector<Schedule> m2m_, l2l_;                          // per level
  Schedule m2l_;                                             // every level, by level
  std::array<int, kTuneLevels> m2l_mode_taken_{};
  std::vector<uint32_t> m2l_level_task_;                     // task range per level
  Rows m2l_red_;                                             // clustered-mode reductions
  std::vector<uint32_t> m2l_red_level_;
};

// Compensated f64 direct sum on the device (unscaled sums q / r and their gradient):
// inputs and outputs on the host (on_device = false) or device pointers; synchronous.
void direct_sum(int device, const double *sources, const double *charges, int64_t ns, const double *targets,
                int64_t nt, double *phi, double *grad, bool on_device);
// The same on device pointers, queued on `stream` without allocating: the partial sums
// live in a fixed per-GPU array (noalloc.hip), which bounds the targets per call.
void direct_sum_noalloc(int device, const double *sources, const double *charges, int64_t ns, const double *targets,
                        int64_t nt, double *phi, double *grad, hipStream_t stream);
int64_t direct_sum_noalloc_max_targets();

}  // namespace fmmh
