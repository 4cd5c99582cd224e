// tune.cpp -- the autotuner of the device operator's launch shapes: which P2P form, which
// M2L mode per level, which L2L task width per level and which M2M form per level an
// operator runs, timed on the operator's own levels and buffers at build.
//
// The reference's device autotune (fmm-exec/src/tune.rs:1-110) times its M2L strategy,
// the GEMM of each translation kind per pair bucket and the P2P layout per
// points-per-leaf bucket on the Fmm's own levels, keeps the winners in a cache file and
// falls back to a static rule where nothing was tuned.  This tuner takes the decisions of
// this path's kernels (the static rule's candidate first, so a cut budget keeps it):
//   - P2P (mixed precision): the chunk size from which targets read their sources by
//     uniform loads instead of LDS tiles with phases: 33 (static), 17, 49, 65, never;
//   - M2L per level: box mode (tiles of 128 targets of one octant walking their offsets)
//     or clustered mode (scratch rows per target and parent-level neighbour);
//   - L2L per level: 16, 64 or 128 columns per task;
//   - M2M per parent level: the per-tile form or the split form;
//   - the bf16x3 M2L's pipeline stage: one k-step per stage (ring of three) or two (ring
//     of two), timed over every level's M2L.
// Each candidate runs once and is synchronised (first-touch and caches excluded), then
// is timed in batches of launches on the operator's stream, doubling until a batch takes
// at least 2 ms, and its time is the median of 5 batches; the fastest wins, the first
// measured on a tie.  Tuning stops at the budget's deadline, checked before every
// candidate (and before the three candidate operators of the per-level decisions are
// built, from the first one's build time): decisions not reached keep the static rule.
// The cache file of decisions per bucket is the host binding's (fmm_amd/tune.py).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <limits>
#include <memory>
#include <vector>

#include "operator.hpp"
#include "tune.hpp"

namespace fmmh {

namespace {

using Clock = std::chrono::steady_clock;

constexpr double kMinBatchMs = 2.0;
constexpr int kBatches = 5;

// Median ms per call of f (launches on stream s), or NaN past the deadline.
template <class F>
double time_calls(hipStream_t s, F &&f, Clock::time_point deadline) {
  if (Clock::now() > deadline) return std::nan("");
  f();
  hip_check(hipStreamSynchronize(s), "autotune: warm-up");
  hipEvent_t a, b;
  hip_check(hipEventCreate(&a), "hipEventCreate");
  hip_check(hipEventCreate(&b), "hipEventCreate");
  auto batch = [&](int reps) {
    hip_check(hipEventRecord(a, s), "hipEventRecord");
    for (int r = 0; r < reps; ++r) f();
    hip_check(hipEventRecord(b, s), "hipEventRecord");
    hip_check(hipEventSynchronize(b), "hipEventSynchronize");
    float ms = 0.f;
    hip_check(hipEventElapsedTime(&ms, a, b), "hipEventElapsedTime");
    return double(ms);
  };
  int reps = 1;
  for (double ms = batch(reps); ms < kMinBatchMs && reps < (1 << 14) && Clock::now() < deadline;
       ms = batch(reps))
    reps *= 2;
  std::vector<double> per;
  for (int k = 0; k < kBatches; ++k) per.push_back(batch(reps) / reps);
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  std::sort(per.begin(), per.end());
  return per[kBatches / 2];
}

void prepare(DeviceOperator &op, size_t ns) {
  // unit charges and one full evaluation, so every buffer holds finite values
  std::vector<double> q(ns, 1.0);
  op.load_charges(q.data(), false);
  op.reset();
  op.evaluate();
  op.sync();
}

}  // namespace

void autotune(const Plan &plan, const OperatorOptions &options, double budget_s, Tuning &out, TuneReport &report) {
  const auto deadline = Clock::now() + std::chrono::microseconds(int64_t(budget_s * 1e6));
  report = TuneReport{};
  out = Tuning{};
  const int L = plan.nlevels;
  // candidate operators: A the static rule; B every M2L level in the other mode, L2L 16
  // columns, M2M per-tile; C L2L 64, M2M split; D L2L 128
  OperatorOptions oa = options;
  oa.tuning = Tuning{};
  oa.tuning.p2p_uniform_min = 0;
  const auto built0 = Clock::now();
  auto A = std::make_unique<DeviceOperator>(plan, oa);
  prepare(*A, size_t(plan.ns));
  const auto build_time = Clock::now() - built0;  // one candidate operator's build
  hipStream_t sa = A->stream();

  // P2P form (the mixed path only has two forms)
  if (options.precision == kMixed) {
    const int cand[kTuneCandidates] = {33, 17, 49, 65, 129};
    double best = std::numeric_limits<double>::infinity();
    for (int c = 0; c < kTuneCandidates; ++c) {
      A->set_p2p_uniform_min(cand[c]);
      const double ms = time_calls(sa, [&] {
        for (int l = 0; l < L; ++l) A->p2p(l);
      }, deadline);
      report.ms[kTuneP2p][0][c] = ms;
      if (ms < best) {
        best = ms;
        out.p2p_uniform_min = cand[c];
      }
    }
    A->set_p2p_uniform_min(0);
  }
  // the bf16x3 M2L's stage shape (dense M2L in the mixed precision)
  if (options.precision == kMixed && options.m2l == kM2lDense) {
    const int cand[2] = {0, 2};
    double best = std::numeric_limits<double>::infinity();
    for (int c = 0; c < 2; ++c) {
      A->set_m2l_stage(cand[c]);
      const double ms = time_calls(sa, [&] {
        for (int l = 2; l < L; ++l) A->m2l(l);
      }, deadline);
      report.ms[kTuneStage][0][c] = ms;
      if (ms < best) {
        best = ms;
        out.m2l_stage = cand[c];
      }
    }
    A->set_m2l_stage(0);
  }
  // the per-level decisions need three more operators: skip them when their builds alone
  // would overrun the budget (the static rule stays)
  if (Clock::now() + 3 * build_time > deadline) return;

  OperatorOptions ob = oa, oc = oa, od = oa;
  for (int l = 0; l < std::min(L, kTuneLevels); ++l) {
    const int m = A->m2l_mode_of(l);
    ob.tuning.m2l_mode[l] = m == 1 ? 2 : m == 2 ? 1 : 0;
    ob.tuning.l2l_width[l] = 16;
    ob.tuning.m2m_form[l] = 1;
    oc.tuning.l2l_width[l] = 64;
    oc.tuning.m2m_form[l] = 2;
    od.tuning.l2l_width[l] = 128;
  }
  std::unique_ptr<DeviceOperator> ops[4];
  ops[0] = std::move(A);
  ops[1] = std::make_unique<DeviceOperator>(plan, ob);
  ops[2] = std::make_unique<DeviceOperator>(plan, oc);
  ops[3] = std::make_unique<DeviceOperator>(plan, od);
  for (int k = 1; k < 4; ++k) prepare(*ops[k], size_t(plan.ns));

  for (int l = 0; l < std::min(L, kTuneLevels); ++l) {
    // M2L: the static mode (A) against the other (B)
    if (options.m2l == kM2lDense && ops[0]->m2l_mode_of(l) != 0) {
      double best = std::numeric_limits<double>::infinity();
      for (int k = 0; k < 2; ++k) {
        DeviceOperator &op = *ops[k];
        const double ms = time_calls(op.stream(), [&] { op.m2l(l); }, deadline);
        report.ms[kTuneM2l][l][k] = ms;
        if (ms < best) {
          best = ms;
          out.m2l_mode[l] = op.m2l_mode_of(l);
        }
      }
    }
    // L2L: the static width first, then the other widths
    if (l >= 1) {
      double best = std::numeric_limits<double>::infinity();
      const int order[4] = {0, 1, 2, 3};
      for (int k : order) {
        DeviceOperator &op = *ops[k];
        if (k > 0 && op.l2l_width_of(l) == ops[0]->l2l_width_of(l)) continue;  // the static width again
        const double ms = time_calls(op.stream(), [&] { op.l2l(l); }, deadline);
        report.ms[kTuneL2l][l][k] = ms;
        if (ms < best) {
          best = ms;
          out.l2l_width[l] = op.l2l_width_of(l);
        }
      }
    }
    // M2M into the parents of level l: static (A), per-tile (B), split (C)
    if (l + 1 < L) {
      double best = std::numeric_limits<double>::infinity();
      for (int k = 0; k < 3; ++k) {
        DeviceOperator &op = *ops[k];
        const double ms = time_calls(op.stream(), [&] { op.m2m(l); }, deadline);
        report.ms[kTuneM2m][l][k] = ms;
        if (ms < best) {
          best = ms;
          out.m2m_form[l] = k == 0 ? 0 : k;
        }
      }
    }
  }
}

}  // namespace fmmh
