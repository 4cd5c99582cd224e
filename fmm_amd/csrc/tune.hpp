// tune.hpp -- the autotuner of the device operator's launch shapes (tune.cpp).
#pragma once

#include "operator.hpp"
#include "plan.hpp"

namespace fmmh {

constexpr int kTuneKinds = 5;       // P2P, M2L, L2L, M2M, M2L stage
constexpr int kTuneCandidates = 5;  // at most per decision
enum TuneKind : int { kTuneP2p = 0, kTuneM2l = 1, kTuneL2l = 2, kTuneM2m = 3, kTuneStage = 4 };

// ms per call of every candidate timed, NaN where not timed: [kind][level][candidate]
// (P2P: level 0, candidates 33, 17, 49, 65, never; M2L: the static mode, the other; L2L:
// the static width, 16, 64, 128; M2M: static, per-tile, split; M2L stage: level 0, one
// k-step per stage, two k-steps in a ring of two).
struct TuneReport {
  double ms[kTuneKinds][kTuneLevels][kTuneCandidates];
  TuneReport() {
    for (auto &k : ms)
      for (auto &l : k)
        for (double &c : l) c = __builtin_nan("");
  }
};

// Times the candidates on operators built over `plan` with `options` (whose own tuning
// is ignored) within budget_s seconds; out holds the winners (0 where the static rule
// was kept or nothing was timed).
void autotune(const Plan &plan, const OperatorOptions &options, double budget_s, Tuning &out, TuneReport &report);

}  // namespace fmmh
