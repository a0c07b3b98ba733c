// HIP-event phase timing of the optimizer's device work (enabled on request; the events are
// recorded on the optimizer's own stream, so they bracket exactly the kernels of each phase).
#pragma once

#include <hip/hip_runtime.h>

#include <utility>
#include <vector>

#include "../../include/symhip.h"
#include "./common.h"

namespace symhip {

class PhaseTimer {
 public:
  static constexpr int kNumPhases = SYMHIP_NUM_PHASES;
  ~PhaseTimer() {
    for (auto& e : pool_) {
      (void)hipEventDestroy(e.first);
      (void)hipEventDestroy(e.second);
    }
  }
  void enable(bool on) {
    enabled_ = on;
    reset();
  }
  bool enabled() const { return enabled_; }
  void reset() {
    for (int i = 0; i < kNumPhases; ++i) {
      total_ms_[i] = 0;
      count_[i] = 0;
    }
    used_.clear();
  }
  int begin(int phase, hipStream_t s) {
    if (!enabled_) return -1;
    if (used_.size() == pool_.size()) {
      hipEvent_t a, b;
      HIP_OK(hipEventCreate(&a));
      HIP_OK(hipEventCreate(&b));
      pool_.push_back({a, b});
    }
    const int idx = (int)used_.size();
    used_.push_back({phase, idx});
    HIP_OK(hipEventRecord(pool_[idx].first, s));
    return idx;
  }
  void end(int idx, hipStream_t s) {
    if (!enabled_ || idx < 0) return;
    HIP_OK(hipEventRecord(pool_[idx].second, s));
  }
  // Call after the stream has been synchronized.
  void collect() {
    if (!enabled_) return;
    for (const auto& u : used_) {
      float ms = 0;
      HIP_OK(hipEventElapsedTime(&ms, pool_[u.second].first, pool_[u.second].second));
      total_ms_[u.first] += ms;
      count_[u.first]++;
    }
    used_.clear();
  }
  double total_ms(int p) const { return total_ms_[p]; }
  int64_t count(int p) const { return count_[p]; }

 private:
  bool enabled_ = false;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> pool_;
  std::vector<std::pair<int, int>> used_;
  double total_ms_[kNumPhases] = {};
  int64_t count_[kNumPhases] = {};
};

// Scope guard: records a phase on the stream when the timer is enabled.
class PhaseScope {
 public:
  PhaseScope(PhaseTimer* t, int phase, hipStream_t s) : t_(t), s_(s) {
    if (t_) idx_ = t_->begin(phase, s_);
  }
  ~PhaseScope() {
    if (t_) t_->end(idx_, s_);
  }

 private:
  PhaseTimer* t_;
  hipStream_t s_;
  int idx_ = -1;
};

}  // namespace symhip
