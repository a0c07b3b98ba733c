// Batched Levenberg-Marquardt for many small robot-3D-localization problems (Pose3 chains with
// odometry factors between consecutive poses and landmark matching factors) in one persistent
// kernel: one wavefront runs the whole LM loop of one problem.
#pragma once

#include <vector>

#include "./common.h"
#include "../../include/symhip.h"

namespace symhip {

// Problem tables of a batch in the kernel's layout, built on the host (no device memory).
struct PackedBatch {
  std::vector<int32_t> pose_begin;  // [n_problems + 1] first global pose of each problem
  std::vector<int32_t> pose_soff;   // [total_poses] offset of each pose's storage in cvals
  std::vector<int32_t> match_ptr;   // [total_poses + 1] matching factors of each pose
  std::vector<int32_t> match_args;  // 3 per match: landmark, measurement, sigma offsets
  std::vector<int32_t> odo_args;    // 3 per pose: measurement, sigmas, epsilon (-1: last pose)
  std::vector<double> cvals;        // every problem's Values storage, concatenated
};
PackedBatch pack_batch(const symhip_problem_desc* descs, int32_t n_problems,
                       const symhip_optimizer_params& params);

// symhip_batch_solve_noalloc: the batched LM kernel on caller-provided device tables, with its
// working state in a file-scope device array (no allocation). Launches on `stream`.
void batch_solve_noalloc(const symhip_batch_tables& t, const symhip_optimizer_params& p,
                         int32_t precision, double epsilon, double* values_out,
                         symhip_optimization_stats* stats_out, symhip_iteration_stats* iters_out,
                         hipStream_t stream);

class BatchLocalization {
 public:
  // Every problem: optimized keys are Pose3 only, factors are MATCHING (one pose) and ODOMETRY
  // (poses at consecutive state positions); at most 64 poses. compute_f32 evaluates the factor
  // bodies in FP32 (values, Hessian blocks, the solve and the cost stay FP64).
  BatchLocalization(const symhip_problem_desc* descs, int32_t n_problems,
                    const symhip_optimizer_params& params, bool compute_f32, double epsilon);
  ~BatchLocalization();
  BatchLocalization(const BatchLocalization&) = delete;
  BatchLocalization& operator=(const BatchLocalization&) = delete;

  // One solve of every problem from its initial Values. Outputs may be null: values_out gets
  // each problem's Values storage with the best poses (concatenated in problem order),
  // iters_out params.iterations + 1 records per problem (rows past num_iterations unused),
  // solve_ms the per-problem wall time inside the launch.
  void solve(double* values_out, symhip_optimization_stats* stats_out,
             symhip_iteration_stats* iters_out, double* solve_ms, double* total_ms);
  // Times every launch configuration (occupancy: one or two wavefronts per SIMD; problem order:
  // as given or most poses first) over `reps` complete solves each and keeps the fastest for
  // later solves; config_ms (kConfigs entries, may be null) receives each one's best time.
  static constexpr int kConfigs = 4;
  void tune(int reps, double* config_ms);
  int config() const { return config_; }
  void set_config(int c) {
    SYMHIP_CHECK(c >= 0 && c < kConfigs, "batch: launch configuration out of range");
    build_graph(c);
    config_ = c;
  }
  int32_t n_problems() const { return n_problems_; }
  int64_t storage_total() const { return (int64_t)cvals_host_size_; }

 private:
  int32_t n_problems_ = 0, total_poses_ = 0;
  symhip_optimizer_params p_;
  bool f32_ = false;
  double eps_ = 0;
  size_t cvals_host_size_ = 0;
  hipStream_t stream_ = nullptr;
  hipGraphExec_t graphs_[kConfigs] = {};
  int config_ = 0;
  DBuf<int32_t> order_;
  hipEvent_t ev0_ = nullptr, ev1_ = nullptr;
  DBuf<int32_t> pose_begin_, pose_soff_, match_ptr_, match_args_, odo_args_;
  DBuf<double> cvals_, vout_, vals_, lin_, maxd_, err_;
  DBuf<symhip_iteration_stats> recs_;
  DBuf<symhip_optimization_stats> stats_;
  DBuf<int64_t> ticks_;
  void build_graph(int c);
};

}  // namespace symhip
