// Batched Levenberg-Marquardt for many small robot-3D-localization problems (Pose3 chains with
// odometry factors between consecutive poses and landmark matching factors) in one persistent
// kernel: one wavefront runs the whole LM loop of one problem.
#pragma once

#include <vector>

#include "./common.h"

namespace symhip {

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
  int32_t n_problems() const { return n_problems_; }
  int64_t storage_total() const { return (int64_t)cvals_host_size_; }

 private:
  int32_t n_problems_ = 0, total_poses_ = 0;
  symhip_optimizer_params p_;
  bool f32_ = false;
  double eps_ = 0;
  size_t cvals_host_size_ = 0;
  hipStream_t stream_ = nullptr;
  hipGraphExec_t graph_ = nullptr;
  hipEvent_t ev0_ = nullptr, ev1_ = nullptr;
  DBuf<int32_t> pose_begin_, pose_soff_, match_ptr_, match_args_, odo_args_;
  DBuf<int64_t> storage_begin_;
  DBuf<double> cvals_, vout_, vals_, lin_, maxd_, err_;
  DBuf<symhip_iteration_stats> recs_;
  DBuf<symhip_optimization_stats> stats_;
  DBuf<int64_t> ticks_;
  void build_graph();
};

}  // namespace symhip
