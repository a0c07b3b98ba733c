// Device-resident Levenberg-Marquardt optimizer (sym::Optimizer + LevenbergMarquardtSolver).
#pragma once

#include <chrono>
#include <memory>
#include <string>
#include <vector>

#include "./chol_numeric.h"
#include "./common.h"
#include "./linearize.h"
#include "./problem.h"
#include "./timer.h"

namespace symhip {

enum LinearSolverKind { kSolverCholesky = 0, kSolverSchur = 1 };

struct SolverOptions {
  int linear_solver = kSolverCholesky;
  int ordering = 1;  // 0 natural, 1 minimum degree
  int precision = 0; // 0 FP64, 1 FP32 Jacobians with FP64 residual/cost (fast form)
  int use_pcg = 0;   // Schur path: preconditioned CG on the reduced camera system
  int pcg_max_iters = 100;
  double pcg_rel_tol = 1e-6;
  int pcg_preconditioner = 0;  // 0 block Jacobi, 1 learned correction
  double pcg_exact_below_reduction = 0;
  double pcg_forcing = 0;
};

// A linear system backend owns the linearization storage of the three LM state blocks and the
// linear solver. The LM logic (optimizer.hip) is independent of the storage format.
class SystemBackend {
 public:
  virtual ~SystemBackend() = default;
  // Relinearize block b at device values; writes residual(b), the system and err (device scalar).
  virtual void relinearize(int b, const double* values, double* err_out, hipStream_t s) = 0;
  // Damp block b's Hessian diagonal (DampHessian, levenberg_marquardt_solver.tcc:22-54); writes
  // damping (N) and keeps the undamped diagonal for restore().
  virtual void damp(int b, const double* lambda_dev, const symhip_optimizer_params& p,
                    bool have_max_diagonal, hipStream_t s) = 0;
  virtual void restore(int b, hipStream_t s) = 0;
  // update = -H^-1 rhs for block b (factorize + solve), in the reference state order.
  virtual void solve(int b, double* update, hipStream_t s) = 0;
  virtual const double* rhs(int b) const = 0;
  virtual const double* residual(int b) const = 0;
  virtual const double* damping() const = 0;
  // The Hessian diagonal before the last damp() (N, reference state order).
  virtual const double* undamped_diagonal() const = 0;
  // Hessian in reference CSC order (for best-linearization and parity output); may be null.
  virtual void hessian_csc(int b, double* out_host, hipStream_t s) = 0;
  virtual int32_t N() const = 0;
  virtual int32_t M() const = 0;
  virtual int64_t nnz() const = 0;
  // out[0]=factorized dim, out[1]=supernodes, out[2]=panel doubles, out[3]=levels, out[4]=flops,
  // out[5]=RCS blocks, out[6]=observation pairs, out[7]=observations, out[8]=diagonal-block
  // and panel flops, out[9]=trailing-update flops
  virtual void info(int64_t* out10) const = 0;
  // Iterative linear solver statistics: out[0]=solves, out[1]=total iterations, out[2]=fallbacks
  // to the direct factorization, out[3]=most iterations in one solve (zeros for direct solvers).
  virtual void linear_stats(int64_t* out4) const { out4[0] = out4[1] = out4[2] = out4[3] = 0; }
  // The linear solver's permutation (Permutation().indices(): old -> new) and the reference's
  // strictly lower pattern of its factor L (debug_stats; sparse Cholesky path only, as the
  // reference's SparseSchurSolver exposes neither).
  virtual void factor_sparsity(std::vector<int32_t>& ordering, std::vector<int32_t>& l_colptr,
                               std::vector<int32_t>& l_rowind) const {
    throw Error("symhip: the factor sparsity is available on the sparse Cholesky path only");
  }
  // Mixed-precision sparse Cholesky only: track >= 0 turns the recording of each solve's final
  // relative residual on (1) or off (0) and clears the record; returns the number recorded and
  // copies up to max of them into out (when non-null).
  virtual int refinement_residuals(int track, double* out, int max, hipStream_t s) {
    throw Error("symhip: refinement residuals need SYMHIP_PRECISION_F32_UPDATES");
  }
  // True while the solver's steps are inexact and it will switch to exact ones (truncated PCG
  // with an exact switch): a small relative reduction then signals the switch (observe), not
  // convergence, and the LM's early exit on it is deferred to the exact steps.
  virtual bool defers_early_exit() const { return false; }
  // Called when a deferred early exit would have ended the run: switch to the exact steps.
  virtual void switch_to_exact() {}
  // Called at the start of every optimize run and after every LM iteration record.
  virtual void begin_run() {}
  virtual void observe(const symhip_iteration_stats& rec) { (void)rec; }
  // Learned PCG preconditioner (Schur path with PCG only; the defaults reject the call).
  // Logging keeps copies of the next max reduced camera systems (full block storage) and their
  // right-hand sides as they are assembled.
  virtual void log_systems(int max) { unsupported(); }
  virtual int logged_systems() const { return 0; }
  // Block pattern of the reduced camera system: returns the node count; fills row_ptr (n + 1)
  // and col (row_ptr[n]) when non-null.
  virtual int rcs_pattern(int32_t* row_ptr, int32_t* col) const { return unsupported(), 0; }
  virtual void logged_system(int idx, double* blocks, double* rhs) const { unsupported(); }
  virtual void precond_params(float* get, const float* set) { unsupported(); }
  // Loss and gradient of the network on logged system idx for host probes X (P columns).
  virtual float precond_loss_grad(int idx, const double* X, int P, float* grad) {
    return unsupported(), 0.0f;
  }
  virtual void train_precond(int steps, float lr, int P, uint64_t seed, int n_systems,
                             float* losses) {
    unsupported();
  }
  // PCG on logged system idx (its own rhs) with block Jacobi (learned == 0) or the network:
  // iterations to rel_tol and the relative error of the step against the supernodal LDL^T solve.
  virtual void eval_precond(int idx, double rel_tol, int max_iters, int learned, int* iters,
                            double* step_err) {
    unsupported();
  }
  void set_timer(PhaseTimer* t) { timer_ = t; }

 protected:
  [[noreturn]] static void unsupported() {
    throw Error("symhip: the learned preconditioner needs the Schur solver with use_pcg");
  }
  PhaseTimer* timer_ = nullptr;
};

std::unique_ptr<SystemBackend> make_csc_backend(const HostProblem& p, const SolverOptions& o,
                                                hipStream_t s);
std::unique_ptr<SystemBackend> make_schur_backend(const HostProblem& p, const SolverOptions& o,
                                                  hipStream_t s);

struct LmDevState;  // device-side scalars (optimizer.hip)

class DeviceOptimizer {
 public:
  DeviceOptimizer(const symhip_problem_desc* d, const symhip_optimizer_params& p,
                  const SolverOptions& o, double epsilon);
  ~DeviceOptimizer();
  // Optimizer::Optimize (opt/optimizer.tcc:70-89 + internal/optimizer_utils.h:36-104).
  void optimize(double* values_inout, int32_t num_iterations, bool populate_best,
                std::vector<symhip_iteration_stats>& iters, symhip_optimization_stats& stats);
  // Same, with Values already on the device (benchmark path: no host transfers in the loop).
  void optimize_device(int32_t num_iterations, std::vector<symhip_iteration_stats>& iters,
                       symhip_optimization_stats& stats);
  // GncOptimizer::Optimize (opt/gnc_optimizer.h:49-118): stages of Levenberg-Marquardt to
  // convergence with the scalar key mu_key (a constant of the factors) stepped from mu_initial to
  // mu_max between them; the iteration records and the budget run across the stages.
  void optimize_gnc(double* values_inout, int32_t num_iterations, int32_t mu_key,
                    const symhip_gnc_params& g, std::vector<symhip_iteration_stats>& iters,
                    symhip_optimization_stats& stats);
  void upload_values(const double* values);
  void download_best_values(double* values);
  void best_linearization(double* residual, double* hval, double* rhs);
  void update_params(const symhip_optimizer_params& p) {
    SYMHIP_CHECK(!p.check_derivatives || p.include_jacobians,
                 "check_derivatives requires include_jacobians");
    p_ = p;
  }
  const HostProblem& host() const { return host_; }
  const symhip_optimizer_params& params() const { return p_; }
  SystemBackend& backend() { return *backend_; }
  hipStream_t stream() const { return stream_; }
  double last_iteration_ms() const { return last_iter_ms_; }
  PhaseTimer& timer() { return timer_; }
  const std::vector<double>& iteration_times() const { return iter_times_; }
  // debug_stats vectors of the last optimize run, one per iteration record (record 0 is the
  // iteration -1 record and has no update): the update, the New block's Values storage and its
  // residual (PopulateIterationStats, levenberg_marquardt_solver.tcc:115-122).
  struct DebugRecord {
    std::vector<double> update, values, residual, jacobian;
  };
  // include_jacobians: the Jacobian's sparsity (stats.jacobian_sparsity, lm.tcc:172-175) and the
  // values of the best block's Jacobian (Linearization::jacobian of GetBestLinearization).
  const JacobianStructure& jacobian_structure();
  void best_jacobian(double* values);
  const std::vector<DebugRecord>& debug_records() const { return debug_; }
  // Warnings of the last optimize run, in the reference's wording (spdlog::warn in
  // levenberg_marquardt_solver.tcc:56-90, 180-183, 225-227, 257-262 and, under debug_checks,
  // opt/internal/linearizer_utils.h:307-367); each is also written to stderr.
  const std::vector<std::string>& warnings() const { return warnings_; }

 private:
  bool iterate(std::vector<symhip_iteration_stats>& iters, int32_t& best_index, int& status,
               int& reason);
  // LevenbergMarquardtSolver::Reset (levenberg_marquardt_solver.h:163-168) with the start Values.
  void reset_run();
  // IterateToConvergenceImpl (opt/internal/optimizer_utils.h:36-78) minus the Values copy.
  void iterate_to_convergence(int32_t n, std::vector<symhip_iteration_stats>& iters,
                              int32_t& best_index, int& status, int& reason);
  void record_debug(int block, bool with_update);
  void warn(const std::string& msg);
  void check_factors(int block);
  void check_derivatives(int block);
  void upload_check_tables();
  void check_iteration();
  std::vector<std::string> warnings_;
  DBuf<int32_t> check_type_, check_ptr_, check_bad_, check_atype_;
  DBuf<int64_t> check_off_;
  void fill_jacobian(int block);
  bool have_jac_structure_ = false;
  JacobianStructure jac_;
  DBuf<double> jval_[3];
  DBuf<int64_t> jstart_;
  DBuf<int32_t> jsptr_;
  std::vector<DebugRecord> debug_;
  void relinearize(int b);

  HostProblem host_;
  symhip_optimizer_params p_;
  SolverOptions o_;
  double epsilon_;
  hipStream_t stream_ = nullptr;
  std::unique_ptr<SystemBackend> backend_;
  DBuf<double> values_[3];
  DBuf<double> update_, last_update_, partials_;
  DBuf<int32_t> retract_type_, retract_soff_, retract_toff_;
  LmDevState* dstate_ = nullptr;  // device
  LmDevState* hstate_ = nullptr;  // pinned host mirror
  int init_ = 0, new_ = 1, best_ = 0, free_ = 2;
  bool valid_[3] = {false, false, false};
  bool linearized_[3] = {false, false, false};
  bool have_max_diagonal_ = false, have_last_update_ = false;
  int iteration_ = -1;
  double last_iter_ms_ = 0;
  std::chrono::steady_clock::time_point run_start_;
  std::vector<double> iter_times_;
  DBuf<double> start_values_;
  bool have_start_ = false;
  PhaseTimer timer_;
};

}  // namespace symhip
