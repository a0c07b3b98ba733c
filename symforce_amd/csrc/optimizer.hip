// Device-resident Levenberg-Marquardt optimizer.
//
// Follows LevenbergMarquardtSolver::Iterate (opt/levenberg_marquardt_solver.tcc:138-343) and the
// three-block state of LevenbergMarquardtStateBase (opt/internal/levenberg_marquardt_state.h:
// 33-270). Damping, the linear solve, the retraction, the cost / linear-error reductions and the
// accept / lambda decision all run on the device; the host reads back one small record per
// iteration (the iteration statistics) and mirrors the state-block bookkeeping from it.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstring>

#include "./geo.h"
#include "./lm_logic.h"
#include "./optimizer.h"
#include "./util.h"

namespace symhip {

struct LmDevState {
  double lambda, nu;
  double err[3];
  double lde;
  double nl2, nu2, dotlu;
  symhip_iteration_stats rec;
  int32_t status, reason, accept, set_best;
};

namespace {

__global__ void lm_reset_kernel(LmDevState* st, double lambda, double nu) {
  st->lambda = lambda;
  st->nu = nu;
  st->status = 0;
  st->reason = 0;
  st->accept = 0;
  st->set_best = 0;
}

// RelaxDampingToInitial (levenberg_marquardt_solver.h:171-174).
__global__ void lm_relax_kernel(LmDevState* st, double initial_lambda, double beta) {
  st->lambda = fmin(st->lambda, initial_lambda);
  st->nu = beta;
}

__global__ void set_value_kernel(double* p, double v) { *p = v; }

// Record the iteration -1 statistics (levenberg_marquardt_solver.tcc:158-186).
__global__ void lm_first_record_kernel(LmDevState* st, int init) {
  st->rec = symhip_iteration_stats{};
  st->rec.iteration = -1;
  st->rec.new_error = st->err[init];
  st->rec.current_lambda = st->lambda;
  st->status = isfinite(st->err[init]) ? 0 : SYMHIP_STATUS_FAILED;
  st->reason = isfinite(st->err[init]) ? 0 : SYMHIP_FAILURE_INITIAL_ERROR_NOT_FINITE;
}

// The accept / reject / lambda logic of Iterate (levenberg_marquardt_solver.tcc:239-340).
__global__ void lm_decide_kernel(LmDevState* st, int init, int nw, int best, int iteration,
                                 symhip_optimizer_params p, int have_last_update, double eps) {
  const LmDecision d = lm_decide(st->err[init], st->err[nw], st->err[best], st->lde, st->lambda,
                                 st->nu, iteration, p, have_last_update, st->nl2, st->nu2,
                                 st->dotlu, eps);
  st->rec = d.rec;
  st->lambda = d.lambda;
  st->nu = d.nu;
  st->status = d.status;
  st->reason = d.reason;
  st->accept = d.accept;
  st->set_best = d.set_best;
}

// Values::Retract over the optimized keys (opt/values.cc:314-327).
__global__ void retract_kernel(int32_t n, const int32_t* __restrict__ type,
                               const int32_t* __restrict__ soff, const int32_t* __restrict__ toff,
                               const double* __restrict__ update, double* __restrict__ values,
                               double eps) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  retract_value(type[i], values + soff[i], update + toff[i], eps);
}

// sum_i u_i (rhs_i - damping_i u_i)  (Linearization::LinearDeltaError, linearization.h:62-67)
__global__ void __launch_bounds__(256) lde_partial_kernel(const double* __restrict__ u,
                                                          const double* __restrict__ rhs,
                                                          const double* __restrict__ damp,
                                                          int64_t n, double* __restrict__ part) {
  __shared__ double ws[4];
  double acc = 0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    acc += u[i] * (rhs[i] - damp[i] * u[i]);
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o, 64);
  if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = ws[0] + ws[1] + ws[2] + ws[3];
}
__global__ void __launch_bounds__(256) sum_into_kernel(const double* __restrict__ part, int nb,
                                                       double* __restrict__ out) {
  __shared__ double ws[4];
  double acc = 0;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) acc += part[i];
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o, 64);
  if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) *out = ws[0] + ws[1] + ws[2] + ws[3];
}

// ---- CSC backend kernels ----
__global__ void damp_csc_kernel(int32_t n, const int32_t* __restrict__ colptr,
                                double* __restrict__ hval, const double* __restrict__ lambda_p,
                                int use_diag, int keep_max, int have_max, double dmin, int unit,
                                double* __restrict__ undamped, double* __restrict__ damping,
                                double* __restrict__ max_diag) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const double lambda = *lambda_p;
  const double u = hval[colptr[j]];
  undamped[j] = u;
  double d = 0;
  if (use_diag) {
    if (keep_max) {
      const double md = have_max ? fmax(max_diag[j], u) : fmax(u, dmin);
      max_diag[j] = md;
      d = md * lambda;
    } else {
      d = fmax(u, dmin) * lambda;
    }
  }
  if (unit) d += lambda;
  damping[j] = d;
  hval[colptr[j]] = u + d;
}
__global__ void restore_csc_kernel(int32_t n, const int32_t* __restrict__ colptr,
                                   double* __restrict__ hval, const double* __restrict__ undamped) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n) hval[colptr[j]] = undamped[j];
}
__global__ void negate_kernel(int32_t n, const double* __restrict__ x, double* __restrict__ y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = -x[i];
}

}  // namespace

// -------------------------------------------------------------------------------------------
// CSC backend: reference hessian_lower layout + supernodal LDL^T
// -------------------------------------------------------------------------------------------
class CscBackend final : public SystemBackend {
 public:
  CscBackend(const HostProblem& p, const SolverOptions& o, hipStream_t s) {
    lin_.setup(p, s);
    for (int b = 0; b < 3; ++b) {
      residual_[b].resize(std::max(1, p.M));
      hval_[b].resize(std::max<int64_t>(1, p.nnz()));
      rhs_[b].resize(std::max(1, p.N));
    }
    undamped_.resize(std::max(1, p.N));
    damping_.resize(std::max(1, p.N));
    max_diag_.resize(std::max(1, p.N));
    x_.resize(std::max(1, p.N));
    partials_.resize(kReduceBlocks);
    colptr_.upload(p.colptr, s);
    colptr_host_ = p.colptr;
    rowind_host_ = p.rowind;
    // Symbolic analysis (the reference's AnalyzeSparsityPattern, done once): nodes are the
    // optimized keys.
    std::vector<int32_t> node_ptr{0};
    for (int32_t k : p.state_keys) node_ptr.push_back(node_ptr.back() + tangent_dim(p.types[k]));
    CholSymbolic sym;
    chol_symbolic(p.N, p.colptr.data(), p.rowind.data(), node_ptr,
                  ordering_from_code(o.ordering), nullptr, 64, sym);
    chol_.setup(sym, p.colptr.data(), p.rowind.data(), s);
    N_ = p.N;
    M_ = p.M;
    nnz_ = p.nnz();
  }
  void relinearize(int b, const double* values, double* err_out, hipStream_t s) override {
    lin_.run(values, residual_[b].get(), hval_[b].get(), rhs_[b].get(), s);
    half_squared_norm(residual_[b].get(), M_, partials_.get(), err_out, s);
  }
  void damp(int b, const double* lambda_dev, const symhip_optimizer_params& p,
            bool have_max_diagonal, hipStream_t s) override {
    hipLaunchKernelGGL(damp_csc_kernel, dim3(grid_for(N_, 256)), dim3(256), 0, s, N_,
                       colptr_.get(), hval_[b].get(), lambda_dev, p.use_diagonal_damping,
                       p.keep_max_diagonal_damping, have_max_diagonal ? 1 : 0,
                       p.diagonal_damping_min, p.use_unit_damping, undamped_.get(),
                       damping_.get(), max_diag_.get());
  }
  void restore(int b, hipStream_t s) override {
    hipLaunchKernelGGL(restore_csc_kernel, dim3(grid_for(N_, 256)), dim3(256), 0, s, N_,
                       colptr_.get(), hval_[b].get(), undamped_.get());
  }
  void solve(int b, double* update, hipStream_t s) override {
    {
      PhaseScope ps(timer_, SYMHIP_PHASE_FACTOR, s);
      // the forward solve runs inside the factorization
      chol_.factorize(hval_[b].get(), s, rhs_[b].get());
    }
    PhaseScope ps(timer_, SYMHIP_PHASE_SOLVE, s);
    chol_.solve_fused(x_.get(), s);
    hipLaunchKernelGGL(negate_kernel, dim3(grid_for(N_, 256)), dim3(256), 0, s, N_, x_.get(),
                       update);
  }
  const double* rhs(int b) const override { return rhs_[b].get(); }
  const double* residual(int b) const override { return residual_[b].get(); }
  const double* damping() const override { return damping_.get(); }
  void hessian_csc(int b, double* out, hipStream_t s) override {
    hval_[b].download(out, nnz_, s);
  }
  void factor_sparsity(std::vector<int32_t>& ordering, std::vector<int32_t>& l_colptr,
                       std::vector<int32_t>& l_rowind) const override {
    const CholSymbolic& y = chol_.symbolic();
    ordering = y.perm;
    scalar_factor_pattern(N_, colptr_host_.data(), rowind_host_.data(), y.perm, l_colptr,
                          l_rowind);
  }
  int32_t N() const override { return N_; }
  int32_t M() const override { return M_; }
  int64_t nnz() const override { return nnz_; }
  SupernodalLDLT& chol() { return chol_; }
  void info(int64_t* out) const override {
    const CholSymbolic& y = chol_.symbolic();
    out[0] = y.n;
    out[1] = y.n_snodes;
    out[2] = y.panel_size();
    out[3] = y.n_levels;
    out[4] = (int64_t)y.factor_flops;
    out[8] = (int64_t)(y.factor_flops - y.update_flops);
    out[9] = (int64_t)y.update_flops;
    out[5] = out[6] = out[7] = 0;
  }

 private:
  CscLinearizer lin_;
  SupernodalLDLT chol_;
  DBuf<double> residual_[3], hval_[3], rhs_[3];
  DBuf<double> undamped_, damping_, max_diag_, x_, partials_;
  DBuf<int32_t> colptr_;
  std::vector<int32_t> colptr_host_, rowind_host_;
  int32_t N_ = 0, M_ = 0;
  int64_t nnz_ = 0;
};

std::unique_ptr<SystemBackend> make_csc_backend(const HostProblem& p, const SolverOptions& o,
                                                hipStream_t s) {
  return std::make_unique<CscBackend>(p, o, s);
}

// -------------------------------------------------------------------------------------------
// DeviceOptimizer
// -------------------------------------------------------------------------------------------
DeviceOptimizer::DeviceOptimizer(const symhip_problem_desc* d, const symhip_optimizer_params& p,
                                 const SolverOptions& o, double epsilon)
    : p_(p), o_(o), epsilon_(epsilon) {
  SYMHIP_CHECK(o.precision == SYMHIP_PRECISION_F64 ||
                   (o.precision == SYMHIP_PRECISION_F32_JAC && o.linear_solver == kSolverSchur),
               "symhip: SYMHIP_PRECISION_F32_JAC is available on the Schur path only");
  SYMHIP_CHECK(!o.use_pcg || o.linear_solver == kSolverSchur,
               "symhip: PCG is available on the Schur path only");
  build_host_problem(d, host_);
  HIP_OK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
  backend_ = o.linear_solver == kSolverSchur ? make_schur_backend(host_, o, stream_)
                                             : make_csc_backend(host_, o, stream_);
  backend_->set_timer(&timer_);
  for (auto& v : values_) v.resize(std::max<int64_t>(1, host_.storage_total));
  update_.resize(std::max(1, host_.N));
  last_update_.resize(std::max(1, host_.N));
  partials_.resize(kReduceBlocks);
  std::vector<int32_t> rt, rs, rto;
  for (int32_t k : host_.state_keys) {
    rt.push_back(host_.types[k]);
    rs.push_back(host_.storage_offset[k]);
    rto.push_back(host_.tan_off[k]);
  }
  retract_type_.upload(rt, stream_);
  retract_soff_.upload(rs, stream_);
  retract_toff_.upload(rto, stream_);
  HIP_OK(hipMalloc(&dstate_, sizeof(LmDevState)));
  HIP_OK(hipHostMalloc(&hstate_, sizeof(LmDevState), hipHostMallocDefault));
  std::memset(hstate_, 0, sizeof(LmDevState));
  upload_values(host_.values.data());
}

DeviceOptimizer::~DeviceOptimizer() {
  if (dstate_) (void)hipFree(dstate_);
  if (hstate_) (void)hipHostFree(hstate_);
  backend_.reset();
  if (stream_) (void)hipStreamDestroy(stream_);
}

void DeviceOptimizer::upload_values(const double* values) {
  start_values_.upload(values, host_.storage_total, stream_);
  have_start_ = true;
  HIP_OK(hipStreamSynchronize(stream_));
}

void DeviceOptimizer::download_best_values(double* values) {
  values_[best_].download(values, host_.storage_total, stream_);
  HIP_OK(hipStreamSynchronize(stream_));
}

void DeviceOptimizer::best_linearization(double* residual, double* hval, double* rhs) {
  if (residual) HIP_OK(hipMemcpyAsync(residual, backend_->residual(best_), sizeof(double) * host_.M,
                                      hipMemcpyDeviceToHost, stream_));
  if (rhs) HIP_OK(hipMemcpyAsync(rhs, backend_->rhs(best_), sizeof(double) * host_.N,
                                 hipMemcpyDeviceToHost, stream_));
  if (hval) backend_->hessian_csc(best_, hval, stream_);
  HIP_OK(hipStreamSynchronize(stream_));
}

void DeviceOptimizer::relinearize(int b) {
  PhaseScope ps(&timer_, SYMHIP_PHASE_LINEARIZE, stream_);
  backend_->relinearize(b, values_[b].get(), &dstate_->err[b], stream_);
  linearized_[b] = true;
}

void DeviceOptimizer::optimize(double* values_inout, int32_t num_iterations, bool populate_best,
                               std::vector<symhip_iteration_stats>& iters,
                               symhip_optimization_stats& stats) {
  // Reset: the input Values go into the New block (LevenbergMarquardtState::ResetValues).
  init_ = 0;
  new_ = 1;
  best_ = 0;
  free_ = 2;
  upload_values(values_inout);
  optimize_device(num_iterations, iters, stats);
  download_best_values(values_inout);
  (void)populate_best;
}

void DeviceOptimizer::reset_run() {
  SYMHIP_CHECK(have_start_, "optimize called before the Values were set");
  HIP_OK(hipMemcpyAsync(values_[1].get(), start_values_.get(), sizeof(double) * host_.storage_total,
                        hipMemcpyDeviceToDevice, stream_));
  init_ = 0;
  new_ = 1;
  best_ = 0;
  free_ = 2;
  iteration_ = -1;
  have_max_diagonal_ = false;
  have_last_update_ = false;
  for (int b = 0; b < 3; ++b) linearized_[b] = false;
  valid_[new_] = true;
  valid_[init_] = valid_[free_] = false;
  hipLaunchKernelGGL(lm_reset_kernel, dim3(1), dim3(1), 0, stream_, dstate_, p_.initial_lambda,
                     p_.dynamic_lambda_update_beta);
  backend_->begin_run();
  debug_.clear();
  iter_times_.clear();
  run_start_ = std::chrono::steady_clock::now();
}

void DeviceOptimizer::iterate_to_convergence(int32_t n, std::vector<symhip_iteration_stats>& iters,
                                             int32_t& best_index, int& status, int& reason) {
  SYMHIP_CHECK(n > 0, "num_iterations must be positive");
  int i;
  for (i = 0; i < n; ++i) {
    const auto t0 = std::chrono::steady_clock::now();
    const size_t before = iters.size();
    const bool done = iterate(iters, best_index, status, reason);
    const auto t1 = std::chrono::steady_clock::now();
    last_iter_ms_ = std::chrono::duration<double, std::milli>(t1 - t0).count();
    const double since = std::chrono::duration<double, std::milli>(t1 - run_start_).count();
    for (size_t k = before; k < iters.size(); ++k) iter_times_.push_back(since);
    if (done) break;
  }
  if (i == n) {
    status = SYMHIP_STATUS_HIT_ITERATION_LIMIT;
    reason = 0;
  }
}

void DeviceOptimizer::optimize_device(int32_t num_iterations,
                                      std::vector<symhip_iteration_stats>& iters,
                                      symhip_optimization_stats& stats) {
  const int32_t n = num_iterations < 0 ? p_.iterations : num_iterations;
  SYMHIP_CHECK(n > 0, "num_iterations must be positive");
  // Reset (levenberg_marquardt_solver.h:163-183): the start Values go into the New block.
  reset_run();
  iters.clear();
  int32_t best_index = 0;
  int status = 0, reason = 0;
  iterate_to_convergence(n, iters, best_index, status, reason);
  stats.num_iterations = (int32_t)iters.size();
  stats.best_index = best_index;
  stats.status = status;
  stats.failure_reason = reason;
}

void DeviceOptimizer::optimize_gnc(double* values_inout, int32_t num_iterations, int32_t mu_key,
                                   const symhip_gnc_params& g,
                                   std::vector<symhip_iteration_stats>& iters,
                                   symhip_optimization_stats& stats) {
  SYMHIP_CHECK(mu_key >= 0 && mu_key < host_.n_keys, "gnc: mu key index out of range");
  SYMHIP_CHECK(host_.types[mu_key] == SYMHIP_TYPE_SCALAR, "gnc: the mu key must be a Scalar");
  SYMHIP_CHECK(host_.tan_off[mu_key] < 0, "gnc: the mu key must not be optimized");
  const int32_t n = num_iterations < 0 ? p_.iterations : num_iterations;
  SYMHIP_CHECK(n > 0, "num_iterations must be positive");
  const int64_t mu_off = host_.storage_offset[mu_key];
  bool updating = g.mu_initial < g.mu_max && g.mu_step > 0.0;
  values_inout[mu_off] = g.mu_initial;
  const symhip_optimizer_params user = p_;
  if (updating) p_.early_exit_min_reduction = g.gnc_update_min_reduction;
  upload_values(values_inout);
  init_ = 0;
  new_ = 1;
  best_ = 0;
  free_ = 2;
  reset_run();
  iters.clear();
  int32_t best_index = 0;
  int status = 0, reason = 0;
  try {
    iterate_to_convergence(n, iters, best_index, status, reason);
    double mu = g.mu_initial;  // the stage's best Values hold it
    while ((int32_t)iters.size() < n) {
      if (status != SYMHIP_STATUS_SUCCESS || !updating) break;
      mu += g.mu_step;
      // RelaxDampingToInitial (levenberg_marquardt_solver.h:171-174)
      hipLaunchKernelGGL(lm_relax_kernel, dim3(1), dim3(1), 0, stream_, dstate_,
                         user.initial_lambda, user.dynamic_lambda_update_beta);
      if (mu >= g.mu_max) {
        mu = g.mu_max;
        p_.early_exit_min_reduction = user.early_exit_min_reduction;
        updating = false;
      }
      // OptimizeContinue: ResetState(best Values with the new mu) (:121-131,
      // levenberg_marquardt_solver.h:178-183): the Values go into the New block, every
      // linearization is dropped, lambda and the iteration count stay.
      if (best_ != new_)
        HIP_OK(hipMemcpyAsync(values_[new_].get(), values_[best_].get(),
                              sizeof(double) * host_.storage_total, hipMemcpyDeviceToDevice,
                              stream_));
      hipLaunchKernelGGL(set_value_kernel, dim3(1), dim3(1), 0, stream_,
                         values_[new_].get() + mu_off, mu);
      have_max_diagonal_ = false;
      have_last_update_ = false;
      for (int b = 0; b < 3; ++b) linearized_[b] = false;
      valid_[new_] = true;
      valid_[init_] = valid_[free_] = false;
      iterate_to_convergence(n - (int32_t)iters.size(), iters, best_index, status, reason);
    }
  } catch (...) {
    p_ = user;
    throw;
  }
  p_ = user;
  download_best_values(values_inout);
  stats.num_iterations = (int32_t)iters.size();
  stats.best_index = best_index;
  stats.status = status;
  stats.failure_reason = reason;
}

void DeviceOptimizer::record_debug(int block, bool with_update) {
  DebugRecord r;
  r.values.resize(host_.storage_total);
  r.residual.resize(host_.M);
  values_[block].download(r.values.data(), r.values.size(), stream_);
  if (host_.M)
    HIP_OK(hipMemcpyAsync(r.residual.data(), backend_->residual(block), sizeof(double) * host_.M,
                          hipMemcpyDeviceToHost, stream_));
  if (with_update) {
    r.update.resize(host_.N);
    update_.download(r.update.data(), r.update.size(), stream_);
  }
  HIP_OK(hipStreamSynchronize(stream_));
  debug_.push_back(std::move(r));
}

bool DeviceOptimizer::iterate(std::vector<symhip_iteration_stats>& iters, int32_t& best_index,
                              int& status, int& reason) {
  std::swap(init_, new_);  // state_.Step()
  iteration_++;
  if (!linearized_[init_]) {
    relinearize(init_);
    // SetBestToInit
    if (best_ != init_) {
      if (best_ != new_) free_ = best_;
      best_ = init_;
    }
  }
  if (iteration_ == 0) {
    hipLaunchKernelGGL(lm_first_record_kernel, dim3(1), dim3(1), 0, stream_, dstate_, init_);
    HIP_OK(hipMemcpyAsync(hstate_, dstate_, sizeof(LmDevState), hipMemcpyDeviceToHost, stream_));
    HIP_OK(hipStreamSynchronize(stream_));
    iters.push_back(hstate_->rec);
    if (p_.debug_stats) record_debug(init_, false);
    if (hstate_->status == SYMHIP_STATUS_FAILED) {
      status = hstate_->status;
      reason = hstate_->reason;
      return true;
    }
  }
  const int32_t N = host_.N;
  {
    PhaseScope ps(&timer_, SYMHIP_PHASE_DAMP, stream_);
    backend_->damp(init_, &dstate_->lambda, p_, have_max_diagonal_, stream_);
  }
  if (p_.use_diagonal_damping && p_.keep_max_diagonal_damping) have_max_diagonal_ = true;
  backend_->solve(init_, update_.get(), stream_);
  backend_->restore(init_, stream_);
  // UpdateNewFromInit: copy Init values into New, then retract the optimized keys.
  {
    PhaseScope ps(&timer_, SYMHIP_PHASE_RETRACT, stream_);
    HIP_OK(hipMemcpyAsync(values_[new_].get(), values_[init_].get(),
                          sizeof(double) * host_.storage_total, hipMemcpyDeviceToDevice,
                          stream_));
    valid_[new_] = true;
    const int32_t nk = (int32_t)host_.state_keys.size();
    hipLaunchKernelGGL(retract_kernel, dim3(grid_for(nk, 256)), dim3(256), 0, stream_, nk,
                       retract_type_.get(), retract_soff_.get(), retract_toff_.get(),
                       update_.get(), values_[new_].get(), epsilon_);
  }
  relinearize(new_);
  const int decide_ev = timer_.begin(SYMHIP_PHASE_DECIDE, stream_);
  {
    const int nb = std::max(1, std::min(kReduceBlocks, grid_for(N, 256)));
    hipLaunchKernelGGL(lde_partial_kernel, dim3(nb), dim3(256), 0, stream_, update_.get(),
                       backend_->rhs(init_), backend_->damping(), (int64_t)N, partials_.get());
    hipLaunchKernelGGL(sum_into_kernel, dim3(1), dim3(256), 0, stream_, partials_.get(), nb,
                       &dstate_->lde);
  }
  if (p_.enable_bold_updates && have_last_update_) {
    dot(last_update_.get(), last_update_.get(), N, partials_.get(), &dstate_->nl2, stream_);
    dot(update_.get(), update_.get(), N, partials_.get(), &dstate_->nu2, stream_);
    dot(last_update_.get(), update_.get(), N, partials_.get(), &dstate_->dotlu, stream_);
  }
  hipLaunchKernelGGL(lm_decide_kernel, dim3(1), dim3(1), 0, stream_, dstate_, init_, new_, best_,
                     iteration_, p_, have_last_update_ ? 1 : 0, epsilon_);
  timer_.end(decide_ev, stream_);
  HIP_OK(hipMemcpyAsync(hstate_, dstate_, sizeof(LmDevState), hipMemcpyDeviceToHost, stream_));
  HIP_OK(hipStreamSynchronize(stream_));
  timer_.collect();
  HIP_OK(hipGetLastError());
  iters.push_back(hstate_->rec);
  if (p_.debug_stats) record_debug(new_, true);
  backend_->observe(hstate_->rec);
  if (!hstate_->accept) {
    std::swap(new_, init_);  // SwapNewAndInit
  } else {
    have_last_update_ = true;
    HIP_OK(hipMemcpyAsync(last_update_.get(), update_.get(), sizeof(double) * N,
                          hipMemcpyDeviceToDevice, stream_));
    if (hstate_->set_best) {
      // SetBestToNew
      if (best_ != new_) {
        if (best_ != init_) free_ = best_;
        best_ = new_;
      }
      best_index = (int32_t)iters.size() - 1;
    }
    // SetInitToNotBest
    if (best_ == init_) {
      init_ = free_;
      free_ = best_;
    }
  }
  status = hstate_->status;
  reason = hstate_->reason;
  return status != 0;
}

}  // namespace symhip
