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
#include <cstdio>
#include <cstring>
#include <limits>
#include <string>

#include "./factor_traits.h"
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
  double zero_diag;        // entries of the damped diagonal below epsilon (CheckHessianDiagonal)
  double nonfinite_update; // non-finite entries of the update (debug_checks)
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

// Key as the reference prints it (opt/key.cc:23-50): letter, then _sub and _super when valid, a
// negative index written with an n.
std::string key_string(int8_t letter, int64_t sub, int64_t super) {
  std::string s(1, (char)letter);
  for (int64_t v : {sub, super}) {
    if (v == std::numeric_limits<int64_t>::min()) continue;
    s += '_';
    if (v < 0) s += 'n';
    s += std::to_string(v < 0 ? -v : v);
  }
  return s;
}
std::string fmt_double(double v) {
  if (std::isnan(v)) return "nan";
  char b[32];
  std::snprintf(b, sizeof(b), "%.6g", v);
  return b;
}

// RelaxDampingToInitial (levenberg_marquardt_solver.h:171-174).
__global__ void lm_relax_kernel(LmDevState* st, double initial_lambda, double beta) {
  st->lambda = fmin(st->lambda, initial_lambda);
  st->nu = beta;
}

__global__ void set_value_kernel(double* p, double v) { *p = v; }

// debug_checks: is the factor's linearization finite (CheckLinearizedFactor,
// opt/internal/linearizer_utils.h:307-367; J finite makes its Hessian and rhs finite)?
template <int Code>
__device__ bool factor_finite(const double* const* a) {
  using F = FT<Code>;
  double r[F::D];
  double J[F::D][F::T];
  F::template eval<double>(a, r, J);
  bool ok = true;
  for (int i = 0; i < F::D; ++i) {
    ok = ok && isfinite(r[i]);
    for (int c = 0; c < F::T; ++c) ok = ok && isfinite(J[i][c]);
  }
  return ok;
}
template <int Code>
__device__ bool factor_finite_upto(int code, const double* const* a) {
  if constexpr (Code >= kNumFactorTypes) {
    return true;
  } else {
    return code == Code ? factor_finite<Code>(a) : factor_finite_upto<Code + 1>(code, a);
  }
}
// check_derivatives (internal/derivative_checker.h:32-73, NumericalDerivative with delta =
// sqrt(eps) through the retraction): the factor's analytic Jacobian against central differences
// of its residual, as three sums of squares (difference, analytic, numerical) for the whole
// problem's Frobenius-norm comparison.
template <int Code>
__device__ void factor_derivative_sums(const double* const* a, const int32_t* atype, double eps,
                                       double* out) {
  using F = FT<Code>;
  double r[F::D], J[F::D][F::T], rp[F::D], rm[F::D], Jx[F::D][F::T];
  F::template eval<double>(a, r, J);
  const double delta = sqrt(eps);
  double sd = 0, sj = 0, sn = 0;
  for (int li = 0; li < F::NL; ++li) {
    const int td = F::td(li);
    const int sdim = atype[li] == SYMHIP_TYPE_POSE3 ? 7 : atype[li] == SYMHIP_TYPE_ROT3 ? 4 : td;
    for (int k = 0; k < td; ++k) {
      double vp[8], vm[8], dp[6] = {0, 0, 0, 0, 0, 0}, dm[6] = {0, 0, 0, 0, 0, 0};
      for (int s = 0; s < sdim; ++s) vp[s] = vm[s] = a[li][s];
      dp[k] = delta;
      dm[k] = -delta;
      retract_value(atype[li], vp, dp, eps);
      retract_value(atype[li], vm, dm, eps);
      const double* ap[6];
      for (int q = 0; q < 6; ++q) ap[q] = a[q];
      ap[li] = vp;
      F::template eval<double>(ap, rp, Jx);
      ap[li] = vm;
      F::template eval<double>(ap, rm, Jx);
      const int c = F::foff(li) + k;
      for (int i = 0; i < F::D; ++i) {
        const double n = (rp[i] - rm[i]) / (2 * delta);
        sd += (J[i][c] - n) * (J[i][c] - n);
        sj += J[i][c] * J[i][c];
        sn += n * n;
      }
    }
  }
  out[0] = sd;
  out[1] = sj;
  out[2] = sn;
}
template <int Code>
__device__ void factor_derivative_upto(int code, const double* const* a, const int32_t* atype,
                                       double eps, double* out) {
  if constexpr (Code < kNumFactorTypes) {
    if (code == Code) factor_derivative_sums<Code>(a, atype, eps, out);
    else factor_derivative_upto<Code + 1>(code, a, atype, eps, out);
  }
}
__global__ void factor_derivative_kernel(int32_t n, const int32_t* __restrict__ type,
                                         const int32_t* __restrict__ ptr,
                                         const int64_t* __restrict__ off,
                                         const int32_t* __restrict__ atype,
                                         const double* __restrict__ values, double eps,
                                         double* __restrict__ sums) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= n) return;
  const double* a[6];
  const int32_t p0 = ptr[f], na = ptr[f + 1] - p0;
  for (int k = 0; k < 6; ++k) a[k] = values + off[p0 + (k < na ? k : 0)];
  factor_derivative_upto<1>(type[f], a, atype + p0, eps, sums + 3 * (int64_t)f);
}

// include_jacobians: each factor's Jacobian into the sparse Jacobian's values (the columns of
// its optimized arguments; starts as built by build_jacobian_structure).
template <int Code>
__device__ void factor_jacobian_into(const double* const* a, const int64_t* starts,
                                     double* jval) {
  using F = FT<Code>;
  double r[F::D], J[F::D][F::T];
  F::template eval<double>(a, r, J);
  for (int t = 0; t < F::T; ++t) {
    const int64_t q = starts[t];
    if (q < 0) continue;
    for (int i = 0; i < F::D; ++i) jval[q + i] = J[i][t];
  }
}
template <int Code>
__device__ void factor_jacobian_upto(int code, const double* const* a, const int64_t* starts,
                                     double* jval) {
  if constexpr (Code < kNumFactorTypes) {
    if (code == Code) factor_jacobian_into<Code>(a, starts, jval);
    else factor_jacobian_upto<Code + 1>(code, a, starts, jval);
  }
}
__global__ void jacobian_fill_kernel(int32_t n, const int32_t* __restrict__ type,
                                     const int32_t* __restrict__ ptr,
                                     const int64_t* __restrict__ off,
                                     const int32_t* __restrict__ sptr,
                                     const int64_t* __restrict__ starts,
                                     const double* __restrict__ values, double* __restrict__ jval) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= n) return;
  const double* a[6];
  const int32_t p0 = ptr[f], na = ptr[f + 1] - p0;
  for (int k = 0; k < 6; ++k) a[k] = values + off[p0 + (k < na ? k : 0)];
  factor_jacobian_upto<1>(type[f], a, starts + sptr[f], jval);
}

// One thread per factor: bad[f] = 1 when factor f linearizes to a non-finite value at `values`.
__global__ void factor_check_kernel(int32_t n, const int32_t* __restrict__ type,
                                    const int32_t* __restrict__ ptr,
                                    const int64_t* __restrict__ off,
                                    const double* __restrict__ values, int32_t* __restrict__ bad) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= n) return;
  const double* a[6];
  const int32_t p0 = ptr[f], na = ptr[f + 1] - p0;
  for (int k = 0; k < 6; ++k) a[k] = values + off[p0 + (k < na ? k : 0)];
  bad[f] = factor_finite_upto<1>(type[f], a) ? 0 : 1;
}

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

// sum_i u_i (rhs_i - damping_i u_i)  (Linearization::LinearDeltaError, linearization.h:62-67),
// and in the same pass the two counts of the iteration's checks: entries of the damped diagonal
// (undamped + damping) below eps in magnitude (CheckHessianDiagonal, lm.tcc:56-90) and
// non-finite entries of the update (lm.tcc:225-227). Partials: part[b], part[nb + b],
// part[2 nb + b].
__global__ void __launch_bounds__(256) lde_partial_kernel(const double* __restrict__ u,
                                                          const double* __restrict__ rhs,
                                                          const double* __restrict__ damp,
                                                          const double* __restrict__ und,
                                                          double eps, int64_t n,
                                                          double* __restrict__ part) {
  __shared__ double ws[3][4];
  double acc = 0, zeros = 0, bad = 0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double ui = u[i], di = damp[i];
    acc += ui * (rhs[i] - di * ui);
    zeros += fabs(und[i] + di) < eps ? 1.0 : 0.0;
    bad += isfinite(ui) ? 0.0 : 1.0;
  }
  for (int o = 32; o > 0; o >>= 1) {
    acc += __shfl_down(acc, o, 64);
    zeros += __shfl_down(zeros, o, 64);
    bad += __shfl_down(bad, o, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    ws[0][threadIdx.x >> 6] = acc;
    ws[1][threadIdx.x >> 6] = zeros;
    ws[2][threadIdx.x >> 6] = bad;
  }
  __syncthreads();
  if (threadIdx.x < 3)
    part[threadIdx.x * gridDim.x + blockIdx.x] =
        ws[threadIdx.x][0] + ws[threadIdx.x][1] + ws[threadIdx.x][2] + ws[threadIdx.x][3];
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

// out = b - A x (b non-null) or out = A x, for the symmetric A held as its lower CSC (colptr,
// rowind, vals): row i of A is column i's entries (rows >= i) and the entries (i, j), j < i,
// which tptr / tcol / tidx list per row (column j and the entry's index in vals). One thread
// per row, FP64, no atomics: the residuals and products of the mixed-precision form's
// refinement.
__global__ void __launch_bounds__(256) sym_matvec_kernel(
    int32_t n, const int32_t* __restrict__ colptr, const int32_t* __restrict__ rowind,
    const double* __restrict__ vals, const int32_t* __restrict__ tptr,
    const int32_t* __restrict__ tcol, const int64_t* __restrict__ tidx,
    const double* __restrict__ b, const double* __restrict__ x, double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double acc = 0.0;
  for (int32_t e = colptr[i]; e < colptr[i + 1]; ++e) acc += vals[e] * x[rowind[e]];
  for (int32_t e = tptr[i]; e < tptr[i + 1]; ++e) acc += vals[tidx[e]] * x[tcol[e]];
  out[i] = b ? b[i] - acc : acc;
}

// CG step with the scalars on the device: alpha = rz / pq; x += alpha p; r -= alpha q.
__global__ void cg_step_kernel(int32_t n, const double* __restrict__ rz,
                               const double* __restrict__ pq, const double* __restrict__ p,
                               const double* __restrict__ q, double* __restrict__ x,
                               double* __restrict__ r) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double alpha = *rz / *pq;
  x[i] += alpha * p[i];
  r[i] -= alpha * q[i];
}

// p = z + beta p with beta = rz_new / rz_old (p = z when rz_old is null).
__global__ void cg_direction_kernel(int32_t n, const double* __restrict__ rz_new,
                                    const double* __restrict__ rz_old,
                                    const double* __restrict__ z, double* __restrict__ p) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  p[i] = rz_old ? z[i] + (*rz_new / *rz_old) * p[i] : z[i];
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
    if (o.precision == SYMHIP_PRECISION_F32_UPDATES) setup_refinement(p, s);
  }
  ~CscBackend() override {
    if (host_sums_) (void)hipHostFree(host_sums_);
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
    if (chol_.f32_updates()) refine(b, s);
    hipLaunchKernelGGL(negate_kernel, dim3(grid_for(N_, 256)), dim3(256), 0, s, N_, x_.get(),
                       update);
  }
  // Mixed-precision form: out[0] solves, out[1] refinement (CG) iterations, out[2] fallbacks to
  // the FP64 factorization, out[3] most iterations in one solve.
  void linear_stats(int64_t* out4) const override {
    out4[0] = solves_;
    out4[1] = refine_steps_;
    out4[2] = fallbacks_;
    out4[3] = max_refine_steps_;
  }
  // With tracking on, every mixed-precision solve also records its final relative residual.
  int refinement_residuals(int track, double* out, int max, hipStream_t s) override {
    SYMHIP_CHECK(chol_.f32_updates(),
                 "symhip: refinement residuals need SYMHIP_PRECISION_F32_UPDATES");
    if (track >= 0) {
      track_ = track != 0;
      hist_.clear();
    }
    const int n = std::min<int>((int)hist_.size(), max);
    if (out) std::copy(hist_.begin(), hist_.begin() + n, out);
    return (int)hist_.size();
  }
  const double* rhs(int b) const override { return rhs_[b].get(); }
  const double* residual(int b) const override { return residual_[b].get(); }
  const double* damping() const override { return damping_.get(); }
  const double* undamped_diagonal() const override { return undamped_.get(); }
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
  // Refinement of x = A^-1 b against the FP64 damped Hessian A: conjugate gradients
  // preconditioned by the mixed factor, from the mixed solve's x, with FP64 products, residuals
  // and scalars, until |b - A x| <= kRefineTol |b| (at most kRefineMax iterations). Plain
  // refinement (x += M^-1 (b - A x)) contracts only while cond(A) times the FP32 error of the
  // updates is below one; CG on M^-1 A still converges where it is not. Each iteration reads
  // two scalars on the host (one small copy and a stream synchronization). CG stops at the
  // tolerance or where it stagnates (the FP64 floor of an ill-conditioned system, ~1e-11 here
  // for the FP64 factorization too); a breakdown, or a true residual above kFallbackTol at the
  // end, refactors A with FP64 updates and solves directly (counted as a fallback).
  void refine(int b, hipStream_t s) {
    const dim3 g(grid_for(N_, 256)), blk(256);
    double *x = x_.get(), *r = r_.get(), *z = dx_.get(), *p = p_.get(), *q = q_.get();
    double* sc = sums_.get();  // 0: b.b, 1: r.r, 2: p.q, 3 + (k & 1): r.z of iteration k
    auto matvec = [&](const double* rhs, const double* v, double* out) {
      hipLaunchKernelGGL(sym_matvec_kernel, g, blk, 0, s, N_, colptr_.get(), rowind_.get(),
                         hval_[b].get(), tptr_.get(), tcol_.get(), tidx_.get(), rhs, v, out);
    };
    auto read = [&]() {
      HIP_OK(hipMemcpyAsync(host_sums_, sc, 5 * sizeof(double), hipMemcpyDeviceToHost, s));
      HIP_OK(hipStreamSynchronize(s));
      return host_sums_[0] > 0 ? std::sqrt(host_sums_[1] / host_sums_[0])
                               : std::sqrt(host_sums_[1]);
    };
    dot(rhs_[b].get(), rhs_[b].get(), N_, partials_.get(), sc, s);
    matvec(rhs_[b].get(), x, r);
    dot(r, r, N_, partials_.get(), sc + 1, s);
    double rel = read();
    int k = 0;
    bool ok = !(rel > kRefineTol);
    if (!ok) {
      chol_.solve(r, z, s);
      dot(r, z, N_, partials_.get(), sc + 3, s);
      hipLaunchKernelGGL(cg_direction_kernel, g, blk, 0, s, N_, sc + 3, nullptr, z, p);
      double best = rel;
      int since_best = 0;
      for (k = 1; k <= kRefineMax; ++k) {
        double* rz = sc + 3 + ((k - 1) & 1);
        matvec(nullptr, p, q);
        dot(p, q, N_, partials_.get(), sc + 2, s);
        hipLaunchKernelGGL(cg_step_kernel, g, blk, 0, s, N_, rz, sc + 2, p, q, x, r);
        dot(r, r, N_, partials_.get(), sc + 1, s);
        rel = read();
        if (!(host_sums_[2] > 0) || !std::isfinite(rel)) break;  // breakdown
        // converged, or stagnating at the FP64 floor of this system (no halving of the best
        // residual in 3 iterations; accepted below kFallbackTol)
        since_best = rel < 0.5 * best ? 0 : since_best + 1;
        best = std::min(best, rel);
        if (rel <= kRefineTol || since_best == 3) {
          ok = true;
          break;
        }
        if (k == kRefineMax) break;
        chol_.solve(r, z, s);
        double* rz_new = sc + 3 + (k & 1);
        dot(r, z, N_, partials_.get(), rz_new, s);
        hipLaunchKernelGGL(cg_direction_kernel, g, blk, 0, s, N_, rz_new, rz, z, p);
      }
      k = std::min(k, kRefineMax);
      // the recurrence's residual drifts from the true one: check the true residual
      matvec(rhs_[b].get(), x, r);
      dot(r, r, N_, partials_.get(), sc + 1, s);
      rel = read();
      ok = ok && rel <= kFallbackTol;
    }
    if (!ok) {
      chol_.set_f32_updates(false);
      chol_.factorize(hval_[b].get(), s, rhs_[b].get());
      chol_.solve_fused(x, s);
      chol_.set_f32_updates(true);
      matvec(rhs_[b].get(), x, r);
      dot(r, r, N_, partials_.get(), sc + 1, s);
      rel = read();
      ++fallbacks_;
    }
    ++solves_;
    refine_steps_ += k;
    max_refine_steps_ = std::max<int64_t>(max_refine_steps_, k);
    if (track_ && hist_.size() < 65536) hist_.push_back(rel);
  }
  static constexpr int kRefineMax = 30;
  static constexpr double kRefineTol = 1e-13;
  static constexpr double kFallbackTol = 1e-10;
  // The row lists of the strictly lower entries (for sym_matvec_kernel) and the refinement
  // work vectors; the factorization switches to FP32 updates.
  void setup_refinement(const HostProblem& p, hipStream_t s) {
    std::vector<int32_t> tptr(p.N + 1, 0), tcol;
    std::vector<int64_t> tidx;
    for (int j = 0; j < p.N; ++j)
      for (int32_t e = p.colptr[j]; e < p.colptr[j + 1]; ++e)
        if (p.rowind[e] > j) ++tptr[p.rowind[e] + 1];
    for (int i = 0; i < p.N; ++i) tptr[i + 1] += tptr[i];
    tcol.resize(std::max(1, tptr[p.N]));
    tidx.resize(std::max(1, tptr[p.N]));
    std::vector<int32_t> fill(tptr.begin(), tptr.end() - 1);
    for (int j = 0; j < p.N; ++j)
      for (int32_t e = p.colptr[j]; e < p.colptr[j + 1]; ++e)
        if (p.rowind[e] > j) {
          const int32_t q = fill[p.rowind[e]]++;
          tcol[q] = j;
          tidx[q] = e;
        }
    rowind_.upload(p.rowind, s);
    tptr_.upload(tptr, s);
    tcol_.upload(tcol, s);
    tidx_.upload(tidx, s);
    r_.resize(std::max(1, p.N));
    dx_.resize(std::max(1, p.N));
    p_.resize(std::max(1, p.N));
    q_.resize(std::max(1, p.N));
    sums_.resize(5);
    HIP_OK(hipHostMalloc((void**)&host_sums_, 5 * sizeof(double), hipHostMallocDefault));
    chol_.set_f32_updates(true);
    HIP_OK(hipStreamSynchronize(s));
  }

  CscLinearizer lin_;
  SupernodalLDLT chol_;
  DBuf<double> r_, dx_, p_, q_, sums_;
  double* host_sums_ = nullptr;  // pinned
  DBuf<int32_t> rowind_, tptr_, tcol_;
  DBuf<int64_t> tidx_;
  int64_t solves_ = 0, refine_steps_ = 0, max_refine_steps_ = 0, fallbacks_ = 0;
  bool track_ = false;
  std::vector<double> hist_;
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
  // Optimizer ctor (optimizer.tcc:39): the derivative check compares against the Jacobian
  SYMHIP_CHECK(!p.check_derivatives || p.include_jacobians,
               "check_derivatives requires include_jacobians");
  SYMHIP_CHECK(o.precision != SYMHIP_PRECISION_F32_JAC || o.linear_solver == kSolverSchur,
               "symhip: SYMHIP_PRECISION_F32_JAC is available on the Schur path only");
  SYMHIP_CHECK(o.precision != SYMHIP_PRECISION_F32_UPDATES || o.linear_solver != kSolverSchur,
               "symhip: SYMHIP_PRECISION_F32_UPDATES is available on the sparse Cholesky path "
               "only");
  SYMHIP_CHECK(o.precision >= SYMHIP_PRECISION_F64 && o.precision <= SYMHIP_PRECISION_F32_UPDATES,
               "symhip: unknown precision");
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
  partials_.resize(3 * kReduceBlocks);
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
  {
    PhaseScope ps(&timer_, SYMHIP_PHASE_LINEARIZE, stream_);
    backend_->relinearize(b, values_[b].get(), &dstate_->err[b], stream_);
  }
  linearized_[b] = true;
  if (p_.include_jacobians) fill_jacobian(b);
  if (p_.debug_checks) check_factors(b);
  if (p_.check_derivatives) check_derivatives(b);
}

// The iteration's checks, from the counts the linear-error pass left in the device state:
// CheckHessianDiagonal (lm.tcc:56-90; the indices are found on the host, only when there are
// any), the non-finite update under debug_checks (:225-227) and the non-finite new error
// (:257-262).
void DeviceOptimizer::check_iteration() {
  const symhip_iteration_stats& rec = hstate_->rec;
  if (p_.verbose) {
    // PopulateIterationStats under verbose (lm.tcc:105-113)
    const double prev = hstate_->err[init_];
    const double gain = (prev - rec.new_error) / (prev - rec.new_error_linear);
    std::fprintf(stderr,
                 "[symhip] LM<symhip> [iter %4d] lambda: %.3e, error prev/linear/new: "
                 "%.3e/%.3e/%.3e, rel reduction: %.5e, gain ratio: %.5e\n",
                 rec.iteration, rec.current_lambda, prev, rec.new_error_linear, rec.new_error,
                 rec.relative_reduction, gain);
  }
  if (hstate_->zero_diag > 0) {
    const int32_t N = host_.N;
    std::vector<double> u(N), d(N);
    HIP_OK(hipMemcpyAsync(u.data(), backend_->undamped_diagonal(), sizeof(double) * N,
                          hipMemcpyDeviceToHost, stream_));
    HIP_OK(hipMemcpyAsync(d.data(), backend_->damping(), sizeof(double) * N,
                          hipMemcpyDeviceToHost, stream_));
    HIP_OK(hipStreamSynchronize(stream_));
    std::vector<int32_t> idx;
    for (int32_t i = 0; i < N; ++i)
      if (std::fabs(u[i] + d[i]) < epsilon_) idx.push_back(i);
    char head[160];
    std::snprintf(head, sizeof(head),
                  "LM<symhip> Zero on diagonal after damping (with lambda = %.2e, epsilon = "
                  "%.2e) at indices: ",
                  rec.current_lambda, epsilon_);
    constexpr size_t kShow = 15;
    std::string m = std::string(head) + "[";
    for (size_t i = 0; i < std::min(kShow, idx.size()); ++i)
      m += (i ? ", " : "") + std::to_string(idx[i]);
    if (idx.size() > kShow) m += ", ... (" + std::to_string(idx.size() - kShow) + " omitted)";
    warn(m + "]");
  }
  if (p_.debug_checks && hstate_->nonfinite_update > 0)
    warn("LM<symhip> Non-finite update: " + std::to_string((int64_t)hstate_->nonfinite_update) +
         " of " + std::to_string(host_.N) + " entries");
  if (!std::isfinite(rec.new_error)) {
    warn("LM<symhip> Encountered non-finite error: " + fmt_double(rec.new_error));
    if (!p_.debug_checks)
      warn("LM<symhip> Turn on debug_checks to see which factor is causing this");
  }
}

void DeviceOptimizer::warn(const std::string& msg) {
  std::fprintf(stderr, "[symhip] %s\n", msg.c_str());
  warnings_.push_back(msg);
}

const JacobianStructure& DeviceOptimizer::jacobian_structure() {
  if (!have_jac_structure_) {
    build_jacobian_structure(host_, jac_);
    have_jac_structure_ = true;
  }
  return jac_;
}

void DeviceOptimizer::fill_jacobian(int b) {
  const int32_t nf = host_.n_factors;
  if (nf == 0) return;
  upload_check_tables();
  if (jstart_.size() == 0) {
    const JacobianStructure& j = jacobian_structure();
    jstart_.upload(j.starts, stream_);
    jsptr_.upload(j.sptr, stream_);
    for (auto& v : jval_) v.resize(std::max<size_t>(1, j.rowind.size()));
  }
  hipLaunchKernelGGL(jacobian_fill_kernel, dim3(grid_for(nf, 64)), dim3(64), 0, stream_, nf,
                     check_type_.get(), check_ptr_.get(), check_off_.get(), jsptr_.get(),
                     jstart_.get(), values_[b].get(), jval_[b].get());
}

void DeviceOptimizer::best_jacobian(double* values) {
  SYMHIP_CHECK(p_.include_jacobians && jstart_.size() != 0,
               "no Jacobian: create the optimizer with include_jacobians and optimize first");
  jval_[best_].download(values, jac_.rowind.size(), stream_);
  HIP_OK(hipStreamSynchronize(stream_));
}

void DeviceOptimizer::upload_check_tables() {
  if (check_type_.size() != 0) return;
  std::vector<int64_t> off;
  std::vector<int32_t> at;
  for (int32_t k : host_.fargs) {
    off.push_back(host_.storage_offset[k]);
    at.push_back(host_.types[k]);
  }
  check_type_.upload(host_.ftype, stream_);
  check_ptr_.upload(host_.fargs_ptr, stream_);
  check_off_.upload(off, stream_);
  check_atype_.upload(at, stream_);
  check_bad_.resize(std::max(1, host_.n_factors));
}

// check_derivatives: the Jacobian of the whole problem at block b's Values against its central
// differences, Eigen's isApprox with precision 10 sqrt(eps) (|J - Jn| <= p min(|J|, |Jn|),
// Frobenius norms; derivative_checker.h:53-58), a failure reported with the worst factors as
// the reference's assertion fails the optimization (optimizer.tcc:261-270).
void DeviceOptimizer::check_derivatives(int b) {
  const int32_t nf = host_.n_factors;
  if (nf == 0) return;
  upload_check_tables();
  DBuf<double> sums;
  sums.resize((size_t)3 * nf);
  hipLaunchKernelGGL(factor_derivative_kernel, dim3(grid_for(nf, 64)), dim3(64), 0, stream_, nf,
                     check_type_.get(), check_ptr_.get(), check_off_.get(), check_atype_.get(),
                     values_[b].get(), epsilon_, sums.get());
  std::vector<double> s((size_t)3 * nf);
  sums.download(s.data(), s.size(), stream_);
  HIP_OK(hipStreamSynchronize(stream_));
  double sd = 0, sj = 0, sn = 0;
  for (int32_t f = 0; f < nf; ++f) {
    sd += s[3 * f];
    sj += s[3 * f + 1];
    sn += s[3 * f + 2];
  }
  const double prec = 10 * std::sqrt(epsilon_);
  if (std::sqrt(sd) <= prec * std::min(std::sqrt(sj), std::sqrt(sn))) return;
  std::vector<int32_t> worst(nf);
  for (int32_t f = 0; f < nf; ++f) worst[f] = f;
  const size_t k = std::min<size_t>(5, worst.size());
  std::partial_sort(worst.begin(), worst.begin() + k, worst.end(),
                    [&](int32_t x, int32_t y) { return s[3 * x] > s[3 * y]; });
  std::string m = "check_derivatives: symbolic and numerical jacobians don't match (|J - Jn| = " +
                  fmt_double(std::sqrt(sd)) + ", |J| = " + fmt_double(std::sqrt(sj)) +
                  ", |Jn| = " + fmt_double(std::sqrt(sn)) + "); largest differences in factors";
  for (size_t i = 0; i < k; ++i)
    m += " " + std::to_string(worst[i]) + " (" + fmt_double(std::sqrt(s[3 * worst[i]])) + ")";
  throw Error(m);
}

// debug_checks: the factors whose linearization at block b's Values is not finite, reported as
// the reference's linearizer reports them (opt/internal/linearizer_utils.h:307-367): the factor,
// each argument's key, offset, size and value, and its residual.
void DeviceOptimizer::check_factors(int b) {
  const int32_t nf = host_.n_factors;
  if (nf == 0) return;
  upload_check_tables();
  hipLaunchKernelGGL(factor_check_kernel, dim3(grid_for(nf, 128)), dim3(128), 0, stream_, nf,
                     check_type_.get(), check_ptr_.get(), check_off_.get(), values_[b].get(),
                     check_bad_.get());
  std::vector<int32_t> bad(nf);
  check_bad_.download(bad.data(), nf, stream_);
  HIP_OK(hipStreamSynchronize(stream_));
  std::vector<double> vals, res;
  for (int32_t f = 0; f < nf; ++f) {
    if (!bad[f]) continue;
    if (vals.empty()) {
      vals.resize(host_.storage_total);
      res.resize(host_.M);
      values_[b].download(vals.data(), vals.size(), stream_);
      if (host_.M)
        HIP_OK(hipMemcpyAsync(res.data(), backend_->residual(b), sizeof(double) * host_.M,
                              hipMemcpyDeviceToHost, stream_));
      HIP_OK(hipStreamSynchronize(stream_));
    }
    const FactorSpec sp = factor_spec(host_.ftype[f]);
    std::string m = "LM<symhip> Non-finite linearization for factor " + std::to_string(f) +
                    " (type " + std::to_string(host_.ftype[f]) + "):";
    for (int32_t q = host_.fargs_ptr[f]; q < host_.fargs_ptr[f + 1]; ++q) {
      const int32_t k = host_.fargs[q];
      const int32_t o = host_.storage_offset[k], w = storage_dim(host_.types[k]);
      m += "\n  " + key_string(host_.letters[k], host_.subs[k], host_.supers[k]) +
           " (offset=" + std::to_string(o) + ", size=" + std::to_string(w) + ") = [";
      for (int32_t i = 0; i < w; ++i) m += (i ? ", " : "") + fmt_double(vals[o + i]);
      m += "]";
    }
    m += "\nResidual: [";
    for (int i = 0; i < sp.res_dim; ++i)
      m += (i ? ", " : "") + fmt_double(res[host_.res_off[f] + i]);
    m += "]";
    warn(m);
  }
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
  warnings_.clear();
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
  if (p_.include_jacobians && jstart_.size() != 0) {
    r.jacobian.resize(jac_.rowind.size());
    jval_[block].download(r.jacobian.data(), r.jacobian.size(), stream_);
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
    if (!std::isfinite(hstate_->rec.new_error)) {
      warn("LM<symhip> Encountered non-finite initial error: " +
           fmt_double(hstate_->rec.new_error));
      if (!p_.debug_checks)
        warn("LM<symhip> Turn on debug_checks to see which factor is causing this");
    }
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
                       backend_->rhs(init_), backend_->damping(), backend_->undamped_diagonal(),
                       epsilon_, (int64_t)N, partials_.get());
    hipLaunchKernelGGL(sum_into_kernel, dim3(1), dim3(256), 0, stream_, partials_.get(), nb,
                       &dstate_->lde);
    hipLaunchKernelGGL(sum_into_kernel, dim3(1), dim3(256), 0, stream_, partials_.get() + nb, nb,
                       &dstate_->zero_diag);
    hipLaunchKernelGGL(sum_into_kernel, dim3(1), dim3(256), 0, stream_,
                       partials_.get() + 2 * nb, nb, &dstate_->nonfinite_update);
  }
  if (p_.enable_bold_updates && have_last_update_) {
    dot(last_update_.get(), last_update_.get(), N, partials_.get(), &dstate_->nl2, stream_);
    dot(update_.get(), update_.get(), N, partials_.get(), &dstate_->nu2, stream_);
    dot(last_update_.get(), update_.get(), N, partials_.get(), &dstate_->dotlu, stream_);
  }
  // An inexact (truncated PCG) step whose reduction falls below the early-exit threshold switches
  // the solver to exact steps instead of ending the run; the early exit applies to those.
  symhip_optimizer_params pd = p_;
  const bool deferred = backend_->defers_early_exit();
  if (deferred) pd.early_exit_min_reduction = 0.0;
  hipLaunchKernelGGL(lm_decide_kernel, dim3(1), dim3(1), 0, stream_, dstate_, init_, new_, best_,
                     iteration_, pd, have_last_update_ ? 1 : 0, epsilon_);
  timer_.end(decide_ev, stream_);
  HIP_OK(hipMemcpyAsync(hstate_, dstate_, sizeof(LmDevState), hipMemcpyDeviceToHost, stream_));
  HIP_OK(hipStreamSynchronize(stream_));
  timer_.collect();
  HIP_OK(hipGetLastError());
  iters.push_back(hstate_->rec);
  check_iteration();
  if (p_.debug_stats) record_debug(new_, true);
  backend_->observe(hstate_->rec);
  if (deferred) {
    const double rel = hstate_->rec.relative_reduction, ee = p_.early_exit_min_reduction;
    if (rel > -ee / 10 && rel < ee) backend_->switch_to_exact();
    // At the switch the damping that made the inexact steps acceptable is relaxed to the initial
    // one (lambda = min(lambda, initial), nu = beta): it says nothing about exact steps, and a
    // lambda grown there would make the first exact steps small enough for the early exit.
    if (!backend_->defers_early_exit())
      hipLaunchKernelGGL(lm_relax_kernel, dim3(1), dim3(1), 0, stream_, dstate_,
                         p_.initial_lambda, p_.dynamic_lambda_update_beta);
  }
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
