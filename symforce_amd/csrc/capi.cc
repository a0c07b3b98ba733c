// extern "C" entry points of libsymhip (include/symhip.h).
#include <hip/hip_runtime.h>

#include <cstring>
#include <memory>
#include <string>

#include "../../include/symhip.h"
#include "./common.h"
#include "./chol_numeric.h"
#include "./linearize.h"
#include "./batch.h"
#include "./factor_traits.h"
#include "./geo.h"
#include "./lm_logic.h"
#include "./optimizer.h"
#include "./problem.h"

namespace {
thread_local std::string g_last_error;

template <typename F>
int guarded(F&& f) {
  try {
    f();
    return 0;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return -1;
  } catch (...) {
    g_last_error = "unknown error";
    return -1;
  }
}
}  // namespace

struct symhip_problem {
  symhip::HostProblem host;
  hipStream_t stream = nullptr;
  symhip::CscLinearizer lin;
  symhip::DBuf<double> values, residual, hval, rhs;
  bool device_ready = false;
  // Device state is created on first use so the structure queries work without a GPU.
  void ensure_device() {
    if (device_ready) return;
    HIP_OK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    lin.setup(host, stream);
    device_ready = true;
  }
  ~symhip_problem() {
    if (stream) (void)hipStreamDestroy(stream);
  }
};

struct symhip_optimizer {
  std::unique_ptr<symhip::DeviceOptimizer> opt;
  std::vector<symhip_iteration_stats> iters;
};

struct symhip_cholesky {
  symhip::CholSymbolic sym;
  std::vector<int32_t> colptr, rowind;
  hipStream_t stream = nullptr;
  symhip::SupernodalLDLT ldlt;
  symhip::DBuf<double> vals, b, x;
  ~symhip_cholesky() {
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {
void copy_stats(const std::vector<symhip_iteration_stats>& v, symhip_iteration_stats* out,
                int32_t max_iters) {
  if (!out) return;
  for (size_t i = 0; i < v.size() && (int32_t)i < max_iters; ++i) out[i] = v[i];
}
}  // namespace

extern "C" {

const char* symhip_last_error(void) { return g_last_error.c_str(); }

int symhip_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

int symhip_problem_create(const symhip_problem_desc* desc, symhip_problem_t* out) {
  return guarded([&] {
    SYMHIP_CHECK(out != nullptr, "null output handle");
    auto p = std::make_unique<symhip_problem>();
    symhip::build_host_problem(desc, p->host);
    *out = p.release();
  });
}

int symhip_factor_info(int32_t code, int32_t* out) {
  return guarded([&] {
    SYMHIP_CHECK(out, "invalid arguments");
    symhip::dispatch_factor_type(code, [&]<int C>() {
      using F = symhip::FT<C>;
      out[0] = F::NA;
      out[1] = F::NL;
      out[2] = F::D;
      out[3] = F::T;
    });
  });
}

int symhip_factor_linearize_host(int32_t code, const double* const* args, int32_t precision,
                                 double* residual, double* jacobian) {
  return guarded([&] {
    SYMHIP_CHECK(args && residual && jacobian, "invalid arguments");
    SYMHIP_CHECK(precision == SYMHIP_PRECISION_F64 || precision == SYMHIP_PRECISION_F32_JAC,
                 "symhip: precision must be SYMHIP_PRECISION_F64 or SYMHIP_PRECISION_F32_JAC");
    symhip::dispatch_factor_type(code, [&]<int C>() {
      using F = symhip::FT<C>;
      auto run = [&](auto zero) {
        using S = decltype(zero);
        S r[F::D], J[F::D][F::T];
        F::template eval<S>(args, r, J);
        for (int i = 0; i < F::D; ++i) {
          residual[i] = double(r[i]);
          for (int c = 0; c < F::T; ++c) jacobian[i * F::T + c] = double(J[i][c]);
        }
      };
      if (precision == SYMHIP_PRECISION_F64)
        run(0.0);
      else
        run(0.0f);
    });
  });
}

int symhip_problem_destroy(symhip_problem_t p) {
  return guarded([&] { delete p; });
}

int symhip_problem_dims(symhip_problem_t p, int64_t* out) {
  return guarded([&] {
    out[0] = p->host.N;
    out[1] = p->host.M;
    out[2] = p->host.nnz();
    out[3] = p->host.storage_total;
    out[4] = (int64_t)p->host.state_keys.size();
  });
}

int symhip_problem_hessian_structure(symhip_problem_t p, int32_t* colptr, int32_t* rowind) {
  return guarded([&] {
    std::memcpy(colptr, p->host.colptr.data(), sizeof(int32_t) * p->host.colptr.size());
    std::memcpy(rowind, p->host.rowind.data(), sizeof(int32_t) * p->host.rowind.size());
  });
}

int64_t symhip_problem_col_starts(symhip_problem_t p, int32_t* out) {
  if (out) std::memcpy(out, p->host.col_starts.data(), sizeof(int32_t) * p->host.col_starts.size());
  return (int64_t)p->host.col_starts.size();
}

int symhip_problem_state_keys(symhip_problem_t p, int32_t* key_index, int32_t* tangent_offset) {
  return guarded([&] {
    for (size_t i = 0; i < p->host.state_keys.size(); ++i) {
      key_index[i] = p->host.state_keys[i];
      tangent_offset[i] = p->host.tan_off[p->host.state_keys[i]];
    }
  });
}

int symhip_linearize(symhip_problem_t p, const double* values, double* residual,
                     double* hessian_values, double* rhs) {
  return guarded([&] {
    p->ensure_device();
    const auto& h = p->host;
    p->values.upload(values, h.storage_total, p->stream);
    p->residual.resize(h.M);
    p->hval.resize(h.nnz());
    p->rhs.resize(h.N);
    p->lin.run(p->values.get(), p->residual.get(), p->hval.get(), p->rhs.get(), p->stream);
    if (residual) p->residual.download(residual, h.M, p->stream);
    if (hessian_values) p->hval.download(hessian_values, h.nnz(), p->stream);
    if (rhs) p->rhs.download(rhs, h.N, p->stream);
    HIP_OK(hipStreamSynchronize(p->stream));
  });
}

int symhip_optimizer_create(const symhip_problem_desc* desc, const symhip_optimizer_params* params,
                            const symhip_solver_options* options, double epsilon,
                            symhip_optimizer_t* out) {
  return guarded([&] {
    SYMHIP_CHECK(desc && params && out, "null argument");
    symhip::SolverOptions o;
    if (options) {
      o.linear_solver = options->linear_solver;
      o.ordering = options->ordering;
      o.precision = options->precision;
      o.use_pcg = options->use_pcg;
      o.pcg_max_iters = options->pcg_max_iters;
      o.pcg_rel_tol = options->pcg_rel_tol;
      o.pcg_preconditioner = options->pcg_preconditioner;
      o.pcg_exact_below_reduction = options->pcg_exact_below_reduction;
      o.pcg_forcing = options->pcg_forcing;
    }
    SYMHIP_CHECK(params->lambda_update_type == SYMHIP_LAMBDA_STATIC ||
                     params->lambda_update_type == SYMHIP_LAMBDA_DYNAMIC,
                 "Invalid lambda update type");
    auto h = std::make_unique<symhip_optimizer>();
    h->opt = std::make_unique<symhip::DeviceOptimizer>(desc, *params, o, epsilon);
    *out = h.release();
  });
}

int symhip_optimizer_destroy(symhip_optimizer_t o) {
  return guarded([&] { delete o; });
}

int symhip_optimizer_optimize(symhip_optimizer_t o, double* values, int32_t num_iterations,
                              symhip_iteration_stats* iters, int32_t max_iters,
                              symhip_optimization_stats* stats) {
  return guarded([&] {
    symhip_optimization_stats st{};
    o->opt->optimize(values, num_iterations, false, o->iters, st);
    copy_stats(o->iters, iters, max_iters);
    if (stats) *stats = st;
  });
}

int symhip_optimizer_optimize_gnc(symhip_optimizer_t o, double* values, int32_t num_iterations,
                                  const symhip_gnc_params* gnc, int32_t mu_key_index,
                                  symhip_iteration_stats* iters, int32_t max_iters,
                                  symhip_optimization_stats* stats) {
  return guarded([&] {
    SYMHIP_CHECK(gnc != nullptr && values != nullptr, "null argument");
    symhip_optimization_stats st{};
    o->opt->optimize_gnc(values, num_iterations, mu_key_index, *gnc, o->iters, st);
    copy_stats(o->iters, iters, max_iters);
    if (stats) *stats = st;
  });
}

int symhip_optimizer_best_linearization(symhip_optimizer_t o, double* residual,
                                        double* hessian_values, double* rhs) {
  return guarded([&] { o->opt->best_linearization(residual, hessian_values, rhs); });
}

int symhip_optimizer_update_params(symhip_optimizer_t o, const symhip_optimizer_params* p) {
  return guarded([&] { o->opt->update_params(*p); });
}

int symhip_optimizer_set_values(symhip_optimizer_t o, const double* values) {
  return guarded([&] { o->opt->upload_values(values); });
}

int symhip_optimizer_optimize_resident(symhip_optimizer_t o, int32_t num_iterations,
                                       symhip_iteration_stats* iters, int32_t max_iters,
                                       symhip_optimization_stats* stats) {
  return guarded([&] {
    symhip_optimization_stats st{};
    o->opt->optimize_device(num_iterations, o->iters, st);
    copy_stats(o->iters, iters, max_iters);
    if (stats) *stats = st;
  });
}

int symhip_optimizer_get_values(symhip_optimizer_t o, double* values) {
  return guarded([&] { o->opt->download_best_values(values); });
}

int32_t symhip_optimizer_iteration_times(symhip_optimizer_t o, double* ms, int32_t max) {
  const auto& t = o->opt->iteration_times();
  int32_t n = 0;
  for (; n < (int32_t)t.size() && n < max; ++n) ms[n] = t[n];
  return n;
}

int symhip_optimizer_enable_timing(symhip_optimizer_t o, int32_t enable) {
  return guarded([&] { o->opt->timer().enable(enable != 0); });
}

int symhip_optimizer_phase_times(symhip_optimizer_t o, double* total_ms, int64_t* counts) {
  return guarded([&] {
    for (int i = 0; i < SYMHIP_NUM_PHASES; ++i) {
      if (total_ms) total_ms[i] = o->opt->timer().total_ms(i);
      if (counts) counts[i] = o->opt->timer().count(i);
    }
  });
}

int symhip_optimizer_debug_record(symhip_optimizer_t o, int32_t index, double* update,
                                  double* values, double* residual) {
  return guarded([&] {
    SYMHIP_CHECK(o, "invalid arguments");
    const auto& recs = o->opt->debug_records();
    SYMHIP_CHECK(!recs.empty(), "no debug records: create the optimizer with debug_stats set");
    SYMHIP_CHECK(index >= 0 && index < (int32_t)recs.size(), "debug record index out of range");
    const auto& r = recs[index];
    if (update) std::copy(r.update.begin(), r.update.end(), update);
    if (values) std::copy(r.values.begin(), r.values.end(), values);
    if (residual) std::copy(r.residual.begin(), r.residual.end(), residual);
  });
}

int64_t symhip_optimizer_jacobian_sparsity(symhip_optimizer_t o, int32_t* colptr,
                                           int32_t* rowind) {
  int64_t nnz = -1;
  const int rc = guarded([&] {
    SYMHIP_CHECK(o, "invalid arguments");
    SYMHIP_CHECK(o->opt->params().include_jacobians,
                 "no Jacobian: create the optimizer with include_jacobians");
    const symhip::JacobianStructure& j = o->opt->jacobian_structure();
    nnz = (int64_t)j.rowind.size();
    if (colptr) std::copy(j.colptr.begin(), j.colptr.end(), colptr);
    if (rowind) std::copy(j.rowind.begin(), j.rowind.end(), rowind);
  });
  return rc < 0 ? rc : nnz;
}

int symhip_optimizer_best_jacobian(symhip_optimizer_t o, double* values) {
  return guarded([&] {
    SYMHIP_CHECK(o && values, "invalid arguments");
    o->opt->best_jacobian(values);
  });
}

int symhip_optimizer_debug_jacobian(symhip_optimizer_t o, int32_t index, double* values) {
  return guarded([&] {
    SYMHIP_CHECK(o && values, "invalid arguments");
    const auto& recs = o->opt->debug_records();
    SYMHIP_CHECK(!recs.empty(), "no debug records: create the optimizer with debug_stats set");
    SYMHIP_CHECK(index >= 0 && index < (int32_t)recs.size(), "debug record index out of range");
    SYMHIP_CHECK(!recs[index].jacobian.empty(),
                 "no Jacobian values: create the optimizer with include_jacobians set");
    std::copy(recs[index].jacobian.begin(), recs[index].jacobian.end(), values);
  });
}

int64_t symhip_optimizer_warnings(symhip_optimizer_t o, char* buf, int64_t cap) {
  int64_t len = -1;
  const int rc = guarded([&] {
    SYMHIP_CHECK(o && cap >= 0 && (buf || cap == 0), "invalid arguments");
    std::string all;
    for (const auto& w : o->opt->warnings()) all += w + "\x1e";
    len = (int64_t)all.size();
    if (cap > 0) {
      const int64_t n = std::min<int64_t>(cap - 1, len);
      std::copy(all.begin(), all.begin() + n, buf);
      buf[n] = 0;
    }
  });
  return rc < 0 ? rc : len;
}

int64_t symhip_optimizer_factor_sparsity(symhip_optimizer_t o, int32_t* ordering,
                                         int32_t* colptr, int32_t* rowind) {
  int64_t nnz = -1;
  const int rc = guarded([&] {
    SYMHIP_CHECK(o, "invalid arguments");
    std::vector<int32_t> ord, lp, li;
    o->opt->backend().factor_sparsity(ord, lp, li);
    nnz = (int64_t)li.size();
    if (ordering) std::copy(ord.begin(), ord.end(), ordering);
    if (colptr) std::copy(lp.begin(), lp.end(), colptr);
    if (rowind) std::copy(li.begin(), li.end(), rowind);
  });
  return rc == 0 ? nnz : rc;
}

int symhip_lm_decide(double e_init, double e_new, double e_best, double linear_delta,
                     double lambda, double nu, int32_t iteration,
                     const symhip_optimizer_params* params, int32_t have_last_update,
                     double last_norm2, double update_norm2, double last_dot_update,
                     double epsilon, symhip_lm_decision* out) {
  return guarded([&] {
    SYMHIP_CHECK(params && out, "invalid arguments");
    const symhip::LmDecision d =
        symhip::lm_decide(e_init, e_new, e_best, linear_delta, lambda, nu, iteration, *params,
                          have_last_update, last_norm2, update_norm2, last_dot_update, epsilon);
    *out = symhip_lm_decision{d.lambda, d.nu, d.rec, d.status, d.reason, d.accept, d.set_best};
  });
}

int symhip_retract(int32_t n_keys, const int32_t* types, const int32_t* storage_offsets,
                   const int32_t* tangent_offsets, double* values, const double* delta,
                   double epsilon) {
  return guarded([&] {
    SYMHIP_CHECK(n_keys >= 0 && (n_keys == 0 || (types && storage_offsets && tangent_offsets &&
                                                 values && delta)),
                 "invalid arguments");
    for (int32_t k = 0; k < n_keys; ++k)
      symhip::retract_value(types[k], values + storage_offsets[k], delta + tangent_offsets[k],
                            epsilon);
  });
}

int symhip_optimizer_solver_info(symhip_optimizer_t o, int64_t* out) {
  return guarded([&] { o->opt->backend().info(out); });
}

int symhip_optimizer_linear_solver_stats(symhip_optimizer_t o, int64_t* out) {
  return guarded([&] { o->opt->backend().linear_stats(out); });
}

int32_t symhip_optimizer_refinement_residuals(symhip_optimizer_t o, int32_t track, double* out,
                                              int32_t max) {
  int32_t n = -1;
  guarded([&] { n = o->opt->backend().refinement_residuals(track, out, max, o->opt->stream()); });
  return n;
}

int symhip_optimizer_log_systems(symhip_optimizer_t o, int32_t max) {
  return guarded([&] { o->opt->backend().log_systems(max); });
}

int32_t symhip_optimizer_logged_systems(symhip_optimizer_t o) {
  int32_t n = -1;
  guarded([&] { n = o->opt->backend().logged_systems(); });
  return n;
}

int32_t symhip_optimizer_rcs_pattern(symhip_optimizer_t o, int32_t* row_ptr, int32_t* col) {
  int32_t n = -1;
  guarded([&] { n = o->opt->backend().rcs_pattern(row_ptr, col); });
  return n;
}

int symhip_optimizer_logged_system(symhip_optimizer_t o, int32_t idx, double* blocks,
                                   double* rhs) {
  return guarded([&] { o->opt->backend().logged_system(idx, blocks, rhs); });
}

int symhip_optimizer_preconditioner_params(symhip_optimizer_t o, float* get, const float* set) {
  return guarded([&] { o->opt->backend().precond_params(get, set); });
}

int symhip_optimizer_preconditioner_loss_grad(symhip_optimizer_t o, int32_t idx, const double* X,
                                              int32_t P, float* loss, float* grad) {
  return guarded([&] {
    SYMHIP_CHECK(X && loss, "null argument");
    *loss = o->opt->backend().precond_loss_grad(idx, X, P, grad);
  });
}

int symhip_optimizer_train_preconditioner(symhip_optimizer_t o, int32_t steps, float lr,
                                          int32_t P, uint64_t seed, int32_t n_systems,
                                          float* losses) {
  return guarded([&] { o->opt->backend().train_precond(steps, lr, P, seed, n_systems, losses); });
}

int symhip_optimizer_eval_preconditioner(symhip_optimizer_t o, int32_t idx, double rel_tol,
                                         int32_t max_iters, int32_t learned, int32_t* iterations,
                                         double* step_err) {
  return guarded([&] {
    o->opt->backend().eval_precond(idx, rel_tol, max_iters, learned, iterations, step_err);
  });
}

int symhip_cholesky_create(int32_t n, const int32_t* colptr, const int32_t* rowind,
                           const int32_t* node_ptr, int32_t n_nodes, int32_t ordering,
                           const int32_t* given_order, symhip_cholesky_t* out) {
  return guarded([&] {
    SYMHIP_CHECK(n >= 0 && colptr && rowind && out, "invalid arguments");
    auto c = std::make_unique<symhip_cholesky>();
    c->colptr.assign(colptr, colptr + n + 1);
    c->rowind.assign(rowind, rowind + colptr[n]);
    std::vector<int32_t> nodes;
    if (node_ptr) {
      nodes.assign(node_ptr, node_ptr + n_nodes + 1);
    } else {
      nodes.resize(n + 1);
      for (int32_t i = 0; i <= n; ++i) nodes[i] = i;
    }
    std::vector<int32_t> given;
    symhip::Ordering ord = symhip::ordering_from_code(ordering);
    if (given_order) {
      given.assign(given_order, given_order + (int32_t)nodes.size() - 1);
      ord = symhip::Ordering::kGiven;
    }
    symhip::chol_symbolic(n, colptr, rowind, nodes, ord, given_order ? &given : nullptr, 64,
                          c->sym);
    *out = c.release();
  });
}

int symhip_cholesky_destroy(symhip_cholesky_t c) {
  return guarded([&] { delete c; });
}

static void ensure_cholesky_device(symhip_cholesky_t c) {
  if (c->stream) return;
  HIP_OK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  c->ldlt.setup(c->sym, c->colptr.data(), c->rowind.data(), c->stream);
}

int symhip_cholesky_factorize(symhip_cholesky_t c, const double* values) {
  return guarded([&] {
    ensure_cholesky_device(c);
    c->vals.upload(values, c->rowind.size(), c->stream);
    c->ldlt.factorize(c->vals.get(), c->stream);
    HIP_OK(hipStreamSynchronize(c->stream));
  });
}

int symhip_cholesky_set_f32_updates(symhip_cholesky_t c, int32_t on) {
  return guarded([&] {
    ensure_cholesky_device(c);
    c->ldlt.set_f32_updates(on != 0);
  });
}

int symhip_cholesky_solve(symhip_cholesky_t c, const double* b, double* x, int32_t nrhs) {
  return guarded([&] {
    ensure_cholesky_device(c);
    const int32_t n = c->sym.n;
    c->b.resize(std::max(1, n));
    c->x.resize(std::max(1, n));
    for (int32_t k = 0; k < nrhs; ++k) {
      c->b.upload(b + (int64_t)k * n, n, c->stream);
      c->ldlt.solve(c->b.get(), c->x.get(), c->stream);
      c->x.download(x + (int64_t)k * n, n, c->stream);
    }
    HIP_OK(hipStreamSynchronize(c->stream));
  });
}

int symhip_cholesky_permutation(symhip_cholesky_t c, int32_t* inv_perm) {
  return guarded([&] { std::memcpy(inv_perm, c->sym.inv_perm.data(), sizeof(int32_t) * c->sym.n); });
}

int symhip_cholesky_etree(symhip_cholesky_t c, int32_t* parent, int32_t* counts) {
  return guarded([&] {
    std::vector<int32_t> p, k;
    symhip::scalar_etree_counts(c->sym.n, c->colptr.data(), c->rowind.data(), c->sym.perm, p, k);
    std::memcpy(parent, p.data(), sizeof(int32_t) * p.size());
    std::memcpy(counts, k.data(), sizeof(int32_t) * k.size());
  });
}

int symhip_cholesky_d(symhip_cholesky_t c, double* d) {
  return guarded([&] {
    std::vector<double> panels, dd;
    c->ldlt.download_factor(panels, dd, c->stream);
    std::memcpy(d, dd.data(), sizeof(double) * dd.size());
  });
}

int64_t symhip_cholesky_factor(symhip_cholesky_t c, int32_t* colptr, int32_t* rowind,
                               double* values) {
  int64_t nnz = -1;
  const int rc = guarded([&] {
    SYMHIP_CHECK(c, "invalid arguments");
    std::vector<int32_t> lp, li;
    symhip::scalar_factor_pattern(c->sym.n, c->colptr.data(), c->rowind.data(), c->sym.perm, lp,
                                  li);
    nnz = (int64_t)li.size();
    if (colptr) std::copy(lp.begin(), lp.end(), colptr);
    if (rowind) std::copy(li.begin(), li.end(), rowind);
    if (values) {
      SYMHIP_CHECK(c->stream, "symhip_cholesky_factor: no factorization yet");
      c->ldlt.factor_values(lp, li, values, c->stream);
    }
  });
  return rc == 0 ? nnz : rc;
}

int symhip_cholesky_info(symhip_cholesky_t c, int64_t* out) {
  return guarded([&] {
    out[0] = c->sym.n;
    out[1] = c->sym.n_snodes;
    out[2] = c->sym.panel_size();
    out[3] = c->sym.n_levels;
    out[4] = (int64_t)c->sym.factor_flops;
  });
}

}  // extern "C"

#include "./gnn.h"
static_assert(SYMHIP_LEARNED_PRECOND_PARAMS == symhip::LearnedPrecond::kParams,
              "parameter count in the public header");

struct symhip_batch {
  std::unique_ptr<symhip::BatchLocalization> b;
};

extern "C" {

int symhip_batch_create(const symhip_problem_desc* descs, int32_t n_problems,
                        const symhip_optimizer_params* params, int32_t precision, double epsilon,
                        symhip_batch_t* out) {
  return guarded([&] {
    SYMHIP_CHECK(descs && params && out && n_problems > 0, "invalid arguments");
    SYMHIP_CHECK(precision == SYMHIP_PRECISION_F64 || precision == SYMHIP_PRECISION_F32_JAC,
                 "unknown precision");
    auto h = std::make_unique<symhip_batch>();
    h->b = std::make_unique<symhip::BatchLocalization>(
        descs, n_problems, *params, precision == SYMHIP_PRECISION_F32_JAC, epsilon);
    *out = h.release();
  });
}

int symhip_batch_destroy(symhip_batch_t batch) {
  return guarded([&] { delete batch; });
}

int64_t symhip_batch_storage_size(symhip_batch_t batch) {
  return batch ? batch->b->storage_total() : -1;
}

int symhip_batch_pack_sizes(const symhip_problem_desc* descs, int32_t n_problems,
                            const symhip_optimizer_params* params, int64_t* out4) {
  return guarded([&] {
    SYMHIP_CHECK(params && out4, "invalid arguments");
    const symhip::PackedBatch pk = symhip::pack_batch(descs, n_problems, *params);
    out4[0] = n_problems;
    out4[1] = (int64_t)pk.pose_soff.size();
    out4[2] = (int64_t)pk.match_args.size() / 3;
    out4[3] = (int64_t)pk.cvals.size();
  });
}

int symhip_batch_pack(const symhip_problem_desc* descs, int32_t n_problems,
                      const symhip_optimizer_params* params, int32_t* pose_begin,
                      int32_t* pose_soff, int32_t* match_ptr, int32_t* match_args,
                      int32_t* odo_args, double* values) {
  return guarded([&] {
    SYMHIP_CHECK(params && pose_begin && pose_soff && match_ptr && odo_args && values,
                 "invalid arguments");
    const symhip::PackedBatch pk = symhip::pack_batch(descs, n_problems, *params);
    SYMHIP_CHECK(match_args || pk.match_args.empty(), "invalid arguments: match_args");
    std::copy(pk.pose_begin.begin(), pk.pose_begin.end(), pose_begin);
    std::copy(pk.pose_soff.begin(), pk.pose_soff.end(), pose_soff);
    std::copy(pk.match_ptr.begin(), pk.match_ptr.end(), match_ptr);
    std::copy(pk.match_args.begin(), pk.match_args.end(), match_args);
    std::copy(pk.odo_args.begin(), pk.odo_args.end(), odo_args);
    std::copy(pk.cvals.begin(), pk.cvals.end(), values);
  });
}

int symhip_batch_solve_noalloc(const symhip_batch_tables* tables,
                               const symhip_optimizer_params* params, int32_t precision,
                               double epsilon, double* values_out,
                               symhip_optimization_stats* stats_out,
                               symhip_iteration_stats* iters_out, void* stream) {
  return guarded([&] {
    SYMHIP_CHECK(tables && params, "invalid arguments");
    symhip::batch_solve_noalloc(*tables, *params, precision, epsilon, values_out, stats_out,
                                iters_out, static_cast<hipStream_t>(stream));
  });
}

int symhip_batch_tune(symhip_batch_t batch, int32_t reps, double* config_ms, int32_t* chosen) {
  return guarded([&] {
    SYMHIP_CHECK(batch, "invalid arguments");
    batch->b->tune(reps, config_ms);
    if (chosen) *chosen = batch->b->config();
  });
}

int symhip_batch_set_config(symhip_batch_t batch, int32_t config) {
  return guarded([&] {
    SYMHIP_CHECK(batch, "invalid arguments");
    batch->b->set_config(config);
  });
}

int symhip_batch_solve(symhip_batch_t batch, double* values_out,
                       symhip_optimization_stats* stats_out, symhip_iteration_stats* iters_out,
                       double* solve_ms, double* total_ms) {
  return guarded([&] {
    SYMHIP_CHECK(batch, "invalid arguments");
    batch->b->solve(values_out, stats_out, iters_out, solve_ms, total_ms);
  });
}

}  // extern "C"
