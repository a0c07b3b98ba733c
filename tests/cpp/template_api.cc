// The templated C++ surface (include/sym/symforce_amd.h) used as the reference's tests use the
// reference API. One mode per run, selected by argv[1]; each prints one JSON line that
// tests/test_cpp_api_gpu.py checks against the CPU oracle or the mathematics:
//
//   lm_linear <double|float>  sym::LevenbergMarquardtSolver<Scalar> with a caller LinearizeFunc:
//                             "Converges for a linear problem in one iteration"
//                             (test/symforce_levenberg_marquardt_test.cc:17-87)
//   ba_schur                  sym::Optimizer<double, LevenbergMarquardtSolver<double,
//                             hip::SparseSchurSolver<>>> on a generated Snavely BA problem
//   debug_stats               sym::Optimizerd with debug_stats and populate_best_linearization
//   covariance                Linearize, ComputeFullCovariance, ComputeAllCovariances and the
//                             Schur-complement ComputeCovariances on a pose graph
//   optimizerf                sym::Optimizerf on the reference pose-smoothing problem
//   gnc                       sym::GncOptimizer<sym::Optimizerd> on the reference GNC problem
//                             (test/symforce_gnc_test.cc:13-80), the same random draws
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <sstream>
#include <string>

#include "sym/symforce_amd.h"
#include "symhip_problems.h"

namespace {

std::string num(double v) {
  char b[64];
  std::snprintf(b, sizeof(b), "%.17g", v);
  return b;
}
template <typename V>
std::string arr(const V& v, size_t n) {
  std::string s = "[";
  for (size_t i = 0; i < n; ++i) s += (i ? "," : "") + num((double)v[i]);
  return s + "]";
}
template <typename Stats>
std::string iterations_json(const Stats& st) {
  std::string s = "[";
  for (size_t i = 0; i < st.iterations.size(); ++i) {
    const auto& it = st.iterations[i];
    s += (i ? "," : "");
    s += "{\"iteration\":" + std::to_string(it.iteration) + ",\"update_accepted\":" +
         (it.update_accepted ? "true" : "false") + ",\"current_lambda\":" +
         num(it.current_lambda) + ",\"new_error\":" + num(it.new_error) +
         ",\"new_error_linear\":" + num(it.new_error_linear) + "}";
  }
  return s + "]";
}

// A generated problem as sym::Values and sym::Factors (keys in the generator's order).
struct Generated {
  sym::Valuesd values;
  std::vector<sym::Factord> factors;
};
Generated from_generator(symhip_generated* g) {
  symhip_problem_desc d{};
  sym::check(symhip_gen_desc(g, &d));
  Generated out;
  std::vector<sym::Key> keys;
  int64_t off = 0;
  for (int32_t k = 0; k < d.n_keys; ++k) {
    keys.emplace_back((char)d.key_letters[k], d.key_subs[k], d.key_supers[k]);
    out.values.SetFromStorage(keys.back(), d.key_types[k], d.values + off);
    off += sym::StorageDim(d.key_types[k]);
  }
  for (int32_t f = 0; f < d.n_factors; ++f) {
    std::vector<sym::Key> fk;
    for (int32_t q = d.factor_arg_ptr[f]; q < d.factor_arg_ptr[f + 1]; ++q)
      fk.push_back(keys[d.factor_args[q]]);
    out.factors.push_back(sym::Factord::FromCode(d.factor_types[f], fk));
  }
  return out;
}

template <typename Scalar>
int lm_linear() {
  constexpr int M = 9, N = 6;
  std::mt19937 gen(12345);
  std::uniform_real_distribution<double> u(-1.0, 1.0);
  std::vector<double> J(M * N);  // column-major
  for (double& x : J) x = u(gen);
  sym::optimizer_params_t params{};
  params.debug_stats = 1;
  params.initial_lambda = 1.0;
  params.lambda_update_type = sym::lambda_update_type_t::STATIC;
  params.lambda_up_factor = 3.0;
  params.lambda_down_factor = 1.0 / 3.0;
  params.lambda_lower_bound = 0.0;
  params.lambda_upper_bound = 0.0;
  params.iterations = 1;
  params.use_diagonal_damping = 0;
  params.use_unit_damping = 0;
  const Scalar kEpsilon = Scalar(1e-10);
  using Solver = sym::LevenbergMarquardtSolver<Scalar>;
  Solver solver(params, "", kEpsilon);
  using StateVector = sym::VectorN<Scalar, N, SYMHIP_TYPE_VECTOR6>;
  // r = J v; hessian_lower = lower(J^T J) (dense lower pattern); rhs = J^T r
  auto residual_func = [&](const sym::Values<Scalar>& values,
                           sym::SparseLinearization<Scalar>& lin) {
    const auto v = values.template At<StateVector>('v');
    lin.residual = sym::VectorX<Scalar>(M);
    for (int i = 0; i < M; ++i) {
      double acc = 0;
      for (int j = 0; j < N; ++j) acc += J[j * M + i] * (double)v.data[j];
      lin.residual[i] = (Scalar)acc;
    }
    std::vector<int32_t> cp{0}, ri;
    std::vector<Scalar> hv;
    for (int c = 0; c < N; ++c) {
      for (int r = c; r < N; ++r) {
        double acc = 0;
        for (int i = 0; i < M; ++i) acc += J[r * M + i] * J[c * M + i];
        ri.push_back(r);
        hv.push_back((Scalar)acc);
      }
      cp.push_back((int32_t)ri.size());
    }
    lin.hessian_lower = sym::SparseMatrix<Scalar>(N, N, cp, ri, hv);
    lin.rhs = sym::VectorX<Scalar>(N);
    for (int j = 0; j < N; ++j) {
      double acc = 0;
      for (int i = 0; i < M; ++i) acc += J[j * M + i] * (double)lin.residual[i];
      lin.rhs[j] = (Scalar)acc;
    }
  };
  sym::Values<Scalar> values_init;
  StateVector start;
  start.data.fill(Scalar(100));
  values_init.Set('v', start);
  const sym::index_t index = values_init.CreateIndex();
  sym::SparseLinearization<Scalar> lin;
  residual_func(values_init, lin);
  lin.SetInitialized();
  const double error_init = (double)lin.Error();
  solver.SetIndex(index);
  solver.Reset(values_init);
  sym::OptimizationStats<typename Solver::MatrixType> stats;
  const auto status = solver.Iterate(residual_func, stats);
  const auto& res = stats.iterations.back().residual;
  double error_final = 0;
  for (int i = 0; i < res.size(); ++i) error_final += 0.5 * res[i] * res[i];
  const auto best = solver.GetBestValues().template At<StateVector>('v');
  double norm = 0;
  for (Scalar x : best.data) norm += (double)x * x;
  std::printf("{\"error_init\": %s, \"error_final\": %s, \"best_norm\": %s, \"records\": %zu, "
              "\"exited\": %s, \"ordering_size\": %zu, \"L_shape\": [%lld, %lld]}\n",
              num(error_init).c_str(), num(error_final).c_str(), num(std::sqrt(norm)).c_str(),
              stats.iterations.size(), status ? "true" : "false",
              stats.linear_solver_ordering.size(),
              (long long)stats.cholesky_factor_sparsity.shape[0],
              (long long)stats.cholesky_factor_sparsity.shape[1]);
  return 0;
}

int ba_schur() {
  symhip_generated* g = symhip_gen_ba_snavely(12, 300, 5);
  Generated p = from_generator(g);
  symhip_gen_free(g);
  auto params = sym::DefaultOptimizerParams();
  params.lambda_update_type = sym::lambda_update_type_t::DYNAMIC;
  params.iterations = 30;
  using NLS = sym::LevenbergMarquardtSolver<double, sym::hip::SparseSchurSolver<>>;
  sym::Optimizer<double, NLS> optimizer(params, p.factors, "ba", {}, sym::kDefaultEpsilond,
                                        sym::hip::SparseSchurSolver<>(SYMHIP_ORDERING_MIN_DEGREE));
  sym::Optimizer<double, NLS>::Stats stats;
  optimizer.Optimize(p.values, -1, false, stats);  // the virtual overload
  std::printf("{\"status\": %d, \"failure_reason\": %d, \"best_index\": %d, \"iterations\": %s, "
              "\"values\": %s, \"num_keys\": %zu}\n",
              stats.status, stats.failure_reason, stats.best_index,
              iterations_json(stats).c_str(), arr(p.values.Data(), p.values.Data().size()).c_str(),
              optimizer.Keys().size());
  return 0;
}

int debug_stats() {
  symhip_generated* g = symhip_gen_pose_smoothing(10, 42);
  Generated p = from_generator(g);
  symhip_gen_free(g);
  auto params = sym::DefaultOptimizerParams();
  params.iterations = 50;
  params.early_exit_min_reduction = 0.0001;
  params.debug_stats = 1;
  sym::Optimizerd optimizer(params, p.factors);
  const sym::Valuesd start = p.values;
  const auto stats = optimizer.Optimize(p.values, -1, true);
  std::ostringstream it;
  it << "[";
  for (size_t i = 0; i < stats.iterations.size(); ++i) {
    const auto& r = stats.iterations[i];
    it << (i ? "," : "") << "{\"iteration\":" << r.iteration
       << ",\"update\":" << arr(r.update, r.update.size())
       << ",\"residual\":" << arr(r.residual, r.residual.size())
       << ",\"values\":" << arr(r.values.data, r.values.data.size())
       << ",\"new_error\":" << num(r.new_error) << "}";
  }
  it << "]";
  const auto& L = stats.cholesky_factor_sparsity;
  const auto& bl = *stats.best_linearization;
  std::printf("{\"iterations\": %s, \"ordering\": %s, \"L_colptr\": %s, \"L_rowind\": %s, "
              "\"L_shape\": [%lld, %lld], \"best_residual\": %s, \"best_rhs\": %s, "
              "\"best_index\": %d, \"start\": %s}\n",
              it.str().c_str(),
              arr(stats.linear_solver_ordering, stats.linear_solver_ordering.size()).c_str(),
              arr(L.column_pointers, L.column_pointers.size()).c_str(),
              arr(L.row_indices, L.row_indices.size()).c_str(), (long long)L.shape[0],
              (long long)L.shape[1], arr(bl.residual, bl.residual.size()).c_str(),
              arr(bl.rhs, bl.rhs.size()).c_str(), stats.best_index,
              arr(start.Data(), start.Data().size()).c_str());
  return 0;
}

int covariance() {
  symhip_generated* g = symhip_gen_pose_graph(60, 15, 3);
  Generated p = from_generator(g);
  symhip_gen_free(g);
  auto params = sym::DefaultOptimizerParams();
  sym::Optimizerd optimizer(params, p.factors);
  optimizer.Optimize(p.values);
  const auto lin = optimizer.Linearize(p.values);
  sym::MatrixX<double> full;
  optimizer.ComputeFullCovariance(lin, full);
  std::unordered_map<sym::Key, sym::MatrixX<double>> all, first;
  optimizer.ComputeAllCovariances(lin, all);
  const std::vector<sym::Key> prefix(optimizer.Keys().begin(), optimizer.Keys().begin() + 5);
  const int info = optimizer.ComputeCovariances(lin, prefix, first);
  const auto& H = lin.hessian_lower;
  double max_all = 0, max_first = 0;
  int off = 0;
  for (const sym::Key& k : optimizer.Keys()) {
    const auto& b = all.at(k);
    for (int c = 0; c < b.cols(); ++c)
      for (int r = 0; r < b.rows(); ++r)
        max_all = std::max(max_all, std::abs(b(r, c) - full(off + r, off + c)));
    off += b.rows();
  }
  std::string blocks = "[";
  for (size_t i = 0; i < prefix.size(); ++i) {
    const auto& b = first.at(prefix[i]);
    blocks += (i ? "," : "") + arr(b.data(), b.size());
  }
  blocks += "]";
  (void)max_first;
  std::printf("{\"n\": %d, \"colptr\": %s, \"rowind\": %s, \"values\": %s, \"full\": %s, "
              "\"all_vs_full\": %s, \"prefix_blocks\": %s, \"prefix_info\": %d, "
              "\"keys\": %zu}\n",
              H.cols(), arr(H.outerIndexPtr(), H.cols() + 1).c_str(),
              arr(H.innerIndexPtr(), H.nonZeros()).c_str(),
              arr(H.valuePtr(), H.nonZeros()).c_str(), arr(full.data(), full.size()).c_str(),
              num(max_all).c_str(), blocks.c_str(), info, optimizer.Keys().size());
  return 0;
}

int optimizerf() {
  symhip_generated* g = symhip_gen_pose_smoothing(10, 42);
  symhip_problem_desc d{};
  sym::check(symhip_gen_desc(g, &d));
  sym::Valuesf values;
  std::vector<sym::Key> keys;
  int64_t off = 0;
  for (int32_t k = 0; k < d.n_keys; ++k) {
    keys.emplace_back((char)d.key_letters[k], d.key_subs[k], d.key_supers[k]);
    values.SetFromStorage(keys.back(), d.key_types[k], d.values + off);
    off += sym::StorageDim(d.key_types[k]);
  }
  std::vector<sym::Factorf> factors;
  for (int32_t f = 0; f < d.n_factors; ++f) {
    std::vector<sym::Key> fk;
    for (int32_t q = d.factor_arg_ptr[f]; q < d.factor_arg_ptr[f + 1]; ++q)
      fk.push_back(keys[d.factor_args[q]]);
    factors.push_back(sym::Factorf::FromCode(d.factor_types[f], fk));
  }
  symhip_gen_free(g);
  auto params = sym::DefaultOptimizerParams();
  params.iterations = 50;
  params.early_exit_min_reduction = 0.0001;
  const auto stats = sym::Optimize(params, factors, values);
  const auto& last = stats.iterations.back();
  std::printf("{\"iteration\": %d, \"lambda\": %s, \"error\": %s, \"status\": %d}\n",
              last.iteration, num(last.current_lambda).c_str(), num(last.new_error).c_str(),
              stats.status);
  return 0;
}

// "Test GNC" (test/symforce_gnc_test.cc:22-80): 20 Barron residuals on a Vector5 x, 3 of them
// outliers, mu raised from 0 by 0.33 to 0.99; the samples are drawn as the reference draws them
// (std::mt19937(42), a fresh std::normal_distribution per Random<Vector5d> call).
int gnc() {
  constexpr double kEpsilon = 1e-12;
  const int n_residuals = 20, n_outliers = 3;
  std::mt19937 gen(42);
  auto random5 = [&]() {
    std::normal_distribution<double> dist{};
    sym::Vector5d v;
    for (int i = 0; i < 5; ++i) v.data[i] = dist(gen);
    return v;
  };
  sym::Valuesd initial_values;
  sym::Vector5d ones;
  ones.data.fill(1.0);
  initial_values.Set('x', ones);
  initial_values.Set('e', sym::kDefaultEpsilond);
  std::vector<double> ys;
  for (int i = 0; i < n_residuals; ++i) {
    sym::Vector5d y = random5();
    for (int k = 0; k < 5; ++k) y.data[k] = (i < n_outliers ? 10.0 : 0.0) + 0.1 * y.data[k];
    initial_values.Set(sym::Key('y', i), y);
    ys.insert(ys.end(), y.data.begin(), y.data.end());
  }
  std::vector<sym::Factord> factors;
  for (int i = 0; i < n_residuals; ++i)
    factors.push_back(sym::Factord::Hessian(sym::BarronFactor<double>,
                                            {'x', sym::Key('y', i), 'u', 'e'}, {'x'}));
  auto params = sym::DefaultOptimizerParams();
  sym::optimizer_gnc_params_t gnc_params{};
  gnc_params.mu_initial = 0;
  gnc_params.mu_max = 0.99;
  gnc_params.mu_step = 0.33;
  gnc_params.gnc_update_min_reduction = 1e-3;
  sym::GncOptimizer<sym::Optimizerd> gnc_optimizer(params, gnc_params, 'u', factors,
                                                   "sym::Optimize", std::vector<sym::Key>{},
                                                   kEpsilon);
  sym::Valuesd gnc_values = initial_values;
  const auto gnc_stats = gnc_optimizer.Optimize(gnc_values);
  sym::Valuesd regular_values = initial_values;
  regular_values.Set('u', 0.0);
  sym::Optimize(params, factors, regular_values, kEpsilon);
  const auto xg = gnc_values.At<sym::Vector5d>('x');
  const auto xr = regular_values.At<sym::Vector5d>('x');
  std::printf("{\"iterations\": %s, \"status\": %d, \"best_index\": %d, \"x_gnc\": %s, "
              "\"x_regular\": %s, \"mu_final\": %s, \"ys\": %s, \"values\": %s}\n",
              iterations_json(gnc_stats).c_str(), gnc_stats.status, gnc_stats.best_index,
              arr(xg.data, 5).c_str(), arr(xr.data, 5).c_str(),
              num(gnc_values.AtScalar('u')).c_str(), arr(ys, ys.size()).c_str(),
              arr(gnc_values.Data(), gnc_values.Data().size()).c_str());
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  const std::string mode = argc > 1 ? argv[1] : "";
  try {
    if (mode == "lm_linear")
      return (argc > 2 && std::strcmp(argv[2], "float") == 0) ? lm_linear<float>()
                                                               : lm_linear<double>();
    if (mode == "ba_schur") return ba_schur();
    if (mode == "debug_stats") return debug_stats();
    if (mode == "covariance") return covariance();
    if (mode == "optimizerf") return optimizerf();
    if (mode == "gnc") return gnc();
  } catch (const std::exception& e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
  std::fprintf(stderr, "unknown mode '%s'\n", mode.c_str());
  return 2;
}
