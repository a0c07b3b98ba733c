// Drop-in check of the C++ host surface (include/sym/symforce_amd.h): the reference test
// "Test pose smoothing" (test/symforce_optimizer_test.cc:79-176), written against sym:: types and
// run on the device optimizer. Prints one JSON line with the last iteration's statistics.
#include <cmath>
#include <cstdio>
#include <random>

#include "sym/symforce_amd.h"

namespace {

std::array<double, 4> qmul(const std::array<double, 4>& a, const std::array<double, 4>& b) {
  return {a[3] * b[0] + a[0] * b[3] + a[1] * b[2] - a[2] * b[1],
          a[3] * b[1] - a[0] * b[2] + a[1] * b[3] + a[2] * b[0],
          a[3] * b[2] + a[0] * b[1] - a[1] * b[0] + a[2] * b[3],
          a[3] * b[3] - a[0] * b[0] - a[1] * b[1] - a[2] * b[2]};
}

// Pose3::Retract with the default epsilon: rotation right-multiplied, translation added.
sym::Pose3d retract(const sym::Pose3d& p, const double* d) {
  const double eps = sym::kDefaultEpsilond;
  const double t = std::sqrt(eps * eps + d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  const double s = std::sin(t / 2) / t;
  const auto q = qmul({p.data[0], p.data[1], p.data[2], p.data[3]},
                      {s * d[0], s * d[1], s * d[2], std::cos(t / 2)});
  sym::Pose3d r;
  r.data = {q[0], q[1], q[2], q[3], p.data[4] + d[3], p.data[5] + d[4], p.data[6] + d[5]};
  return r;
}

sym::Matrix66d diag(double v) {
  sym::Matrix66d m;
  for (int i = 0; i < 6; ++i) m.data[i * 6 + i] = v;
  return m;
}

}  // namespace

int main() {
  const int num_keys = 10;
  sym::Pose3d prior_start;
  sym::Pose3d prior_last;
  prior_last.data = {0, 0, std::sin(M_PI / 2), std::cos(M_PI / 2), 5, 0, 0};
  sym::Valuesd values;
  std::mt19937 gen(42);
  for (int i = 0; i < num_keys; ++i) {
    std::normal_distribution<double> dist{};
    double v[6];
    for (double& x : v) x = 0.4 * dist(gen);
    values.Set(sym::Key('P', i), retract(prior_start, v));
  }
  values.Set(sym::Key('a', 0), prior_start);
  values.Set(sym::Key('a', 1), prior_last);
  values.Set(sym::Key('a', 2), sym::Pose3d{});
  values.Set(sym::Key('s', 0), diag(1 / 0.1));
  values.Set(sym::Key('s', 1), diag(1 / 0.5));
  values.Set(sym::Key('e'), 1e-10);

  std::vector<sym::Factord> factors;
  factors.push_back(sym::Factord::Jacobian(sym::PriorFactorPose3<double>,
                                           {{'P', 0}, {'a', 0}, {'s', 0}, {'e'}}));
  factors.push_back(sym::Factord::Jacobian(sym::PriorFactorPose3<double>,
                                           {{'P', num_keys - 1}, {'a', 1}, {'s', 0}, {'e'}}));
  for (int i = 0; i < num_keys - 1; ++i)
    factors.push_back(sym::Factord::Jacobian(sym::BetweenFactorPose3<double>,
                                             {{'P', i}, {'P', i + 1}, {'a', 2}, {'s', 1}, {'e'}}));

  sym::optimizer_params_t params = sym::DefaultOptimizerParams();
  params.iterations = 50;
  params.early_exit_min_reduction = 0.0001;
  sym::Optimizerd optimizer(params, factors);
  const auto stats = optimizer.Optimize(values);
  const auto& last = stats.iterations.back();
  std::printf("{\"iteration\": %d, \"lambda\": %.17g, \"error\": %.17g, \"status\": %d}\n",
              last.iteration, last.current_lambda, last.new_error, stats.status);
  return 0;
}
