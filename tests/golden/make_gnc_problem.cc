// Writes gnc_problem.json: the 20 Vector5 samples of the reference's "Test GNC"
// (test/symforce_gnc_test.cc:26-41), drawn as it draws them: std::mt19937(42), and for each
// sym::Random<Vector5d> call (gen/cpp/sym/ops/matrix/storage_ops.h:212-217) a fresh
// std::normal_distribution<double> filling the five entries in order; the first 3 samples are
// outliers around 10. Build and run: g++ -O2 -o /tmp/make_gnc_problem make_gnc_problem.cc &&
// /tmp/make_gnc_problem > gnc_problem.json
#include <cstdio>
#include <random>

int main() {
  const int n_residuals = 20, n_outliers = 3;
  std::mt19937 gen(42);
  std::printf("{\"source\": \"test/symforce_gnc_test.cc:26-41\", \"n_outliers\": %d, \"ys\": [",
              n_outliers);
  for (int i = 0; i < n_residuals; ++i) {
    std::normal_distribution<double> dist{};
    std::printf("%s[", i ? ", " : "");
    for (int k = 0; k < 5; ++k)
      std::printf("%s%.17g", k ? ", " : "", (i < n_outliers ? 10.0 : 0.0) + 0.1 * dist(gen));
    std::printf("]");
  }
  std::printf("]}\n");
  return 0;
}
