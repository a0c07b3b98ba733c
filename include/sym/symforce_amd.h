// C++ host surface of symforce_amd: the reference's sym:: optimizer API over the C ABI
// (include/symhip.h). Header-only; link against libsymhip.so.
//
//   reference type / function                      here
//   sym::Key (opt/key.h:26-103)                    sym::Key
//   sym::Values<double> (opt/values.h:30-324)      sym::Valuesd (flat storage + typed Set/At)
//   sym::Factor<double>::Hessian (opt/factor.h:194-234)
//                                                  sym::Factord::Hessian(<factor tag>, keys, opt)
//   sym::optimizer_params_t / DefaultOptimizerParams (opt/optimizer.cc:8-56)
//                                                  sym::optimizer_params_t / DefaultOptimizerParams
//   sym::Optimizer<double> (opt/optimizer.h:72-325) sym::Optimizerd
//   sym::OptimizationStats (opt/optimization_stats.h:22-94)
//                                                  sym::OptimizationStats
//   sym::SparseCholeskySolver (sparse_cholesky_solver.h:27-160)
//                                                  sym::hip::SparseCholeskySolver
//
// The reference's Factor stores an opaque std::function; the device needs to know which generated
// factor it is, so factors are constructed from tag objects named after the generated functions
// (sym::BetweenFactorPose3<double>, sym::SnavelyReprojectionFactor<double>, ...). Errors throw
// std::runtime_error, as SYM_ASSERT does (opt/assert.h:41-47).
#pragma once

#include <array>
#include <cstdint>
#include <limits>
#include <stdexcept>
#include <string>
#include <tuple>
#include <unordered_map>
#include <vector>

#include "../symhip.h"

namespace sym {

inline void check(int rc) {
  if (rc != 0) throw std::runtime_error(symhip_last_error());
}

// ------------------------------------------------------------------------------------------------
class Key {
 public:
  using letter_t = char;
  using subscript_t = int64_t;
  using superscript_t = int64_t;
  static constexpr subscript_t kInvalidSub = std::numeric_limits<subscript_t>::min();
  static constexpr superscript_t kInvalidSuper = std::numeric_limits<superscript_t>::min();
  Key(letter_t letter, subscript_t sub = kInvalidSub, superscript_t super = kInvalidSuper)
      : letter_(letter), sub_(sub), super_(super) {}
  Key WithSuper(superscript_t super) const { return Key(letter_, sub_, super); }
  letter_t Letter() const { return letter_; }
  subscript_t Sub() const { return sub_; }
  superscript_t Super() const { return super_; }
  bool operator==(const Key& o) const {
    return letter_ == o.letter_ && sub_ == o.sub_ && super_ == o.super_;
  }
  static bool LexicalLessThan(const Key& a, const Key& b) {
    return std::make_tuple(a.letter_, a.sub_, a.super_) < std::make_tuple(b.letter_, b.sub_, b.super_);
  }

 private:
  letter_t letter_;
  subscript_t sub_;
  superscript_t super_;
};

struct KeyHash {
  size_t operator()(const Key& k) const {
    size_t h = std::hash<int64_t>()(k.Sub()) * 1000003u ^ std::hash<int64_t>()(k.Super());
    return h * 31u + (unsigned char)k.Letter();
  }
};

// ------------------------------------------------------------------------------------------------
// Value types (storage layouts as the generated runtime: Pose3 [qx qy qz qw tx ty tz]).
struct Rot3d {
  std::array<double, 4> data{0, 0, 0, 1};
  static constexpr int32_t kType = SYMHIP_TYPE_ROT3;
};
struct Pose3d {
  std::array<double, 7> data{0, 0, 0, 1, 0, 0, 0};
  static constexpr int32_t kType = SYMHIP_TYPE_POSE3;
};
template <int N, int32_t T>
struct VectorNd {
  std::array<double, N> data{};
  static constexpr int32_t kType = T;
};
using Vector2d = VectorNd<2, SYMHIP_TYPE_VECTOR2>;
using Vector3d = VectorNd<3, SYMHIP_TYPE_VECTOR3>;
using Vector4d = VectorNd<4, SYMHIP_TYPE_VECTOR4>;
using Vector6d = VectorNd<6, SYMHIP_TYPE_VECTOR6>;
using Matrix33d = VectorNd<9, SYMHIP_TYPE_MATRIX33>;  // column-major
using Matrix66d = VectorNd<36, SYMHIP_TYPE_MATRIX66>;  // column-major

// sym::Values<double>: keys in append order over one flat storage vector.
class Valuesd {
 public:
  template <typename T>
  void Set(const Key& key, const T& value) {
    SetRaw(key, T::kType, value.data.data(), (int)value.data.size());
  }
  void Set(const Key& key, double value) { SetRaw(key, SYMHIP_TYPE_SCALAR, &value, 1); }
  template <typename T>
  T At(const Key& key) const {
    T out;
    const auto& e = entries_.at(index_.at(key));
    if (e.type != T::kType) throw std::runtime_error("Values::At: type mismatch");
    std::copy(data_.begin() + e.offset, data_.begin() + e.offset + out.data.size(),
              out.data.begin());
    return out;
  }
  double AtScalar(const Key& key) const { return data_.at(entries_.at(index_.at(key)).offset); }
  const std::vector<double>& Data() const { return data_; }
  std::vector<double>& Data() { return data_; }
  size_t NumEntries() const { return entries_.size(); }
  int32_t IndexOf(const Key& key) const {
    auto it = index_.find(key);
    if (it == index_.end()) throw std::runtime_error("Values: key not found");
    return it->second;
  }
  struct Entry {
    Key key;
    int32_t type;
    int32_t offset;
  };
  const std::vector<Entry>& Entries() const { return entries_; }

 private:
  void SetRaw(const Key& key, int32_t type, const double* v, int n) {
    auto it = index_.find(key);
    if (it != index_.end()) {
      const Entry& e = entries_[it->second];
      if (e.type != type) throw std::runtime_error("Values::Set: type mismatch");
      std::copy(v, v + n, data_.begin() + e.offset);
      return;
    }
    index_[key] = (int32_t)entries_.size();
    entries_.push_back({key, type, (int32_t)data_.size()});
    data_.insert(data_.end(), v, v + n);
  }
  std::vector<Entry> entries_;
  std::unordered_map<Key, int32_t, KeyHash> index_;
  std::vector<double> data_;
};

// ------------------------------------------------------------------------------------------------
// Factor tags named after the generated functions they stand for.
template <int32_t Code>
struct FactorTag {
  static constexpr int32_t kCode = Code;
};
template <typename Scalar>
constexpr FactorTag<SYMHIP_FACTOR_PRIOR_POSE3> PriorFactorPose3{};
template <typename Scalar>
constexpr FactorTag<SYMHIP_FACTOR_BETWEEN_POSE3> BetweenFactorPose3{};
template <typename Scalar>
constexpr FactorTag<SYMHIP_FACTOR_PRIOR_ROT3> PriorFactorRot3{};
template <typename Scalar>
constexpr FactorTag<SYMHIP_FACTOR_BETWEEN_ROT3> BetweenFactorRot3{};
template <typename Scalar>
constexpr FactorTag<SYMHIP_FACTOR_SNAVELY> SnavelyReprojectionFactor{};
template <typename Scalar>
constexpr FactorTag<SYMHIP_FACTOR_SNAVELY_HUBER> SnavelyReprojectionHuberFactor{};
template <typename Scalar>
constexpr FactorTag<SYMHIP_FACTOR_REPROJ_CAUCHY> LinearCameraReprojectionCauchyFactor{};
template <typename Scalar>
constexpr FactorTag<SYMHIP_FACTOR_MATCHING> MatchingFactor{};
template <typename Scalar>
constexpr FactorTag<SYMHIP_FACTOR_ODOMETRY> OdometryFactor{};
template <typename Scalar>
constexpr FactorTag<SYMHIP_FACTOR_BETWEEN_POSE3_CAUCHY> BetweenFactorPose3Cauchy{};

class Factord {
 public:
  // Factor::Hessian(func, keys_to_func, keys_to_optimize): the optimized keys of a registered
  // factor are its linearized arguments; passing a subset is allowed only as the Optimizer's key
  // list (frozen keys), as in the reference.
  template <int32_t Code>
  static Factord Hessian(FactorTag<Code>, std::vector<Key> keys,
                         std::vector<Key> /*keys_to_optimize*/ = {}) {
    Factord f;
    f.type_ = Code;
    f.keys_ = std::move(keys);
    return f;
  }
  template <int32_t Code>
  static Factord Jacobian(FactorTag<Code> tag, std::vector<Key> keys) {
    return Hessian(tag, std::move(keys));
  }
  int32_t Type() const { return type_; }
  const std::vector<Key>& AllKeys() const { return keys_; }

 private:
  int32_t type_ = 0;
  std::vector<Key> keys_;
};

// ------------------------------------------------------------------------------------------------
using optimizer_params_t = symhip_optimizer_params;
using lambda_update_type_t = symhip_lambda_update_type;

inline optimizer_params_t DefaultOptimizerParams() {
  optimizer_params_t p{};
  p.initial_lambda = 1.0;
  p.lambda_lower_bound = 0.0;
  p.lambda_upper_bound = 1000000.0;
  p.lambda_update_type = SYMHIP_LAMBDA_STATIC;
  p.lambda_up_factor = 4.0;
  p.lambda_down_factor = 1 / 4.0;
  p.dynamic_lambda_update_beta = 2.0;
  p.dynamic_lambda_update_gamma = 3.0;
  p.dynamic_lambda_update_p = 3;
  p.use_diagonal_damping = 0;
  p.use_unit_damping = 1;
  p.keep_max_diagonal_damping = 0;
  p.diagonal_damping_min = 1e-6;
  p.iterations = 50;
  p.early_exit_min_reduction = 1e-6;
  p.early_exit_min_absolute_error = 0.0;
  p.enable_bold_updates = 0;
  return p;
}

struct OptimizationStats {
  std::vector<symhip_iteration_stats> iterations;
  int32_t best_index = 0;
  int32_t status = SYMHIP_STATUS_INVALID;
  int32_t failure_reason = SYMHIP_FAILURE_INVALID;
};

constexpr double kDefaultEpsilond = 10.0 * std::numeric_limits<double>::epsilon();

// sym::Optimizer<double> with the LM loop on the device.
class Optimizerd {
 public:
  Optimizerd(const optimizer_params_t& params, std::vector<Factord> factors,
             const std::string& name = "sym::Optimize", std::vector<Key> keys = {},
             double epsilon = kDefaultEpsilond,
             symhip_solver_options options = DefaultSolverOptions())
      : params_(params), factors_(std::move(factors)), name_(name), keys_(std::move(keys)),
        epsilon_(epsilon), options_(options) {}
  Optimizerd(const Optimizerd&) = delete;
  Optimizerd& operator=(const Optimizerd&) = delete;
  ~Optimizerd() {
    if (h_) symhip_optimizer_destroy(h_);
  }

  static symhip_solver_options DefaultSolverOptions() {
    return symhip_solver_options{SYMHIP_SOLVER_CHOLESKY, SYMHIP_ORDERING_MIN_DEGREE,
                                 SYMHIP_PRECISION_F64, 0, 100, 1e-6};
  }

  // Optimizer::Optimize (opt/optimizer.tcc:70-89): values updated in place with the best values.
  OptimizationStats Optimize(Valuesd& values, int num_iterations = -1) {
    Build(values);
    OptimizationStats stats;
    stats.iterations.resize(4096);
    symhip_optimization_stats st{};
    check(symhip_optimizer_optimize(h_, values.Data().data(), num_iterations,
                                    stats.iterations.data(), (int32_t)stats.iterations.size(),
                                    &st));
    stats.iterations.resize(st.num_iterations);
    stats.best_index = st.best_index;
    stats.status = st.status;
    stats.failure_reason = st.failure_reason;
    return stats;
  }
  const optimizer_params_t& Params() const { return params_; }
  void UpdateParams(const optimizer_params_t& p) {
    params_ = p;
    if (h_) check(symhip_optimizer_update_params(h_, &params_));
  }

 private:
  void Build(const Valuesd& values) {
    if (h_) return;  // structure is fixed after the first call, as the reference Linearizer's
    std::vector<int8_t> letters;
    std::vector<int64_t> subs, supers;
    std::vector<int32_t> types;
    for (const auto& e : values.Entries()) {
      letters.push_back((int8_t)e.key.Letter());
      subs.push_back(e.key.Sub());
      supers.push_back(e.key.Super());
      types.push_back(e.type);
    }
    std::vector<int32_t> ftypes, fptr{0}, fargs, opt;
    for (const auto& f : factors_) {
      ftypes.push_back(f.Type());
      for (const auto& k : f.AllKeys()) fargs.push_back(values.IndexOf(k));
      fptr.push_back((int32_t)fargs.size());
    }
    for (const auto& k : keys_) opt.push_back(values.IndexOf(k));
    symhip_problem_desc d{};
    d.n_keys = (int32_t)letters.size();
    d.key_letters = letters.data();
    d.key_subs = subs.data();
    d.key_supers = supers.data();
    d.key_types = types.data();
    d.values = values.Data().data();
    d.n_factors = (int32_t)ftypes.size();
    d.factor_types = ftypes.data();
    d.factor_arg_ptr = fptr.data();
    d.factor_args = fargs.data();
    d.n_opt_keys = (int32_t)opt.size();
    d.opt_keys = opt.empty() ? nullptr : opt.data();
    check(symhip_optimizer_create(&d, &params_, &options_, epsilon_, &h_));
  }

  optimizer_params_t params_;
  std::vector<Factord> factors_;
  std::string name_;
  std::vector<Key> keys_;
  double epsilon_;
  symhip_solver_options options_;
  symhip_optimizer_t h_ = nullptr;
};

// Free function Optimize(params, factors, values) (opt/optimizer.h:334-340).
inline OptimizationStats Optimize(const optimizer_params_t& params, std::vector<Factord> factors,
                                  Valuesd& values, double epsilon = kDefaultEpsilond) {
  Optimizerd opt(params, std::move(factors), "sym::Optimize", {}, epsilon);
  return opt.Optimize(values);
}

namespace hip {

// SparseCholeskySolver drop-in: AnalyzeSparsityPattern / Factorize / Solve on a lower CSC matrix.
class SparseCholeskySolver {
 public:
  explicit SparseCholeskySolver(int32_t ordering = SYMHIP_ORDERING_MIN_DEGREE)
      : ordering_(ordering) {}
  ~SparseCholeskySolver() {
    if (h_) symhip_cholesky_destroy(h_);
  }
  SparseCholeskySolver(const SparseCholeskySolver&) = delete;
  SparseCholeskySolver& operator=(const SparseCholeskySolver&) = delete;
  void AnalyzeSparsityPattern(int32_t n, const int32_t* colptr, const int32_t* rowind) {
    if (h_) symhip_cholesky_destroy(h_);
    check(symhip_cholesky_create(n, colptr, rowind, nullptr, 0, ordering_, nullptr, &h_));
    n_ = n;
  }
  bool Factorize(const double* values) {
    check(symhip_cholesky_factorize(h_, values));
    return true;
  }
  std::vector<double> Solve(const std::vector<double>& b) const {
    std::vector<double> x(b.size());
    check(symhip_cholesky_solve(h_, b.data(), x.data(), (int32_t)(b.size() / n_)));
    return x;
  }
  std::vector<int32_t> InversePermutation() const {
    std::vector<int32_t> p(n_);
    check(symhip_cholesky_permutation(h_, p.data()));
    return p;
  }

 private:
  int32_t ordering_;
  int32_t n_ = 0;
  symhip_cholesky_t h_ = nullptr;
};

}  // namespace hip
}  // namespace sym
