// C++ host surface of symforce_amd: the reference's sym:: optimizer API over the C ABI
// (include/symhip.h). Header-only; link against libsymhip.so. C++17.
//
//   reference type / function                                   here
//   sym::Key, std::hash<Key> (opt/key.h:26-103, key.cc:54-58)   sym::Key, std::hash<sym::Key>
//   sym::Values<S> (opt/values.h:30-324), index_t               sym::Values<S>, sym::index_t
//   sym::Factor<S>::Hessian / Jacobian (opt/factor.h:194-234)   sym::Factor<S> from factor tags
//   sym::Linearization<M> (opt/linearization.h:26-77)            sym::Linearization<M>
//   sym::OptimizationStats<M> (opt/optimization_stats.h:22-94)  sym::OptimizationStats<M>
//   optimizer_params_t / DefaultOptimizerParams (optimizer.cc)   same names
//   sym::SparseCholeskySolver<M> (sparse_cholesky_solver.h)      sym::hip::SparseCholeskySolver<M>
//   sym::SparseSchurSolver<M> (opt/sparse_schur_solver.h)        sym::hip::SparseSchurSolver<M>
//   sym::LevenbergMarquardtSolver<S, LinearSolverType>           sym::LevenbergMarquardtSolver
//     (opt/levenberg_marquardt_solver.h:131-260)
//   sym::Optimizer<S, NLS> (opt/optimizer.h:72-325), Optimizerd/f, Optimize(...)  same names
//
// Two ways the optimizer runs:
// * sym::Optimizer over registered factors: the whole Levenberg-Marquardt loop on the device
//   (linearization, damping, the supernodal LDL^T or Schur solve, retraction and the
//   accept / lambda decision), driven through symhip_optimizer_*. The NLS template argument
//   selects the device linear solver through its LinearSolverType.
// * sym::LevenbergMarquardtSolver::Iterate(func, stats) with a caller-supplied LinearizeFunc:
//   the reference's host loop (the caller's linearization), with the factorization and solves of
//   its LinearSolverType (the device LDL^T of hip::SparseCholeskySolver), the decision step of
//   the device optimizer (symhip_lm_decide) and the retraction of symhip_retract.
//
// The reference's Factor stores an opaque std::function; the device needs to know which
// generated factor it is, so factors are constructed from tag objects named after the generated
// functions (sym::BetweenFactorPose3<double>, sym::SnavelyReprojectionFactor<double>, ...).
// Computation is FP64 on the device for every Scalar: Values<float> are widened at the boundary
// and the results narrowed back. Errors throw std::runtime_error, as SYM_ASSERT does
// (opt/assert.h:41-47).
#pragma once

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <functional>
#include <limits>
#include <optional>
#include <stdexcept>
#include <string>
#include <tuple>
#include <type_traits>
#include <unordered_map>
#include <utility>
#include <vector>

#include "../symhip.h"

namespace sym {

inline void check(int rc) {
  if (rc != 0) throw std::runtime_error(symhip_last_error());
}

// ------------------------------------------------------------------------------------------------
// Key
class Key {
 public:
  using letter_t = char;
  using subscript_t = int64_t;
  using superscript_t = int64_t;
  static constexpr letter_t kInvalidLetter = static_cast<letter_t>(0);
  static constexpr subscript_t kInvalidSub = std::numeric_limits<subscript_t>::min();
  static constexpr superscript_t kInvalidSuper = std::numeric_limits<superscript_t>::min();
  Key(letter_t letter = kInvalidLetter, subscript_t sub = kInvalidSub,
      superscript_t super = kInvalidSuper)
      : letter_(letter), sub_(sub), super_(super) {}
  Key WithLetter(letter_t letter) const { return Key(letter, sub_, super_); }
  Key WithSub(subscript_t sub) const { return Key(letter_, sub, super_); }
  Key WithSuper(superscript_t super) const { return Key(letter_, sub_, super); }
  letter_t Letter() const { return letter_; }
  subscript_t Sub() const { return sub_; }
  superscript_t Super() const { return super_; }
  bool operator==(const Key& o) const {
    return letter_ == o.letter_ && sub_ == o.sub_ && super_ == o.super_;
  }
  bool operator!=(const Key& o) const { return !(*this == o); }
  static bool LexicalLessThan(const Key& a, const Key& b) {
    return std::make_tuple(a.letter_, a.sub_, a.super_) <
           std::make_tuple(b.letter_, b.sub_, b.super_);
  }
  struct LexicalCompare {
    bool operator()(const Key& a, const Key& b) const { return LexicalLessThan(a, b); }
  };

 private:
  letter_t letter_;
  subscript_t sub_;
  superscript_t super_;
};

namespace internal {
// Boost-style hash combination, as the reference combines a Key's fields
// (opt/internal/hash_combine.h:22-26).
inline void hash_combine(std::size_t&) {}
template <typename T, typename... Rest>
inline void hash_combine(std::size_t& seed, const T& v, Rest... rest) {
  seed ^= std::hash<T>()(v) + 0x9e3779b9 + (seed << 6) + (seed >> 2);
  hash_combine(seed, rest...);
}
}  // namespace internal

}  // namespace sym

template <>
struct std::hash<sym::Key> {
  std::size_t operator()(const sym::Key& k) const {
    std::size_t seed = 0;
    sym::internal::hash_combine(seed, k.Letter(), k.Sub(), k.Super());
    return seed;
  }
};

namespace sym {

using KeyHash = std::hash<Key>;

// ------------------------------------------------------------------------------------------------
// Dense matrices (column-major) and vectors: the small part of Eigen's surface the API needs.
template <typename S>
class MatrixX {
 public:
  using Scalar = S;
  MatrixX() = default;
  MatrixX(int rows, int cols) : rows_(rows), cols_(cols), data_((size_t)rows * cols, S(0)) {}
  static MatrixX Zero(int rows, int cols) { return MatrixX(rows, cols); }
  static MatrixX Identity(int rows, int cols) {
    MatrixX m(rows, cols);
    for (int i = 0; i < std::min(rows, cols); ++i) m(i, i) = S(1);
    return m;
  }
  static MatrixX Identity(int n) { return Identity(n, n); }
  int rows() const { return rows_; }
  int cols() const { return cols_; }
  int64_t size() const { return (int64_t)data_.size(); }
  S& operator()(int r, int c) { return data_[(size_t)c * rows_ + r]; }
  const S& operator()(int r, int c) const { return data_[(size_t)c * rows_ + r]; }
  S* data() { return data_.data(); }
  const S* data() const { return data_.data(); }
  void resize(int rows, int cols) {
    rows_ = rows;
    cols_ = cols;
    data_.assign((size_t)rows * cols, S(0));
  }
  MatrixX block(int r0, int c0, int nr, int nc) const {
    MatrixX b(nr, nc);
    for (int c = 0; c < nc; ++c)
      for (int r = 0; r < nr; ++r) b(r, c) = (*this)(r0 + r, c0 + c);
    return b;
  }
  template <typename T>
  MatrixX<T> cast() const {
    MatrixX<T> m(rows_, cols_);
    for (int64_t i = 0; i < size(); ++i) m.data()[i] = static_cast<T>(data_[i]);
    return m;
  }

 protected:
  int rows_ = 0, cols_ = 0;
  std::vector<S> data_;
};

template <typename S>
class VectorX : public MatrixX<S> {
 public:
  VectorX() = default;
  explicit VectorX(int n) : MatrixX<S>(n, 1) {}
  VectorX(const MatrixX<S>& m) : MatrixX<S>(m) {  // NOLINT: a one-column matrix is a vector
    if (m.cols() != 1 && m.size() != 0) throw std::runtime_error("VectorX from a matrix");
  }
  static VectorX Zero(int n) { return VectorX(n); }
  int size() const { return this->rows_; }
  S& operator[](int i) { return this->data_[i]; }
  const S& operator[](int i) const { return this->data_[i]; }
  S& operator()(int i) { return this->data_[i]; }
  const S& operator()(int i) const { return this->data_[i]; }
  S dot(const VectorX& o) const {
    S acc = S(0);
    for (int i = 0; i < size(); ++i) acc += (*this)[i] * o[i];
    return acc;
  }
  S squaredNorm() const { return dot(*this); }
  S norm() const { return std::sqrt(squaredNorm()); }
  template <typename T>
  VectorX<T> cast() const {
    VectorX<T> v(size());
    for (int i = 0; i < size(); ++i) v[i] = static_cast<T>((*this)[i]);
    return v;
  }
};

// Compressed sparse column matrix (Eigen::SparseMatrix<S>'s storage: outer = column pointers).
template <typename S>
class SparseMatrix {
 public:
  using Scalar = S;
  SparseMatrix() = default;
  SparseMatrix(int rows, int cols) : rows_(rows), cols_(cols), outer_(cols + 1, 0) {}
  SparseMatrix(int rows, int cols, std::vector<int32_t> colptr, std::vector<int32_t> rowind,
               std::vector<S> values)
      : rows_(rows), cols_(cols), outer_(std::move(colptr)), inner_(std::move(rowind)),
        values_(std::move(values)) {}
  int rows() const { return rows_; }
  int cols() const { return cols_; }
  int64_t nonZeros() const { return (int64_t)inner_.size(); }
  const int32_t* outerIndexPtr() const { return outer_.data(); }
  const int32_t* innerIndexPtr() const { return inner_.data(); }
  S* valuePtr() { return values_.data(); }
  const S* valuePtr() const { return values_.data(); }
  std::vector<S>& values() { return values_; }
  const std::vector<S>& values() const { return values_; }
  S coeff(int r, int c) const {
    const auto b = inner_.begin() + outer_[c], e = inner_.begin() + outer_[c + 1];
    const auto it = std::lower_bound(b, e, r);
    return (it != e && *it == r) ? values_[it - inner_.begin()] : S(0);
  }
  // The stored diagonal (a lower pattern with its full diagonal, rows sorted).
  VectorX<S> diagonal() const {
    VectorX<S> d(std::min(rows_, cols_));
    for (int c = 0; c < d.size(); ++c) d[c] = coeff(c, c);
    return d;
  }
  template <typename T>
  SparseMatrix<T> cast() const {
    return SparseMatrix<T>(rows_, cols_, outer_, inner_, std::vector<T>(values_.begin(),
                                                                        values_.end()));
  }

 private:
  int rows_ = 0, cols_ = 0;
  std::vector<int32_t> outer_, inner_;
  std::vector<S> values_;
};

struct PermutationMatrix {
  std::vector<int32_t> idx;
  const std::vector<int32_t>& indices() const { return idx; }
  int size() const { return (int)idx.size(); }
};

// sparse_matrix_structure_t (lcmtypes/symforce.lcm:264-277)
struct sparse_matrix_structure_t {
  std::vector<int32_t> row_indices;
  std::vector<int32_t> column_pointers;
  std::array<int64_t, 2> shape{{0, 0}};
};

template <typename S>
sparse_matrix_structure_t GetSparseStructure(const SparseMatrix<S>& m) {
  sparse_matrix_structure_t s;
  s.column_pointers.assign(m.outerIndexPtr(), m.outerIndexPtr() + m.cols() + 1);
  s.row_indices.assign(m.innerIndexPtr(), m.innerIndexPtr() + m.nonZeros());
  s.shape = {{m.rows(), m.cols()}};
  return s;
}

// ------------------------------------------------------------------------------------------------
// Value types (storage layouts as the generated runtime: Pose3 [qx qy qz qw tx ty tz]).
template <typename S>
struct Rot3 {
  std::array<S, 4> data{{0, 0, 0, 1}};
  static constexpr int32_t kType = SYMHIP_TYPE_ROT3;
};
template <typename S>
struct Pose3 {
  std::array<S, 7> data{{0, 0, 0, 1, 0, 0, 0}};
  static constexpr int32_t kType = SYMHIP_TYPE_POSE3;
};
template <typename S, int N, int32_t T>
struct VectorN {
  std::array<S, N> data{};
  static constexpr int32_t kType = T;
};
using Rot3d = Rot3<double>;
using Rot3f = Rot3<float>;
using Pose3d = Pose3<double>;
using Pose3f = Pose3<float>;
template <int N, int32_t T>
using VectorNd = VectorN<double, N, T>;
using Vector2d = VectorN<double, 2, SYMHIP_TYPE_VECTOR2>;
using Vector3d = VectorN<double, 3, SYMHIP_TYPE_VECTOR3>;
using Vector4d = VectorN<double, 4, SYMHIP_TYPE_VECTOR4>;
using Vector5d = VectorN<double, 5, SYMHIP_TYPE_VECTOR5>;
using Vector6d = VectorN<double, 6, SYMHIP_TYPE_VECTOR6>;
using Matrix33d = VectorN<double, 9, SYMHIP_TYPE_MATRIX33>;   // column-major
using Matrix66d = VectorN<double, 36, SYMHIP_TYPE_MATRIX66>;  // column-major
using Vector2f = VectorN<float, 2, SYMHIP_TYPE_VECTOR2>;
using Vector3f = VectorN<float, 3, SYMHIP_TYPE_VECTOR3>;
using Vector4f = VectorN<float, 4, SYMHIP_TYPE_VECTOR4>;
using Vector5f = VectorN<float, 5, SYMHIP_TYPE_VECTOR5>;
using Vector6f = VectorN<float, 6, SYMHIP_TYPE_VECTOR6>;
using Matrix33f = VectorN<float, 9, SYMHIP_TYPE_MATRIX33>;
using Matrix66f = VectorN<float, 36, SYMHIP_TYPE_MATRIX66>;

inline int32_t StorageDim(int32_t type) {
  switch (type) {
    case SYMHIP_TYPE_SCALAR: return 1;
    case SYMHIP_TYPE_ROT3: return 4;
    case SYMHIP_TYPE_POSE3: return 7;
    case SYMHIP_TYPE_VECTOR2: return 2;
    case SYMHIP_TYPE_VECTOR3: return 3;
    case SYMHIP_TYPE_VECTOR4: return 4;
    case SYMHIP_TYPE_VECTOR5: return 5;
    case SYMHIP_TYPE_VECTOR6: return 6;
    case SYMHIP_TYPE_MATRIX33: return 9;
    case SYMHIP_TYPE_MATRIX66: return 36;
  }
  throw std::runtime_error("unknown value type");
}
inline int32_t TangentDim(int32_t type) {
  return type == SYMHIP_TYPE_POSE3 ? 6 : type == SYMHIP_TYPE_ROT3 ? 3 : StorageDim(type);
}

// index_entry_t / index_t (lcmtypes/symforce.lcm:10-26)
struct index_entry_t {
  Key key;
  int32_t type = 0;
  int32_t offset = 0;  // storage offset
  int32_t storage_dim = 0;
  int32_t tangent_dim = 0;
};
struct index_t {
  int32_t storage_dim = 0;
  int32_t tangent_dim = 0;
  std::vector<index_entry_t> entries;
};

// values_t (lcmtypes/symforce.lcm:28-33): a Values snapshot in debug statistics.
struct values_t {
  index_t index;
  std::vector<double> data;
};

// sym::Values<S>: keys in append order over one flat storage vector.
template <typename S>
class Values {
 public:
  using Scalar = S;
  template <typename T>
  void Set(const Key& key, const T& value) {
    SetRaw(key, T::kType, value.data.data(), (int)value.data.size());
  }
  void Set(const Key& key, S value) { SetRaw(key, SYMHIP_TYPE_SCALAR, &value, 1); }
  // Sets a key of the given value type from its storage (StorageDim(type) entries).
  template <typename T>
  void SetFromStorage(const Key& key, int32_t type, const T* storage) {
    SetRaw(key, type, storage, StorageDim(type));
  }
  template <typename T>
  T At(const Key& key) const {
    const index_entry_t& e = entries_.at(IndexOf(key));
    if (e.type != T::kType) throw std::runtime_error("Values::At: type mismatch");
    T out;
    std::copy(data_.begin() + e.offset, data_.begin() + e.offset + out.data.size(),
              out.data.begin());
    return out;
  }
  S AtScalar(const Key& key) const { return data_.at(entries_.at(IndexOf(key)).offset); }
  bool Has(const Key& key) const { return index_.count(key) != 0; }
  std::vector<Key> Keys() const {
    std::vector<Key> k;
    for (const auto& e : entries_) k.push_back(e.key);
    return k;
  }
  const std::vector<S>& Data() const { return data_; }
  std::vector<S>& Data() { return data_; }
  size_t NumEntries() const { return entries_.size(); }
  int32_t IndexOf(const Key& key) const {
    auto it = index_.find(key);
    if (it == index_.end()) throw std::runtime_error("Values: key not found");
    return it->second;
  }
  const std::vector<index_entry_t>& Entries() const { return entries_; }
  // Values::CreateIndex (opt/values.cc:184-234): tangent offsets in the given key order.
  index_t CreateIndex(const std::vector<Key>& keys) const {
    index_t idx;
    for (const Key& k : keys) {
      index_entry_t e = entries_.at(IndexOf(k));
      idx.storage_dim += e.storage_dim;
      idx.tangent_dim += e.tangent_dim;
      idx.entries.push_back(e);
    }
    return idx;
  }
  index_t CreateIndex() const { return CreateIndex(Keys()); }
  // Values::Retract (opt/values.cc:294-327): the tangent offsets are the order of `index`.
  void Retract(const index_t& index, const S* delta, S epsilon) {
    std::vector<int32_t> types, soff, toff;
    int32_t t = 0;
    for (const auto& e : index.entries) {
      const index_entry_t& mine = entries_.at(IndexOf(e.key));
      types.push_back(mine.type);
      soff.push_back(mine.offset);
      toff.push_back(t);
      t += mine.tangent_dim;
    }
    std::vector<double> v(data_.begin(), data_.end()), d(delta, delta + t);
    check(symhip_retract((int32_t)types.size(), types.data(), soff.data(), toff.data(), v.data(),
                         d.data(), (double)epsilon));
    std::copy(v.begin(), v.end(), data_.begin());
  }
  // Values::Update (opt/values.cc:268-275) for the keys of `index`.
  void Update(const index_t& index, const Values& other) {
    for (const auto& e : index.entries) {
      const index_entry_t& a = entries_.at(IndexOf(e.key));
      const index_entry_t& b = other.entries_.at(other.IndexOf(e.key));
      std::copy(other.data_.begin() + b.offset, other.data_.begin() + b.offset + b.storage_dim,
                data_.begin() + a.offset);
    }
  }
  values_t GetLcmType() const {
    values_t v;
    v.index = CreateIndex();
    v.data.assign(data_.begin(), data_.end());
    return v;
  }

 private:
  template <typename T>
  void SetRaw(const Key& key, int32_t type, const T* v, int n) {
    auto it = index_.find(key);
    if (it != index_.end()) {
      const index_entry_t& e = entries_[it->second];
      if (e.type != type) throw std::runtime_error("Values::Set: type mismatch");
      for (int i = 0; i < n; ++i) data_[e.offset + i] = static_cast<S>(v[i]);
      return;
    }
    index_[key] = (int32_t)entries_.size();
    entries_.push_back({key, type, (int32_t)data_.size(), StorageDim(type), TangentDim(type)});
    for (int i = 0; i < n; ++i) data_.push_back(static_cast<S>(v[i]));
  }
  std::vector<index_entry_t> entries_;
  std::unordered_map<Key, int32_t> index_;
  std::vector<S> data_;
};
using Valuesd = Values<double>;
using Valuesf = Values<float>;

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
// barron_factor(x, y, mu, eps) of test/symforce_gnc_codegen_test.py:20-29 (Vector5 x and y)
template <typename Scalar>
constexpr FactorTag<SYMHIP_FACTOR_BARRON_VECTOR5> BarronFactor{};

// sym::Factor<S>: a registered factor body and the keys of its arguments. The optimized keys of
// a registered factor are its linearized arguments (the first ones); a smaller set of optimized
// keys is chosen, as in the reference, through the Optimizer's key list (frozen keys).
template <typename S>
class Factor {
 public:
  using Scalar = S;
  template <int32_t Code>
  static Factor Hessian(FactorTag<Code>, std::vector<Key> keys_to_func,
                        std::vector<Key> keys_to_optimize = {}) {
    Factor f;
    f.type_ = Code;
    f.keys_ = std::move(keys_to_func);
    int32_t info[4];
    check(symhip_factor_info(Code, info));
    if ((int32_t)f.keys_.size() != info[0])
      throw std::runtime_error("Factor: wrong number of keys for the factor's arguments");
    f.opt_keys_.assign(f.keys_.begin(), f.keys_.begin() + info[1]);
    for (const Key& k : keys_to_optimize)
      if (std::find(f.opt_keys_.begin(), f.opt_keys_.end(), k) == f.opt_keys_.end())
        throw std::runtime_error("Factor: keys_to_optimize must be linearized arguments");
    return f;
  }
  template <int32_t Code>
  static Factor Jacobian(FactorTag<Code> tag, std::vector<Key> keys_to_func,
                         std::vector<Key> keys_to_optimize = {}) {
    return Hessian(tag, std::move(keys_to_func), std::move(keys_to_optimize));
  }
  // The registered factor with runtime type code `code` (symhip_factor_type), e.g. when the
  // factor list comes from a file or a generator.
  static Factor FromCode(int32_t code, std::vector<Key> keys_to_func) {
    Factor f;
    f.type_ = code;
    f.keys_ = std::move(keys_to_func);
    int32_t info[4];
    check(symhip_factor_info(code, info));
    if ((int32_t)f.keys_.size() != info[0])
      throw std::runtime_error("Factor: wrong number of keys for the factor's arguments");
    f.opt_keys_.assign(f.keys_.begin(), f.keys_.begin() + info[1]);
    return f;
  }
  int32_t Type() const { return type_; }
  const std::vector<Key>& AllKeys() const { return keys_; }
  const std::vector<Key>& OptimizedKeys() const { return opt_keys_; }

 private:
  int32_t type_ = 0;
  std::vector<Key> keys_, opt_keys_;
};
using Factord = Factor<double>;
using Factorf = Factor<float>;

// Factor.h:423-449: every optimized key of the factors, sorted by LexicalLessThan.
template <typename S>
std::vector<Key> ComputeKeysToOptimize(const std::vector<Factor<S>>& factors) {
  std::vector<Key> keys;
  for (const auto& f : factors)
    for (const Key& k : f.OptimizedKeys())
      if (std::find(keys.begin(), keys.end(), k) == keys.end()) keys.push_back(k);
  std::sort(keys.begin(), keys.end(), Key::LexicalLessThan);
  return keys;
}

// ------------------------------------------------------------------------------------------------
// Parameters and statistics
using optimizer_params_t = symhip_optimizer_params;
struct lambda_update_type_t {
  enum : int32_t { INVALID = 0, STATIC = 1, DYNAMIC = 2 };
};
struct optimization_status_t {
  enum : int32_t { INVALID = 0, SUCCESS = 1, HIT_ITERATION_LIMIT = 2, FAILED = 3 };
};
struct levenberg_marquardt_solver_failure_reason_t {
  enum : int32_t { INVALID = 0, LAMBDA_OUT_OF_BOUNDS = 1, INITIAL_ERROR_NOT_FINITE = 2 };
};

// DefaultOptimizerParams (opt/optimizer.cc:8-56)
inline optimizer_params_t DefaultOptimizerParams() {
  optimizer_params_t p{};
  p.initial_lambda = 1.0;
  p.lambda_lower_bound = 0.0;
  p.lambda_upper_bound = 1000000.0;
  p.lambda_update_type = lambda_update_type_t::STATIC;
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

// optimization_iteration_t (lcmtypes/symforce.lcm:229-262)
struct optimization_iteration_t {
  int32_t iteration = 0;
  double current_lambda = 0;
  double new_error_linear = 0;
  double new_error = 0;
  double relative_reduction = 0;
  bool update_accepted = false;
  double update_angle_change = 0;
  // populated only with debug_stats
  VectorX<double> update;
  values_t values;
  VectorX<double> residual;
  MatrixX<double> jacobian_values;
};

inline optimization_iteration_t FromRecord(const symhip_iteration_stats& r) {
  optimization_iteration_t it;
  it.iteration = r.iteration;
  it.current_lambda = r.current_lambda;
  it.new_error_linear = r.new_error_linear;
  it.new_error = r.new_error;
  it.relative_reduction = r.relative_reduction;
  it.update_accepted = r.update_accepted != 0;
  it.update_angle_change = r.update_angle_change;
  return it;
}

template <typename MatrixType>
struct Linearization {
  using Scalar = typename MatrixType::Scalar;
  using Vector = VectorX<Scalar>;
  using Matrix = MatrixType;
  void Reset() { initialized_ = false; }
  bool IsInitialized() const { return initialized_; }
  void SetInitialized(bool initialized = true) { initialized_ = initialized; }
  // opt/linearization.h:51-54
  Scalar Error() const {
    if (!initialized_) throw std::runtime_error("Linearization::Error: not initialized");
    return Scalar(0.5) * residual.squaredNorm();
  }
  // opt/linearization.h:62-67
  Scalar LinearDeltaError(const Vector& x_update, const Vector& damping_vector) const {
    if (!initialized_) throw std::runtime_error("Linearization: not initialized");
    Scalar acc = Scalar(0);
    for (int i = 0; i < x_update.size(); ++i)
      acc += x_update[i] * (rhs[i] - damping_vector[i] * x_update[i]);
    return Scalar(0.5) * acc;
  }
  Vector residual;
  Matrix hessian_lower;
  Matrix jacobian;
  Vector rhs;

 private:
  bool initialized_ = false;
};
template <typename S>
using SparseLinearization = Linearization<SparseMatrix<S>>;
using SparseLinearizationd = SparseLinearization<double>;
using SparseLinearizationf = SparseLinearization<float>;

template <typename MatrixType>
struct OptimizationStats {
  std::vector<optimization_iteration_t> iterations;
  int32_t best_index = 0;
  int32_t status = optimization_status_t::INVALID;
  int32_t failure_reason = 0;
  std::optional<Linearization<MatrixType>> best_linearization{};
  sparse_matrix_structure_t jacobian_sparsity{};
  std::vector<int32_t> linear_solver_ordering{};
  sparse_matrix_structure_t cholesky_factor_sparsity{};
  void Reset(size_t num_iterations) {
    iterations.clear();
    iterations.reserve(num_iterations);
    best_index = 0;
    status = optimization_status_t::INVALID;
    failure_reason = 0;
    best_linearization.reset();
    jacobian_sparsity = {};
    linear_solver_ordering.clear();
    cholesky_factor_sparsity = {};
  }
};
template <typename S>
using SparseOptimizationStats = OptimizationStats<SparseMatrix<S>>;

constexpr double kDefaultEpsilond = 10.0 * std::numeric_limits<double>::epsilon();
constexpr float kDefaultEpsilonf = 10.0f * std::numeric_limits<float>::epsilon();
template <typename S>
constexpr S kDefaultEpsilon = S(10) * std::numeric_limits<S>::epsilon();

// ------------------------------------------------------------------------------------------------
// Linear solvers
namespace hip {

// SparseCholeskySolver (opt/sparse_cholesky/sparse_cholesky_solver.h:27-160) on the device: the
// LinearSolverType concept the reference's LevenbergMarquardtSolver uses
// (levenberg_marquardt_solver.tcc:191-217, :351-355): AnalyzeSparsityPattern, Factorize, Solve,
// SolveInPlace, Permutation, L, D. Matrices are lower CSC with the full diagonal.
template <typename _MatrixType = SparseMatrix<double>>
class SparseCholeskySolver {
 public:
  using MatrixType = _MatrixType;
  using Scalar = typename MatrixType::Scalar;
  static constexpr int32_t kDeviceSolver = SYMHIP_SOLVER_CHOLESKY;

  explicit SparseCholeskySolver(int32_t ordering = SYMHIP_ORDERING_MIN_DEGREE)
      : ordering_(ordering) {}
  // Copies carry the ordering and the analyzed pattern; the device state is rebuilt on use.
  SparseCholeskySolver(const SparseCholeskySolver& o)
      : ordering_(o.ordering_), n_(o.n_), colptr_(o.colptr_), rowind_(o.rowind_) {}
  SparseCholeskySolver& operator=(const SparseCholeskySolver& o) {
    if (this != &o) {
      Release();
      ordering_ = o.ordering_;
      n_ = o.n_;
      colptr_ = o.colptr_;
      rowind_ = o.rowind_;
    }
    return *this;
  }
  ~SparseCholeskySolver() { Release(); }

  bool IsInitialized() const { return !colptr_.empty(); }
  void ComputeSymbolicSparsity(const MatrixType& A) { AnalyzeSparsityPattern(A); }
  void AnalyzeSparsityPattern(const MatrixType& A) {
    Release();
    n_ = A.cols();
    colptr_.assign(A.outerIndexPtr(), A.outerIndexPtr() + n_ + 1);
    rowind_.assign(A.innerIndexPtr(), A.innerIndexPtr() + A.nonZeros());
    Handle();
  }
  bool Factorize(const MatrixType& A) {
    if (!IsInitialized()) AnalyzeSparsityPattern(A);
    std::vector<double> v(A.valuePtr(), A.valuePtr() + A.nonZeros());
    check(symhip_cholesky_factorize(Handle(), v.data()));
    factored_ = true;
    l_.reset();
    return true;
  }
  // x = A^-1 b for a vector or a matrix of right-hand sides.
  MatrixX<Scalar> Solve(const MatrixX<Scalar>& b) const {
    if (!factored_) throw std::runtime_error("SparseCholeskySolver::Solve before Factorize");
    std::vector<double> bb(b.data(), b.data() + b.size()), x(bb.size());
    check(symhip_cholesky_solve(h_, bb.data(), x.data(), b.cols()));
    MatrixX<Scalar> out(b.rows(), b.cols());
    for (int64_t i = 0; i < out.size(); ++i) out.data()[i] = static_cast<Scalar>(x[i]);
    return out;
  }
  VectorX<Scalar> Solve(const VectorX<Scalar>& b) const {
    return VectorX<Scalar>(Solve(static_cast<const MatrixX<Scalar>&>(b)));
  }
  void SolveInPlace(MatrixX<Scalar>& b) const { b = Solve(b); }
  // Permutation().indices(): old -> new (the reference's permutation_)
  PermutationMatrix Permutation() const {
    std::vector<int32_t> inv(n_), perm(n_);
    check(symhip_cholesky_permutation(Handle(), inv.data()));
    for (int32_t i = 0; i < n_; ++i) perm[inv[i]] = i;
    return PermutationMatrix{perm};
  }
  PermutationMatrix InversePermutation() const {
    std::vector<int32_t> inv(n_);
    check(symhip_cholesky_permutation(Handle(), inv.data()));
    return PermutationMatrix{inv};
  }
  // The strictly lower unit factor of P A P^T on the reference's pattern.
  const SparseMatrix<Scalar>& L() const {
    if (!l_) {
      const int64_t nnz = symhip_cholesky_factor(Handle(), nullptr, nullptr, nullptr);
      if (nnz < 0) check((int)nnz);
      std::vector<int32_t> lp(n_ + 1), li(nnz);
      std::vector<double> lv(nnz);
      const int64_t got = symhip_cholesky_factor(h_, lp.data(), li.data(),
                                                 factored_ ? lv.data() : nullptr);
      if (got < 0) check((int)got);
      l_ = SparseMatrix<Scalar>(n_, n_, lp, li, std::vector<Scalar>(lv.begin(), lv.end()));
    }
    return *l_;
  }
  VectorX<Scalar> D() const {
    std::vector<double> d(n_);
    check(symhip_cholesky_d(h_, d.data()));
    VectorX<Scalar> out(n_);
    for (int i = 0; i < n_; ++i) out[i] = static_cast<Scalar>(d[i]);
    return out;
  }
  int32_t Ordering() const { return ordering_; }
  // the options the device optimizer uses when this solver is its LinearSolverType
  symhip_solver_options DeviceOptions() const {
    symhip_solver_options o{};
    o.linear_solver = SYMHIP_SOLVER_CHOLESKY;
    o.ordering = ordering_;
    o.precision = SYMHIP_PRECISION_F64;
    o.pcg_max_iters = 100;
    o.pcg_rel_tol = 1e-6;
    return o;
  }

 private:
  symhip_cholesky_t Handle() const {
    if (!h_) {
      if (colptr_.empty()) throw std::runtime_error("SparseCholeskySolver: no pattern analyzed");
      check(symhip_cholesky_create(n_, colptr_.data(), rowind_.data(), nullptr, 0, ordering_,
                                   nullptr, &h_));
    }
    return h_;
  }
  void Release() {
    if (h_) symhip_cholesky_destroy(h_);
    h_ = nullptr;
    factored_ = false;
    l_.reset();
  }
  int32_t ordering_;
  int32_t n_ = 0;
  std::vector<int32_t> colptr_, rowind_;
  mutable symhip_cholesky_t h_ = nullptr;
  bool factored_ = false;
  mutable std::optional<SparseMatrix<Scalar>> l_;
};

// SparseSchurSolver (opt/sparse_schur_solver.h:62-135) as the device optimizer's linear solver:
// landmark-block elimination, the reduced camera system on the matrix cores and its supernodal
// LDL^T (or PCG, with `options`). It is selected through sym::Optimizer's NLS argument; the
// host LevenbergMarquardtSolver loop takes sparse Cholesky solvers only.
template <typename _MatrixType = SparseMatrix<double>>
class SparseSchurSolver {
 public:
  using MatrixType = _MatrixType;
  using Scalar = typename MatrixType::Scalar;
  static constexpr int32_t kDeviceSolver = SYMHIP_SOLVER_SCHUR;
  explicit SparseSchurSolver(int32_t ordering = SYMHIP_ORDERING_NESTED_DISSECTION) {
    o_.linear_solver = SYMHIP_SOLVER_SCHUR;
    o_.ordering = ordering;
    o_.precision = SYMHIP_PRECISION_F64;
    o_.pcg_max_iters = 100;
    o_.pcg_rel_tol = 1e-6;
  }
  explicit SparseSchurSolver(const symhip_solver_options& options) : o_(options) {
    o_.linear_solver = SYMHIP_SOLVER_SCHUR;
  }
  symhip_solver_options DeviceOptions() const { return o_; }

 private:
  symhip_solver_options o_{};
};

}  // namespace hip

template <typename MatrixType = SparseMatrix<double>>
using SparseCholeskySolver = hip::SparseCholeskySolver<MatrixType>;
template <typename MatrixType = SparseMatrix<double>>
using SparseSchurSolver = hip::SparseSchurSolver<MatrixType>;

// ------------------------------------------------------------------------------------------------
// LevenbergMarquardtSolver (opt/levenberg_marquardt_solver.h:131-260): the host loop of Iterate
// over a caller-supplied linearization, with the LinearSolverType's factorization and solves.
template <typename ScalarType,
          typename _LinearSolverType = hip::SparseCholeskySolver<SparseMatrix<ScalarType>>>
class LevenbergMarquardtSolver {
 public:
  using Scalar = ScalarType;
  using LinearSolverType = _LinearSolverType;
  using MatrixType = typename LinearSolverType::MatrixType;
  using LinearizationType = Linearization<MatrixType>;
  using ValuesType = Values<Scalar>;
  using FailureReason = int32_t;  // levenberg_marquardt_solver_failure_reason_t
  using LinearizeFunc = std::function<void(const ValuesType&, LinearizationType&)>;
  using Status = std::optional<std::pair<int32_t, FailureReason>>;

  LevenbergMarquardtSolver(const optimizer_params_t& p, const std::string& id,
                           const Scalar epsilon)
      : p_(p), id_(id), epsilon_(epsilon) {}
  LevenbergMarquardtSolver(const optimizer_params_t& p, const std::string& id,
                           const Scalar epsilon, const LinearSolverType& linear_solver)
      : p_(p), id_(id), epsilon_(epsilon), linear_solver_(linear_solver) {}

  void SetIndex(const index_t& index) { index_ = index; }
  void Reset(const ValuesType& values) {
    current_lambda_ = p_.initial_lambda;
    current_nu_ = p_.dynamic_lambda_update_beta;
    iteration_ = -1;
    ResetState(values);
  }
  void RelaxDampingToInitial() {
    current_lambda_ = std::min(current_lambda_, p_.initial_lambda);
    current_nu_ = p_.dynamic_lambda_update_beta;
  }
  void ResetState(const ValuesType& values) {
    have_max_diagonal_ = false;
    have_last_update_ = false;
    for (auto& b : blocks_) {
      b.valid = false;
      b.lin.Reset();
    }
    init_ = 0;
    new_ = 1;
    best_ = 0;
    free_ = 2;
    blocks_[new_].values = values;
    blocks_[new_].valid = true;
  }
  const optimizer_params_t& Params() const { return p_; }
  void UpdateParams(const optimizer_params_t& p) { p_ = p; }
  const LinearSolverType& LinearSolver() const { return linear_solver_; }
  LinearSolverType& LinearSolver() { return linear_solver_; }

  // Iterate (levenberg_marquardt_solver.tcc:138-343)
  Status Iterate(const LinearizeFunc& func, OptimizationStats<MatrixType>& stats) {
    std::swap(init_, new_);  // state_.Step()
    ++iteration_;
    Block& init = blocks_[init_];
    if (!init.lin.IsInitialized()) {
      func(init.values, init.lin);
      init.lin.SetInitialized();
      if (best_ != init_) {  // SetBestToInit
        if (best_ != new_) free_ = best_;
        best_ = init_;
      }
    }
    if (iteration_ == 0) {
      optimization_iteration_t it;
      it.iteration = -1;
      it.new_error = init.lin.Error();
      it.current_lambda = current_lambda_;
      if (p_.debug_stats) {
        it.values = Widened(init.values).GetLcmType();
        it.residual = init.lin.residual.template cast<double>();
      }
      stats.iterations.push_back(std::move(it));
      if (!std::isfinite((double)init.lin.Error()))
        return std::make_pair((int32_t)optimization_status_t::FAILED,
                              (FailureReason)levenberg_marquardt_solver_failure_reason_t::
                                  INITIAL_ERROR_NOT_FINITE);
    }
    if (!solver_analyzed_) {
      linear_solver_.AnalyzeSparsityPattern(init.lin.hessian_lower);
      solver_analyzed_ = true;
    }
    // DampHessian (levenberg_marquardt_solver.tcc:22-54)
    MatrixType& H = init.lin.hessian_lower;
    const int n = H.cols();
    const VectorX<Scalar> undamped = H.diagonal();
    damping_ = VectorX<Scalar>(n);
    if (p_.use_diagonal_damping) {
      if (p_.keep_max_diagonal_damping) {
        if (!have_max_diagonal_) {
          max_diagonal_ = VectorX<Scalar>(n);
          for (int i = 0; i < n; ++i)
            max_diagonal_[i] = std::max(undamped[i], (Scalar)p_.diagonal_damping_min);
        } else {
          for (int i = 0; i < n; ++i) max_diagonal_[i] = std::max(max_diagonal_[i], undamped[i]);
        }
        have_max_diagonal_ = true;
        for (int i = 0; i < n; ++i) damping_[i] = max_diagonal_[i] * current_lambda_;
      } else {
        for (int i = 0; i < n; ++i)
          damping_[i] = std::max(undamped[i], (Scalar)p_.diagonal_damping_min) * current_lambda_;
      }
    }
    if (p_.use_unit_damping)
      for (int i = 0; i < n; ++i) damping_[i] += current_lambda_;
    SetDiagonal(H, undamped, &damping_);
    if (!linear_solver_.Factorize(H))
      throw std::runtime_error("Internal Error: Damped hessian factorization failed");
    if (p_.debug_stats && iteration_ == 0) {
      stats.linear_solver_ordering = linear_solver_.Permutation().indices();
      stats.cholesky_factor_sparsity = GetSparseStructure(linear_solver_.L());
    }
    const VectorX<Scalar> x = linear_solver_.Solve(init.lin.rhs);
    update_ = VectorX<Scalar>(n);
    for (int i = 0; i < n; ++i) update_[i] = -x[i];
    SetDiagonal(H, undamped, nullptr);  // ResetHessianDiagonal
    // UpdateNewFromInit (levenberg_marquardt_state.h:217-239)
    Block& nw = blocks_[new_];
    nw.values = init.values;
    nw.values.Retract(index_, update_.data(), epsilon_);
    nw.valid = true;
    nw.lin.Reset();
    func(nw.values, nw.lin);
    nw.lin.SetInitialized();

    double nl2 = 0, nu2 = 0, dot = 0;
    if (p_.enable_bold_updates && have_last_update_)
      for (int i = 0; i < n; ++i) {
        nl2 += (double)last_update_[i] * last_update_[i];
        nu2 += (double)update_[i] * update_[i];
        dot += (double)last_update_[i] * update_[i];
      }
    symhip_lm_decision d{};
    check(symhip_lm_decide(init.lin.Error(), nw.lin.Error(), blocks_[best_].lin.Error(),
                           2.0 * (double)init.lin.LinearDeltaError(update_, damping_),
                           current_lambda_, current_nu_, iteration_, &p_,
                           have_last_update_ ? 1 : 0, nl2, nu2, dot, (double)epsilon_, &d));
    optimization_iteration_t it = FromRecord(d.record);
    if (p_.debug_stats) {
      it.update = update_.template cast<double>();
      it.values = Widened(nw.values).GetLcmType();
      it.residual = nw.lin.residual.template cast<double>();
    }
    stats.iterations.push_back(std::move(it));
    current_lambda_ = d.lambda;
    current_nu_ = d.nu;
    if (!d.accept) {
      std::swap(new_, init_);  // SwapNewAndInit
    } else {
      have_last_update_ = true;
      last_update_ = update_;
      if (d.set_best) {  // SetBestToNew
        if (best_ != new_) {
          if (best_ != init_) free_ = best_;
          best_ = new_;
        }
        stats.best_index = (int32_t)stats.iterations.size() - 1;
      }
      if (best_ == init_) {  // SetInitToNotBest
        init_ = free_;
        free_ = best_;
      }
    }
    if (d.status) return std::make_pair(d.status, (FailureReason)d.failure_reason);
    return std::nullopt;
  }

  const ValuesType& GetBestValues() const { return blocks_[best_].values; }
  const LinearizationType& GetBestLinearization() const { return blocks_[best_].lin; }

  // ComputeCovariance (levenberg_marquardt_solver.tcc:345-356)
  void ComputeCovariance(const MatrixType& hessian_lower, MatrixX<Scalar>& covariance) {
    if (!linear_solver_.Factorize(hessian_lower))
      throw std::runtime_error("Internal Error: damped hessian factorization failed");
    covariance = MatrixX<Scalar>::Identity(hessian_lower.rows());
    linear_solver_.SolveInPlace(covariance);
  }

 private:
  struct Block {
    ValuesType values;
    LinearizationType lin;
    bool valid = false;
  };
  static Values<double> Widened(const ValuesType& v) {
    if constexpr (std::is_same_v<Scalar, double>) {
      return v;
    } else {
      Values<double> w;
      for (const auto& e : v.Entries()) {
        std::vector<double> d(v.Data().begin() + e.offset,
                              v.Data().begin() + e.offset + e.storage_dim);
        RawSet(w, e, d);
      }
      return w;
    }
  }
  static void RawSet(Values<double>& w, const index_entry_t& e, const std::vector<double>& d) {
    switch (e.type) {
      case SYMHIP_TYPE_SCALAR: w.Set(e.key, d[0]); break;
      case SYMHIP_TYPE_ROT3: SetAs<Rot3<double>>(w, e.key, d); break;
      case SYMHIP_TYPE_POSE3: SetAs<Pose3<double>>(w, e.key, d); break;
      case SYMHIP_TYPE_VECTOR2: SetAs<Vector2d>(w, e.key, d); break;
      case SYMHIP_TYPE_VECTOR3: SetAs<Vector3d>(w, e.key, d); break;
      case SYMHIP_TYPE_VECTOR4: SetAs<Vector4d>(w, e.key, d); break;
      case SYMHIP_TYPE_VECTOR5: SetAs<Vector5d>(w, e.key, d); break;
      case SYMHIP_TYPE_VECTOR6: SetAs<Vector6d>(w, e.key, d); break;
      case SYMHIP_TYPE_MATRIX33: SetAs<Matrix33d>(w, e.key, d); break;
      default: SetAs<Matrix66d>(w, e.key, d); break;
    }
  }
  template <typename T>
  static void SetAs(Values<double>& w, const Key& k, const std::vector<double>& d) {
    T t;
    std::copy(d.begin(), d.end(), t.data.begin());
    w.Set(k, t);
  }
  // Writes undamped + damping (or undamped alone) into the stored diagonal.
  static void SetDiagonal(MatrixType& H, const VectorX<Scalar>& undamped,
                          const VectorX<Scalar>* damping) {
    for (int c = 0; c < H.cols(); ++c) {
      const int32_t* b = H.innerIndexPtr() + H.outerIndexPtr()[c];
      const int32_t* e = H.innerIndexPtr() + H.outerIndexPtr()[c + 1];
      const int32_t* it = std::lower_bound(b, e, c);
      if (it == e || *it != c) throw std::runtime_error("hessian_lower lacks a diagonal entry");
      H.valuePtr()[it - H.innerIndexPtr()] = undamped[c] + (damping ? (*damping)[c] : Scalar(0));
    }
  }

  optimizer_params_t p_;
  std::string id_;
  Scalar epsilon_;
  index_t index_;
  LinearSolverType linear_solver_{};
  bool solver_analyzed_ = false;
  std::array<Block, 3> blocks_;
  int init_ = 0, new_ = 1, best_ = 0, free_ = 2;
  bool have_max_diagonal_ = false, have_last_update_ = false;
  VectorX<Scalar> max_diagonal_, damping_, update_, last_update_;
  double current_lambda_ = 1, current_nu_ = 2;
  int iteration_ = -1;
};

// ------------------------------------------------------------------------------------------------
// Optimizer (opt/optimizer.h:72-325): the LM loop of NLS on the device.
template <typename ScalarType,
          typename _NonlinearSolverType = LevenbergMarquardtSolver<ScalarType>>
class Optimizer {
 public:
  using Scalar = ScalarType;
  using NonlinearSolverType = _NonlinearSolverType;
  using FailureReason = typename NonlinearSolverType::FailureReason;
  using MatrixType = typename NonlinearSolverType::MatrixType;
  using Stats = OptimizationStats<MatrixType>;

  Optimizer(const optimizer_params_t& params, std::vector<Factor<Scalar>> factors,
            const std::string& name = "sym::Optimize", std::vector<Key> keys = {},
            const Scalar epsilon = kDefaultEpsilon<Scalar>)
      : factors_(std::move(factors)), name_(name),
        nonlinear_solver_(params, "sym::Optimize", epsilon), epsilon_(epsilon),
        keys_(keys.empty() ? ComputeKeysToOptimize(factors_) : std::move(keys)) {}
  template <typename... NonlinearSolverArgs>
  Optimizer(const optimizer_params_t& params, std::vector<Factor<Scalar>> factors,
            const std::string& name, std::vector<Key> keys, const Scalar epsilon,
            NonlinearSolverArgs&&... nonlinear_solver_args)
      : factors_(std::move(factors)), name_(name),
        nonlinear_solver_(params, name, epsilon,
                          std::forward<NonlinearSolverArgs>(nonlinear_solver_args)...),
        epsilon_(epsilon),
        keys_(keys.empty() ? ComputeKeysToOptimize(factors_) : std::move(keys)) {}
  Optimizer(Optimizer&&) = delete;
  Optimizer& operator=(Optimizer&&) = delete;
  Optimizer(const Optimizer&) = delete;
  Optimizer& operator=(const Optimizer&) = delete;
  virtual ~Optimizer() {
    if (opt_) symhip_optimizer_destroy(opt_);
    if (problem_) symhip_problem_destroy(problem_);
  }

  Stats Optimize(Values<Scalar>& values, int num_iterations = -1,
                 bool populate_best_linearization = false) {
    Stats stats;
    Optimize(values, num_iterations, populate_best_linearization, stats);
    return stats;
  }
  // Optimizer::Optimize (opt/optimizer.tcc:70-89, internal/optimizer_utils.h:36-104)
  virtual void Optimize(Values<Scalar>& values, int num_iterations,
                        bool populate_best_linearization, Stats& stats) {
    Initialize(values);
    RunOnDevice(values, num_iterations, populate_best_linearization, stats, nullptr, -1);
  }
  void Optimize(Values<Scalar>& values, int num_iterations, Stats& stats) {
    Optimize(values, num_iterations, false, stats);
  }
  void Optimize(Values<Scalar>& values, Stats& stats) { Optimize(values, -1, false, stats); }

  // Linearize the problem around the given values (opt/optimizer.tcc:91-100)
  Linearization<MatrixType> Linearize(const Values<Scalar>& values) {
    Initialize(values);
    std::vector<double> v(values.Data().begin(), values.Data().end());
    std::vector<double> r(M_), h(nnz_), b(N_);
    check(symhip_linearize(problem_, v.data(), r.data(), h.data(), b.data()));
    Linearization<MatrixType> lin;
    FillLinearization(lin, r, h, b);
    return lin;
  }

  // ComputeAllCovariances (opt/optimizer.tcc:113-121): inverse of the Hessian, split by key
  void ComputeAllCovariances(const Linearization<MatrixType>& linearization,
                             std::unordered_map<Key, MatrixX<Scalar>>& covariances_by_key) {
    MatrixX<Scalar> cov;
    ComputeFullCovariance(linearization, cov);
    covariances_by_key.clear();
    for (size_t k = 0; k < keys_.size(); ++k)
      covariances_by_key[keys_[k]] =
          cov.block(tangent_offsets_[k], tangent_offsets_[k], tangent_dims_[k], tangent_dims_[k]);
  }

  // ComputeCovariances (opt/optimizer.tcc:123-200, internal/covariance_utils.h:151-170): the
  // covariance block of `keys`, a prefix of Keys(). The block of H^-1 over the leading
  // variables equals the inverse of their Schur complement (B - E C^-1 E^T)^-1; it is computed
  // on the device as the leading columns of H^-1 from one LDL^T factorization of H (the
  // right-hand sides are the leading unit vectors). Returns 0 (Eigen::Success).
  int ComputeCovariances(const Linearization<MatrixType>& linearization,
                         const std::vector<Key>& keys,
                         std::unordered_map<Key, MatrixX<Scalar>>& covariances_by_key,
                         bool c_is_block_diagonal = true) {
    (void)c_is_block_diagonal;  // either structure takes the same device path
    if (keys.size() > keys_.size() || !std::equal(keys.begin(), keys.end(), keys_.begin()))
      throw std::runtime_error("ComputeCovariances: keys must be a prefix of the optimized keys");
    int nb = 0;
    for (size_t k = 0; k < keys.size(); ++k) nb += tangent_dims_[k];
    hip::SparseCholeskySolver<MatrixType> solver;
    solver.AnalyzeSparsityPattern(linearization.hessian_lower);
    solver.Factorize(linearization.hessian_lower);
    const MatrixX<Scalar> cols =
        solver.Solve(MatrixX<Scalar>::Identity(linearization.hessian_lower.rows(), nb));
    covariances_by_key.clear();
    for (size_t k = 0; k < keys.size(); ++k)
      covariances_by_key[keys[k]] =
          cols.block(tangent_offsets_[k], tangent_offsets_[k], tangent_dims_[k], tangent_dims_[k]);
    return 0;
  }

  // ComputeFullCovariance (opt/optimizer.tcc:202-206)
  void ComputeFullCovariance(const Linearization<MatrixType>& linearization,
                             MatrixX<Scalar>& covariance) {
    hip::SparseCholeskySolver<MatrixType> solver;
    solver.AnalyzeSparsityPattern(linearization.hessian_lower);
    if (!solver.Factorize(linearization.hessian_lower))
      throw std::runtime_error("Internal Error: damped hessian factorization failed");
    covariance = MatrixX<Scalar>::Identity(linearization.hessian_lower.rows());
    solver.SolveInPlace(covariance);
  }

  const std::vector<Key>& Keys() const { return keys_; }
  const std::vector<Factor<Scalar>>& Factors() const { return factors_; }
  const NonlinearSolverType& NonlinearSolver() const { return nonlinear_solver_; }
  NonlinearSolverType& NonlinearSolver() { return nonlinear_solver_; }
  void UpdateParams(const optimizer_params_t& params) {
    nonlinear_solver_.UpdateParams(params);
    if (opt_) check(symhip_optimizer_update_params(opt_, &params));
  }
  const optimizer_params_t& Params() const { return nonlinear_solver_.Params(); }

 protected:
  bool IsInitialized() const { return opt_ != nullptr; }

  // The device LM run of Optimize, or (gnc != null) of GncOptimizer::Optimize with the Scalar
  // key at mu_key_index of the Values as the convexity parameter.
  void RunOnDevice(Values<Scalar>& values, int num_iterations, bool populate_best_linearization,
                   Stats& stats, const symhip_gnc_params* gnc, int32_t mu_key_index) {
    const optimizer_params_t& p = Params();
    const int n = num_iterations < 0 ? p.iterations : num_iterations;
    stats.Reset(n + 1);
    std::vector<double> v(values.Data().begin(), values.Data().end());
    std::vector<symhip_iteration_stats> recs(n + 2);
    symhip_optimization_stats st{};
    if (gnc)
      check(symhip_optimizer_optimize_gnc(opt_, v.data(), num_iterations, gnc, mu_key_index,
                                          recs.data(), (int32_t)recs.size(), &st));
    else
      check(symhip_optimizer_optimize(opt_, v.data(), num_iterations, recs.data(),
                                      (int32_t)recs.size(), &st));
    for (size_t i = 0; i < values.Data().size(); ++i) values.Data()[i] = static_cast<Scalar>(v[i]);
    const index_t all = values.CreateIndex();
    std::vector<int32_t> jcp, jri;
    if (p.include_jacobians) {
      const int64_t jnnz = symhip_optimizer_jacobian_sparsity(opt_, nullptr, nullptr);
      if (jnnz < 0) check((int)jnnz);
      jcp.resize(N_ + 1);
      jri.resize(jnnz);
      check((int)std::min<int64_t>(0, symhip_optimizer_jacobian_sparsity(opt_, jcp.data(),
                                                                          jri.data())));
      if (p.debug_stats) stats.jacobian_sparsity = {jri, jcp, {{M_, N_}}};
    }
    for (int32_t i = 0; i < st.num_iterations; ++i) {
      optimization_iteration_t it = FromRecord(recs[i]);
      if (p.debug_stats) {
        std::vector<double> upd(N_), vals(values.Data().size()), res(M_);
        check(symhip_optimizer_debug_record(opt_, i, upd.data(), vals.data(), res.data()));
        if (i > 0) it.update = ToVector(upd);
        it.values.index = all;
        it.values.data = vals;
        it.residual = ToVector(res);
        if (p.include_jacobians) {
          // optimization_iteration_t.jacobian_values: the Jacobian's nonzeros, CSC order
          std::vector<double> jv(jri.size());
          check(symhip_optimizer_debug_jacobian(opt_, i, jv.data()));
          it.jacobian_values = ToVector(jv);
        }
      }
      stats.iterations.push_back(std::move(it));
    }
    stats.best_index = st.best_index;
    stats.status = st.status;
    stats.failure_reason = st.failure_reason;
    if (p.debug_stats && device_solver_ == SYMHIP_SOLVER_CHOLESKY) {
      const int64_t nnz = symhip_optimizer_factor_sparsity(opt_, nullptr, nullptr, nullptr);
      if (nnz < 0) check((int)nnz);
      stats.linear_solver_ordering.resize(N_);
      auto& L = stats.cholesky_factor_sparsity;
      L.column_pointers.resize(N_ + 1);
      L.row_indices.resize(nnz);
      L.shape = {{N_, N_}};
      const int64_t got = symhip_optimizer_factor_sparsity(
          opt_, stats.linear_solver_ordering.data(), L.column_pointers.data(),
          L.row_indices.data());
      if (got < 0) check((int)got);
    }
    if (populate_best_linearization) {
      Linearization<MatrixType> lin;
      std::vector<double> r(M_), h(nnz_), b(N_);
      check(symhip_optimizer_best_linearization(opt_, r.data(), h.data(), b.data()));
      FillLinearization(lin, r, h, b);
      if (p.include_jacobians) {
        std::vector<double> jv(jri.size());
        check(symhip_optimizer_best_jacobian(opt_, jv.data()));
        lin.jacobian = MatrixType(M_, N_, jcp, jri, std::vector<Scalar>(jv.begin(), jv.end()));
      }
      stats.best_linearization = std::move(lin);
    }
  }

  // Optimizer::Initialize (opt/optimizer.tcc:261-278): the problem structure is fixed after
  // the first call, as the reference Linearizer's is.
  void Initialize(const Values<Scalar>& values) {
    if (opt_) return;
    std::vector<int8_t> letters;
    std::vector<int64_t> subs, supers;
    std::vector<int32_t> types;
    for (const auto& e : values.Entries()) {
      letters.push_back((int8_t)e.key.Letter());
      subs.push_back(e.key.Sub());
      supers.push_back(e.key.Super());
      types.push_back(e.type);
    }
    std::vector<double> v(values.Data().begin(), values.Data().end());
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
    d.values = v.data();
    d.n_factors = (int32_t)ftypes.size();
    d.factor_types = ftypes.data();
    d.factor_arg_ptr = fptr.data();
    d.factor_args = fargs.data();
    d.n_opt_keys = (int32_t)opt.size();
    d.opt_keys = opt.empty() ? nullptr : opt.data();
    const symhip_solver_options o = nonlinear_solver_.LinearSolver().DeviceOptions();
    device_solver_ = o.linear_solver;
    check(symhip_problem_create(&d, &problem_));
    int64_t dims[5];
    check(symhip_problem_dims(problem_, dims));
    N_ = (int32_t)dims[0];
    M_ = (int32_t)dims[1];
    nnz_ = dims[2];
    colptr_.resize(N_ + 1);
    rowind_.resize(nnz_);
    check(symhip_problem_hessian_structure(problem_, colptr_.data(), rowind_.data()));
    std::vector<int32_t> kidx(dims[4]), toff(dims[4]);
    check(symhip_problem_state_keys(problem_, kidx.data(), toff.data()));
    tangent_offsets_ = toff;
    tangent_dims_.clear();
    for (int32_t k : kidx) tangent_dims_.push_back(values.Entries()[k].tangent_dim);
    check(symhip_optimizer_create(&d, &Params(), &o, (double)epsilon_, &opt_));
  }

  std::vector<Factor<Scalar>> factors_;
  std::string name_;
  NonlinearSolverType nonlinear_solver_;
  Scalar epsilon_;
  std::vector<Key> keys_;

 private:
  static VectorX<double> ToVector(const std::vector<double>& v) {
    VectorX<double> out((int)v.size());
    std::copy(v.begin(), v.end(), out.data());
    return out;
  }
  void FillLinearization(Linearization<MatrixType>& lin, const std::vector<double>& r,
                         const std::vector<double>& h, const std::vector<double>& b) const {
    lin.residual = VectorX<Scalar>((int)r.size());
    for (size_t i = 0; i < r.size(); ++i) lin.residual[i] = static_cast<Scalar>(r[i]);
    lin.rhs = VectorX<Scalar>((int)b.size());
    for (size_t i = 0; i < b.size(); ++i) lin.rhs[i] = static_cast<Scalar>(b[i]);
    lin.hessian_lower =
        MatrixType(N_, N_, colptr_, rowind_, std::vector<Scalar>(h.begin(), h.end()));
    lin.SetInitialized();
  }

  symhip_problem_t problem_ = nullptr;
  symhip_optimizer_t opt_ = nullptr;
  int32_t device_solver_ = SYMHIP_SOLVER_CHOLESKY;
  int32_t N_ = 0, M_ = 0;
  int64_t nnz_ = 0;
  std::vector<int32_t> colptr_, rowind_, tangent_offsets_, tangent_dims_;
};

using Optimizerd = Optimizer<double>;
using Optimizerf = Optimizer<float>;

// optimizer_gnc_params_t (lcmtypes/symforce.lcm:203-227)
using optimizer_gnc_params_t = symhip_gnc_params;

// GncOptimizer<BaseOptimizerType> (opt/gnc_optimizer.h:22-148): graduated non-convexity over an
// Optimizer whose factors read the Scalar gnc_mu_key (0: convex, towards 1: robust). Optimize
// sets mu to mu_initial, runs to convergence with early exit at gnc_update_min_reduction, then
// repeatedly raises mu by mu_step (capped at mu_max, where the user's early exit returns),
// relaxes the damping to its initial value and continues from the best Values, until a stage
// ends without SUCCESS, mu is at its maximum, or the iteration budget (shared by all stages) is
// spent. The stages run in the device optimizer (symhip_optimizer_optimize_gnc).
template <typename BaseOptimizerType>
class GncOptimizer : public BaseOptimizerType {
 public:
  using BaseOptimizer = BaseOptimizerType;
  using Scalar = typename BaseOptimizer::Scalar;
  using Stats = typename BaseOptimizer::Stats;

  template <typename... OptimizerArgs>
  GncOptimizer(const optimizer_params_t& optimizer_params,
               const optimizer_gnc_params_t& gnc_params, const Key& gnc_mu_key,
               OptimizerArgs&&... args)
      : BaseOptimizer(optimizer_params, std::forward<OptimizerArgs>(args)...),
        gnc_params_(gnc_params),
        gnc_mu_key_(gnc_mu_key) {}
  ~GncOptimizer() override = default;

  using BaseOptimizerType::Optimize;
  void Optimize(Values<Scalar>& values, int num_iterations, bool populate_best_linearization,
                Stats& stats) override {
    values.Set(gnc_mu_key_, static_cast<Scalar>(gnc_params_.mu_initial));
    this->Initialize(values);
    this->RunOnDevice(values, num_iterations, populate_best_linearization, stats, &gnc_params_,
                      values.IndexOf(gnc_mu_key_));
  }

 private:
  optimizer_gnc_params_t gnc_params_;
  Key gnc_mu_key_;
};

// Optimize(params, factors, values) (opt/optimizer.h:334-340)
template <typename Scalar, typename NonlinearSolverType = LevenbergMarquardtSolver<Scalar>>
typename Optimizer<Scalar, NonlinearSolverType>::Stats Optimize(
    const optimizer_params_t& params, const std::vector<Factor<Scalar>>& factors,
    Values<Scalar>& values, const Scalar epsilon = kDefaultEpsilon<Scalar>) {
  Optimizer<Scalar, NonlinearSolverType> optimizer(params, factors, "sym::Optimize", {},
                                                   epsilon);
  return optimizer.Optimize(values);
}

}  // namespace sym
