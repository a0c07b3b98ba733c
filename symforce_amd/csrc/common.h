// Shared host-side definitions: error handling, device buffers, value/factor type tables.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/symhip_types.h"

namespace symhip {

struct Error : std::runtime_error {
  using std::runtime_error::runtime_error;
};

#define SYMHIP_CHECK(cond, msg)                                  \
  do {                                                           \
    if (!(cond)) throw ::symhip::Error(std::string("symhip: ") + (msg)); \
  } while (0)

#define HIP_OK(expr)                                                                        \
  do {                                                                                      \
    hipError_t _e = (expr);                                                                 \
    if (_e != hipSuccess)                                                                   \
      throw ::symhip::Error(std::string("HIP error ") + hipGetErrorString(_e) + " at " +   \
                            __FILE__ + ":" + std::to_string(__LINE__) + " in " #expr);      \
  } while (0)

// Owning device allocation (no copy, movable).
template <typename T>
class DBuf {
 public:
  DBuf() = default;
  explicit DBuf(size_t n) { resize(n); }
  ~DBuf() { release(); }
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  DBuf(DBuf&& o) noexcept : p_(o.p_), n_(o.n_) {
    o.p_ = nullptr;
    o.n_ = 0;
  }
  DBuf& operator=(DBuf&& o) noexcept {
    if (this != &o) {
      release();
      p_ = o.p_;
      n_ = o.n_;
      o.p_ = nullptr;
      o.n_ = 0;
    }
    return *this;
  }
  void resize(size_t n) {
    if (n == n_) return;
    release();
    if (n) HIP_OK(hipMalloc(&p_, n * sizeof(T)));
    n_ = n;
  }
  void release() {
    if (p_) (void)hipFree(p_);
    p_ = nullptr;
    n_ = 0;
  }
  void upload(const T* h, size_t n, hipStream_t s) {
    resize(n);
    if (n) HIP_OK(hipMemcpyAsync(p_, h, n * sizeof(T), hipMemcpyHostToDevice, s));
  }
  void upload(const std::vector<T>& h, hipStream_t s) { upload(h.data(), h.size(), s); }
  void download(T* h, size_t n, hipStream_t s) const {
    if (n) HIP_OK(hipMemcpyAsync(h, p_, n * sizeof(T), hipMemcpyDeviceToHost, s));
  }
  void zero(hipStream_t s) {
    if (n_) HIP_OK(hipMemsetAsync(p_, 0, n_ * sizeof(T), s));
  }
  T* get() const { return p_; }
  size_t size() const { return n_; }

 private:
  T* p_ = nullptr;
  size_t n_ = 0;
};

inline int storage_dim(int type) {
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
  throw Error("symhip: unknown value type " + std::to_string(type));
}
inline int tangent_dim(int type) {
  if (type == SYMHIP_TYPE_ROT3) return 3;
  if (type == SYMHIP_TYPE_POSE3) return 6;
  return storage_dim(type);
}

// Argument signature of each registered factor type; bit k of lin_mask marks argument k as
// linearized. The order of linearized arguments defines the factor's tangent layout.
struct FactorSpec {
  int nargs;
  int arg_types[6];
  int lin_mask;
  int res_dim;
  int tdim;
};
inline FactorSpec factor_spec(int type) {
  const int P = SYMHIP_TYPE_POSE3, R = SYMHIP_TYPE_ROT3, S = SYMHIP_TYPE_SCALAR,
            V2 = SYMHIP_TYPE_VECTOR2, V3 = SYMHIP_TYPE_VECTOR3, V4 = SYMHIP_TYPE_VECTOR4,
            V5 = SYMHIP_TYPE_VECTOR5, V6 = SYMHIP_TYPE_VECTOR6, M3 = SYMHIP_TYPE_MATRIX33,
            M6 = SYMHIP_TYPE_MATRIX66;
  switch (type) {
    case SYMHIP_FACTOR_PRIOR_POSE3: return {4, {P, P, M6, S}, 0b1, 6, 6};
    case SYMHIP_FACTOR_BETWEEN_POSE3: return {5, {P, P, P, M6, S}, 0b11, 6, 12};
    case SYMHIP_FACTOR_PRIOR_ROT3: return {4, {R, R, M3, S}, 0b1, 3, 3};
    case SYMHIP_FACTOR_BETWEEN_ROT3: return {5, {R, R, R, M3, S}, 0b11, 3, 6};
    case SYMHIP_FACTOR_SNAVELY: return {5, {P, V3, V3, V2, S}, 0b111, 2, 12};
    case SYMHIP_FACTOR_SNAVELY_HUBER: return {6, {P, V3, V3, V2, S, S}, 0b111, 2, 12};
    case SYMHIP_FACTOR_REPROJ_CAUCHY: return {6, {P, V3, V2, V4, S, S}, 0b11, 2, 9};
    case SYMHIP_FACTOR_MATCHING: return {4, {P, V3, V3, S}, 0b1, 3, 6};
    case SYMHIP_FACTOR_ODOMETRY: return {5, {P, P, P, V6, S}, 0b11, 6, 12};
    case SYMHIP_FACTOR_BETWEEN_POSE3_CAUCHY: return {6, {P, P, P, M6, S, S}, 0b11, 6, 12};
    case SYMHIP_FACTOR_BARRON_VECTOR5: return {4, {V5, V5, S, S}, 0b1, 5, 5};
  }
  throw Error("symhip: unknown factor type " + std::to_string(type));
}
constexpr int kNumFactorTypes = 12;  // valid codes are 1..11

}  // namespace symhip
