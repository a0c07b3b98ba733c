// Small device utilities shared by the kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "./common.h"

namespace symhip {

// Zero-fill with a kernel instead of hipMemsetAsync so the call is safe inside captured graphs.
void fill_zero(double* p, int64_t n, hipStream_t s);
void fill_zero(float* p, int64_t n, hipStream_t s);
void fill_zero_i32(int32_t* p, int64_t n, hipStream_t s);

// Deterministic two-pass FP64 reductions into a device scalar (out[0]).
// sum_sq: out = 0.5 * sum(x^2)  (Linearization::Error, opt/linearization.h:51-54)
void half_squared_norm(const double* x, int64_t n, double* partials, double* out, hipStream_t s);
// out = sum(a .* b)
void dot(const double* a, const double* b, int64_t n, double* partials, double* out, hipStream_t s);
constexpr int kReduceBlocks = 1024;

inline int grid_for(int64_t n, int block) { return (int)((n + block - 1) / block); }

}  // namespace symhip
