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

// XCD-aware workgroup order. The dispatcher hands workgroup i to XCD i % 8, each XCD with its own
// L2; kernels whose neighbouring tiles share inputs launch xcd_grid(n) workgroups and process
// logical tile xcd_tile(blockIdx.x, n): XCD k then owns the contiguous logical range
// [k * per, (k + 1) * per). Tiles past n (padding) return at once.
constexpr int kXcds = 8;
inline int xcd_grid(int n_wg) { return ((n_wg + kXcds - 1) / kXcds) * kXcds; }
__device__ __forceinline__ int xcd_tile(int wg, int n_wg) {
  const int per = (n_wg + kXcds - 1) / kXcds;
  return (wg % kXcds) * per + wg / kXcds;
}

}  // namespace symhip
