#include "./util.h"

namespace symhip {

namespace {

template <typename T>
__global__ void fill_zero_kernel(T* p, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = T(0);
}

// Block-level sum of one double per thread (wave64 shuffles + LDS across waves).
__device__ double block_sum(double v) {
  __shared__ double warp_sums[16];
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) warp_sums[wid] = v;
  __syncthreads();
  double t = 0;
  if (threadIdx.x == 0) {
    const int nw = (blockDim.x + 63) >> 6;
    for (int i = 0; i < nw; ++i) t += warp_sums[i];
  }
  __syncthreads();
  return t;
}

__global__ void __launch_bounds__(256) sq_partial_kernel(const double* __restrict__ x, int64_t n,
                                                         double* __restrict__ partials) {
  double acc = 0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    acc += x[i] * x[i];
  acc = block_sum(acc);
  if (threadIdx.x == 0) partials[blockIdx.x] = acc;
}

__global__ void __launch_bounds__(256) dot_partial_kernel(const double* __restrict__ a,
                                                          const double* __restrict__ b, int64_t n,
                                                          double* __restrict__ partials) {
  double acc = 0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    acc += a[i] * b[i];
  acc = block_sum(acc);
  if (threadIdx.x == 0) partials[blockIdx.x] = acc;
}

__global__ void __launch_bounds__(256) finish_kernel(const double* __restrict__ partials, int nb,
                                                     double scale, double* __restrict__ out) {
  double acc = 0;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) acc += partials[i];
  acc = block_sum(acc);
  if (threadIdx.x == 0) out[0] = scale * acc;
}

}  // namespace

void fill_zero(double* p, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  const int grid = (int)std::min<int64_t>(grid_for(n, 256), 4096);
  hipLaunchKernelGGL(fill_zero_kernel<double>, dim3(grid), dim3(256), 0, s, p, n);
  HIP_OK(hipGetLastError());
}
void fill_zero(float* p, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  const int grid = (int)std::min<int64_t>(grid_for(n, 256), 4096);
  hipLaunchKernelGGL(fill_zero_kernel<float>, dim3(grid), dim3(256), 0, s, p, n);
  HIP_OK(hipGetLastError());
}
void fill_zero_i32(int32_t* p, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  const int grid = (int)std::min<int64_t>(grid_for(n, 256), 4096);
  hipLaunchKernelGGL(fill_zero_kernel<int32_t>, dim3(grid), dim3(256), 0, s, p, n);
  HIP_OK(hipGetLastError());
}

void half_squared_norm(const double* x, int64_t n, double* partials, double* out, hipStream_t s) {
  const int nb = (int)std::max<int64_t>(1, std::min<int64_t>(kReduceBlocks, grid_for(n, 256)));
  hipLaunchKernelGGL(sq_partial_kernel, dim3(nb), dim3(256), 0, s, x, n, partials);
  hipLaunchKernelGGL(finish_kernel, dim3(1), dim3(256), 0, s, partials, nb, 0.5, out);
  HIP_OK(hipGetLastError());
}

void dot(const double* a, const double* b, int64_t n, double* partials, double* out,
         hipStream_t s) {
  const int nb = (int)std::max<int64_t>(1, std::min<int64_t>(kReduceBlocks, grid_for(n, 256)));
  hipLaunchKernelGGL(dot_partial_kernel, dim3(nb), dim3(256), 0, s, a, b, n, partials);
  hipLaunchKernelGGL(finish_kernel, dim3(1), dim3(256), 0, s, partials, nb, 1.0, out);
  HIP_OK(hipGetLastError());
}

}  // namespace symhip
