// Learned block-Jacobi correction (see gnn.h): feature extraction, BF16 matrix-core GEMMs,
// graph aggregation, the per-camera preconditioner head with its loss gradient, and Adam.
#include "./gnn.h"

#include <algorithm>
#include <cmath>
#include <random>

#include "./util.h"

namespace symhip {

namespace {

constexpr int kNd = LearnedPrecond::kNd, kF = LearnedPrecond::kF, kH = LearnedPrecond::kH,
              kO = LearnedPrecond::kO;
constexpr int kNe = kNd * kNd;
constexpr double kScale = 0.1;
// parameter offsets (row-major weights, W[k][j] = W[k * width + j])
constexpr int kOffW1 = 0, kOffB1 = kOffW1 + 2 * kF * kH, kOffW2 = kOffB1 + kH,
              kOffB2 = kOffW2 + 2 * kH * kH, kOffW3 = kOffB2 + kH, kOffB3 = kOffW3 + 2 * kH * kH,
              kOffWo = kOffB3 + kH, kOffBo = kOffWo + kH * kO;
static_assert(kOffBo + kO == LearnedPrecond::kParams, "parameter layout");

__device__ __forceinline__ int tri(int r, int c) { return r * (r + 1) / 2 + c; }

// Round to nearest even bfloat16 (finite inputs).
__device__ __forceinline__ short to_bf16(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (short)(u >> 16);
}

__global__ void block_norm_kernel(int64_t nblk, const double* __restrict__ S,
                                  double* __restrict__ nrm) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= nblk) return;
  double s = 0;
  for (int e = 0; e < kNe; ++e) {
    const double v = S[t * kNe + e];
    s += v * v;
  }
  nrm[t] = sqrt(s);
}

__device__ __forceinline__ double wsum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ double wmax(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ double wmin(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
  return v;
}

// Node features and row-normalized edge weights (one wavefront per camera).
__global__ void __launch_bounds__(64) features_kernel(int32_t n, const int32_t* __restrict__ row_ptr,
                                                      const int32_t* __restrict__ col,
                                                      const int32_t* __restrict__ diag,
                                                      const double* __restrict__ nrm,
                                                      const double* __restrict__ S,
                                                      float* __restrict__ x, float* __restrict__ w) {
  const int a = blockIdx.x, lane = threadIdx.x;
  const int t0 = row_ptr[a], t1 = row_ptr[a + 1];
  const double naa = nrm[diag[a]];
  double deg = 0, osum = 0, omax = 0, rsum = 0;
  for (int t = t0 + lane; t < t1; t += 64) {
    const int b = col[t];
    double wr = 0;
    if (b != a) {
      deg += 1;
      osum += nrm[t];
      omax = fmax(omax, nrm[t]);
      wr = nrm[t] / sqrt(naa * nrm[diag[b]]);
    }
    rsum += wr;
    w[t] = (float)wr;
  }
  deg = wsum(deg);
  osum = wsum(osum);
  omax = wmax(omax);
  rsum = wsum(rsum);
  for (int t = t0 + lane; t < t1; t += 64) {
    const int b = col[t];
    const double wr = b != a ? nrm[t] / sqrt(naa * nrm[diag[b]]) : 0.0;
    w[t] = rsum > 0 ? (float)(wr / rsum) : 0.0f;
  }
  const double d = lane < kNd ? S[(int64_t)diag[a] * kNe + lane * (kNd + 1)] : 1.0;
  const double ld = lane < kNd ? log(d) : 0.0;
  const double mean = wsum(ld) / kNd;
  const double dmax = wmax(lane < kNd ? d : 0.0), dmin = wmin(lane < kNd ? d : INFINITY);
  float* xa = x + (int64_t)a * kF;
  if (lane < kNd) xa[lane] = (float)(ld - mean);
  if (lane == 0) {
    xa[9] = (float)log1p(deg);
    xa[10] = (float)log(osum / naa + 1e-12);
    xa[11] = (float)log(omax / naa + 1e-12);
    xa[12] = (float)log(dmax / dmin);
    xa[13] = 1.0f;
    xa[14] = 0.0f;
    xa[15] = 0.0f;
  }
}

// z[a] = [h_a, sum_b w_ab h_b] (one wavefront per camera, lane = feature).
__global__ void __launch_bounds__(64) concat_aggregate_kernel(
    int32_t n, const int32_t* __restrict__ row_ptr, const int32_t* __restrict__ col,
    const float* __restrict__ w, const float* __restrict__ h, int K, float* __restrict__ z) {
  const int a = blockIdx.x, j = threadIdx.x;
  if (j >= K) return;
  float acc = 0;
  for (int t = row_ptr[a]; t < row_ptr[a + 1]; ++t) acc += w[t] * h[(int64_t)col[t] * K + j];
  z[(int64_t)a * 2 * K + j] = h[(int64_t)a * K + j];
  z[(int64_t)a * 2 * K + K + j] = acc;
}

// dh[a] = dz[a, :K] + sum_b w_ba dz[b, K:] (A^T through the mirror block index).
__global__ void __launch_bounds__(64) aggregate_transpose_kernel(
    int32_t n, const int32_t* __restrict__ row_ptr, const int32_t* __restrict__ col,
    const int32_t* __restrict__ mirror, const float* __restrict__ w, const float* __restrict__ dz,
    int K, float* __restrict__ dh) {
  const int a = blockIdx.x, j = threadIdx.x;
  if (j >= K) return;
  float acc = 0;
  for (int t = row_ptr[a]; t < row_ptr[a + 1]; ++t)
    acc += w[mirror[t]] * dz[(int64_t)col[t] * 2 * K + K + j];
  dh[(int64_t)a * K + j] = dz[(int64_t)a * 2 * K + j] + acc;
}

// C = A B (+ bias) with A(i, k) = A[i a_rs + k a_cs], B(k, j) = B[k b_rs + j b_cs], both rounded
// to bfloat16, FP32 accumulation on v_mfma_f32_16x16x16bf16 (one wavefront per 16 x 16 tile of C:
// lane l feeds A(i = l % 16, k = 4 (l / 16) + u) and B(k = 4 (l / 16) + u, j = l % 16), u < 4, and
// holds C(4 (l / 16) + r, l % 16) in register r). R, if given, receives relu(C).
__global__ void __launch_bounds__(256) gemm_bf16_kernel(int M, int N, int K, const float* __restrict__ A,
                                                        int64_t a_rs, int64_t a_cs,
                                                        const float* __restrict__ B, int64_t b_rs,
                                                        int64_t b_cs, const float* __restrict__ bias,
                                                        float* __restrict__ C, int ldc,
                                                        float* __restrict__ R) {
  typedef short s4 __attribute__((ext_vector_type(4)));
  typedef float f4 __attribute__((ext_vector_type(4)));
  const int tiles_n = (N + 15) / 16;
  const int tile = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  if (tile >= ((M + 15) / 16) * tiles_n) return;
  const int lane = threadIdx.x & 63, row = lane & 15, kb = lane >> 4;
  const int ti = tile / tiles_n, tj = tile % tiles_n;
  const int i = ti * 16 + row, j = tj * 16 + row;
  const int ic = min(i, M - 1), jc = min(j, N - 1);
  f4 acc = {0, 0, 0, 0};
  for (int k0 = 0; k0 < K; k0 += 16) {
    s4 av, bv;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int k = k0 + 4 * kb + u;
      const int kc = min(k, K - 1);
      const float a = A[ic * a_rs + kc * a_cs];
      const float b = B[kc * b_rs + jc * b_cs];
      av[u] = to_bf16((i < M && k < K) ? a : 0.0f);
      bv[u] = to_bf16((j < N && k < K) ? b : 0.0f);
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc, 0, 0, 0);
  }
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int oi = ti * 16 + 4 * kb + r, oj = tj * 16 + row;
    if (oi < M && oj < N) {
      const float v = acc[r] + (bias ? bias[oj] : 0.0f);
      C[(int64_t)oi * ldc + oj] = v;
      if (R) R[(int64_t)oi * ldc + oj] = fmaxf(v, 0.0f);
    }
  }
}

// out[j] = sum_a v[a * ld + j] (one block per column).
__global__ void __launch_bounds__(256) colsum_kernel(int32_t n, const float* __restrict__ v, int ld,
                                                     float* __restrict__ out) {
  __shared__ float red[4];
  const int j = blockIdx.x;
  float s = 0;
  for (int a = threadIdx.x; a < n; a += 256) s += v[(int64_t)a * ld + j];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) out[j] = (red[0] + red[1]) + (red[2] + red[3]);
}

__global__ void relu_mask_kernel(int64_t n, const float* __restrict__ pre,
                                 const float* __restrict__ g, float* __restrict__ out) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t < n) out[t] = pre[t] > 0 ? g[t] : 0.0f;
}

// Per camera (one wavefront, FP64 in LDS): L = chol(S_aa), Li = L^-1, t = 2 tanh(o / 20),
// U = lowtri(t) with exp on the diagonal, K = U^T Li, M = K^T K -> Minv. With probes (Y = S X), pass 1 (AB null)
// writes the camera's parts of A = <M Y, X> and B = |M Y|^2; pass 2 (AB = {A, B}) forms
// E = (A / B) M Y_a - X_a, G = 2 A / (B Z) E Y_a^T, dU = Li (G + G^T) Li^T U and
// dout = 0.1 lowtri(dU) (times U_rr on the exponential diagonal).
__global__ void __launch_bounds__(64) head_kernel(int32_t n, const int32_t* __restrict__ diag,
                                                  const double* __restrict__ S,
                                                  const float* __restrict__ o,
                                                  double* __restrict__ Minv,
                                                  const double* __restrict__ X,
                                                  const double* __restrict__ Y, int P, double invZ,
                                                  const double* __restrict__ AB,
                                                  float* __restrict__ dout,
                                                  double* __restrict__ partA,
                                                  double* __restrict__ partB) {
  __shared__ double L[kNd][kNd + 1], Li[kNd][kNd + 1], U[kNd][kNd + 1], Km[kNd][kNd + 1],
      Mm[kNd][kNd + 1], E[kNd][8], G[kNd][kNd + 1], T[kNd][kNd + 1];
  const int a = blockIdx.x, lane = threadIdx.x;
  const int64_t n9 = (int64_t)n * kNd;
  for (int e = lane; e < kNe; e += 64) {
    const int r = e / kNd, c = e % kNd;
    L[r][c] = c <= r ? S[(int64_t)diag[a] * kNe + e] : 0.0;
    const double oc = c <= r ? 2.0 * tanh((double)o[(int64_t)a * kO + tri(r, c)] / 20.0) : 0.0;
    U[r][c] = r == c ? exp(oc) : oc;
  }
  __syncthreads();
  for (int k = 0; k < kNd; ++k) {
    if (lane == 0) L[k][k] = sqrt(L[k][k]);
    __syncthreads();
    if (lane > k && lane < kNd) L[lane][k] /= L[k][k];
    __syncthreads();
    for (int e = lane; e < kNe; e += 64) {
      const int r = e / kNd, c = e % kNd;
      if (c > k && c <= r) L[r][c] -= L[r][k] * L[c][k];
    }
    __syncthreads();
  }
  // Li column by column (lane = column): forward substitution L Li(:, c) = e_c
  if (lane < kNd) {
    const int c = lane;
    for (int r = 0; r < kNd; ++r) {
      double s = r == c ? 1.0 : 0.0;
      for (int k = 0; k < r; ++k) s -= L[r][k] * Li[k][c];
      Li[r][c] = s / L[r][r];
    }
  }
  __syncthreads();
  for (int e = lane; e < kNe; e += 64) {
    const int r = e / kNd, c = e % kNd;
    double s = 0;
    for (int k = 0; k < kNd; ++k) s += U[k][r] * Li[k][c];
    Km[r][c] = s;
  }
  __syncthreads();
  for (int e = lane; e < kNe; e += 64) {
    const int r = e / kNd, c = e % kNd;
    double s = 0;
    for (int k = 0; k < kNd; ++k) s += Km[k][r] * Km[k][c];
    Mm[r][c] = s;
    if (Minv) Minv[(int64_t)a * kNe + e] = s;
  }
  if (!X) return;
  __syncthreads();
  // u = M Y_a (lane per (row, probe))
  double u = 0, xv = 0;
  const bool act = lane < kNd * P;
  const int i = lane % kNd, p = lane / kNd;
  if (act) {
    xv = X[p * n9 + (int64_t)a * kNd + i];
    for (int k = 0; k < kNd; ++k) u += Mm[i][k] * Y[p * n9 + (int64_t)a * kNd + k];
  }
  if (!AB) {
    const double pa = wsum(act ? u * xv : 0.0), pb = wsum(act ? u * u : 0.0);
    if (lane == 0) {
      partA[a] = pa;
      partB[a] = pb;
    }
    return;
  }
  const double A = AB[0], B = AB[1];
  const double c = A / B, scale = 2.0 * A / B * invZ;
  if (act) E[i][p] = c * u - xv;
  __syncthreads();
  for (int e = lane; e < kNe; e += 64) {
    const int r = e / kNd, cc = e % kNd;
    double s = 0;
    for (int q = 0; q < P; ++q) s += E[r][q] * Y[q * n9 + (int64_t)a * kNd + cc];
    G[r][cc] = scale * s;
  }
  __syncthreads();
  // T = Li (G + G^T)
  for (int e = lane; e < kNe; e += 64) {
    const int r = e / kNd, cc = e % kNd;
    double s = 0;
    for (int k = 0; k < kNd; ++k) s += Li[r][k] * (G[k][cc] + G[cc][k]);
    T[r][cc] = s;
  }
  __syncthreads();
  // Km <- T Li^T
  for (int e = lane; e < kNe; e += 64) {
    const int r = e / kNd, cc = e % kNd;
    double s = 0;
    for (int k = 0; k < kNd; ++k) s += T[r][k] * Li[cc][k];
    Km[r][cc] = s;
  }
  __syncthreads();
  // dU = Km U, packed lower part into dout
  for (int e = lane; e < kNe; e += 64) {
    const int r = e / kNd, cc = e % kNd;
    if (cc > r) continue;
    double s = 0;
    for (int k = 0; k < kNd; ++k) s += Km[r][k] * U[k][cc];
    const double th = tanh((double)o[(int64_t)a * kO + tri(r, cc)] / 20.0);
    dout[(int64_t)a * kO + tri(r, cc)] =
        (float)(kScale * (1.0 - th * th) * s * (r == cc ? U[r][r] : 1.0));
  }
  if (lane < kO - 45) dout[(int64_t)a * kO + 45 + lane] = 0.0f;
}

__global__ void adam_kernel(int n, float* __restrict__ p, const float* __restrict__ g,
                            float* __restrict__ m, float* __restrict__ v, float lr, float c1,
                            float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = 0.9f * m[i] + 0.1f * gi, vi = 0.999f * v[i] + 0.001f * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + 1e-8f);
}

__global__ void sum_parts_kernel(int32_t n, const double* __restrict__ part, double* out) {
  __shared__ double red[4];
  double s = 0;
  for (int a = threadIdx.x; a < n; a += 256) s += part[a];
  s = wsum(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) out[0] = (red[0] + red[1]) + (red[2] + red[3]);
}

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__global__ void probe_kernel(int64_t n, uint64_t seed, double* __restrict__ out) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t h1 = mix64(seed ^ mix64((uint64_t)i * 2)), h2 = mix64(seed ^ mix64((uint64_t)i * 2 + 1));
  const double u1 = ((h1 >> 11) + 0.5) * 0x1.0p-53, u2 = (h2 >> 11) * 0x1.0p-53;
  out[i] = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

void gemm(int M, int N, int K, const float* A, int64_t a_rs, int64_t a_cs, const float* B,
          int64_t b_rs, int64_t b_cs, const float* bias, float* C, int ldc, float* R,
          hipStream_t s) {
  const int tiles = ((M + 15) / 16) * ((N + 15) / 16);
  hipLaunchKernelGGL(gemm_bf16_kernel, dim3(grid_for((int64_t)tiles * 64, 256)), dim3(256), 0, s,
                     M, N, K, A, a_rs, a_cs, B, b_rs, b_cs, bias, C, ldc, R);
}

}  // namespace

void LearnedPrecond::setup(int32_t n_nodes, const std::vector<int32_t>& row_ptr,
                           const std::vector<int32_t>& col, const std::vector<int32_t>& diag,
                           hipStream_t s) {
  n_ = n_nodes;
  nblk_ = (int64_t)col.size();
  std::vector<int32_t> mirror(nblk_);
  for (int32_t a = 0; a < n_; ++a)
    for (int32_t t = row_ptr[a]; t < row_ptr[a + 1]; ++t) {
      const int32_t b = col[t];
      const auto it = std::lower_bound(col.begin() + row_ptr[b], col.begin() + row_ptr[b + 1], a);
      SYMHIP_CHECK(it != col.begin() + row_ptr[b + 1] && *it == a,
                   "learned preconditioner: block pattern is not symmetric");
      mirror[t] = (int32_t)(it - col.begin());
    }
  row_ptr_.upload(row_ptr, s);
  col_.upload(col, s);
  diag_.upload(diag, s);
  mirror_.upload(mirror, s);
  const size_t n = std::max(1, n_);
  w_.resize(std::max<int64_t>(1, nblk_));
  nrm_.resize(std::max<int64_t>(1, nblk_));
  x_.resize(n * kF);
  z1_.resize(n * 2 * kF);
  for (DBuf<float>* b : {&p1_, &h1_, &p2_, &h2_, &p3_, &h3_, &dh_, &dp_}) b->resize(n * kH);
  z2_.resize(n * 2 * kH);
  z3_.resize(n * 2 * kH);
  dz_.resize(n * 2 * kH);
  o_.resize(n * kO);
  do_.resize(n * kO);
  part_.resize(2 * n + 4);
  red_.resize(kReduceBlocks);
  params_.resize(kParams);
  grad_.resize(kParams);
  m_.resize(kParams);
  v_.resize(kParams);
  init_params(0, s);
}

void LearnedPrecond::init_params(uint64_t seed, hipStream_t s) {
  // He-normal hidden layers, zero head: the untrained network is exactly block Jacobi.
  std::vector<float> p(kParams, 0.0f);
  std::mt19937_64 gen(seed);
  std::normal_distribution<float> nrm(0.0f, 1.0f);
  auto fill = [&](int off, int fan_in, int cols) {
    const float sd = std::sqrt(2.0f / fan_in);
    for (int i = 0; i < fan_in * cols; ++i) p[off + i] = sd * nrm(gen);
  };
  fill(kOffW1, 2 * kF, kH);
  fill(kOffW2, 2 * kH, kH);
  fill(kOffW3, 2 * kH, kH);
  set_params(p.data(), s);
}

void LearnedPrecond::set_params(const float* host, hipStream_t s) {
  params_.upload(host, kParams, s);
  fill_zero(m_.get(), kParams, s);
  fill_zero(v_.get(), kParams, s);
  HIP_OK(hipStreamSynchronize(s));
}

void LearnedPrecond::get_params(float* host, hipStream_t s) const {
  params_.download(host, kParams, s);
  HIP_OK(hipStreamSynchronize(s));
}

void LearnedPrecond::forward(const double* S, hipStream_t s) {
  const int n = n_;
  hipLaunchKernelGGL(block_norm_kernel, dim3(grid_for(nblk_, 256)), dim3(256), 0, s, nblk_, S,
                     nrm_.get());
  hipLaunchKernelGGL(features_kernel, dim3(n), dim3(64), 0, s, n, row_ptr_.get(), col_.get(),
                     diag_.get(), nrm_.get(), S, x_.get(), w_.get());
  const float* P = params_.get();
  hipLaunchKernelGGL(concat_aggregate_kernel, dim3(n), dim3(64), 0, s, n, row_ptr_.get(),
                     col_.get(), w_.get(), x_.get(), kF, z1_.get());
  gemm(n, kH, 2 * kF, z1_.get(), 2 * kF, 1, P + kOffW1, kH, 1, P + kOffB1, p1_.get(), kH,
       h1_.get(), s);
  hipLaunchKernelGGL(concat_aggregate_kernel, dim3(n), dim3(64), 0, s, n, row_ptr_.get(),
                     col_.get(), w_.get(), h1_.get(), kH, z2_.get());
  gemm(n, kH, 2 * kH, z2_.get(), 2 * kH, 1, P + kOffW2, kH, 1, P + kOffB2, p2_.get(), kH,
       h2_.get(), s);
  hipLaunchKernelGGL(concat_aggregate_kernel, dim3(n), dim3(64), 0, s, n, row_ptr_.get(),
                     col_.get(), w_.get(), h2_.get(), kH, z3_.get());
  gemm(n, kH, 2 * kH, z3_.get(), 2 * kH, 1, P + kOffW3, kH, 1, P + kOffB3, p3_.get(), kH,
       h3_.get(), s);
  gemm(n, kO, kH, h3_.get(), kH, 1, P + kOffWo, kO, 1, P + kOffBo, o_.get(), kO, nullptr, s);
}

void LearnedPrecond::preconditioner(const double* S, double* Minv, hipStream_t s) {
  SYMHIP_CHECK(ready(), "learned preconditioner not set up");
  forward(S, s);
  hipLaunchKernelGGL(head_kernel, dim3(n_), dim3(64), 0, s, n_, diag_.get(), S, o_.get(), Minv,
                     nullptr, nullptr, 0, 0.0, nullptr, nullptr, nullptr, nullptr);
  HIP_OK(hipGetLastError());
}

float LearnedPrecond::loss_grad(const double* S, const double* X, const double* Y, int P,
                                hipStream_t s) {
  SYMHIP_CHECK(ready(), "learned preconditioner not set up");
  SYMHIP_CHECK(P >= 1 && P <= 7, "learned preconditioner: 1..7 probe vectors");
  const int n = n_;
  const int64_t n9 = (int64_t)n * kNd;
  // Z = |X|^2 on the host side of the boundary: X is small (P n 9 doubles)
  dot(X, X, n9 * P, red_.get(), part_.get() + 2 * n + 2, s);
  double Z = 0;
  HIP_OK(hipMemcpyAsync(&Z, part_.get() + 2 * n + 2, sizeof(double), hipMemcpyDeviceToHost, s));
  HIP_OK(hipStreamSynchronize(s));
  SYMHIP_CHECK(Z > 0, "learned preconditioner: zero probes");
  forward(S, s);
  // pass 1: parts of A = <M Y, X>, B = |M Y|^2; reduce; pass 2: gradient at the optimal scale
  double* pA = part_.get();
  double* pB = part_.get() + n;
  double* AB = part_.get() + 2 * n;
  hipLaunchKernelGGL(head_kernel, dim3(n), dim3(64), 0, s, n, diag_.get(), S, o_.get(), nullptr,
                     X, Y, P, 1.0 / Z, nullptr, nullptr, pA, pB);
  hipLaunchKernelGGL(sum_parts_kernel, dim3(1), dim3(256), 0, s, n, pA, AB);
  hipLaunchKernelGGL(sum_parts_kernel, dim3(1), dim3(256), 0, s, n, pB, AB + 1);
  hipLaunchKernelGGL(head_kernel, dim3(n), dim3(64), 0, s, n, diag_.get(), S, o_.get(), nullptr,
                     X, Y, P, 1.0 / Z, AB, do_.get(), nullptr, nullptr);
  float* Gp = grad_.get();
  const float* Pp = params_.get();
  // head
  gemm(kH, kO, n, h3_.get(), 1, kH, do_.get(), kO, 1, nullptr, Gp + kOffWo, kO, nullptr, s);
  hipLaunchKernelGGL(colsum_kernel, dim3(kO), dim3(256), 0, s, n, do_.get(), kO, Gp + kOffBo);
  gemm(n, kH, kO, do_.get(), kO, 1, Pp + kOffWo, 1, kO, nullptr, dh_.get(), kH, nullptr, s);
  struct Layer {
    const float *z, *pre;
    int kin, offW, offB;
  };
  const Layer layers[3] = {{z1_.get(), p1_.get(), 2 * kF, kOffW1, kOffB1},
                           {z2_.get(), p2_.get(), 2 * kH, kOffW2, kOffB2},
                           {z3_.get(), p3_.get(), 2 * kH, kOffW3, kOffB3}};
  for (int l = 2; l >= 0; --l) {
    const Layer& L = layers[l];
    hipLaunchKernelGGL(relu_mask_kernel, dim3(grid_for((int64_t)n * kH, 256)), dim3(256), 0, s,
                       (int64_t)n * kH, L.pre, dh_.get(), dp_.get());
    gemm(L.kin, kH, n, L.z, 1, L.kin, dp_.get(), kH, 1, nullptr, Gp + L.offW, kH, nullptr, s);
    hipLaunchKernelGGL(colsum_kernel, dim3(kH), dim3(256), 0, s, n, dp_.get(), kH, Gp + L.offB);
    if (l > 0) {
      gemm(n, L.kin, kH, dp_.get(), kH, 1, Pp + L.offW, 1, kH, nullptr, dz_.get(), L.kin, nullptr,
           s);
      hipLaunchKernelGGL(aggregate_transpose_kernel, dim3(n), dim3(64), 0, s, n, row_ptr_.get(),
                         col_.get(), mirror_.get(), w_.get(), dz_.get(), kH, dh_.get());
    }
  }
  double ab[2] = {0, 0};
  HIP_OK(hipMemcpyAsync(ab, AB, sizeof(ab), hipMemcpyDeviceToHost, s));
  HIP_OK(hipStreamSynchronize(s));
  HIP_OK(hipGetLastError());
  const double loss = ab[1] > 0 ? 1.0 - ab[0] * ab[0] / (ab[1] * Z) : 1.0;
  return (float)loss;
}

void LearnedPrecond::probes(double* out, int64_t n, uint64_t seed, hipStream_t s) {
  hipLaunchKernelGGL(probe_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, n, seed, out);
  HIP_OK(hipGetLastError());
}

void LearnedPrecond::adam(float lr, int step, hipStream_t s) {
  const float c1 = 1.0f - std::pow(0.9f, (float)step), c2 = 1.0f - std::pow(0.999f, (float)step);
  hipLaunchKernelGGL(adam_kernel, dim3(grid_for(kParams, 256)), dim3(256), 0, s, kParams,
                     params_.get(), grad_.get(), m_.get(), v_.get(), lr, c1, c2);
  HIP_OK(hipGetLastError());
}

}  // namespace symhip
