// Block-Jacobi preconditioned CG on the reduced camera system, with all scalars on the device.
//
// One CG iteration is two kernels. pcg_spmv (one wavefront per block row) first reduces the
// previous update kernel's partial sums (r.z, r.r) redundantly in every wavefront, takes
// beta = rz_k / rz_{k-1}, checks ||r_k|| <= tol ||r_0|| (and stops every later kernel through a
// device flag), then forms p_k = z_k + beta p_{k-1} on the fly for the block columns it touches
// and q_k = S p_k, with per-row partial sums of p.q. pcg_update (one thread per entry) reduces
// those partials, takes alpha = rz_k / p.q, updates x, r, applies the block-Jacobi
// preconditioner z = M r and writes the next partial sums. Scalars live in a small device array:
//   [0], [1] rz by iteration parity, [2] rr_0, [3] state (0 running, 1 converged,
//   2 breakdown), [4] iteration count at the stop, [5] last rr, [6], [7] the iteration index k
//   by slot, [8] tol^2, [9] the iteration limit. The index is read from the device (kernel s of
//   an iteration pair uses slot s and the spmv kernel writes k + 1 into the other slot) and the
//   tolerance and limit are set by the (uncaptured) init kernel, so a chunk of iterations is the
//   same launch sequence for every solve and is captured once in a hipGraph and replayed.
#include <algorithm>

#include "./pcg.h"
#include "./util.h"

namespace symhip {

namespace {

constexpr int kScal = 10;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Gauss-Jordan inverse of the SPD diagonal block (one wavefront per node), on [S_aa | I] in LDS.
// A non-positive pivot (the block is not SPD) falls back to the inverse diagonal (Jacobi).
template <int ND, typename TS>
__global__ void __launch_bounds__(64) precond_kernel(int32_t n, const int32_t* __restrict__ diag,
                                                     const TS* __restrict__ S,
                                                     const double* __restrict__ corr,
                                                     double* __restrict__ Minv) {
  constexpr int W = 2 * ND, NA = ND * W, NE = ND * ND, PER = (NA + 63) / 64;
  __shared__ double A[ND][W + 1];
  const int a = blockIdx.x, l = threadIdx.x;
  const TS* B = S + (int64_t)diag[a] * NE;
  for (int e = l; e < NA; e += 64) {
    const int i = e / W, j = e % W;
    A[i][j] = j < ND ? (double)B[i * ND + j] : (j - ND == i ? 1.0 : 0.0);
  }
  __syncthreads();
  bool bad = false;
  for (int k = 0; k < ND; ++k) {
    double piv = A[k][k];
    if (!(piv > 0)) {
      bad = true;
      piv = 1.0;
    }
    const double ip = 1.0 / piv;
    double nv[PER];
#pragma unroll
    for (int t = 0; t < PER; ++t) {
      const int e = l + 64 * t;
      nv[t] = 0;
      if (e < NA) {
        const int i = e / W, j = e % W;
        nv[t] = i == k ? A[k][j] * ip : A[i][j] - A[i][k] * A[k][j] * ip;
      }
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < PER; ++t) {
      const int e = l + 64 * t;
      if (e < NA) A[e / W][e % W] = nv[t];
    }
    __syncthreads();
  }
  for (int e = l; e < NE; e += 64) {
    const int i = e / ND, j = e % ND;
    double v;
    if (bad) {
      const double d = fabs((double)B[i * ND + i]);
      v = i == j ? 1.0 / (d > 0 ? d : 1.0) : 0.0;
    } else {
      v = A[i][ND + j];
    }
    if (corr) v += corr[(int64_t)a * NE + e];
    Minv[(int64_t)a * NE + e] = v;
  }
}

// Block reduction of two values over a 256-thread block; thread 0 gets the totals.
__device__ __forceinline__ void block_sum2(double& u, double& v) {
  __shared__ double red[2][4];
  u = wave_sum(u);
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = u;
    red[1][threadIdx.x >> 6] = v;
  }
  __syncthreads();
  u = red[0][0] + red[0][1] + red[0][2] + red[0][3];
  v = red[1][0] + red[1][1] + red[1][2] + red[1][3];
}

// x = 0, r = b, p_{-1} = 0, z = M b; partials of r.z and r.r.
template <int ND>
__global__ void __launch_bounds__(256) pcg_init_kernel(int32_t n, const double* __restrict__ b,
                                                       const double* __restrict__ Minv,
                                                       double* __restrict__ x,
                                                       double* __restrict__ r,
                                                       double* __restrict__ z,
                                                       double* __restrict__ p_old,
                                                       double* __restrict__ partB,
                                                       double* __restrict__ scal, double tol2,
                                                       int max_iters) {
  constexpr int RPB = 256 / ND, NE = ND * ND;
  __shared__ double rs[RPB * ND];
  const int t = threadIdx.x, rloc = t / ND, i = t - rloc * ND;
  const int a = blockIdx.x * RPB + rloc;
  const bool act = t < RPB * ND && a < n;
  const int64_t e = (int64_t)a * ND + i;
  const double bi = act ? b[e] : 0.0;
  if (act) {
    x[e] = 0.0;
    r[e] = bi;
    p_old[e] = 0.0;
  }
  if (t < RPB * ND) rs[t] = bi;
  __syncthreads();
  double zi = 0;
  if (act) {
#pragma unroll
    for (int j = 0; j < ND; ++j) zi += Minv[(int64_t)a * NE + i * ND + j] * rs[rloc * ND + j];
    z[e] = zi;
  }
  double rz = bi * zi, rr = bi * bi;
  block_sum2(rz, rr);
  if (t == 0) {
    partB[2 * blockIdx.x] = rz;
    partB[2 * blockIdx.x + 1] = rr;
  }
  if (blockIdx.x == 0 && t == 0) {
    for (int k = 0; k < 8; ++k) scal[k] = 0.0;
    scal[8] = tol2;
    scal[9] = (double)max_iters;
  }
}

// q_a = sum_b S_ab p_b for row a, lanes split as G groups of ND (group g takes every G-th block
// of the row, lane i of a group the block's row i); groups are summed with shuffles. S is read as
// TS (FP32 in the FP32 form) and widened; the sums are FP64.
template <int ND, typename TS>
__device__ __forceinline__ double row_times(const int32_t* __restrict__ row_ptr,
                                            const int32_t* __restrict__ col,
                                            const TS* __restrict__ S, const double* z,
                                            const double* p_old, double beta, int a, int lane) {
  constexpr int G = 64 / ND, NE = ND * ND;
  const int g = lane / ND, i = lane - g * ND;
  double acc = 0;
  if (g < G) {
    const int t1 = row_ptr[a + 1];
    for (int t = row_ptr[a] + g; t < t1; t += G) {
      const int bc = col[t];
      const TS* Sr = S + (int64_t)t * NE + i * ND;
      const double* zb = z + (int64_t)bc * ND;
      const double* pb = p_old + (int64_t)bc * ND;
#pragma unroll
      for (int j = 0; j < ND; ++j) acc += (double)Sr[j] * (zb[j] + beta * pb[j]);
    }
  }
  double tot = acc;
#pragma unroll
  for (int gg = 1; gg < G; ++gg) tot += __shfl(acc, gg * ND + (lane < ND ? lane : 0), 64);
  return tot;  // valid in lanes < ND
}

template <int ND, typename TS>
__global__ void __launch_bounds__(256) pcg_spmv_kernel(
    int slot, int32_t n, const int32_t* __restrict__ row_ptr,
    const int32_t* __restrict__ col, const TS* __restrict__ S, const double* __restrict__ z,
    const double* __restrict__ p_old, double* __restrict__ p_new, double* __restrict__ q,
    const double* __restrict__ partB, int nbB, double* __restrict__ partA,
    double* __restrict__ scal) {
  if (scal[3] != 0.0) return;
  const int k = (int)scal[6 + slot];
  const double tol2 = scal[8];
  const int max_iters = (int)scal[9];
  if (k > max_iters) return;
  const int lane = threadIdx.x & 63;
  const int a = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + (threadIdx.x >> 6));
  double rz = 0, rr = 0;
  for (int t = lane; t < nbB; t += 64) {
    rz += partB[2 * t];
    rr += partB[2 * t + 1];
  }
  rz = wave_sum(rz);
  rr = wave_sum(rr);
  const double rr0 = k == 0 ? rr : scal[2];
  const double beta = k == 0 ? 0.0 : rz / scal[(k - 1) & 1];
  const bool finite = rz == rz && rr == rr;
  const bool conv = finite && !(rr > tol2 * rr0);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    scal[k & 1] = rz;
    if (k == 0) scal[2] = rr;
    scal[5] = rr;
    scal[6 + (slot ^ 1)] = k + 1;
    if (conv || !finite) {
      scal[3] = conv ? 1.0 : 2.0;
      scal[4] = k;
    }
  }
  // k == max_iters only checks the last residual
  if (conv || !finite || a >= n || k == max_iters) return;
  const double qi = row_times<ND, TS>(row_ptr, col, S, z, p_old, beta, a, lane);
  double pq = 0;
  if (lane < ND) {
    const int64_t e = (int64_t)a * ND + lane;
    const double pn = z[e] + beta * p_old[e];
    p_new[e] = pn;
    q[e] = qi;
    pq = pn * qi;
  }
  pq = wave_sum(pq);
  if (lane == 0) partA[a] = pq;
}

template <int ND>
__global__ void __launch_bounds__(256) pcg_update_kernel(
    int slot, int32_t n, const double* __restrict__ Minv,
    const double* __restrict__ p_new, const double* __restrict__ q,
    const double* __restrict__ partA, double* __restrict__ x, double* __restrict__ r,
    double* __restrict__ z, double* __restrict__ partB, double* __restrict__ scal) {
  if (scal[3] != 0.0) return;
  const int k = (int)scal[6 + slot];
  if (k >= (int)scal[9]) return;
  constexpr int RPB = 256 / ND, NE = ND * ND;
  __shared__ double rs[RPB * ND];
  double pq = 0, dummy = 0;
  for (int t = threadIdx.x; t < n; t += 256) pq += partA[t];
  block_sum2(pq, dummy);
  const bool ok = pq > 0;
  const double alpha = ok ? scal[k & 1] / pq : 0.0;
  if (!ok && blockIdx.x == 0 && threadIdx.x == 0) {
    scal[3] = 2.0;
    scal[4] = k;
  }
  const int t = threadIdx.x, rloc = t / ND, i = t - rloc * ND;
  const int a = blockIdx.x * RPB + rloc;
  const bool act = t < RPB * ND && a < n;
  const int64_t e = (int64_t)a * ND + i;
  double ri = 0;
  if (act) {
    x[e] += alpha * p_new[e];
    ri = r[e] - alpha * q[e];
    r[e] = ri;
  }
  __syncthreads();  // block_sum2 reuses its shared totals below
  if (t < RPB * ND) rs[t] = ri;
  __syncthreads();
  double zi = 0;
  if (act) {
#pragma unroll
    for (int j = 0; j < ND; ++j) zi += Minv[(int64_t)a * NE + i * ND + j] * rs[rloc * ND + j];
    z[e] = zi;
  }
  double rz = ri * zi, rr = ri * ri;
  block_sum2(rz, rr);
  if (t == 0) {
    partB[2 * blockIdx.x] = rz;
    partB[2 * blockIdx.x + 1] = rr;
  }
}

template <int ND, typename TS>
__global__ void __launch_bounds__(256) bspmv_kernel(int32_t n, const int32_t* __restrict__ row_ptr,
                                                    const int32_t* __restrict__ col,
                                                    const TS* __restrict__ S,
                                                    const double* __restrict__ x,
                                                    double* __restrict__ y) {
  const int lane = threadIdx.x & 63;
  const int a = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + (threadIdx.x >> 6));
  if (a >= n) return;
  const double v = row_times<ND, TS>(row_ptr, col, S, x, x, 0.0, a, lane);
  if (lane < ND) y[(int64_t)a * ND + lane] = v;
}

template <typename F>
void dispatch_nd(int nd, F&& f) {
  switch (nd) {
    case 6: f.template operator()<6>(); break;
    case 9: f.template operator()<9>(); break;
    default: throw Error("BlockPcg: unsupported block size " + std::to_string(nd));
  }
}

}  // namespace

BlockPcg::~BlockPcg() {
  if (chunk_) (void)hipGraphExecDestroy(chunk_);
  if (host_scal_) (void)hipHostFree(host_scal_);
}

void BlockPcg::setup(int nd, int32_t n_nodes, const std::vector<int32_t>& row_ptr,
                     const std::vector<int32_t>& col, const std::vector<int32_t>& diag,
                     hipStream_t s, bool f32) {
  SYMHIP_CHECK((int)row_ptr.size() == n_nodes + 1 && (int)diag.size() == n_nodes,
               "BlockPcg: pattern size mismatch");
  SYMHIP_CHECK(row_ptr.back() == (int32_t)col.size(), "BlockPcg: row_ptr does not cover col");
  for (int32_t a = 0; a < n_nodes; ++a)
    SYMHIP_CHECK(diag[a] >= row_ptr[a] && diag[a] < row_ptr[a + 1] && col[diag[a]] == a,
                 "BlockPcg: diagonal block index outside its row");
  for (int32_t c : col) SYMHIP_CHECK(c >= 0 && c < n_nodes, "BlockPcg: column out of range");
  nd_ = nd;
  n_ = n_nodes;
  n_blk_ = (int64_t)col.size();
  dispatch_nd(nd, [&]<int ND>() {
    rows_per_block_ = 256 / ND;
  });
  n_upd_blocks_ = std::max(1, grid_for(n_, rows_per_block_));
  row_ptr_.upload(row_ptr, s);
  col_.upload(col, s);
  diag_.upload(diag, s);
  const size_t nv = std::max<size_t>(1, (size_t)n_ * nd);
  f32_ = f32;
  if (f32_) S32_.resize(std::max<size_t>(1, (size_t)n_blk_ * nd * nd));
  else S_.resize(std::max<size_t>(1, (size_t)n_blk_ * nd * nd));
  Minv_.resize(std::max<size_t>(1, (size_t)n_ * nd * nd));
  r_.resize(nv);
  z_.resize(nv);
  p_[0].resize(nv);
  p_[1].resize(nv);
  q_.resize(nv);
  partA_.resize(std::max(1, n_));
  partB_.resize(2 * (size_t)n_upd_blocks_);
  scal_.resize(kScal);
  if (!host_scal_) HIP_OK(hipHostMalloc((void**)&host_scal_, kScal * sizeof(double)));
  HIP_OK(hipStreamSynchronize(s));
}

void BlockPcg::build_preconditioner(const double* correction, hipStream_t s) {
  if (n_ == 0) return;
  dispatch_nd(nd_, [&]<int ND>() {
    if (f32_)
      hipLaunchKernelGGL((precond_kernel<ND, float>), dim3(n_), dim3(64), 0, s, n_, diag_.get(),
                         S32_.get(), correction, Minv_.get());
    else
      hipLaunchKernelGGL((precond_kernel<ND, double>), dim3(n_), dim3(64), 0, s, n_, diag_.get(),
                         S_.get(), correction, Minv_.get());
  });
  HIP_OK(hipGetLastError());
}

void BlockPcg::multiply_with(const double* S, const double* x, double* y, hipStream_t s) {
  if (n_ == 0) return;
  dispatch_nd(nd_, [&]<int ND>() {
    hipLaunchKernelGGL((bspmv_kernel<ND, double>), dim3(grid_for(n_, 4)), dim3(256), 0, s, n_,
                       row_ptr_.get(), col_.get(), S, x, y);
  });
  HIP_OK(hipGetLastError());
}

void BlockPcg::capture_chunk(double* x, hipStream_t s) {
  if (chunk_ && chunk_x_ == x) return;
  if (chunk_) (void)hipGraphExecDestroy(chunk_);
  chunk_ = nullptr;
  hipGraph_t g;
  HIP_OK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  dispatch_nd(nd_, [&]<int ND>() {
    for (int i = 0; i < kChunk; ++i) {
      const int slot = i & 1;
      auto spmv = [&](auto* S) {
        using TS = std::remove_pointer_t<decltype(S)>;
        hipLaunchKernelGGL((pcg_spmv_kernel<ND, TS>), dim3(grid_for(n_, 4)), dim3(256), 0, s,
                           slot, n_, row_ptr_.get(), col_.get(), S, z_.get(),
                           p_[slot ^ 1].get(), p_[slot].get(), q_.get(), partB_.get(),
                           n_upd_blocks_, partA_.get(), scal_.get());
      };
      if (f32_) spmv(S32_.get()); else spmv(S_.get());
      hipLaunchKernelGGL((pcg_update_kernel<ND>), dim3(n_upd_blocks_), dim3(256), 0, s, slot,
                         n_, Minv_.get(), p_[slot].get(), q_.get(), partA_.get(), x, r_.get(),
                         z_.get(), partB_.get(), scal_.get());
    }
  });
  HIP_OK(hipStreamEndCapture(s, &g));
  HIP_OK(hipGraphInstantiate(&chunk_, g, nullptr, nullptr, 0));
  HIP_OK(hipGraphDestroy(g));
  chunk_x_ = x;
}

PcgResult BlockPcg::solve(const double* b, double* x, int max_iters, double rel_tol,
                          hipStream_t s, int check_every) {
  PcgResult res;
  if (n_ == 0) {
    res.converged = true;
    return res;
  }
  SYMHIP_CHECK(max_iters >= 1, "BlockPcg: bad iteration limit");
  (void)check_every;
  const double tol2 = rel_tol * rel_tol;
  capture_chunk(x, s);
  dispatch_nd(nd_, [&]<int ND>() {
    hipLaunchKernelGGL((pcg_init_kernel<ND>), dim3(n_upd_blocks_), dim3(256), 0, s, n_, b,
                       Minv_.get(), x, r_.get(), z_.get(), p_[1].get(), partB_.get(), scal_.get(),
                       tol2, max_iters);
  });
  HIP_OK(hipGetLastError());
  // chunks of kChunk iterations cover k = 0 .. max_iters (the last spmv only checks)
  const int chunks = max_iters / kChunk + 1;
  for (int c = 0; c < chunks; ++c) {
    HIP_OK(hipGraphLaunch(chunk_, s));
    if (c + 1 < chunks) {
      HIP_OK(hipMemcpyAsync(host_scal_, scal_.get(), kScal * sizeof(double),
                            hipMemcpyDeviceToHost, s));
      HIP_OK(hipStreamSynchronize(s));
      if (host_scal_[3] != 0.0) break;
    }
  }
  HIP_OK(hipMemcpyAsync(host_scal_, scal_.get(), kScal * sizeof(double), hipMemcpyDeviceToHost,
                        s));
  HIP_OK(hipStreamSynchronize(s));
  res.converged = host_scal_[3] == 1.0;
  res.breakdown = host_scal_[3] == 2.0;
  res.iterations = host_scal_[3] != 0.0 ? (int)host_scal_[4] : max_iters;
  res.rel_residual = host_scal_[2] > 0 ? std::sqrt(host_scal_[5] / host_scal_[2]) : 0.0;
  return res;
}

}  // namespace symhip
