// Device numeric supernodal LDL^T: assembly, per-level diagonal-block factorization, panel
// triangular solve, right-looking update (SYRK-like) scattered into ancestor panels, and the
// triangular solves. All FP64.
#include <algorithm>

#include "./chol_numeric.h"
#include "./util.h"

namespace symhip {

namespace {

constexpr int kMaxCols = 64;  // supernode width cap: diag block fits in LDS (32 KB)
constexpr int kTile = 64;     // update tile

struct SnView {
  const int64_t* panel_off;
  const int32_t* col_begin;
  const int32_t* rows_ptr;
  const int32_t* rows;
  __device__ int ncols(int s) const { return col_begin[s + 1] - col_begin[s]; }
  __device__ int ld(int s) const { return rows_ptr[s + 1] - rows_ptr[s]; }
};

__global__ void assemble_kernel(int64_t nnz, const int64_t* __restrict__ map,
                                const double* __restrict__ vals, double* __restrict__ panel) {
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < nnz;
       q += (int64_t)gridDim.x * blockDim.x)
    panel[map[q]] = vals[q];
}

// Dense LDL^T of each listed supernode's diagonal block, in LDS.
__global__ void __launch_bounds__(256) diag_factor_kernel(const int32_t* __restrict__ list,
                                                          SnView v, double* __restrict__ panel,
                                                          double* __restrict__ D) {
  __shared__ double A[kMaxCols * kMaxCols];
  const int s = list[blockIdx.x];
  const int n = v.ncols(s), ld = v.ld(s);
  double* P = panel + v.panel_off[s];
  for (int idx = threadIdx.x; idx < n * n; idx += blockDim.x) {
    const int r = idx % n, c = idx / n;
    A[c * kMaxCols + r] = r >= c ? P[(int64_t)c * ld + r] : 0.0;
  }
  __syncthreads();
  for (int k = 0; k < n; ++k) {
    const double d = A[k * kMaxCols + k];
    const double inv = 1.0 / d;
    const int m = n - k - 1;
    for (int idx = threadIdx.x; idx < m * m; idx += blockDim.x) {
      const int i = k + 1 + idx % m, j = k + 1 + idx / m;
      if (i >= j) A[j * kMaxCols + i] -= A[k * kMaxCols + i] * A[k * kMaxCols + j] * inv;
    }
    __syncthreads();
    for (int i = k + 1 + threadIdx.x; i < n; i += blockDim.x) A[k * kMaxCols + i] *= inv;
  }
  __syncthreads();
  for (int idx = threadIdx.x; idx < n * n; idx += blockDim.x) {
    const int r = idx % n, c = idx / n;
    if (r > c) P[(int64_t)c * ld + r] = A[c * kMaxCols + r];
    if (r == c) {
      D[v.col_begin[s] + c] = A[c * kMaxCols + c];
      P[(int64_t)c * ld + r] = A[c * kMaxCols + c];
    }
  }
}

// Below-diagonal rows: L21 = A21 L11^-T D^-1, one thread per row, items = (supernode, chunk).
__global__ void __launch_bounds__(256) trsm_kernel(const int32_t* __restrict__ items, SnView v,
                                                   double* __restrict__ panel,
                                                   const double* __restrict__ D) {
  __shared__ double G[kMaxCols * kMaxCols];  // G(k, j) = L(k, j) * d_j, j < k
  __shared__ double Dinv[kMaxCols];
  const int s = items[2 * blockIdx.x], chunk = items[2 * blockIdx.x + 1];
  const int n = v.ncols(s), ld = v.ld(s), mb = ld - n;
  double* P = panel + v.panel_off[s];
  const double* Ds = D + v.col_begin[s];
  for (int idx = threadIdx.x; idx < n * n; idx += blockDim.x) {
    const int k = idx % n, j = idx / n;
    G[j * kMaxCols + k] = k > j ? P[(int64_t)j * ld + k] * Ds[j] : 0.0;
  }
  for (int k = threadIdx.x; k < n; k += blockDim.x) Dinv[k] = 1.0 / Ds[k];
  __syncthreads();
  const int i = chunk * blockDim.x + threadIdx.x;
  if (i >= mb) return;
  double* row = P + n + i;  // element (k) at row[k * ld]
  for (int k = 0; k < n; ++k) {
    double acc = row[(int64_t)k * ld];
    for (int j = 0; j < k; ++j) acc -= row[(int64_t)j * ld] * G[j * kMaxCols + k];
    row[(int64_t)k * ld] = acc * Dinv[k];
  }
}

// U = L21 D L21^T tile (ti, tj), scattered as -U into ancestor panels.
__global__ void __launch_bounds__(256) update_kernel(
    const int32_t* __restrict__ items, SnView v, double* __restrict__ panel,
    const double* __restrict__ D, const int32_t* __restrict__ upd_ptr,
    const int32_t* __restrict__ upd_target, const int32_t* __restrict__ upd_seg_begin,
    const int32_t* __restrict__ upd_rel_ptr, const int32_t* __restrict__ upd_rel) {
  __shared__ double As[16][kTile + 1];
  __shared__ double Bs[16][kTile + 1];
  const int s = items[3 * blockIdx.x], ti = items[3 * blockIdx.x + 1],
            tj = items[3 * blockIdx.x + 2];
  const int n = v.ncols(s), ld = v.ld(s), mb = ld - n;
  const double* P = panel + v.panel_off[s];
  const double* Ds = D + v.col_begin[s];
  const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
  double acc[4][4] = {};
  for (int k0 = 0; k0 < n; k0 += 16) {
    for (int idx = threadIdx.x; idx < 16 * kTile; idx += blockDim.x) {
      const int kk = idx / kTile, r = idx % kTile;
      const int k = k0 + kk;
      const int ri = ti * kTile + r, rj = tj * kTile + r;
      As[kk][r] = (k < n && ri < mb) ? P[(int64_t)k * ld + n + ri] : 0.0;
      Bs[kk][r] = (k < n && rj < mb) ? P[(int64_t)k * ld + n + rj] * Ds[k] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < 16; ++kk) {
      double a[4], b[4];
#pragma unroll
      for (int x = 0; x < 4; ++x) {
        a[x] = As[kk][ty + 16 * x];
        b[x] = Bs[kk][tx + 16 * x];
      }
#pragma unroll
      for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < 4; ++y) acc[x][y] += a[x] * b[y];
    }
    __syncthreads();
  }
  const int seg0 = upd_ptr[s], seg1 = upd_ptr[s + 1];
  const int32_t* rows_s = v.rows + v.rows_ptr[s] + n;
#pragma unroll
  for (int y = 0; y < 4; ++y) {
    const int j = tj * kTile + tx + 16 * y;
    if (j >= mb) continue;
    // segment containing below-row j
    int lo = seg0, hi = seg1 - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (upd_seg_begin[mid] <= j) lo = mid; else hi = mid - 1;
    }
    const int t = upd_target[lo];
    const int jb = upd_seg_begin[lo];
    const int col = rows_s[j] - v.col_begin[t];
    const int ldt = v.ld(t);
    double* Pt = panel + v.panel_off[t] + (int64_t)col * ldt;
    const int32_t* rel = upd_rel + upd_rel_ptr[lo];
#pragma unroll
    for (int x = 0; x < 4; ++x) {
      const int i = ti * kTile + ty + 16 * x;
      if (i >= mb || i < j) continue;
      unsafeAtomicAdd(Pt + rel[i - jb], -acc[x][y]);
    }
  }
}

__global__ void permute_kernel(int32_t n, const int32_t* __restrict__ perm,
                               const double* __restrict__ b, double* __restrict__ y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[perm[i]] = b[i];
}
__global__ void unpermute_kernel(int32_t n, const int32_t* __restrict__ perm,
                                 const double* __restrict__ y, double* __restrict__ x) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = y[perm[i]];
}
__global__ void dscale_kernel(int32_t n, const double* __restrict__ D, double* __restrict__ y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] /= D[i];
}

// Forward: y_own = L11^-1 y_own; y_below -= L21 y_own (atomics: siblings share ancestors).
__global__ void __launch_bounds__(256) fwd_kernel(const int32_t* __restrict__ list, SnView v,
                                                  const double* __restrict__ panel,
                                                  double* __restrict__ y) {
  __shared__ double ys[kMaxCols];
  const int s = list[blockIdx.x];
  const int n = v.ncols(s), ld = v.ld(s), mb = ld - n;
  const double* P = panel + v.panel_off[s];
  double* yo = y + v.col_begin[s];
  for (int k = threadIdx.x; k < n; k += blockDim.x) ys[k] = yo[k];
  __syncthreads();
  for (int k = 0; k < n; ++k) {
    for (int i = k + 1 + threadIdx.x; i < n; i += blockDim.x) ys[i] -= P[(int64_t)k * ld + i] * ys[k];
    __syncthreads();
  }
  for (int k = threadIdx.x; k < n; k += blockDim.x) yo[k] = ys[k];
  const int32_t* rows = v.rows + v.rows_ptr[s] + n;
  for (int i = threadIdx.x; i < mb; i += blockDim.x) {
    double acc = 0;
    for (int k = 0; k < n; ++k) acc += P[(int64_t)k * ld + n + i] * ys[k];
    unsafeAtomicAdd(&y[rows[i]], -acc);
  }
}

// Backward: y_own -= L21^T y_below; y_own = L11^-T y_own.
__global__ void __launch_bounds__(256) bwd_kernel(const int32_t* __restrict__ list, SnView v,
                                                  const double* __restrict__ panel,
                                                  double* __restrict__ y) {
  __shared__ double ys[kMaxCols];
  const int s = list[blockIdx.x];
  const int n = v.ncols(s), ld = v.ld(s), mb = ld - n;
  const double* P = panel + v.panel_off[s];
  double* yo = y + v.col_begin[s];
  const int32_t* rows = v.rows + v.rows_ptr[s] + n;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  for (int k = wave; k < n; k += nw) {
    double acc = 0;
    for (int i = lane; i < mb; i += 64) acc += P[(int64_t)k * ld + n + i] * y[rows[i]];
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o, 64);
    if (lane == 0) ys[k] = yo[k] - acc;
  }
  __syncthreads();
  for (int k = n - 1; k >= 0; --k) {
    for (int i = threadIdx.x; i < k; i += blockDim.x) ys[i] -= P[(int64_t)i * ld + k] * ys[k];
    __syncthreads();
  }
  for (int k = threadIdx.x; k < n; k += blockDim.x) yo[k] = ys[k];
}

}  // namespace

void SupernodalLDLT::setup(const CholSymbolic& sym, const int32_t* colptr, const int32_t* rowind,
                           hipStream_t s) {
  SYMHIP_CHECK(sym.max_snode_cols <= kMaxCols, "supernode wider than the kernel tile");
  sym_ = sym;
  std::vector<int64_t> map = chol_assembly_map(sym, colptr, rowind);
  asm_map_.upload(map, s);
  panel_.resize(std::max<int64_t>(1, sym.panel_size()));
  D_.resize(std::max(1, sym.n));
  y_.resize(std::max(1, sym.n));
  panel_off_.upload(sym.sn_panel_off, s);
  col_begin_.upload(sym.sn_col_begin, s);
  rows_ptr_.upload(sym.sn_rows_ptr, s);
  rows_.upload(sym.sn_rows, s);
  perm_.upload(sym.perm, s);
  upd_ptr_.upload(sym.upd_ptr, s);
  upd_target_.upload(sym.upd_target, s);
  upd_seg_begin_.upload(sym.upd_seg_begin, s);
  upd_rel_ptr_.upload(sym.upd_rel_ptr, s);
  upd_rel_.upload(sym.upd_rel, s);
  upd_col_off_.upload(sym.upd_col_off, s);
  // Work lists per level
  std::vector<int32_t> snl, trsm, upd;
  lvl_snode_ptr_.assign(1, 0);
  lvl_trsm_ptr_.assign(1, 0);
  lvl_upd_ptr_.assign(1, 0);
  for (int l = 0; l < sym.n_levels; ++l) {
    for (int q = sym.level_ptr[l]; q < sym.level_ptr[l + 1]; ++q) {
      const int sn = sym.level_snodes[q];
      snl.push_back(sn);
      const int n = sym.sn_col_begin[sn + 1] - sym.sn_col_begin[sn];
      const int mb = sym.sn_rows_ptr[sn + 1] - sym.sn_rows_ptr[sn] - n;
      for (int c = 0; c * 256 < mb; ++c) {
        trsm.push_back(sn);
        trsm.push_back(c);
      }
      const int T = (mb + kTile - 1) / kTile;
      for (int ti = 0; ti < T; ++ti)
        for (int tj = 0; tj <= ti; ++tj) {
          upd.push_back(sn);
          upd.push_back(ti);
          upd.push_back(tj);
        }
    }
    lvl_snode_ptr_.push_back((int32_t)snl.size());
    lvl_trsm_ptr_.push_back((int32_t)trsm.size() / 2);
    lvl_upd_ptr_.push_back((int32_t)upd.size() / 3);
  }
  lvl_snodes_.upload(snl, s);
  trsm_items_.upload(trsm, s);
  upd_items_.upload(upd, s);
  HIP_OK(hipStreamSynchronize(s));
}

void SupernodalLDLT::factorize(const double* a_values, hipStream_t s) {
  fill_zero(panel_.get(), sym_.panel_size(), s);
  const int64_t nnz = (int64_t)asm_map_.size();
  if (nnz) {
    const int grid = (int)std::min<int64_t>(grid_for(nnz, 256), 8192);
    hipLaunchKernelGGL(assemble_kernel, dim3(grid), dim3(256), 0, s, nnz, asm_map_.get(), a_values,
                       panel_.get());
  }
  factorize_assembled(s);
}

void SupernodalLDLT::factorize_assembled(hipStream_t s) {
  const SnView v{panel_off_.get(), col_begin_.get(), rows_ptr_.get(), rows_.get()};
  for (int l = 0; l < sym_.n_levels; ++l) {
    const int ns = lvl_snode_ptr_[l + 1] - lvl_snode_ptr_[l];
    hipLaunchKernelGGL(diag_factor_kernel, dim3(ns), dim3(256), 0, s,
                       lvl_snodes_.get() + lvl_snode_ptr_[l], v, panel_.get(), D_.get());
    const int nt = lvl_trsm_ptr_[l + 1] - lvl_trsm_ptr_[l];
    if (nt)
      hipLaunchKernelGGL(trsm_kernel, dim3(nt), dim3(256), 0, s,
                         trsm_items_.get() + 2 * lvl_trsm_ptr_[l], v, panel_.get(), D_.get());
    const int nu = lvl_upd_ptr_[l + 1] - lvl_upd_ptr_[l];
    if (nu)
      hipLaunchKernelGGL(update_kernel, dim3(nu), dim3(256), 0, s,
                         upd_items_.get() + 3 * lvl_upd_ptr_[l], v, panel_.get(), D_.get(),
                         upd_ptr_.get(), upd_target_.get(), upd_seg_begin_.get(),
                         upd_rel_ptr_.get(), upd_rel_.get());
  }
  HIP_OK(hipGetLastError());
}

void SupernodalLDLT::solve(const double* b, double* x, hipStream_t s) {
  const int n = sym_.n;
  if (n == 0) return;
  const SnView v{panel_off_.get(), col_begin_.get(), rows_ptr_.get(), rows_.get()};
  hipLaunchKernelGGL(permute_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, n, perm_.get(), b,
                     y_.get());
  for (int l = 0; l < sym_.n_levels; ++l) {
    const int ns = lvl_snode_ptr_[l + 1] - lvl_snode_ptr_[l];
    hipLaunchKernelGGL(fwd_kernel, dim3(ns), dim3(256), 0, s,
                       lvl_snodes_.get() + lvl_snode_ptr_[l], v, panel_.get(), y_.get());
  }
  hipLaunchKernelGGL(dscale_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, n, D_.get(), y_.get());
  for (int l = sym_.n_levels - 1; l >= 0; --l) {
    const int ns = lvl_snode_ptr_[l + 1] - lvl_snode_ptr_[l];
    hipLaunchKernelGGL(bwd_kernel, dim3(ns), dim3(256), 0, s,
                       lvl_snodes_.get() + lvl_snode_ptr_[l], v, panel_.get(), y_.get());
  }
  hipLaunchKernelGGL(unpermute_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, n, perm_.get(),
                     y_.get(), x);
  HIP_OK(hipGetLastError());
}

void SupernodalLDLT::download_factor(std::vector<double>& panels, std::vector<double>& d,
                                     hipStream_t s) const {
  panels.resize(sym_.panel_size());
  d.resize(sym_.n);
  panel_.download(panels.data(), panels.size(), s);
  D_.download(d.data(), d.size(), s);
  HIP_OK(hipStreamSynchronize(s));
}

}  // namespace symhip
