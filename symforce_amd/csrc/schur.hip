// Schur-complement backend: landmark-block elimination + reduced-camera-system (RCS) assembly.
//
// Replaces SparseSchurSolver (opt/sparse_schur_solver.tcc:16-163) behind the LM loop for problems
// whose trailing optimized keys are 3-dof landmarks, each touched by factors that otherwise touch
// one "camera node" (the B keys that always appear together, e.g. Pose3 + intrinsics). With
// V_p the 3x3 landmark block, W_o = J_node^T J_point per observation and L_p L_p^T = V_p:
//   Z_o = W_o L_p^-T,  h_p = L_p^-1 g_p,
//   S(a,b) = B(a,b) - sum_{p in a,b} Z_{p,a} Z_{p,b}^T,  v_a = g_a - sum_{o in a} Z_o h_p,
//   S y = v,  z_p = L_p^-T (h_p - sum_o Z_o^T y_node(o)).
// Layout in HBM: observations are stored point-major (points sorted by first camera), with a
// camera-major index for the camera reductions; S is never materialized separately, each S block
// is reduced by one wavefront over its precomputed list of observation pairs and written straight
// into the supernodal LDL^T panels.
#include <algorithm>
#include <numeric>

#include "./chol_numeric.h"
#include "./gnn.h"
#include "./pcg.h"
#include "./factor_traits.h"
#include "./optimizer.h"
#include "./util.h"

namespace symhip {

namespace {

struct ObsView {
  int32_t n_obs;
  const int32_t* arg_off;   // [NA][n_obs]
  const int32_t* res_off;   // [n_obs]
  const int32_t* node;      // [n_obs]
};

// Per-observation stride of Z (ND x 3 row-major in TW) rounded up to whole 16-byte vectors, so
// every observation's Z starts 16-byte aligned (the pad entries stay zero).
template <int ND, typename TW>
constexpr int z_stride() {
  constexpr int v = 16 / (int)sizeof(TW);
  return (ND * 3 + v - 1) / v * v;
}

// Per observation (thread): evaluate the factor, write its residual and W_o = J_node^T J_point
// (ND x 3 row-major, stored as TW), and add its J_point^T J_point (lower 6: 00 10 11 20 21 22) and
// J_point^T r to the point's V_p and g_p. Observations are point-major, so a point's terms sit in
// consecutive lanes: a segmented shuffle scan sums them inside the wavefront and the last lane of
// each segment adds the sum with one atomic per value (V and g start at zero). A wavefront's W
// rows go through LDS and out as 16-byte pieces from consecutive lanes (a lane storing its own
// 216-byte row writes 64 partial cache lines per instruction: 119 -> 101 us per launch at 1000
// cameras, in blocks of two wavefronts so that the staging leaves room for five blocks per CU).
// TW = float is the FP32 form: the Hessian-side products (W, V) come from the FP32 factor body;
// the residual and the gradient J_point^T r stay FP64.
template <int FTC, int ND, typename TW>
__global__ void __launch_bounds__(128) lin_points_kernel(int32_t n_obs,
                                                         const int32_t* __restrict__ obs_point,
                                                         ObsView ov, const double* __restrict__ values,
                                                         double* __restrict__ residual,
                                                         TW* __restrict__ W,
                                                         double* __restrict__ V,
                                                         double* __restrict__ gp) {
  using F = FT<FTC>;
  static_assert(F::T == ND + 3, "landmark factor layout");
  constexpr int kW = ND * 3;
  __shared__ TW wst[2][64 * kW];  // each wavefront's W rows, copied out coalesced at the end
  const int o = blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  TW* ws = wst[threadIdx.x >> 6];
  const bool live = o < n_obs;
  double acc[9] = {};  // v[6], g[3]
  int p = -1 - lane;   // distinct per dead lane so dead lanes never join a segment
  if (live) {
    p = obs_point[o];
    const double* a[F::NA];
#pragma unroll
    for (int k = 0; k < F::NA; ++k) a[k] = values + ov.arg_off[k * ov.n_obs + o];
    double r[F::D], J[F::D][F::T];
    F::template eval<double>(a, r, J);
    const int ro = ov.res_off[o];
#pragma unroll
    for (int i = 0; i < F::D; ++i) residual[ro + i] = r[i];
    TW* Wo = ws + lane * kW;
    // Hessian-side products in the storage precision
    auto hess = [&](const auto (&Jh)[F::D][F::T]) {
      using S = std::remove_cv_t<std::remove_reference_t<decltype(Jh[0][0])>>;
#pragma unroll
      for (int i = 0; i < ND; ++i)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          S h = 0;
#pragma unroll
          for (int q = 0; q < F::D; ++q) h += Jh[q][i] * Jh[q][ND + k];
          Wo[i * 3 + k] = (TW)h;
        }
      int idx = 0;
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j <= i; ++j) {
          S h = 0;
#pragma unroll
          for (int q = 0; q < F::D; ++q) h += Jh[q][ND + i] * Jh[q][ND + j];
          acc[idx++] = (double)h;
        }
    };
    if constexpr (std::is_same_v<TW, float>) {
      float rf[F::D], Jf[F::D][F::T];
      F::template eval<float>(a, rf, Jf);
      hess(Jf);
    } else {
      hess(J);
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      double gi = 0;
#pragma unroll
      for (int q = 0; q < F::D; ++q) gi += J[q][ND + i] * r[q];
      acc[6 + i] = gi;
    }
  }
  // the wavefront's W rows are one contiguous range: 16-byte pieces from consecutive lanes
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  {
    const int o0 = o - lane;
    const int cnt = min(64, n_obs - o0);
    TW* dst = W + (int64_t)o0 * kW;
    constexpr int kVec = 16 / (int)sizeof(TW);
    typedef TW v16 __attribute__((ext_vector_type(kVec)));
    static_assert((64 * kW) % kVec == 0, "a full wavefront's rows are whole pieces");
    if (cnt == 64) {
#pragma unroll 2
      for (int m = lane; m < 64 * kW / kVec; m += 64)
        *reinterpret_cast<v16*>(dst + m * kVec) = *reinterpret_cast<const v16*>(ws + m * kVec);
    } else {
      for (int e = lane; e < cnt * kW; e += 64) dst[e] = ws[e];
    }
  }
  // inclusive segmented scan over lanes with equal p (segments are contiguous)
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int pu = __shfl_up(p, off, 64);
    const bool take = lane >= off && pu == p;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      const double u = __shfl_up(acc[k], off, 64);
      if (take) acc[k] += u;
    }
  }
  const int pn = __shfl_down(p, 1, 64);
  if (live && (lane == 63 || pn != p)) {
#pragma unroll
    for (int k = 0; k < 6; ++k) unsafeAtomicAdd(&V[(int64_t)p * 6 + k], acc[k]);
#pragma unroll
    for (int k = 0; k < 3; ++k) unsafeAtomicAdd(&gp[(int64_t)p * 3 + k], acc[6 + k]);
  }
}

// Sums NV per-lane values over the 64 lanes of a wavefront through LDS, 16 values a round (scr:
// this wavefront's own 16 x 65 doubles): the partials go in as scr[c][lane], lane l adds 16 of
// value l / 4's (quarter l % 4) and two quad shuffles finish; out(i, total) runs in the lane
// with l % 4 == 0 of value i. Six rounds of 64-lane shuffles per value cost several times more.
template <int NV, typename F>
__device__ __forceinline__ void wave_sum_lds(const double (&v)[NV], double (*scr)[65], F&& out) {
  const int lane = threadIdx.x & 63, c = lane >> 2, qtr = lane & 3;
#pragma unroll
  for (int g = 0; g < NV; g += 16) {
#pragma unroll
    for (int k = 0; k < 16; ++k)
      if (g + k < NV) scr[k][lane] = v[g + k];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    double t = 0;
#pragma unroll
    for (int j = 0; j < 16; ++j) t += scr[c][16 * qtr + j];
    t += __shfl_xor(t, 1, 64);
    t += __shfl_xor(t, 2, 64);
    if (qtr == 0 && g + c < NV) out(g + c, t);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();  // the next round overwrites scr
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
}

// Sums NV per-lane values over the 256 threads of a block (within each wavefront through LDS,
// then the four wavefront partials); the totals are valid in threads 0 .. NV - 1 as the return
// value (zero elsewhere).
template <int NV>
__device__ __forceinline__ double block_sum_256(double (&v)[NV], double (*red)[NV],
                                                double (*scr)[16][65]) {
  const int wave = threadIdx.x >> 6;
  wave_sum_lds<NV>(v, scr[wave], [&](int i, double t) { red[wave][i] = t; });
  __syncthreads();
  const int t = threadIdx.x;
  return t < NV ? red[0][t] + red[1][t] + red[2][t] + red[3][t] : 0.0;
}

// Per camera node (a block): recompute its observations' node Jacobians and reduce the node
// diagonal block (full ND x ND) and node rhs. Recomputing is cheaper than storing J.
template <int FTC, int ND>
__global__ void __launch_bounds__(256) lin_cams_kernel(int32_t n_nodes,
                                                       const int32_t* __restrict__ cam_obs_ptr,
                                                       const int32_t* __restrict__ cam_obs,
                                                       ObsView ov, const double* __restrict__ values,
                                                       double* __restrict__ Bdiag,
                                                       double* __restrict__ gB) {
  using F = FT<FTC>;
  constexpr int NH = ND * (ND + 1) / 2;
  // a block per camera (its observations spread over four wavefronts); neighbouring cameras
  // share landmarks: keep consecutive cameras on one XCD
  __shared__ double red[4][NH + ND];
  __shared__ double scr[4][16][65];
  const int cam = xcd_tile(blockIdx.x, n_nodes);
  if (cam >= n_nodes) return;  // block-uniform
  double hg[NH + ND] = {};
  double* h = hg;
  double* g = hg + NH;
  for (int q0 = cam_obs_ptr[cam] + threadIdx.x; q0 < cam_obs_ptr[cam + 1]; q0 += 256) {
    const int o = cam_obs[q0];
    const double* a[F::NA];
#pragma unroll
    for (int k = 0; k < F::NA; ++k) a[k] = values + ov.arg_off[k * ov.n_obs + o];
    double r[F::D], J[F::D][F::T];
    F::template eval<double>(a, r, J);
    int idx = 0;
#pragma unroll
    for (int i = 0; i < ND; ++i) {
#pragma unroll
      for (int q = 0; q < F::D; ++q) g[i] += J[q][i] * r[q];
#pragma unroll
      for (int j = 0; j <= i; ++j) {
        double acc = 0;
#pragma unroll
        for (int q = 0; q < F::D; ++q) acc += J[q][i] * J[q][j];
        h[idx++] += acc;
      }
    }
  }
  const double tot = block_sum_256<NH + ND>(hg, red, scr);
  // thread t < NH holds packed lower entry t of B, thread NH + i entry i of g
  const int t = threadIdx.x;
  if (t < NH) {
    int i = 0;
    while ((i + 1) * (i + 2) / 2 <= t) ++i;
    const int j = t - i * (i + 1) / 2;
    double* Bd = Bdiag + (int64_t)cam * ND * ND;
    Bd[i * ND + j] = tot;
    Bd[j * ND + i] = tot;
  } else if (t < NH + ND) {
    gB[(int64_t)cam * ND + t - NH] = tot;
  }
}

// Factors touching only camera nodes (priors, odometry between cameras): thread per factor,
// H and rhs scattered with atomics into dense per-block storage of the S pattern.
template <int FTC, int ND>
__global__ void lin_bfac_kernel(int32_t n, const int32_t* __restrict__ arg_off,
                                const int32_t* __restrict__ res_off,
                                const int32_t* __restrict__ col_map,  // [T][n]: node*ND+dim or -1
                                const int32_t* __restrict__ blk_of,   // [NL*NL][n] block index
                                const double* __restrict__ values, double* __restrict__ residual,
                                double* __restrict__ Bfac, double* __restrict__ gB) {
  using F = FT<FTC>;
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= n) return;
  const double* a[F::NA];
#pragma unroll
  for (int k = 0; k < F::NA; ++k) a[k] = values + arg_off[k * n + f];
  double r[F::D], J[F::D][F::T];
  F::template eval<double>(a, r, J);
  const int ro = res_off[f];
#pragma unroll
  for (int i = 0; i < F::D; ++i) residual[ro + i] = r[i];
  for (int c = 0; c < F::T; ++c) {
    const int gc = col_map[c * n + f];
    if (gc < 0) continue;
    double acc = 0;
#pragma unroll
    for (int q = 0; q < F::D; ++q) acc += J[q][c] * r[q];
    unsafeAtomicAdd(&gB[gc], acc);
    for (int c2 = 0; c2 < F::T; ++c2) {
      const int gc2 = col_map[c2 * n + f];
      if (gc2 < 0) continue;
      const int na = gc / ND, nb = gc2 / ND;
      if (na < nb || (na == nb && gc < gc2)) continue;  // lower triangle (node-major)
      double hh = 0;
#pragma unroll
      for (int q = 0; q < F::D; ++q) hh += J[q][c] * J[q][c2];
      // block (na, nb) entry (row dim gc%ND, col dim gc2%ND)
      const int blk = blk_of[(c * F::T + c2) * n + f];
      unsafeAtomicAdd(&Bfac[(int64_t)blk * ND * ND + (gc % ND) * ND + (gc2 % ND)], hh);
    }
  }
}

// Damping of camera and point diagonals (DampHessian semantics, tcc:22-54), writing the damping
// vector and the undamped diagonal in reference state order.
template <int ND>
__global__ void damp_schur_kernel(int32_t n_nodes, int32_t n_points,
                                  const double* __restrict__ Bdiag, const double* __restrict__ Bfac,
                                  const int32_t* __restrict__ diag_blk,
                                  const double* __restrict__ V, const int32_t* __restrict__ node_tan,
                                  const int32_t* __restrict__ pt_tan,
                                  const double* __restrict__ lambda_p, int use_diag, int keep_max,
                                  int have_max, double dmin, int unit, double* __restrict__ max_diag,
                                  double* __restrict__ damping, double* __restrict__ undamped,
                                  double* __restrict__ dnode, double* __restrict__ dpoint) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int nb = n_nodes * ND;
  if (i >= nb + 3 * n_points) return;
  const double lambda = *lambda_p;
  double u;
  int tan;
  if (i < nb) {
    const int a = i / ND, k = i % ND;
    u = Bdiag[(int64_t)a * ND * ND + k * ND + k] + Bfac[(int64_t)diag_blk[a] * ND * ND + k * ND + k];
    tan = node_tan[i];
  } else {
    const int p = (i - nb) / 3, k = (i - nb) % 3;
    const int di = k == 0 ? 0 : (k == 1 ? 2 : 5);
    u = V[(int64_t)p * 6 + di];
    tan = pt_tan[p * 3 + k];
  }
  double d = 0;
  if (use_diag) {
    if (keep_max) {
      const double md = have_max ? fmax(max_diag[tan], u) : fmax(u, dmin);
      max_diag[tan] = md;
      d = md * lambda;
    } else {
      d = fmax(u, dmin) * lambda;
    }
  }
  if (unit) d += lambda;
  damping[tan] = d;
  undamped[tan] = u;
  if (i < nb) dnode[i] = d; else dpoint[i - nb] = d;
}

// Per point: damped V_p -> L_p (Cholesky), h_p = L_p^-1 g_p.
__global__ void __launch_bounds__(256) point_factor_kernel(int32_t n_points,
                                                           const double* __restrict__ V,
                                                           const double* __restrict__ dpoint,
                                                           const double* __restrict__ gp,
                                                           double* __restrict__ Lp,
                                                           double* __restrict__ hp) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n_points) return;
  const double* v = V + (int64_t)p * 6;
  const double a00 = v[0] + dpoint[p * 3 + 0], a10 = v[1], a11 = v[2] + dpoint[p * 3 + 1];
  const double a20 = v[3], a21 = v[4], a22 = v[5] + dpoint[p * 3 + 2];
  const double l00 = sqrt(a00);
  const double l10 = a10 / l00, l20 = a20 / l00;
  const double l11 = sqrt(a11 - l10 * l10);
  const double l21 = (a21 - l20 * l10) / l11;
  const double l22 = sqrt(a22 - l20 * l20 - l21 * l21);
  double* L = Lp + (int64_t)p * 6;
  L[0] = l00;
  L[1] = l10;
  L[2] = l11;
  L[3] = l20;
  L[4] = l21;
  L[5] = l22;
  const double* g = gp + (int64_t)p * 3;
  const double h0 = g[0] / l00;
  const double h1 = (g[1] - l10 * h0) / l11;
  const double h2 = (g[2] - l20 * h0 - l21 * h1) / l22;
  hp[p * 3 + 0] = h0;
  hp[p * 3 + 1] = h1;
  hp[p * 3 + 2] = h2;
}

// Per (observation, camera row): Z_o(i, :) = W_o(i, :) L_p^-T (forward substitution with L_p).
// Consecutive threads walk consecutive rows of W and Z, so the 216 B per observation move as
// contiguous runs.
template <int ND, typename TW>
__global__ void __launch_bounds__(256) point_elim_kernel(int32_t n_obs,
                                                         const int32_t* __restrict__ obs_point,
                                                         const double* __restrict__ Lp,
                                                         const TW* __restrict__ W,
                                                         TW* __restrict__ Z) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= (int64_t)n_obs * ND) return;
  const int64_t o = t / ND;
  const double* L = Lp + (int64_t)obs_point[o] * 6;
  const TW* w = W + t * 3;
  constexpr int ZS = z_stride<ND, TW>();
  const double z0 = (double)w[0] / L[0];
  const double z1 = ((double)w[1] - L[1] * z0) / L[2];
  const double z2 = ((double)w[2] - L[3] * z0 - L[4] * z1) / L[5];
  TW* z = Z + o * ZS + (t - o * ND) * 3;
  z[0] = (TW)z0;
  z[1] = (TW)z1;
  z[2] = (TW)z2;
}

// One wavefront per S block (a, b): S = [Bdiag + damping - ZZ_a if a == b] + Bfac
// - sum_pairs Z_x Z_y^T. The pair sum is the product X Y^T with X = [Z_x1 Z_x2 ...] (ND x 3P)
// and Y = [Z_y1 Z_y2 ...], so K = 3P is packed densely into FP64 matrix-core steps
// (v_mfma_f64_16x16x4f64; a batch of four pairs = 12 K entries = three MFMAs): at step h lane l
// feeds A(i = l % 16, k = l / 16) = X(i, 4h + k) and B(k, j = l % 16) = Y(j, 4h + k) (rows
// i >= ND and K entries past the block's pairs are zero) and holds D(l / 16 + 4 r, l % 16) in
// accumulator register r (the FP64 16x16x4 output interleaves rows across the four lane groups).
// The pair list is padded so a batch reaching past the last block stays inside it. Diagonal
// blocks carry no pairs: their sum_o Z_o Z_o^T (ZZ_a) comes from the per-camera reduction. Waves
// take work items from a host-built list: (block, pair range, flags), blocks with many pairs
// split into several items whose partial sums are added atomically (kRcsAtomic; the output is
// zero-filled first) and only the first item of a block adds the B / damping / ZZ terms
// (kRcsFirst); block -1 = idle wave. At 1000 cameras the launch moves ~5 TB/s of Z copies past
// L2 (each Z block is read once per pair it is in, ~10 times, against ~2 for perfect reuse) with
// the matrix cores a third busy: traffic, not a wave's latency, is its bound, and the item order
// trades L2 reuse (row order) against the launch's tail (largest items first).
// Output: into the LDL^T panels at (base + col * ld + row) or transposed, or (sfull != null)
// into full symmetric block storage, block row-major, at f_lo[blk] and its mirror f_up[blk]
// (transposed). Z is read as TW (FP32 in the FP32 form) and widened: the sums are FP64 on the
// matrix cores either way; the full storage is written as TS.
constexpr int kRcsChunk = 512;  // pairs per work item at most
constexpr int kRcsBatch = 4;    // pairs per batch (a multiple of 4: 3 K entries per pair, 4 per MFMA)
constexpr bool kRcsHeaviestFirst = false;  // order each XCD's items by size (else by camera row)
constexpr double kRcsTailFrac = 0.4;       // row order: share of the pairs dealt heaviest first
constexpr int kRcsTailChunk = 256;         // row order: pairs per item at most in that tail
constexpr int kRcsFirst = 1, kRcsAtomic = 2;
template <int ND, typename TW, typename TS>
__global__ void __launch_bounds__(256) rcs_assemble_kernel(
    int32_t n_items, const int4* __restrict__ items, const int32_t* __restrict__ blk_a,
    const int32_t* __restrict__ blk_b, const int32_t* __restrict__ pairs,
    const TW* __restrict__ Z,
    const double* __restrict__ ZZ, const double* __restrict__ Bdiag,
    const double* __restrict__ Bfac, const double* __restrict__ dnode,
    const int64_t* __restrict__ panel_base, const int32_t* __restrict__ panel_ld,
    const int32_t* __restrict__ panel_trans, double* __restrict__ panel,
    TS* __restrict__ sfull, const int32_t* __restrict__ f_lo,
    const int32_t* __restrict__ f_up) {
  static_assert(ND <= 16, "camera node wider than the MFMA tile");
  typedef double d4 __attribute__((ext_vector_type(4)));
  constexpr int NE = ND * ND;
  const int t = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (t >= n_items) return;
  const int4 item = items[t];
  const int blk = __builtin_amdgcn_readfirstlane(item.x);
  if (blk < 0) return;
  const int q_begin = __builtin_amdgcn_readfirstlane(item.y);
  const int q_end = __builtin_amdgcn_readfirstlane(item.z);
  const int flags = __builtin_amdgcn_readfirstlane(item.w);
  const int lane = threadIdx.x & 63;
  const int a = blk_a[blk], b = blk_b[blk];
  const int row = lane & 15, kg = lane >> 4;
  const bool live = row < ND;
  const int rowoff = (live ? row : 0) * 3;
  d4 acc = {};
  // A batch is kRcsBatch pairs = 2 kRcsBatch observations = 3 kRcsBatch K entries =
  // kRcsBatch * 3 / 4 MFMAs. Its Z blocks (ZS elements = CPO 16-byte chunks each) are copied
  // whole into the wavefront's LDS buffer by direct-to-LDS loads (global_load_lds_dwordx4: lane
  // t of load u writes chunk s = 64 u + t at byte 16 s, so observation slot o = s / CPO lands at
  // o * ZS elements): every byte of a fetched cache line is used, where per-lane 8-byte gathers
  // in the MFMA layout touched ~10 lines per instruction for 288 useful bytes. Slot o of pair
  // p = o / 2 is its x (o even) or y (o odd), i.e. pairs[2 q + o]. At step h lane l reads its
  // operands from LDS: K = 4h + l / 16 is pair K / 3, column K % 3. Pairs past the item's end
  // come from the next pairs or the padding (valid observations), masked by ok.
  constexpr int ZS = z_stride<ND, TW>(), NB = kRcsBatch, NH = NB * 3 / 4;
  constexpr int CPO = ZS * (int)sizeof(TW) / 16;
  constexpr int SLOTS = 2 * NB * CPO, NI = (SLOTS + 63) / 64;
  static_assert(ZS * (int)sizeof(TW) % 16 == 0, "Z blocks must be whole 16-byte chunks");
  static_assert(NI <= 15, "vmcnt immediate");
  __shared__ __attribute__((aligned(16))) char lds_z[4][NI * 1024];
  const int w = threadIdx.x >> 6;
  // Every lane of every copy is active (a masked copy would be a branch, and the wait counts
  // below assume NI copies per batch): slots past the batch's chunks re-read chunk 0 into the
  // unused tail of the buffer.
  int slot_o[NI], slot_c[NI];
#pragma unroll
  for (int u = 0; u < NI; ++u) {
    const int sl = 64 * u + lane;
    slot_o[u] = sl < SLOTS ? sl / CPO : 0;
    slot_c[u] = sl < SLOTS ? sl % CPO : 0;
  }
  // Each lane fetches the observation index of its own slots with vector loads, one batch
  // ahead (after the batch's copy, so the wait for the copy leaves them in flight). The index
  // loads are always issued (a batch past the item's end re-reads its first pairs) to keep the
  // wait count fixed.
  auto fetch_idx = [&](int qb, int (&ix)[NI]) {
    const int qs = qb < q_end ? qb : q_begin;
#pragma unroll
    for (int u = 0; u < NI; ++u) ix[u] = pairs[2 * qs + slot_o[u]];
  };
  auto issue = [&](const int (&ix)[NI], char* buf) {
#pragma unroll
    for (int u = 0; u < NI; ++u) {
      const char* src = reinterpret_cast<const char*>(Z + (int64_t)ix[u] * ZS) + slot_c[u] * 16;
      __builtin_amdgcn_global_load_lds(src, (__attribute__((address_space(3))) void*)(buf + u * 1024),
                                       16, 0, 0);
    }
  };
  // per-lane operand offsets (elements) of step h inside a batch buffer
  int off_x[NH], off_y[NH];
#pragma unroll
  for (int h = 0; h < NH; ++h) {
    const int K = 4 * h + kg;
    const int p = K / 3, col = K - 3 * p;
    off_x[h] = (2 * p) * ZS + rowoff + col;
    off_y[h] = (2 * p + 1) * ZS + rowoff + col;
  }
  // The operand reads are written as asm blocks (three steps' A and B operands each) that also
  // wait for them: the compiler would otherwise put a wait for every outstanding direct-to-LDS
  // copy (the next batch's included) before the first read.
  static_assert(NH % 3 == 0, "operand reads go in groups of three MFMA steps");
  auto read3 = [&](uint32_t base, int g, double* av, double* bv) {
    if constexpr (sizeof(TW) == 8) {
      asm volatile(
          "ds_read_b64 %0, %6\n\t"
          "ds_read_b64 %1, %7\n\t"
          "ds_read_b64 %2, %8\n\t"
          "ds_read_b64 %3, %9\n\t"
          "ds_read_b64 %4, %10\n\t"
          "ds_read_b64 %5, %11\n\t"
          "s_waitcnt lgkmcnt(0)"
          : "=&v"(av[0]), "=&v"(av[1]), "=&v"(av[2]), "=&v"(bv[0]), "=&v"(bv[1]), "=&v"(bv[2])
          : "v"(base + 8 * off_x[g]), "v"(base + 8 * off_x[g + 1]), "v"(base + 8 * off_x[g + 2]),
            "v"(base + 8 * off_y[g]), "v"(base + 8 * off_y[g + 1]), "v"(base + 8 * off_y[g + 2])
          : "memory");
    } else {
      float fa[3], fb[3];
      asm volatile(
          "ds_read_b32 %0, %6\n\t"
          "ds_read_b32 %1, %7\n\t"
          "ds_read_b32 %2, %8\n\t"
          "ds_read_b32 %3, %9\n\t"
          "ds_read_b32 %4, %10\n\t"
          "ds_read_b32 %5, %11\n\t"
          "s_waitcnt lgkmcnt(0)"
          : "=&v"(fa[0]), "=&v"(fa[1]), "=&v"(fa[2]), "=&v"(fb[0]), "=&v"(fb[1]), "=&v"(fb[2])
          : "v"(base + 4 * off_x[g]), "v"(base + 4 * off_x[g + 1]), "v"(base + 4 * off_x[g + 2]),
            "v"(base + 4 * off_y[g]), "v"(base + 4 * off_y[g + 1]), "v"(base + 4 * off_y[g + 2])
          : "memory");
#pragma unroll
      for (int h = 0; h < 3; ++h) {
        av[h] = (double)fa[h];
        bv[h] = (double)fb[h];
      }
    }
  };
  auto consume = [&](int q, const char* buf) {
    const uint32_t base = (uint32_t)(uintptr_t)(__attribute__((address_space(3))) const char*)buf;
    double av[NH], bv[NH];
#pragma unroll
    for (int g = 0; g < NH; g += 3) read3(base, g, av + g, bv + g);
#pragma unroll
    for (int h = 0; h < NH; ++h) {
      const int p = (4 * h + kg) / 3;
      const bool ok = live && q + p < q_end;
      av[h] = ok ? av[h] : 0.0;
      bv[h] = ok ? bv[h] : 0.0;
    }
#pragma unroll
    for (int h = 0; h < NH; ++h)
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av[h], bv[h], acc, 0, 0, 0);
  };
  // Per batch: copy batch k, fetch batch k + 1's indices, wait for the copy (the index loads
  // issued after it may stay in flight: vector memory operations complete in order), read,
  // multiply. (A second buffer, batch k + 1's copy overlapping batch k's reads and products,
  // measured no faster: the launch is bound by the copies' HBM and L2 traffic, not by a wave's
  // latency.)
  int ix[NI];
  fetch_idx(q_begin, ix);
  for (int q = q_begin; q < q_end; q += NB) {
    issue(ix, lds_z[w]);
    fetch_idx(q + NB, ix);
    asm volatile("s_waitcnt vmcnt(%0)" ::"n"(NI) : "memory");
    consume(q, lds_z[w]);
  }
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");  // the last index loads
  const d4 sum = acc;
  const int j = row;
  if (j >= ND) return;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int i = kg + 4 * r;
    if (i >= ND) continue;
    if (a == b && i < j) continue;
    double s = -sum[r];
    if (flags & kRcsFirst) {
      s += Bfac[(int64_t)blk * NE + i * ND + j];
      if (a == b)
        s += Bdiag[(int64_t)a * NE + i * ND + j] - ZZ[(int64_t)a * NE + i * ND + j] +
             (i == j ? dnode[a * ND + i] : 0.0);
    }
    const bool atomic = flags & kRcsAtomic;
    if (sfull) {
      TS* lo = sfull + (int64_t)f_lo[blk] * NE + i * ND + j;
      TS* up = sfull + (int64_t)f_up[blk] * NE + j * ND + i;
      if (atomic) {
        unsafeAtomicAdd(lo, (TS)s);
        unsafeAtomicAdd(up, (TS)s);
      } else {
        *lo = (TS)s;
        *up = (TS)s;
      }
    } else {
      // entry (row a*ND+i, col b*ND+j) of S
      const int64_t base = panel_base[blk];
      const int ld = panel_ld[blk];
      const int64_t pos = panel_trans[blk] ? base + (int64_t)i * ld + j : base + (int64_t)j * ld + i;
      if (atomic) unsafeAtomicAdd(panel + pos, s); else panel[pos] = s;
    }
  }
}

// Per camera node (a block of four wavefronts): v_a = g_a - sum_{o in a} Z_o h_p(o) and the
// diagonal-block elimination term ZZ_a = sum_{o in a} Z_o Z_o^T (full ND x ND, row-major). The
// wavefronts take every fourth chunk of 64 of the camera's observations; a chunk's Z rows are
// copied into LDS with coalesced 16-byte loads (a few consecutive lanes per row), each lane sums
// its observation's terms from its LDS row, and the four wavefronts' sums are added at the end.
// (One wavefront per camera, each lane loading its own row: 90 us per launch at 1000 cameras;
// this form 72 us: the chain of passes over a camera's ~1000 observations is a quarter as long.)
template <int ND, typename TW>
__global__ void __launch_bounds__(256) reduced_rhs_kernel(int32_t n_nodes,
                                                          const int32_t* __restrict__ cam_obs_ptr,
                                                          const int32_t* __restrict__ cam_obs,
                                                          const int32_t* __restrict__ obs_point,
                                                          const TW* __restrict__ Z,
                                                          const double* __restrict__ hp,
                                                          const double* __restrict__ gB,
                                                          double* __restrict__ v,
                                                          double* __restrict__ ZZ) {
  constexpr int NH = ND * (ND + 1) / 2;
  constexpr int ZS = z_stride<ND, TW>();
  constexpr int ZL = ZS + 1;                  // LDS row stride (odd: conflict-free row reads)
  constexpr int kVec = 16 / (int)sizeof(TW);  // Z entries per 16-byte piece
  constexpr int kPpo = ZS / kVec;             // pieces per observation
  constexpr int kZl = 64 * ZL;                // one wavefront's staged rows
  static_assert(kZl >= 16 * 65, "the reduction staging fits in the row staging");
  __shared__ double lds[4][kZl];
  __shared__ double red[4][ND + NH];
  const int cam = xcd_tile(blockIdx.x, n_nodes);
  if (cam >= n_nodes) return;  // block-uniform
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  double* zs = lds[wave];
  double acc[ND] = {}, zz[NH] = {};
  const int q_end = cam_obs_ptr[cam + 1];
  for (int q0 = cam_obs_ptr[cam] + 64 * wave; q0 < q_end; q0 += 256) {
    // the 64 observations' Z rows, each 16-byte piece by a lane (rows clamped to the range)
    typedef TW v16 __attribute__((ext_vector_type(kVec)));
#pragma unroll
    for (int m = 0; m < kPpo; ++m) {
      const int piece = lane + 64 * m;
      const int ol = piece / kPpo, within = (piece % kPpo) * kVec;
      const int o = cam_obs[min(q0 + ol, q_end - 1)];
      const v16 zv = *reinterpret_cast<const v16*>(Z + (int64_t)o * ZS + within);
#pragma unroll
      for (int e = 0; e < kVec; ++e) zs[ol * ZL + within + e] = (double)zv[e];
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (q0 + lane < q_end) {
      const int p = obs_point[cam_obs[q0 + lane]];
      const double h0 = hp[p * 3], h1 = hp[p * 3 + 1], h2 = hp[p * 3 + 2];
      double z[ND * 3];
#pragma unroll
      for (int e = 0; e < ND * 3; ++e) z[e] = zs[lane * ZL + e];
#pragma unroll
      for (int i = 0; i < ND; ++i) acc[i] += z[i * 3] * h0 + z[i * 3 + 1] * h1 + z[i * 3 + 2] * h2;
      int idx = 0;
#pragma unroll
      for (int i = 0; i < ND; ++i)
#pragma unroll
        for (int k = 0; k <= i; ++k)
          zz[idx++] += z[i * 3] * z[k * 3] + z[i * 3 + 1] * z[k * 3 + 1] + z[i * 3 + 2] * z[k * 3 + 2];
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();  // the next pass (or the reduction) overwrites the rows
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
  double az[ND + NH];
#pragma unroll
  for (int i = 0; i < ND; ++i) az[i] = acc[i];
#pragma unroll
  for (int i = 0; i < NH; ++i) az[ND + i] = zz[i];
  wave_sum_lds<ND + NH>(az, reinterpret_cast<double (*)[65]>(zs),
                        [&](int e, double t) { red[wave][e] = t; });
  __syncthreads();
  const int e = threadIdx.x;
  if (e < ND + NH) {
    const double t = red[0][e] + red[1][e] + red[2][e] + red[3][e];
    if (e < ND) {
      v[cam * ND + e] = gB[cam * ND + e] - t;
    } else {
      const int p = e - ND;  // packed lower entry (i, k), k <= i
      int i = 0;
      while ((i + 1) * (i + 2) / 2 <= p) ++i;
      const int k = p - i * (i + 1) / 2;
      double* zo = ZZ + (int64_t)cam * ND * ND;
      zo[i * ND + k] = t;
      zo[k * ND + i] = t;
    }
  }
}

// Per point: z_p = L_p^-T (h_p - sum_o Z_o^T y_node(o)); write -z (and -y for the nodes) into the
// update vector in reference state order. A wavefront takes 64 consecutive points, whose
// observations are one contiguous range of Z rows: a chunk of 64 of them is copied into LDS with
// coalesced 16-byte loads (consecutive lanes, consecutive pieces), each lane forms one
// observation's Z_o^T y_node(o) from its LDS row, and each lane then sums the observations of its
// own point. (A lane loading its own observation's 224-byte row touches 64 cache lines per
// instruction and re-requests each line ~8 times: 114 -> 65 us per launch at 1000 cameras.)
constexpr int kBackSubChunk = 64;  // observations staged per wavefront and pass

template <int ND, typename TW>
__global__ void __launch_bounds__(256) back_substitute_kernel(
    int32_t n_points, const int32_t* __restrict__ pt_obs_ptr, const int32_t* __restrict__ obs_node,
    const TW* __restrict__ Z, const double* __restrict__ Lp, const double* __restrict__ hp,
    const double* __restrict__ y, const int32_t* __restrict__ pt_tan, double* __restrict__ update) {
  constexpr int ZS = z_stride<ND, TW>();
  constexpr int ZL = ZS + 1;                   // LDS row stride (odd: conflict-free row reads)
  constexpr int kVec = 16 / (int)sizeof(TW);   // Z entries per 16-byte piece
  constexpr int kPieces = kBackSubChunk * ZS / kVec;
  static_assert(kPieces % 64 == 0 && ZS % kVec == 0, "whole pieces per lane and observation");
  __shared__ double zs[4][kBackSubChunk * ZL];
  __shared__ double ts[4][kBackSubChunk][3];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int p0 = (blockIdx.x * 4 + wave) * 64;
  if (p0 >= n_points) return;  // wave-uniform; the kernel has no block barrier
  const int p = p0 + lane, pe = min(p0 + 64, n_points);
  const int o_begin = pt_obs_ptr[p0], o_end = pt_obs_ptr[pe];
  const int pc = min(p, n_points - 1);
  const int my_b = pt_obs_ptr[pc], my_e = p < n_points ? pt_obs_ptr[pc + 1] : my_b;
  double t0 = 0, t1 = 0, t2 = 0;
  for (int c0 = o_begin; c0 < o_end; c0 += kBackSubChunk) {
    // the chunk's Z rows are contiguous: coalesced 16-byte pieces into LDS (rows clamped)
    typedef TW v16 __attribute__((ext_vector_type(kVec)));
#pragma unroll
    for (int m = 0; m < kPieces / 64; ++m) {
      const int piece = lane + 64 * m;
      const int ol = piece / (ZS / kVec), within = (piece % (ZS / kVec)) * kVec;
      const int o = min(c0 + ol, o_end - 1);
      const v16 zv = *reinterpret_cast<const v16*>(Z + (int64_t)o * ZS + within);
#pragma unroll
      for (int e = 0; e < kVec; ++e) zs[wave][ol * ZL + within + e] = (double)zv[e];
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    {
      const int o = min(c0 + lane, o_end - 1);
      const double* yn = y + obs_node[o] * ND;
      const double* zr = &zs[wave][lane * ZL];
      double s0 = 0, s1 = 0, s2 = 0;
#pragma unroll
      for (int i = 0; i < ND; ++i) {
        const double yi = yn[i];
        s0 += zr[i * 3] * yi;
        s1 += zr[i * 3 + 1] * yi;
        s2 += zr[i * 3 + 2] * yi;
      }
      ts[wave][lane][0] = s0;
      ts[wave][lane][1] = s1;
      ts[wave][lane][2] = s2;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const int lo = max(my_b, c0), hi = min(my_e, c0 + kBackSubChunk);
    for (int o = lo; o < hi; ++o) {
      t0 += ts[wave][o - c0][0];
      t1 += ts[wave][o - c0][1];
      t2 += ts[wave][o - c0][2];
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();  // the next pass overwrites the staging rows
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
  if (p >= n_points) return;
  t0 = hp[p * 3] - t0;
  t1 = hp[p * 3 + 1] - t1;
  t2 = hp[p * 3 + 2] - t2;
  const double* L = Lp + (int64_t)p * 6;
  const double z2 = t2 / L[5];
  const double z1 = (t1 - L[4] * z2) / L[2];
  const double z0 = (t0 - L[1] * z1 - L[3] * z2) / L[0];
  update[pt_tan[p * 3 + 0]] = -z0;
  update[pt_tan[p * 3 + 1]] = -z1;
  update[pt_tan[p * 3 + 2]] = -z2;
}

__global__ void scatter_neg_kernel(int32_t n, const int32_t* __restrict__ idx,
                                   const double* __restrict__ x, double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[idx[i]] = -x[i];
}
__global__ void scatter_kernel(int32_t n, const int32_t* __restrict__ idx,
                               const double* __restrict__ x, double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[idx[i]] = x[i];
}

}  // namespace

// ---------------------------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------------------------
class SchurBackend final : public SystemBackend {
 public:
  SchurBackend(const HostProblem& p, const SolverOptions& o, hipStream_t s);
  void relinearize(int b, const double* values, double* err_out, hipStream_t s) override;
  void damp(int b, const double* lambda_dev, const symhip_optimizer_params& p,
            bool have_max_diagonal, hipStream_t s) override;
  void restore(int, hipStream_t) override {}  // damping is applied out of place
  void solve(int b, double* update, hipStream_t s) override;
  const double* rhs(int b) const override { return rhs_[b].get(); }
  const double* residual(int b) const override { return residual_[b].get(); }
  const double* damping() const override { return damping_.get(); }
  const double* undamped_diagonal() const override { return undamped_.get(); }
  void hessian_csc(int b, double* out, hipStream_t s) override;
  int32_t N() const override { return N_; }
  int32_t M() const override { return M_; }
  int64_t nnz() const override { return host_->nnz(); }
  void info(int64_t* out) const override {
    const CholSymbolic& y = chol_.symbolic();
    out[0] = y.n;
    out[1] = y.n_snodes;
    out[2] = y.panel_size();
    out[3] = y.n_levels;
    out[4] = (int64_t)y.factor_flops;
    out[8] = (int64_t)(y.factor_flops - y.update_flops);
    out[9] = (int64_t)y.update_flops;
    out[5] = n_blocks_;
    out[6] = n_pairs_;
    out[7] = n_obs_;
  }
  void begin_run() override {
    exact_now_ = false;
    last_reduction_ = -1;
  }
  bool defers_early_exit() const override {
    return use_pcg_ && !pcg_strict_ && exact_below_ > 0 && !exact_now_;
  }
  void switch_to_exact() override { exact_now_ = true; }
  void observe(const symhip_iteration_stats& rec) override {
    if (!use_pcg_ || !rec.update_accepted || rec.relative_reduction < 0) return;
    last_reduction_ = rec.relative_reduction;
    if (exact_below_ > 0 && rec.relative_reduction < exact_below_) exact_now_ = true;
  }
  void log_systems(int max) override;
  int logged_systems() const override { return (int)logged_S_.size(); }
  int rcs_pattern(int32_t* row_ptr, int32_t* col) const override;
  void logged_system(int idx, double* blocks, double* rhs) const override;
  void precond_params(float* get, const float* set) override;
  float precond_loss_grad(int idx, const double* X, int P, float* grad) override;
  void train_precond(int steps, float lr, int P, uint64_t seed, int n_systems,
                     float* losses) override;
  void eval_precond(int idx, double rel_tol, int max_iters, int learned, int* iters,
                    double* step_err) override;
  void linear_stats(int64_t* out) const override {
    out[0] = pcg_solves_;
    out[1] = pcg_iters_;
    out[2] = pcg_fallbacks_;
    out[3] = pcg_max_seen_;
  }

 private:
  template <int ND>
  void solve_nd(int b, double* update, hipStream_t s);
  template <int ND>
  void relin_nd(int b, const double* values, hipStream_t s);

  const HostProblem* host_;
  hipStream_t stream_;  // the optimizer's stream (the training / evaluation calls use it too)
  int nd_ = 0, ftype_ = 0;
  int32_t n_nodes_ = 0, n_points_ = 0, n_obs_ = 0, n_blocks_ = 0, N_ = 0, M_ = 0;
  int64_t n_pairs_ = 0;
  // observation tables
  DBuf<int32_t> obs_arg_off_, obs_res_off_, obs_node_, obs_point_, pt_obs_ptr_, cam_obs_ptr_,
      cam_obs_;
  DBuf<int32_t> node_tan_, pt_tan_, blk_a_, blk_b_, pair_ptr_, pairs_, diag_blk_;
  DBuf<uint8_t> rcs_items_;  // int4 work items of the assembly kernel
  int32_t n_rcs_items_ = 0;
  DBuf<int64_t> panel_base_;
  DBuf<int32_t> panel_ld_, panel_trans_;
  // camera-only factors
  struct BGroup {
    int type;
    int32_t n;
    DBuf<int32_t> arg_off, res_off, col_map, blk_of;
  };
  std::vector<BGroup> bgroups_;
  // per state block storage
  DBuf<double> residual_[3], W_[3], V_[3], gp_[3], Bdiag_[3], Bfac_[3], gB_[3], rhs_[3];
  DBuf<double> damping_, undamped_, max_diag_, dnode_, dpoint_, Lp_, hp_, Z_, ZZ_, v_, y_,
      partials_;
  // FP32 form (precision 1): while PCG runs, W and Z are stored in FP32 (W32_, Z32_) and so is S
  // for PCG; residual, gradient, sums, LDL^T and the LM decisions stay FP64. Once the run has
  // switched to the exact LDL^T (exact_now_), linearization returns to plain FP64 storage, the
  // FP64 method finishing the run as its judge. w_f32_[b]: precision of block b's W.
  bool f32_ = false;
  bool w_f32_[3] = {false, false, false};
  DBuf<float> W32_[3], Z32_;
  // Calls f(W, Z) with the typed W of state block b and the Z storage of the same precision.
  template <typename F>
  void with_storage(int b, F&& f) {
    if (w_f32_[b]) f(W32_[b].get(), Z32_.get()); else f(W_[b].get(), Z_.get());
  }
  SupernodalLDLT chol_;
  std::unique_ptr<CscLinearizer> csc_;  // reference-layout H for best-linearization output
  // Iterative form: PCG on the reduced camera system in full block storage, falling back to the
  // supernodal LDL^T when it does not reach the tolerance.
  bool use_pcg_ = false, pcg_strict_ = false, exact_now_ = false;
  double exact_below_ = 0;
  double forcing_ = 0, last_reduction_ = -1;  // inexact-Newton forcing (pcg_forcing)
  int pcg_max_iters_ = 100;
  double pcg_rel_tol_ = 1e-6;
  BlockPcg pcg_;
  DBuf<int32_t> f_lo_, f_up_;
  std::vector<int32_t> full_row_ptr_, full_col_;
  // learned correction of the block-Jacobi preconditioner and the logged training systems
  bool use_learned_ = false;
  LearnedPrecond gnn_;
  int log_max_ = 0;
  std::vector<DBuf<double>> logged_S_, logged_v_;
  DBuf<double> probe_x_, probe_y_, exact_y_;
  template <int ND>
  void panels_from_full(const double* S, hipStream_t s);
  int64_t pcg_solves_ = 0, pcg_iters_ = 0, pcg_fallbacks_ = 0, pcg_max_seen_ = 0;
};

namespace {
struct UnionFind {
  std::vector<int32_t> p;
  explicit UnionFind(int n) : p(n) { std::iota(p.begin(), p.end(), 0); }
  int find(int x) {
    while (p[x] != x) x = p[x] = p[p[x]];
    return x;
  }
  void unite(int a, int b) { p[find(a)] = find(b); }
};
}  // namespace

SchurBackend::SchurBackend(const HostProblem& P, const SolverOptions& o, hipStream_t s)
    : host_(&P), stream_(s) {
  SYMHIP_CHECK(o.precision == SYMHIP_PRECISION_F64 || o.precision == SYMHIP_PRECISION_F32_JAC,
               "symhip: unknown precision");
  f32_ = o.precision == SYMHIP_PRECISION_F32_JAC;
  const int32_t S = (int32_t)P.state_keys.size();
  N_ = P.N;
  M_ = P.M;
  std::vector<int32_t> pos_of(P.n_keys, -1);
  for (int32_t i = 0; i < S; ++i) pos_of[P.state_keys[i]] = i;
  // Landmarks: trailing VECTOR3 state keys that no factor touches twice.
  std::vector<int32_t> touch_count(S, 0);
  std::vector<std::vector<int32_t>> fpos(P.n_factors);  // state positions of linearized args
  for (int32_t f = 0; f < P.n_factors; ++f) {
    const FactorSpec sp = factor_spec(P.ftype[f]);
    for (int a = 0; a < sp.nargs; ++a)
      if ((sp.lin_mask >> a) & 1) {
        const int32_t k = pos_of[P.fargs[P.fargs_ptr[f] + a]];
        SYMHIP_CHECK(k >= 0, "Schur solver requires every linearized argument to be optimized");
        fpos[f].push_back(k);
      }
  }
  int32_t first_c = S;
  {
    std::vector<char> is_v3(S);
    for (int32_t i = 0; i < S; ++i) is_v3[i] = P.types[P.state_keys[i]] == SYMHIP_TYPE_VECTOR3;
    while (first_c > 0 && is_v3[first_c - 1]) --first_c;
    // shrink until no factor touches two landmark keys
    bool ok = false;
    while (!ok) {
      ok = true;
      for (int32_t f = 0; f < P.n_factors && ok; ++f) {
        int cnt = 0, minc = S;
        for (int32_t k : fpos[f])
          if (k >= first_c) {
            ++cnt;
            minc = std::min(minc, k);
          }
        if (cnt > 1) {
          first_c = minc + 1;
          ok = false;
        }
      }
    }
  }
  n_points_ = S - first_c;
  SYMHIP_CHECK(n_points_ > 0, "Schur solver: no landmark block found at the end of the state");
  // Camera nodes: union B keys co-occurring in landmark factors.
  UnionFind uf(first_c);
  std::vector<int32_t> lm_factors, b_factors;
  for (int32_t f = 0; f < P.n_factors; ++f) {
    bool lm = false;
    for (int32_t k : fpos[f]) lm |= k >= first_c;
    if (lm) {
      lm_factors.push_back(f);
      int32_t first_b = -1;
      for (int32_t k : fpos[f])
        if (k < first_c) {
          if (first_b < 0) first_b = k; else uf.unite(k, first_b);
        }
    } else {
      b_factors.push_back(f);
    }
  }
  SYMHIP_CHECK(!lm_factors.empty(), "Schur solver: no landmark factors");
  ftype_ = P.ftype[lm_factors[0]];
  for (int32_t f : lm_factors)
    SYMHIP_CHECK(P.ftype[f] == ftype_, "Schur solver: one landmark factor type per problem");
  SYMHIP_CHECK(ftype_ == SYMHIP_FACTOR_SNAVELY || ftype_ == SYMHIP_FACTOR_SNAVELY_HUBER ||
                   ftype_ == SYMHIP_FACTOR_REPROJ_CAUCHY,
               "Schur solver: unsupported landmark factor type");
  // node ids in order of their first state key; node layout = member keys in state order
  std::vector<int32_t> node_of_pos(first_c, -1), root_node;
  std::vector<int32_t> root_to_node(first_c, -1);
  std::vector<std::vector<int32_t>> node_members;
  for (int32_t k = 0; k < first_c; ++k) {
    const int r = uf.find(k);
    if (root_to_node[r] < 0) {
      root_to_node[r] = (int32_t)node_members.size();
      node_members.push_back({});
    }
    node_of_pos[k] = root_to_node[r];
    node_members[root_to_node[r]].push_back(k);
  }
  n_nodes_ = (int32_t)node_members.size();
  nd_ = 0;
  std::vector<int32_t> pos_off_in_node(first_c, 0);
  for (auto& mem : node_members) {
    int d = 0;
    for (int32_t k : mem) {
      pos_off_in_node[k] = d;
      d += tangent_dim(P.types[P.state_keys[k]]);
    }
    if (nd_ == 0) nd_ = d;
    SYMHIP_CHECK(d == nd_, "Schur solver: camera nodes must all have the same dimension");
  }
  const FactorSpec lsp = factor_spec(ftype_);
  SYMHIP_CHECK(nd_ == lsp.tdim - 3, "Schur solver: camera node does not match the factor layout");
  // each landmark factor: B args in node layout order, landmark last
  for (int32_t f : lm_factors) {
    const auto& ks = fpos[f];
    SYMHIP_CHECK(ks.size() >= 2 && ks[0] < first_c,
                 "Schur solver: landmark factors must also touch a camera node");
    SYMHIP_CHECK(ks.back() >= first_c, "Schur solver: landmark must be the last linearized arg");
    int off = 0;
    for (size_t i = 0; i + 1 < ks.size(); ++i) {
      SYMHIP_CHECK(node_of_pos[ks[i]] == node_of_pos[ks[0]] && pos_off_in_node[ks[i]] == off,
                   "Schur solver: factor arguments do not follow the camera node layout");
      off += tangent_dim(P.types[P.state_keys[ks[i]]]);
    }
  }
  // node -> reference tangent offsets
  std::vector<int32_t> node_tan((size_t)n_nodes_ * nd_);
  for (int32_t a = 0; a < n_nodes_; ++a)
    for (int32_t k : node_members[a]) {
      const int td = tangent_dim(P.types[P.state_keys[k]]);
      for (int d = 0; d < td; ++d)
        node_tan[(size_t)a * nd_ + pos_off_in_node[k] + d] = P.tan_off[P.state_keys[k]] + d;
    }
  // observations grouped by point; points sorted by their first camera node
  std::vector<std::vector<int32_t>> obs_of_pt(n_points_);
  std::vector<int32_t> obs_node_raw(P.n_factors, -1);
  for (int32_t f : lm_factors) {
    obs_of_pt[fpos[f].back() - first_c].push_back(f);
    obs_node_raw[f] = node_of_pos[fpos[f][0]];
  }
  std::vector<int32_t> pt_order(n_points_);
  std::iota(pt_order.begin(), pt_order.end(), 0);
  std::vector<int32_t> first_node(n_points_, INT32_MAX);
  for (int32_t p = 0; p < n_points_; ++p) {
    SYMHIP_CHECK(!obs_of_pt[p].empty(), "Schur solver: landmark without observations");
    std::sort(obs_of_pt[p].begin(), obs_of_pt[p].end(),
              [&](int32_t x, int32_t y) { return obs_node_raw[x] < obs_node_raw[y]; });
    first_node[p] = obs_node_raw[obs_of_pt[p][0]];
  }
  std::stable_sort(pt_order.begin(), pt_order.end(),
                   [&](int32_t x, int32_t y) { return first_node[x] < first_node[y]; });
  std::vector<int32_t> obs_f, pt_ptr{0}, obs_node, obs_point, pt_tan((size_t)n_points_ * 3);
  for (int32_t i = 0; i < n_points_; ++i) {
    const int32_t p = pt_order[i];
    for (int32_t f : obs_of_pt[p]) {
      obs_f.push_back(f);
      obs_node.push_back(obs_node_raw[f]);
      obs_point.push_back(i);
    }
    pt_ptr.push_back((int32_t)obs_f.size());
    for (int d = 0; d < 3; ++d) pt_tan[(size_t)i * 3 + d] = P.tan_off[P.state_keys[first_c + p]] + d;
  }
  n_obs_ = (int32_t)obs_f.size();
  std::vector<int32_t> arg_off((size_t)lsp.nargs * n_obs_), res_off(n_obs_);
  for (int32_t o = 0; o < n_obs_; ++o) {
    const int32_t f = obs_f[o];
    for (int a = 0; a < lsp.nargs; ++a)
      arg_off[(size_t)a * n_obs_ + o] = P.storage_offset[P.fargs[P.fargs_ptr[f] + a]];
    res_off[o] = P.res_off[f];
  }
  std::vector<int32_t> cam_ptr(n_nodes_ + 1, 0), cam_obs(n_obs_);
  for (int32_t o = 0; o < n_obs_; ++o) cam_ptr[obs_node[o] + 1]++;
  for (int32_t a = 0; a < n_nodes_; ++a) cam_ptr[a + 1] += cam_ptr[a];
  {
    std::vector<int32_t> fill(cam_ptr.begin(), cam_ptr.end() - 1);
    for (int32_t o = 0; o < n_obs_; ++o) cam_obs[fill[obs_node[o]]++] = o;
  }
  // S block pattern: diagonal, covisible pairs, camera-factor couplings. Lower: a >= b.
  std::vector<std::vector<int32_t>> row_cols(n_nodes_);
  for (int32_t a = 0; a < n_nodes_; ++a) row_cols[a].push_back(a);
  for (int32_t p = 0; p < n_points_; ++p)
    for (int32_t x = pt_ptr[p]; x < pt_ptr[p + 1]; ++x)
      for (int32_t y = pt_ptr[p]; y < x; ++y) {
        const int32_t a = std::max(obs_node[x], obs_node[y]), b = std::min(obs_node[x], obs_node[y]);
        row_cols[a].push_back(b);
      }
  for (int32_t f : b_factors)
    for (int32_t k1 : fpos[f])
      for (int32_t k2 : fpos[f]) {
        const int32_t a = node_of_pos[k1], b = node_of_pos[k2];
        if (a >= b) row_cols[a].push_back(b);
      }
  std::vector<int32_t> row_ptr(n_nodes_ + 1, 0), blk_a, blk_b;
  for (int32_t a = 0; a < n_nodes_; ++a) {
    auto& rc = row_cols[a];
    std::sort(rc.begin(), rc.end());
    rc.erase(std::unique(rc.begin(), rc.end()), rc.end());
    row_ptr[a + 1] = row_ptr[a] + (int32_t)rc.size();
    for (int32_t b : rc) {
      blk_a.push_back(a);
      blk_b.push_back(b);
    }
  }
  n_blocks_ = (int32_t)blk_a.size();
  auto block_index = [&](int32_t a, int32_t b) -> int32_t {
    const auto& rc = row_cols[a];
    const auto it = std::lower_bound(rc.begin(), rc.end(), b);
    SYMHIP_CHECK(it != rc.end() && *it == b, "internal: S block missing");
    return row_ptr[a] + (int32_t)(it - rc.begin());
  };
  std::vector<int32_t> diag_blk(n_nodes_);
  for (int32_t a = 0; a < n_nodes_; ++a) diag_blk[a] = block_index(a, a);
  // Observation pairs per off-diagonal block (a > b; the diagonal blocks' sum over the camera's
  // own observations is reduced per camera instead). The list is padded with kRcsBatch pairs of
  // observation 0 so the assembly kernel's last batch stays inside it.
  std::vector<int32_t> pcount(n_blocks_ + 1, 0);
  for (int32_t p = 0; p < n_points_; ++p)
    for (int32_t x = pt_ptr[p]; x < pt_ptr[p + 1]; ++x)
      for (int32_t y = pt_ptr[p]; y < pt_ptr[p + 1]; ++y) {
        const int32_t a = obs_node[x], b = obs_node[y];
        if (a > b) pcount[block_index(a, b) + 1] += 1;
      }
  for (int32_t k = 0; k < n_blocks_; ++k) pcount[k + 1] += pcount[k];
  n_pairs_ = pcount[n_blocks_];
  std::vector<int32_t> pairs(2 * ((size_t)pcount[n_blocks_] + kRcsBatch), 0);
  {
    std::vector<int32_t> fill(pcount.begin(), pcount.end() - 1);
    for (int32_t p = 0; p < n_points_; ++p)
      for (int32_t x = pt_ptr[p]; x < pt_ptr[p + 1]; ++x)
        for (int32_t y = pt_ptr[p]; y < pt_ptr[p + 1]; ++y) {
          const int32_t a = obs_node[x], b = obs_node[y];
          if (a <= b) continue;
          const int32_t q = fill[block_index(a, b)]++;
          pairs[2 * (size_t)q] = x;
          pairs[2 * (size_t)q + 1] = y;
        }
  }
  // Assembly work list. Workgroups are dispatched round-robin over the XCDs (XCD = workgroup
  // index mod kXcds), each with its own L2: XCD x takes a contiguous range of camera rows holding
  // about 1/kXcds of the pairs; -1 pads the XCDs to equal length. At 1000 cameras, all items in
  // row order cut the L2 misses to a third of the largest-first order's but end 30% later (a
  // 512-pair item started last is the launch's tail); row order for the first 60% of the pairs
  // and the rest largest first in items of at most 256 pairs is 6% faster than largest first.
  std::vector<int4> rcs_items;
  {
    std::vector<int64_t> before(n_nodes_ + 1, 0);  // pairs (+1 per block) in rows < a
    for (int32_t k = 0; k < n_blocks_; ++k) before[blk_a[k] + 1] += pcount[k + 1] - pcount[k] + 1;
    for (int32_t a = 0; a < n_nodes_; ++a) before[a + 1] += before[a];
    std::vector<std::vector<int32_t>> per(kXcds);
    for (int32_t k = 0; k < n_blocks_; ++k)
      per[std::min<int64_t>(kXcds - 1, before[blk_a[k]] * kXcds / before[n_nodes_])].push_back(k);
    // Each XCD's blocks in row order (consecutive rows share most of their Z blocks, which then
    // stay in the XCD's L2), blocks with more than kRcsChunk pairs split into several items. The
    // blocks of the last kRcsTailFrac of an XCD's pairs are split finer (kRcsTailChunk) and
    // dealt heaviest first, so no long item starts at the end of the launch.
    std::vector<std::vector<int4>> its(kXcds);
    auto add_items = [&](std::vector<int4>& out, int32_t k, int32_t chunk) {
      const int32_t q0 = pcount[k], q1 = pcount[k + 1];
      const int32_t nc = std::max<int32_t>(1, (q1 - q0 + chunk - 1) / chunk);
      for (int32_t c = 0; c < nc; ++c)
        out.push_back(make_int4(k, q0 + c * chunk, std::min(q1, q0 + (c + 1) * chunk),
                                (c == 0 ? kRcsFirst : 0) | (nc > 1 ? kRcsAtomic : 0)));
    };
    for (int x = 0; x < kXcds; ++x) {
      int64_t total = 0, seen = 0;
      for (int32_t k : per[x]) total += pcount[k + 1] - pcount[k];
      std::vector<int4> tail;
      for (int32_t k : per[x]) {
        const bool in_tail = kRcsHeaviestFirst || seen >= (int64_t)((1.0 - kRcsTailFrac) * total);
        if (in_tail) add_items(tail, k, kRcsHeaviestFirst ? kRcsChunk : kRcsTailChunk);
        else add_items(its[x], k, kRcsChunk);
        seen += pcount[k + 1] - pcount[k];
      }
      std::stable_sort(tail.begin(), tail.end(),
                       [](const int4& a, const int4& b) { return a.z - a.y > b.z - b.y; });
      its[x].insert(its[x].end(), tail.begin(), tail.end());
    }
    size_t len = 0;
    for (auto& v : its) len = std::max(len, v.size());
    len = (len + 3) / 4 * 4;
    rcs_items.assign(kXcds * len, make_int4(-1, 0, 0, 0));
    for (int x = 0; x < kXcds; ++x)
      for (size_t i = 0; i < its[x].size(); ++i)
        rcs_items[((i / 4) * kXcds + x) * 4 + i % 4] = its[x][i];
  }
  // RCS symbolic: nodes = cameras (uniform nd), lower CSC pattern of the dense blocks.
  const int32_t nB = n_nodes_ * nd_;
  std::vector<int32_t> scp(nB + 1, 0), sri;
  {
    // column (b, j): rows (b, i>=j) then all rows of a > b in block column b
    std::vector<std::vector<int32_t>> col_rows(n_nodes_);
    for (int32_t k = 0; k < n_blocks_; ++k)
      if (blk_a[k] != blk_b[k]) col_rows[blk_b[k]].push_back(blk_a[k]);
    for (auto& v : col_rows) std::sort(v.begin(), v.end());
    for (int32_t b = 0; b < n_nodes_; ++b)
      for (int j = 0; j < nd_; ++j) {
        for (int i = j; i < nd_; ++i) sri.push_back(b * nd_ + i);
        for (int32_t a : col_rows[b])
          for (int i = 0; i < nd_; ++i) sri.push_back(a * nd_ + i);
        scp[b * nd_ + j + 1] = (int32_t)sri.size();
      }
  }
  std::vector<int32_t> node_ptr(n_nodes_ + 1);
  for (int32_t a = 0; a <= n_nodes_; ++a) node_ptr[a] = a * nd_;
  CholSymbolic sym;
  chol_symbolic(nB, scp.data(), sri.data(), node_ptr,
                ordering_from_code(o.ordering), nullptr, 64, sym);
  // panel position of each S block's (0,0) entry
  std::vector<int32_t> sn_of_col(nB);
  for (int32_t k = 0; k < sym.n_snodes; ++k)
    for (int32_t c = sym.sn_col_begin[k]; c < sym.sn_col_begin[k + 1]; ++c) sn_of_col[c] = k;
  std::vector<int64_t> pbase(n_blocks_);
  std::vector<int32_t> pld(n_blocks_), ptrans(n_blocks_);
  for (int32_t k = 0; k < n_blocks_; ++k) {
    int32_t ra = sym.perm[blk_a[k] * nd_], cb = sym.perm[blk_b[k] * nd_];
    const bool trans = ra < cb;
    if (trans) std::swap(ra, cb);
    const int32_t t = sn_of_col[cb];
    const int32_t c0 = sym.sn_col_begin[t], c1 = sym.sn_col_begin[t + 1];
    int32_t rp;
    if (ra < c1) {
      rp = ra - c0;
    } else {
      const int32_t* b0 = sym.sn_rows.data() + sym.sn_rows_ptr[t] + (c1 - c0);
      const int32_t* e0 = sym.sn_rows.data() + sym.sn_rows_ptr[t + 1];
      const int32_t* it = std::lower_bound(b0, e0, ra);
      SYMHIP_CHECK(it != e0 && *it == ra, "internal: S block outside the factor pattern");
      rp = (int32_t)(it - (sym.sn_rows.data() + sym.sn_rows_ptr[t]));
    }
    const int32_t ld = sym.sn_rows_ptr[t + 1] - sym.sn_rows_ptr[t];
    pbase[k] = sym.sn_panel_off[t] + (int64_t)(cb - c0) * ld + rp;
    pld[k] = ld;
    ptrans[k] = trans ? 1 : 0;
  }
  chol_.setup(sym, scp.data(), sri.data(), s);
  if (o.use_pcg) {
    SYMHIP_CHECK(o.pcg_max_iters >= 1 && o.pcg_rel_tol > 0 && o.pcg_rel_tol < 1,
                 "PCG needs pcg_max_iters >= 1 and 0 < pcg_rel_tol < 1");
    SYMHIP_CHECK(o.use_pcg == 1 || o.use_pcg == 2, "use_pcg must be 0, 1 (truncated) or 2 (strict)");
    use_pcg_ = true;
    pcg_strict_ = o.use_pcg == 2;
    SYMHIP_CHECK(o.pcg_exact_below_reduction >= 0, "pcg_exact_below_reduction must be >= 0");
    exact_below_ = o.pcg_exact_below_reduction;
    SYMHIP_CHECK(o.pcg_forcing >= 0, "pcg_forcing must be >= 0");
    forcing_ = o.pcg_forcing;
    pcg_max_iters_ = o.pcg_max_iters;
    pcg_rel_tol_ = o.pcg_rel_tol;
    // full symmetric block pattern: row a holds b for every lower block (a, b) and (b, a)
    std::vector<std::vector<int32_t>> full(n_nodes_);
    for (int32_t k = 0; k < n_blocks_; ++k) {
      full[blk_a[k]].push_back(blk_b[k]);
      if (blk_a[k] != blk_b[k]) full[blk_b[k]].push_back(blk_a[k]);
    }
    std::vector<int32_t> frp(n_nodes_ + 1, 0), fcol, fdiag(n_nodes_);
    for (int32_t a = 0; a < n_nodes_; ++a) {
      std::sort(full[a].begin(), full[a].end());
      frp[a + 1] = frp[a] + (int32_t)full[a].size();
      fcol.insert(fcol.end(), full[a].begin(), full[a].end());
    }
    auto full_index = [&](int32_t a, int32_t b) {
      const auto it = std::lower_bound(full[a].begin(), full[a].end(), b);
      return frp[a] + (int32_t)(it - full[a].begin());
    };
    std::vector<int32_t> flo(n_blocks_), fup(n_blocks_);
    for (int32_t k = 0; k < n_blocks_; ++k) {
      flo[k] = full_index(blk_a[k], blk_b[k]);
      fup[k] = full_index(blk_b[k], blk_a[k]);
    }
    for (int32_t a = 0; a < n_nodes_; ++a) fdiag[a] = full_index(a, a);
    pcg_.setup(nd_, n_nodes_, frp, fcol, fdiag, s, f32_);
    full_row_ptr_ = frp;
    full_col_ = fcol;
    SYMHIP_CHECK(o.pcg_preconditioner == 0 || o.pcg_preconditioner == 1,
                 "unknown pcg_preconditioner");
    if (nd_ == LearnedPrecond::kNd) gnn_.setup(n_nodes_, frp, fcol, fdiag, s);
    use_learned_ = o.pcg_preconditioner == 1;
    SYMHIP_CHECK(!use_learned_ || !f32_,
                 "the learned preconditioner works on FP64 systems (precision 0)");
    SYMHIP_CHECK(!use_learned_ || gnn_.ready(),
                 "the learned preconditioner needs 9-dof camera nodes");
    f_lo_.upload(flo, s);
    f_up_.upload(fup, s);
  }
  // Camera-only factors
  std::vector<std::vector<int32_t>> bf_by_type(kNumFactorTypes);
  for (int32_t f : b_factors) bf_by_type[P.ftype[f]].push_back(f);
  for (int t = 1; t < kNumFactorTypes; ++t) {
    if (bf_by_type[t].empty()) continue;
    const FactorSpec sp = factor_spec(t);
    const int32_t n = (int32_t)bf_by_type[t].size();
    std::vector<int32_t> ao((size_t)sp.nargs * n), ro(n), cm((size_t)sp.tdim * n),
        bo((size_t)sp.tdim * sp.tdim * n, 0);
    for (int32_t i = 0; i < n; ++i) {
      const int32_t f = bf_by_type[t][i];
      for (int a = 0; a < sp.nargs; ++a)
        ao[(size_t)a * n + i] = P.storage_offset[P.fargs[P.fargs_ptr[f] + a]];
      ro[i] = P.res_off[f];
      int c = 0;
      for (int32_t k : fpos[f]) {
        const int td = tangent_dim(P.types[P.state_keys[k]]);
        for (int d = 0; d < td; ++d, ++c)
          cm[(size_t)c * n + i] = node_of_pos[k] * nd_ + pos_off_in_node[k] + d;
      }
      for (int c1 = 0; c1 < sp.tdim; ++c1)
        for (int c2 = 0; c2 < sp.tdim; ++c2) {
          const int32_t na = cm[(size_t)c1 * n + i] / nd_, nb = cm[(size_t)c2 * n + i] / nd_;
          if (na >= nb) bo[((size_t)c1 * sp.tdim + c2) * n + i] = block_index(na, nb);
        }
    }
    BGroup g;
    g.type = t;
    g.n = n;
    g.arg_off.upload(ao, s);
    g.res_off.upload(ro, s);
    g.col_map.upload(cm, s);
    g.blk_of.upload(bo, s);
    bgroups_.push_back(std::move(g));
  }
  obs_arg_off_.upload(arg_off, s);
  obs_res_off_.upload(res_off, s);
  obs_node_.upload(obs_node, s);
  obs_point_.upload(obs_point, s);
  pt_obs_ptr_.upload(pt_ptr, s);
  cam_obs_ptr_.upload(cam_ptr, s);
  cam_obs_.upload(cam_obs, s);
  node_tan_.upload(node_tan, s);
  pt_tan_.upload(pt_tan, s);
  blk_a_.upload(blk_a, s);
  blk_b_.upload(blk_b, s);
  pair_ptr_.upload(pcount, s);
  pairs_.upload(pairs, s);
  n_rcs_items_ = (int32_t)rcs_items.size();
  rcs_items_.resize(rcs_items.size() * sizeof(int4));
  HIP_OK(hipMemcpyAsync(rcs_items_.get(), rcs_items.data(), rcs_items.size() * sizeof(int4),
                        hipMemcpyHostToDevice, s));
  diag_blk_.upload(diag_blk, s);
  panel_base_.upload(pbase, s);
  panel_ld_.upload(pld, s);
  panel_trans_.upload(ptrans, s);
  const int64_t nde = (int64_t)nd_ * nd_;
  for (int b = 0; b < 3; ++b) {
    residual_[b].resize(std::max(1, M_));
    W_[b].resize((size_t)n_obs_ * nd_ * 3);
    if (f32_) W32_[b].resize((size_t)n_obs_ * nd_ * 3);
    V_[b].resize((size_t)n_points_ * 6);
    gp_[b].resize((size_t)n_points_ * 3);
    Bdiag_[b].resize((size_t)n_nodes_ * nde);
    Bfac_[b].resize((size_t)n_blocks_ * nde);
    gB_[b].resize((size_t)n_nodes_ * nd_);
    rhs_[b].resize(std::max(1, N_));
  }
  damping_.resize(std::max(1, N_));
  undamped_.resize(std::max(1, N_));
  max_diag_.resize(std::max(1, N_));
  dnode_.resize((size_t)n_nodes_ * nd_);
  dpoint_.resize((size_t)n_points_ * 3);
  Lp_.resize((size_t)n_points_ * 6);
  hp_.resize((size_t)n_points_ * 3);
  {
    // Z rows padded to 16-byte multiples (z_stride); the pad entries are zeroed here and never
    // written again
    const size_t zs64 = nd_ == 9 ? z_stride<9, double>() : z_stride<6, double>();
    const size_t zs32 = nd_ == 9 ? z_stride<9, float>() : z_stride<6, float>();
    Z_.resize((size_t)n_obs_ * zs64);
    fill_zero(Z_.get(), (int64_t)Z_.size(), s);
    if (f32_) {
      Z32_.resize((size_t)n_obs_ * zs32);
      fill_zero(Z32_.get(), (int64_t)Z32_.size(), s);
    }
  }
  ZZ_.resize((size_t)n_nodes_ * nde);
  v_.resize((size_t)n_nodes_ * nd_);
  y_.resize((size_t)n_nodes_ * nd_);
  partials_.resize(kReduceBlocks);
  HIP_OK(hipStreamSynchronize(s));
}

template <int ND>
void SchurBackend::relin_nd(int b, const double* values, hipStream_t s) {
  const ObsView ov{n_obs_, obs_arg_off_.get(), obs_res_off_.get(), obs_node_.get()};
  dispatch_factor_type(ftype_, [&]<int C>() {
    if constexpr (FT<C>::T == ND + 3 && FT<C>::NL >= 2) {
      {
        PhaseScope ps(timer_, SYMHIP_PHASE_LIN_POINTS, s);
        fill_zero(V_[b].get(), (int64_t)n_points_ * 6, s);
        fill_zero(gp_[b].get(), (int64_t)n_points_ * 3, s);
        w_f32_[b] = f32_ && !exact_now_;
        with_storage(b, [&](auto* W, auto*) {
          using TW = std::remove_pointer_t<decltype(W)>;
          hipLaunchKernelGGL((lin_points_kernel<C, ND, TW>), dim3(grid_for(n_obs_, 128)),
                             dim3(128), 0, s, n_obs_, obs_point_.get(), ov, values,
                             residual_[b].get(), W, V_[b].get(), gp_[b].get());
        });
      }
      PhaseScope ps(timer_, SYMHIP_PHASE_LIN_CAMS, s);
      hipLaunchKernelGGL((lin_cams_kernel<C, ND>), dim3(xcd_grid(n_nodes_)),
                         dim3(256), 0, s, n_nodes_, cam_obs_ptr_.get(), cam_obs_.get(), ov, values,
                         Bdiag_[b].get(), gB_[b].get());
    }
  });
  fill_zero(Bfac_[b].get(), (int64_t)n_blocks_ * ND * ND, s);
  for (const BGroup& g : bgroups_) {
    dispatch_factor_type(g.type, [&]<int C>() {
      hipLaunchKernelGGL((lin_bfac_kernel<C, ND>), dim3(grid_for(g.n, 64)), dim3(64), 0, s, g.n,
                         g.arg_off.get(), g.res_off.get(), g.col_map.get(), g.blk_of.get(), values,
                         residual_[b].get(), Bfac_[b].get(), gB_[b].get());
    });
  }
  // rhs in reference order
  hipLaunchKernelGGL(scatter_kernel, dim3(grid_for(n_nodes_ * ND, 256)), dim3(256), 0, s,
                     n_nodes_ * ND, node_tan_.get(), gB_[b].get(), rhs_[b].get());
  hipLaunchKernelGGL(scatter_kernel, dim3(grid_for(n_points_ * 3, 256)), dim3(256), 0, s,
                     n_points_ * 3, pt_tan_.get(), gp_[b].get(), rhs_[b].get());
  HIP_OK(hipGetLastError());
}

void SchurBackend::relinearize(int b, const double* values, double* err_out, hipStream_t s) {
  if (nd_ == 9) relin_nd<9>(b, values, s); else relin_nd<6>(b, values, s);
  half_squared_norm(residual_[b].get(), M_, partials_.get(), err_out, s);
}

void SchurBackend::damp(int b, const double* lambda_dev, const symhip_optimizer_params& p,
                        bool have_max, hipStream_t s) {
  const int32_t n = n_nodes_ * nd_ + 3 * n_points_;
  auto launch = [&](auto kernel) {
    hipLaunchKernelGGL(kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, n_nodes_, n_points_,
                       Bdiag_[b].get(), Bfac_[b].get(), diag_blk_.get(), V_[b].get(),
                       node_tan_.get(), pt_tan_.get(), lambda_dev, p.use_diagonal_damping,
                       p.keep_max_diagonal_damping, have_max ? 1 : 0, p.diagonal_damping_min,
                       p.use_unit_damping, max_diag_.get(), damping_.get(), undamped_.get(),
                       dnode_.get(), dpoint_.get());
  };
  if (nd_ == 9) launch(damp_schur_kernel<9>); else launch(damp_schur_kernel<6>);
}

template <int ND>
void SchurBackend::solve_nd(int b, double* update, hipStream_t s) {
  {
    PhaseScope ps(timer_, SYMHIP_PHASE_POINT_ELIM, s);
    hipLaunchKernelGGL(point_factor_kernel, dim3(grid_for(n_points_, 256)), dim3(256), 0, s,
                       n_points_, V_[b].get(), dpoint_.get(), gp_[b].get(), Lp_.get(), hp_.get());
    with_storage(b, [&](auto* W, auto* Z) {
      using TW = std::remove_pointer_t<decltype(W)>;
      hipLaunchKernelGGL((point_elim_kernel<ND, TW>), dim3(grid_for((int64_t)n_obs_ * ND, 256)),
                         dim3(256), 0, s, n_obs_, obs_point_.get(), Lp_.get(), W, Z);
    });
  }
  auto assemble = [&](bool full) {
    // split blocks accumulate atomically: the target storage starts at zero
    const int64_t nfull = pcg_.n_blocks() * ND * ND;
    if (!full) fill_zero(chol_.panels(), chol_.symbolic().panel_size(), s);
    else if (f32_) fill_zero(pcg_.blocks32(), nfull, s);
    else fill_zero(pcg_.blocks(), nfull, s);
    PhaseScope ps(timer_, SYMHIP_PHASE_RCS_ASSEMBLE, s);  // exactly the one assembly kernel
    const int32_t n_items = n_rcs_items_;
    auto launch = [&](auto* Z, auto* sfull) {
      using TW = std::remove_cv_t<std::remove_pointer_t<decltype(Z)>>;
      using TS = std::remove_pointer_t<decltype(sfull)>;
      hipLaunchKernelGGL((rcs_assemble_kernel<ND, TW, TS>), dim3(n_items / 4), dim3(256), 0, s,
                         n_items, reinterpret_cast<const int4*>(rcs_items_.get()), blk_a_.get(),
                         blk_b_.get(), pairs_.get(), Z, ZZ_.get(), Bdiag_[b].get(),
                         Bfac_[b].get(), dnode_.get(),
                         panel_base_.get(), panel_ld_.get(), panel_trans_.get(),
                         chol_.panels(), sfull, f_lo_.get(), f_up_.get());
    };
    // Z in the precision its state block was linearized in; S for PCG in the form's precision
    with_storage(b, [&](auto*, auto* Z) {
      if (full && f32_) launch(Z, pcg_.blocks32());
      else launch(Z, full ? pcg_.blocks() : static_cast<double*>(nullptr));
    });
  };
  auto direct = [&]() {
    {
      // the forward solve runs inside the factorization (v is final by now)
      PhaseScope ps(timer_, SYMHIP_PHASE_FACTOR, s);
      chol_.factorize_assembled(s, v_.get());
    }
    PhaseScope ps(timer_, SYMHIP_PHASE_SOLVE, s);
    chol_.solve_fused(y_.get(), s);
  };
  const bool pcg_now = use_pcg_ && !exact_now_;
  {
    PhaseScope ps(timer_, SYMHIP_PHASE_REDUCED_RHS, s);
    with_storage(b, [&](auto*, auto* Z) {
      using TW = std::remove_pointer_t<decltype(Z)>;
      hipLaunchKernelGGL((reduced_rhs_kernel<ND, TW>), dim3(xcd_grid(n_nodes_)),
                         dim3(256), 0, s, n_nodes_, cam_obs_ptr_.get(), cam_obs_.get(),
                         obs_point_.get(), Z, hp_.get(), gB_[b].get(), v_.get(), ZZ_.get());
    });
  }
  assemble(pcg_now);
  if (pcg_now && (int)logged_S_.size() < log_max_) {
    DBuf<double> Sc, vc;
    Sc.resize((size_t)pcg_.n_blocks() * ND * ND);
    vc.resize((size_t)n_nodes_ * ND);
    HIP_OK(hipMemcpyAsync(Sc.get(), pcg_.blocks(), Sc.size() * sizeof(double),
                          hipMemcpyDeviceToDevice, s));
    HIP_OK(hipMemcpyAsync(vc.get(), v_.get(), vc.size() * sizeof(double), hipMemcpyDeviceToDevice,
                          s));
    logged_S_.push_back(std::move(Sc));
    logged_v_.push_back(std::move(vc));
  }
  if (pcg_now) {
    PcgResult res;
    {
      PhaseScope ps(timer_, SYMHIP_PHASE_SOLVE, s);
      if (use_learned_)
        gnn_.preconditioner(pcg_.blocks(), pcg_.precond(), s);
      else
        pcg_.build_preconditioner(nullptr, s);
      // inexact Newton: the tolerance follows the last accepted step's relative reduction
      double tol = pcg_rel_tol_;
      int max_iters = pcg_max_iters_;
      if (forcing_ > 0 && last_reduction_ >= 0) {
        tol = std::clamp(forcing_ * std::sqrt(last_reduction_), 1e-10, pcg_rel_tol_);
        max_iters = (int)std::min(
            20.0 * pcg_max_iters_,
            std::ceil(pcg_max_iters_ * (1.0 + 2.0 * std::log10(pcg_rel_tol_ / tol))));
      }
      res = pcg_.solve(v_.get(), y_.get(), max_iters, tol, s);
    }
    ++pcg_solves_;
    pcg_iters_ += res.iterations;
    pcg_max_seen_ = std::max<int64_t>(pcg_max_seen_, res.iterations);
    // Truncated mode (use_pcg == 1) keeps a CG iterate that did not reach the tolerance: every
    // CG iterate lowers the quadratic model (the residual 2-norm need not fall monotonically),
    // and the LM gain ratio judges the step. The direct LDL^T takes over when CG breaks down
    // (S not SPD to working precision, non-finite sums) or, in strict mode (use_pcg == 2),
    // whenever the tolerance was not reached.
    const bool stalled = res.breakdown;
    if (stalled || (pcg_strict_ && !res.converged)) {
      ++pcg_fallbacks_;
      assemble(false);
      direct();
    }
  } else {
    direct();
  }
  PhaseScope ps(timer_, SYMHIP_PHASE_BACK_SUB, s);
  hipLaunchKernelGGL(scatter_neg_kernel, dim3(grid_for(n_nodes_ * ND, 256)), dim3(256), 0, s,
                     n_nodes_ * ND, node_tan_.get(), y_.get(), update);
  with_storage(b, [&](auto*, auto* Z) {
    using TW = std::remove_pointer_t<decltype(Z)>;
    hipLaunchKernelGGL((back_substitute_kernel<ND, TW>), dim3(grid_for(n_points_, 256)),
                       dim3(256), 0, s, n_points_, pt_obs_ptr_.get(), obs_node_.get(), Z,
                       Lp_.get(), hp_.get(), y_.get(), pt_tan_.get(), update);
  });
  HIP_OK(hipGetLastError());
}

void SchurBackend::solve(int b, double* update, hipStream_t s) {
  if (nd_ == 9) solve_nd<9>(b, update, s); else solve_nd<6>(b, update, s);
}

void SchurBackend::hessian_csc(int b, double* out, hipStream_t s) {
  (void)b;
  (void)out;
  (void)s;
  throw Error("symhip: reference-layout Hessian output is not available on the Schur path");
}

namespace {
// Full symmetric block storage -> LDL^T panels (the lower blocks, lower half of diagonal ones).
template <int ND>
__global__ void full_to_panels_kernel(int32_t n_blocks, const int32_t* __restrict__ blk_a,
                                      const int32_t* __restrict__ blk_b,
                                      const int32_t* __restrict__ f_lo,
                                      const int64_t* __restrict__ panel_base,
                                      const int32_t* __restrict__ panel_ld,
                                      const int32_t* __restrict__ panel_trans,
                                      const double* __restrict__ S, double* __restrict__ panel) {
  constexpr int NE = ND * ND;
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= (int64_t)n_blocks * NE) return;
  const int blk = (int)(t / NE), e = (int)(t % NE), i = e / ND, j = e % ND;
  if (blk_a[blk] == blk_b[blk] && i < j) return;
  const int64_t base = panel_base[blk];
  const int ld = panel_ld[blk];
  const int64_t pos = panel_trans[blk] ? base + (int64_t)i * ld + j : base + (int64_t)j * ld + i;
  panel[pos] = S[(int64_t)f_lo[blk] * NE + e];
}
}  // namespace

template <int ND>
void SchurBackend::panels_from_full(const double* S, hipStream_t s) {
  fill_zero(chol_.panels(), chol_.symbolic().panel_size(), s);
  hipLaunchKernelGGL((full_to_panels_kernel<ND>), dim3(grid_for((int64_t)n_blocks_ * ND * ND, 256)),
                     dim3(256), 0, s, n_blocks_, blk_a_.get(), blk_b_.get(), f_lo_.get(),
                     panel_base_.get(), panel_ld_.get(), panel_trans_.get(), S, chol_.panels());
  HIP_OK(hipGetLastError());
}

void SchurBackend::log_systems(int max) {
  SYMHIP_CHECK(use_pcg_, "logging reduced camera systems needs use_pcg");
  SYMHIP_CHECK(!f32_, "logging reduced camera systems needs FP64 systems (precision 0)");
  SYMHIP_CHECK(max >= 0, "max must be non-negative");
  log_max_ = max;
  logged_S_.clear();
  logged_v_.clear();
}

int SchurBackend::rcs_pattern(int32_t* row_ptr, int32_t* col) const {
  SYMHIP_CHECK(use_pcg_, "the block pattern is built with use_pcg");
  if (row_ptr) std::copy(full_row_ptr_.begin(), full_row_ptr_.end(), row_ptr);
  if (col) std::copy(full_col_.begin(), full_col_.end(), col);
  return n_nodes_;
}

void SchurBackend::logged_system(int idx, double* blocks, double* rhs) const {
  SYMHIP_CHECK(idx >= 0 && idx < (int)logged_S_.size(), "no such logged system");
  hipStream_t s = stream_;
  if (blocks) logged_S_[idx].download(blocks, logged_S_[idx].size(), s);
  if (rhs) logged_v_[idx].download(rhs, logged_v_[idx].size(), s);
  HIP_OK(hipStreamSynchronize(s));
}

void SchurBackend::precond_params(float* get, const float* set) {
  SYMHIP_CHECK(gnn_.ready(), "the learned preconditioner needs the Schur solver with use_pcg "
                             "and 9-dof camera nodes");
  if (set) gnn_.set_params(set, stream_);
  if (get) gnn_.get_params(get, stream_);
}

float SchurBackend::precond_loss_grad(int idx, const double* X, int P, float* grad) {
  SYMHIP_CHECK(gnn_.ready(), "learned preconditioner unavailable");
  SYMHIP_CHECK(idx >= 0 && idx < (int)logged_S_.size(), "no such logged system");
  SYMHIP_CHECK(P >= 1 && P <= 7, "1..7 probe vectors");
  hipStream_t s = stream_;
  const int64_t n9 = (int64_t)n_nodes_ * nd_;
  probe_x_.upload(X, (size_t)n9 * P, s);
  probe_y_.resize((size_t)n9 * P);
  for (int p = 0; p < P; ++p)
    pcg_.multiply_with(logged_S_[idx].get(), probe_x_.get() + p * n9, probe_y_.get() + p * n9, s);
  const float loss = gnn_.loss_grad(logged_S_[idx].get(), probe_x_.get(), probe_y_.get(), P, s);
  if (grad) {
    HIP_OK(hipMemcpyAsync(grad, gnn_.grad(), sizeof(float) * LearnedPrecond::kParams,
                          hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
  }
  return loss;
}

void SchurBackend::train_precond(int steps, float lr, int P, uint64_t seed, int n_systems,
                                 float* losses) {
  SYMHIP_CHECK(gnn_.ready(), "learned preconditioner unavailable");
  SYMHIP_CHECK(!logged_S_.empty(), "no logged systems to train on");
  const size_t nsys = n_systems > 0 ? std::min<size_t>(n_systems, logged_S_.size())
                                    : logged_S_.size();
  SYMHIP_CHECK(steps >= 1 && P >= 1 && P <= 7 && lr > 0, "bad training arguments");
  hipStream_t s = stream_;
  const int64_t n9 = (int64_t)n_nodes_ * nd_;
  probe_x_.resize((size_t)n9 * P);
  probe_y_.resize((size_t)n9 * P);
  for (int k = 0; k < steps; ++k) {
    const DBuf<double>& S = logged_S_[k % nsys];
    LearnedPrecond::probes(probe_x_.get(), n9 * P, seed * 1000003ull + (uint64_t)k, s);
    for (int p = 0; p < P; ++p)
      pcg_.multiply_with(S.get(), probe_x_.get() + p * n9, probe_y_.get() + p * n9, s);
    const float loss = gnn_.loss_grad(S.get(), probe_x_.get(), probe_y_.get(), P, s);
    if (losses) losses[k] = loss;
    gnn_.adam(lr, k + 1, s);
  }
  HIP_OK(hipStreamSynchronize(s));
}

void SchurBackend::eval_precond(int idx, double rel_tol, int max_iters, int learned, int* iters,
                                double* step_err) {
  SYMHIP_CHECK(idx >= 0 && idx < (int)logged_S_.size(), "no such logged system");
  SYMHIP_CHECK(!learned || gnn_.ready(), "learned preconditioner unavailable");
  hipStream_t s = stream_;
  const int64_t n9 = (int64_t)n_nodes_ * nd_;
  const DBuf<double>& S = logged_S_[idx];
  const double* v = logged_v_[idx].get();
  HIP_OK(hipMemcpyAsync(pcg_.blocks(), S.get(), S.size() * sizeof(double), hipMemcpyDeviceToDevice,
                        s));
  if (learned)
    gnn_.preconditioner(pcg_.blocks(), pcg_.precond(), s);
  else
    pcg_.build_preconditioner(nullptr, s);
  const PcgResult res = pcg_.solve(v, y_.get(), max_iters, rel_tol, s);
  if (iters) *iters = res.converged ? res.iterations : -res.iterations;
  if (step_err) {
    exact_y_.resize((size_t)n9);
    if (nd_ == 9) panels_from_full<9>(S.get(), s); else panels_from_full<6>(S.get(), s);
    chol_.factorize_assembled(s);
    chol_.solve(v, exact_y_.get(), s);
    // relative step error |y - y_exact| / |y_exact| of this evaluation call
    std::vector<double> a(n9), b(n9);
    y_.download(a.data(), n9, s);
    exact_y_.download(b.data(), n9, s);
    HIP_OK(hipStreamSynchronize(s));
    double num = 0, den = 0;
    for (int64_t i = 0; i < n9; ++i) {
      num += (a[i] - b[i]) * (a[i] - b[i]);
      den += b[i] * b[i];
    }
    *step_err = den > 0 ? std::sqrt(num / den) : 0.0;
  }
}

std::unique_ptr<SystemBackend> make_schur_backend(const HostProblem& p, const SolverOptions& o,
                                                  hipStream_t s) {
  return std::make_unique<SchurBackend>(p, o, s);
}

}  // namespace symhip
