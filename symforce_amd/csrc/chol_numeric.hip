// Device numeric supernodal LDL^T: assembly, per-level fused diagonal-block factorization + panel
// triangular solve, right-looking update (SYRK-like) scattered into ancestor panels, and the
// triangular solves. All FP64.
//
// The work per supernode is small (panels are at most 64 columns wide), so every kernel is
// latency-bound: the kernels are written to avoid data-dependent branches (the 64 elimination
// steps are fully unrolled over a panel padded to 64 columns, the padding being an identity
// block) and chains of dependent global loads (each supernode's metadata is one descriptor,
// scatter targets are precomputed per below-row).
#include <algorithm>

#include "./chol_numeric.h"
#include "./util.h"

namespace symhip {

namespace {

constexpr int kMaxCols = 64;  // supernode width cap (panels padded to this width)
constexpr int kTile = 64;     // update tile (rows)

// l11_off: the factored diagonal block L11 (n x n, column-major, strictly lower part used) is
// kept apart from the panel, so the row-chunk blocks of one supernode, which all read the
// unfactored A11 from the panel, never see it half overwritten.
struct SnDesc {
  int64_t panel_off, l11_off;
  int32_t ncols, ld, col_begin, rows_off;
};

__global__ void assemble_kernel(int64_t nnz, const int64_t* __restrict__ map,
                                const double* __restrict__ vals, double* __restrict__ panel) {
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < nnz;
       q += (int64_t)gridDim.x * blockDim.x)
    panel[map[q]] = vals[q];
}

// Right-looking blocked LDL^T of a 64-column panel held as 4 x 4 register patches: rows
// 0..63 are the diagonal block (padded with identity), rows 64..16 RX - 1 are below rows that are
// eliminated in the same sweep (their final values are L21 = A21 L11^-T D^-1). Thread (ty, tx)
// owns rows ty + 16x (x < RX) and cols tx + 16y (y < 4). Step m eliminates columns
// k0 = 4m .. k0 + 3 (register column kb = k0 / 16, owner threads tx = k0 % 16 + u): the owners
// publish the four columns through LDS (one barrier per step, double-buffered), every thread
// factors the 4 x 4 diagonal block redundantly, forms w = c L4^-T for its rows and columns
// (c = the published entries, L(i, k0 + u) = w_u / d_u) and applies
// A(i, j) -= sum_u w_iu w_ju / d_u. The finished columns go to Lout (LDS; D on the diagonal),
// written by the tx == 0 thread of each row. Register patches that are dead for the rest of the
// sweep are skipped at compile time (kb is unrolled): columns y < kb (eliminated) and diagonal
// rows x < kb (above the current block). Within the live patches nothing is masked: entries above
// the diagonal and columns already eliminated pick up products of other such entries, but they
// are never read again (the final values are in Lout), and the entries still to be eliminated
// only ever combine published lower-triangle values.
template <int W>
__device__ __forceinline__ void ldlt4_block(const double (*cb)[W], int k0, double (&d)[4],
                                            double (&dinv)[4], double (&L4)[4][4]) {
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    double dd = cb[u][k0 + u];
#pragma unroll
    for (int v = 0; v < u; ++v) dd -= L4[u][v] * L4[u][v] * d[v];
    d[u] = dd;
    // reciprocal by the hardware estimate and two Newton steps (the IEEE division sequence is
    // twice as long, and these four sit on the sweep's critical path)
    double r0 = __builtin_amdgcn_rcp(dd);
    r0 = fma(r0, fma(-dd, r0, 1.0), r0);
    dinv[u] = fma(r0, fma(-dd, r0, 1.0), r0);
#pragma unroll
    for (int r = u + 1; r < 4; ++r) {
      double x = cb[u][k0 + r];
#pragma unroll
      for (int v = 0; v < u; ++v) x -= L4[r][v] * L4[u][v] * d[v];
      L4[r][u] = x * dinv[u];
    }
  }
}

constexpr int kLoutLd = kMaxCols + 2;  // Lout row stride (doubles): 16-byte rows, fewer conflicts

template <int RX>
__device__ __forceinline__ void ldlt_panel_registers(double (&a)[RX][4],
                                                     double (*cb)[4][16 * RX],
                                                     double (*Lout)[kLoutLd], int ty, int tx) {
#pragma unroll
  for (int kb = 0; kb < 4; ++kb) {
#pragma unroll 1
    for (int mm = 0; mm < 4; ++mm) {
      const int k0 = 16 * kb + 4 * mm, buf = mm & 1, uo = tx - 4 * mm;
      auto row_live = [&](int x) { return x >= 4 || x >= kb; };
      if (uo >= 0 && uo < 4) {
#pragma unroll
        for (int x = 0; x < RX; ++x)
          if (row_live(x)) cb[buf][uo][ty + 16 * x] = a[x][kb];
      }
      __syncthreads();
      double d[4], dinv[4], L4[4][4];
      ldlt4_block<16 * RX>(cb[buf], k0, d, dinv, L4);
      double wr[RX][4], wc[4][4];
#pragma unroll
      for (int x = 0; x < RX; ++x) {
        if (!row_live(x)) continue;
        const int i = ty + 16 * x;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          double c = cb[buf][u][i];
#pragma unroll
          for (int v = 0; v < u; ++v) c -= wr[x][v] * L4[u][v];
          wr[x][u] = c;
        }
      }
#pragma unroll
      for (int y = 0; y < 4; ++y) {
        if (y < kb) continue;
        const int j = tx + 16 * y;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          double c = cb[buf][u][j];
#pragma unroll
          for (int v = 0; v < u; ++v) c -= wc[y][v] * L4[u][v];
          wc[y][u] = c;
        }
      }
      // wr <- L(i, k0 + u) = w_iu / d_u
#pragma unroll
      for (int x = 0; x < RX; ++x)
        if (row_live(x))
#pragma unroll
          for (int u = 0; u < 4; ++u) wr[x][u] *= dinv[u];
#pragma unroll
      for (int x = 0; x < RX; ++x)
#pragma unroll
        for (int y = 0; y < 4; ++y) {
          if (!row_live(x) || y < kb) continue;
#pragma unroll
          for (int u = 0; u < 4; ++u) a[x][y] -= wr[x][u] * wc[y][u];
        }
      if (tx == 0) {
        // rows below the 4 x 4 block (row groups x > kb, and x == kb past it); the block itself
        // (rows k0 .. k0 + 3, all in group kb) gets D and L4
#pragma unroll
        for (int x = kb; x < RX; ++x) {
          const int i = ty + 16 * x;
          if (x == kb && i < k0 + 4) {
            if (i >= k0) {
              const int r = i - k0;
#pragma unroll
              for (int u = 0; u < 4; ++u) {
                double v = 0.0;
#pragma unroll
                for (int rr = 0; rr < 4; ++rr)
                  if (rr == r) v = u < rr ? L4[rr][u] : (u == rr ? d[u] : 0.0);
                Lout[i][k0 + u] = v;
              }
            }
          } else {
#pragma unroll
            for (int u = 0; u < 4; ++u) Lout[i][k0 + u] = wr[x][u];
          }
        }
      }
    }
  }
}

// LDS (doubles) of the two factorization kernels.
constexpr int kDiagRX = 8;  // row groups of the diagonal task: 4 of the diagonal block + 4 below
constexpr int kDiagSmem = 2 * 4 * 16 * kDiagRX + 16 * kDiagRX * kLoutLd;
constexpr int kUpdSmem = 2 * kMaxCols * (kTile + 1);

// Diagonal task (s, chunk), one 256-thread block: eliminates the 64 x 64 diagonal block of
// supernode s together with the chunk's 64 below rows in one right-looking sweep (the diagonal
// block is factored redundantly by every chunk of s; chunk 0 writes L11 to l11 and D), leaving
// L21 = A21 L11^-T D^-1 for its rows, written back into the panel in place.
__device__ __forceinline__ void diag_body(int s, int chunk, const SnDesc* __restrict__ desc,
                                          double* __restrict__ panel, double* __restrict__ l11,
                                          double* __restrict__ D, double* smem) {
  constexpr int RX = kDiagRX;
  auto cb = reinterpret_cast<double (*)[4][16 * RX]>(smem);
  auto Lout = reinterpret_cast<double (*)[kLoutLd]>(smem + 2 * 4 * 16 * RX);
  const SnDesc sd = desc[s];
  const int n = sd.ncols, ld = sd.ld, mb = ld - n;
  double* P = panel + sd.panel_off;
  const int ty = threadIdx.x >> 4, tx = threadIdx.x & 15;
  double a[RX][4];
  // Unconditional loads from clamped (valid) addresses, then selects: a guarded load would be
  // compiled as a branch and wait per element.
#pragma unroll
  for (int x = 0; x < 4; ++x)
#pragma unroll
    for (int y = 0; y < 4; ++y) {
      const int i = ty + 16 * x, j = tx + 16 * y;
      const int ic = min(i, n - 1), jc = min(j, n - 1);
      const double va = P[(int64_t)jc * ld + ic];
      const int r = chunk * 64 + i;
      const double vw = P[(int64_t)jc * ld + (mb > 0 ? n + min(r, mb - 1) : ic)];
      a[x][y] = (i < n && j < n && i >= j) ? va : (i == j ? 1.0 : 0.0);
      a[x + 4][y] = (r < mb && j < n) ? vw : 0.0;
    }
  ldlt_panel_registers<RX>(a, cb, Lout, ty, tx);
  __syncthreads();
  if (chunk == 0) {
#pragma unroll
    for (int x = 0; x < 4; ++x)
#pragma unroll
      for (int y = 0; y < 4; ++y) {
        const int i = ty + 16 * x, j = tx + 16 * y;
        const double v = Lout[i][j];
        if (i < n && j < n && i > j) l11[sd.l11_off + (int64_t)j * n + i] = v;
        if (i == j && i < n) D[sd.col_begin + i] = v;
      }
  }
#pragma unroll
  for (int x = 0; x < 4; ++x)
#pragma unroll
    for (int y = 0; y < 4; ++y) {
      const int r = chunk * 64 + ty + 16 * x, j = tx + 16 * y;
      const double v = Lout[64 + ty + 16 * x][j];
      if (r < mb && j < n) P[(int64_t)j * ld + n + r] = v;
    }
}

__global__ void __launch_bounds__(256) diag_trsm_kernel(const int32_t* __restrict__ items,
                                                        const SnDesc* __restrict__ desc,
                                                        double* __restrict__ panel,
                                                        double* __restrict__ l11,
                                                        double* __restrict__ D) {
  __shared__ double smem[kDiagSmem];
  diag_body(items[2 * blockIdx.x], items[2 * blockIdx.x + 1], desc, panel, l11, D, smem);
}

// Update task (s, ti, tj), one 256-thread block: U = L21 D L21^T tile (ti, tj), scattered as -U
// into ancestor panels. For below-row j of the supernode, scat_col[j] is the panel offset of its
// column in the target supernode and scat_rel[j] + i indexes upd_rel for the target row position
// of below-row i >= j (i < mb).
__device__ __forceinline__ void update_body(int s, int ti, int tj,
                                            const SnDesc* __restrict__ desc,
                                            double* __restrict__ panel,
                                            const double* __restrict__ D,
                                            const int64_t* __restrict__ scat_col,
                                            const int32_t* __restrict__ scat_rel,
                                            const int32_t* __restrict__ upd_rel, double* smem) {
  auto As = reinterpret_cast<double (*)[kTile + 1]>(smem);
  auto Bs = reinterpret_cast<double (*)[kTile + 1]>(smem + kMaxCols * (kTile + 1));
  const SnDesc sd = desc[s];
  const int n = sd.ncols, ld = sd.ld, mb = ld - n;
  const double* P = panel + sd.panel_off;
  const double* Ds = D + sd.col_begin;
  // stage both row tiles (all k at once: the loads are independent)
  // (unconditional loads from clamped addresses; an update item exists only when mb > 0)
#pragma unroll
  for (int t = 0; t < kMaxCols * kTile / 256; ++t) {  // blockDim.x == 256
    const int idx = threadIdx.x + 256 * t;
    const int k = idx / kTile, r = idx % kTile;
    const int ri = ti * kTile + r, rj = tj * kTile + r;
    const int kc = min(k, n - 1);
    const double va = P[(int64_t)kc * ld + n + min(ri, mb - 1)];
    const double vb = P[(int64_t)kc * ld + n + min(rj, mb - 1)] * Ds[kc];
    As[k][r] = (k < n && ri < mb) ? va : 0.0;
    Bs[k][r] = (k < n && rj < mb) ? vb : 0.0;
  }
  // The product on the FP64 matrix cores: wave w computes rows 16w .. 16w + 15 of the tile as
  // four 16 x 16 output blocks c (columns 16c ..), K = 64 in 16 steps of v_mfma_f64_16x16x4f64.
  // Lane l feeds A(i = l % 16, k = l / 16) = As[k][16w + i] and B(k, j = l % 16) = Bs[k][16c + j]
  // and holds D(l / 16 + 4 r, l % 16) of block c in accumulator register r: 5 LDS reads per 4
  // MFMAs (the FMA form needed 8 per 16 FMAs and was bound by them).
  typedef double d4 __attribute__((ext_vector_type(4)));
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // uniform: scalar metadata
  const int lr = lane & 15, lk = lane >> 4;
  // Scatter targets, fetched while the tiles load and before the product loop (the row
  // positions depend on the metadata: two dependent loads, off the critical path here). The
  // scatter runs column by column: wave w takes tile columns w, w + 4, ..., w + 60 and lane l row
  // l, so one atomic instruction covers 64 consecutive rows of one target column (a few cache
  // lines) instead of 16 columns x 4 rows. Row indices are clamped into the column's update
  // segment (rows j .. mb - 1); clamped entries are masked at the scatter.
  constexpr int kCols = kTile / 4;
  int64_t cbase[kCols];
  int32_t rel[kCols];
  {
    const int i = ti * kTile + lane;
#pragma unroll
    for (int q = 0; q < kCols; ++q) {
      const int j = min(tj * kTile + wave + 4 * q, mb - 1);
      cbase[q] = scat_col[sd.rows_off + n + j];
      const int32_t rbase = scat_rel[sd.rows_off + n + j];
      rel[q] = upd_rel[rbase + min(max(i, j), mb - 1)];
    }
  }
  __syncthreads();
  d4 acc[4] = {};
#pragma unroll 4
  for (int k0 = 0; k0 < kMaxCols; k0 += 4) {
    const double av = As[k0 + lk][16 * wave + lr];
    double bv[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) bv[c] = Bs[k0 + lk][16 * c + lr];
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv[c], acc[c], 0, 0, 0);
  }
  // transpose through LDS (the A tile's space): T[column][row]
  __syncthreads();
  auto T = As;
#pragma unroll
  for (int c = 0; c < 4; ++c)
#pragma unroll
    for (int r = 0; r < 4; ++r) T[16 * c + lr][16 * wave + lk + 4 * r] = acc[c][r];
  __syncthreads();
  const int i = ti * kTile + lane;
#pragma unroll
  for (int q = 0; q < kCols; ++q) {
    const int jl = wave + 4 * q, j = tj * kTile + jl;
    const double v = T[jl][lane];
    if (j < mb && i < mb && i >= j) unsafeAtomicAdd(panel + cbase[q] + rel[q], -v);
  }
}

__global__ void __launch_bounds__(256) update_kernel(const int32_t* __restrict__ items,
                                                     const SnDesc* __restrict__ desc,
                                                     double* __restrict__ panel,
                                                     const double* __restrict__ D,
                                                     const int64_t* __restrict__ scat_col,
                                                     const int32_t* __restrict__ scat_rel,
                                                     const int32_t* __restrict__ upd_rel) {
  __shared__ double smem[kUpdSmem];
  update_body(items[3 * blockIdx.x], items[3 * blockIdx.x + 1], items[3 * blockIdx.x + 2], desc,
              panel, D, scat_col, scat_rel, upd_rel, smem);
}

// y = P b; also clears the backward pass's partial sums.
__global__ void permute_kernel(int32_t n, const int32_t* __restrict__ perm,
                               const double* __restrict__ b, double* __restrict__ y,
                               double* __restrict__ part) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    y[perm[i]] = b[i];
    part[i] = 0.0;
  }
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

// Stages the strictly lower part of L11 (n x n column-major at Lc, padded with zeros; unit
// diagonal implied) into LDS: Ls[i][k].
__device__ __forceinline__ void stage_l11(const double* Lc, int n, double (*Ls)[kMaxCols + 1]) {
#pragma unroll
  for (int t = 0; t < kMaxCols * kMaxCols / 256; ++t) {  // blockDim.x == 256
    const int idx = threadIdx.x + 256 * t;
    const int i = idx % kMaxCols, k = idx / kMaxCols;
    const double v = Lc[(int64_t)min(k, n - 1) * n + min(i, n - 1)];  // clamped, then masked
    Ls[i][k] = (i < n && k < n && i > k) ? v : 0.0;
  }
}

// Value of v in lane k (k wave-uniform), through two scalar lane reads: a short dependency for
// the triangular recurrences (an LDS-based shuffle costs a full LDS round trip per step).
__device__ __forceinline__ double lane_value(double v, int k) {
  const int2 u = __builtin_bit_cast(int2, v);
  return __builtin_bit_cast(double, make_int2(__builtin_amdgcn_readlane(u.x, k),
                                              __builtin_amdgcn_readlane(u.y, k)));
}

constexpr int kSolveRows = 256;  // below-rows per solve block (one per thread in the forward pass)

// Forward, block (s, chunk): y_own = L11^-1 y_own (one wavefront, lane r holds y_r, L11 from LDS;
// every chunk block of s solves redundantly, chunk 0 writes the result to yout) and, for the
// chunk's below-rows i, y(row i) -= L21(i, :) y_own with atomics into yacc (siblings at the same
// level share ancestors). A thread takes one below-row and all 64 (padded) columns; its 64 panel
// loads are issued before the triangular solve so they overlap it. Columns past n are clamped to
// column n - 1 and multiplied by ys = 0.
__global__ void __launch_bounds__(256) fwd_kernel(const int32_t* __restrict__ items,
                                                  const SnDesc* __restrict__ desc,
                                                  const int32_t* __restrict__ rows,
                                                  const double* __restrict__ panel,
                                                  const double* __restrict__ l11,
                                                  double* __restrict__ yacc,
                                                  double* __restrict__ yout) {
  __shared__ double Ls[kMaxCols][kMaxCols + 1];
  __shared__ double ys[kMaxCols];
  const int s = items[2 * blockIdx.x], chunk = items[2 * blockIdx.x + 1];
  const SnDesc sd = desc[s];
  const int n = sd.ncols, ld = sd.ld, mb = ld - n;
  const double* P = panel + sd.panel_off;
  const int32_t* rw = rows + sd.rows_off + n;
  const int i = chunk * kSolveRows + threadIdx.x;
  double lrow[kMaxCols];
  {
    // clamped into the panel (the diagonal block when there are no below-rows)
    const int64_t r = mb > 0 ? n + min(i, mb - 1) : 0;
#pragma unroll
    for (int k = 0; k < kMaxCols; ++k) lrow[k] = P[(int64_t)min(k, n - 1) * ld + r];
  }
  stage_l11(l11 + sd.l11_off, n, Ls);
  double yi = 0;
  if (threadIdx.x < n) yi = yacc[sd.col_begin + threadIdx.x];
  __syncthreads();
  if (threadIdx.x < 64) {
    const int r = threadIdx.x;
#pragma unroll
    for (int k = 0; k < kMaxCols; ++k) yi -= Ls[r][k] * lane_value(yi, k);
    if (chunk == 0 && r < n) yout[sd.col_begin + r] = yi;
    ys[r] = r < n ? yi : 0.0;
  }
  __syncthreads();
  if (i < mb) {
    double acc = 0;
#pragma unroll
    for (int k = 0; k < kMaxCols; ++k) acc += lrow[k] * ys[k];
    unsafeAtomicAdd(&yacc[rw[i]], -acc);
  }
}

// Backward, block (s, chunk): part(own) += L21(chunk rows, :)^T y(chunk rows) (wave w takes
// columns w, w + 4, ..., w + 60: 16 accumulators per lane, lanes over the rows, reduced across the
// wave), added atomically into part; the last chunk block of s to finish (a per-supernode arrival
// counter, reset by that block for the next solve) then forms y_own = L11^-T (y_own - part) with
// one wavefront. part must be zero on entry (the permutation kernel clears it).
__global__ void __launch_bounds__(256) bwd_kernel(const int32_t* __restrict__ items,
                                                  const SnDesc* __restrict__ desc,
                                                  const int32_t* __restrict__ rows,
                                                  const double* __restrict__ panel,
                                                  const double* __restrict__ l11,
                                                  double* __restrict__ yout,
                                                  double* __restrict__ part,
                                                  int32_t* __restrict__ counter) {
  __shared__ double Ls[kMaxCols][kMaxCols + 1];
  __shared__ int last;
  const int s = items[2 * blockIdx.x], chunk = items[2 * blockIdx.x + 1];
  const SnDesc sd = desc[s];
  const int n = sd.ncols, ld = sd.ld, mb = ld - n;
  const int nchunks = max(1, (mb + kSolveRows - 1) / kSolveRows);
  const double* P = panel + sd.panel_off + n;
  const int32_t* rw = rows + sd.rows_off + n;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  // Every block stages L11 and reads y_own up front: only the last block of s uses them, but this
  // way their loads overlap the partial sums instead of following the arrival counter (y_own is
  // written by no other block of this launch).
  stage_l11(l11 + sd.l11_off, n, Ls);
  double yown = 0;
  if (threadIdx.x < n) yown = yout[sd.col_begin + threadIdx.x];
  if (mb > 0) {
    constexpr int kPer = kMaxCols / 4;
    double acc[kPer] = {};
    int64_t coff[kPer];
#pragma unroll
    for (int c = 0; c < kPer; ++c) coff[c] = (int64_t)min(wave + 4 * c, n - 1) * ld;
#pragma unroll
    for (int t = 0; t < kSolveRows / 64; ++t) {
      const int i = chunk * kSolveRows + lane + 64 * t;
      const int ic = min(i, mb - 1);
      const double yl = yout[rw[ic]];
      const double yv = i < mb ? yl : 0.0;
#pragma unroll
      for (int c = 0; c < kPer; ++c) acc[c] += P[coff[c] + ic] * yv;
    }
#pragma unroll
    for (int c = 0; c < kPer; ++c) {
      double v = acc[c];
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
      const int k = wave + 4 * c;
      if (lane == 0 && k < n) unsafeAtomicAdd(&part[sd.col_begin + k], v);
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    last = atomicAdd(&counter[s], 1) == nchunks - 1;
  }
  __syncthreads();
  if (!__builtin_amdgcn_readfirstlane(last)) return;  // uniform across the block
  __threadfence();
  if (threadIdx.x < 64) {
    const int r = threadIdx.x;
    double yi = 0;
    if (r < n)
      yi = yown - __hip_atomic_load(&part[sd.col_begin + r], __ATOMIC_RELAXED,
                                    __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll
    for (int k = kMaxCols - 1; k >= 0; --k) yi -= Ls[k][r] * lane_value(yi, k);
    if (r < n) yout[sd.col_begin + r] = yi;
    if (r == 0) counter[s] = 0;
  }
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
  yout_.resize(std::max(1, sym.n));
  part_.resize(std::max(1, sym.n));
  counter_.resize(std::max(1, sym.n_snodes));
  fill_zero_i32(counter_.get(), counter_.size(), s);
  rows_.upload(sym.sn_rows, s);
  perm_.upload(sym.perm, s);
  upd_rel_.upload(sym.upd_rel, s);
  // supernode descriptors
  std::vector<SnDesc> desc(std::max(1, sym.n_snodes));
  int64_t l11_size = 0;
  for (int k = 0; k < sym.n_snodes; ++k) {
    const int n = sym.sn_col_begin[k + 1] - sym.sn_col_begin[k];
    desc[k] = SnDesc{sym.sn_panel_off[k], l11_size, n, sym.sn_rows_ptr[k + 1] - sym.sn_rows_ptr[k],
                     sym.sn_col_begin[k], sym.sn_rows_ptr[k]};
    l11_size += (int64_t)n * n;
  }
  l11_.resize(std::max<int64_t>(1, l11_size));
  fill_zero(l11_.get(), l11_.size(), s);
  desc_.resize(desc.size() * sizeof(SnDesc));
  HIP_OK(hipMemcpyAsync(desc_.get(), desc.data(), desc.size() * sizeof(SnDesc),
                        hipMemcpyHostToDevice, s));
  // scatter targets per below-row
  std::vector<int64_t> scol(std::max<size_t>(1, sym.sn_rows.size()), 0);
  std::vector<int32_t> srel(std::max<size_t>(1, sym.sn_rows.size()), 0);
  for (int k = 0; k < sym.n_snodes; ++k) {
    const int n = sym.sn_col_begin[k + 1] - sym.sn_col_begin[k];
    const int q0 = sym.sn_rows_ptr[k] + n;
    for (int seg = sym.upd_ptr[k]; seg < sym.upd_ptr[k + 1]; ++seg) {
      const int t = sym.upd_target[seg];
      const int jb = sym.upd_seg_begin[seg];
      const int je = seg + 1 < sym.upd_ptr[k + 1] ? sym.upd_seg_begin[seg + 1]
                                                  : sym.sn_rows_ptr[k + 1] - q0;
      const int ldt = sym.sn_rows_ptr[t + 1] - sym.sn_rows_ptr[t];
      for (int j = jb; j < je; ++j) {
        const int col = sym.sn_rows[q0 + j] - sym.sn_col_begin[t];
        scol[q0 + j] = sym.sn_panel_off[t] + (int64_t)col * ldt;
        srel[q0 + j] = sym.upd_rel_ptr[seg] - jb;
      }
    }
  }
  scat_col_.upload(scol, s);
  scat_rel_.upload(srel, s);
  // Work lists per level
  std::vector<int32_t> snl, dt, upd, sol;
  lvl_snode_ptr_.assign(1, 0);
  lvl_solve_ptr_.assign(1, 0);
  lvl_trsm_ptr_.assign(1, 0);
  lvl_upd_ptr_.assign(1, 0);
  for (int l = 0; l < sym.n_levels; ++l) {
    for (int q = sym.level_ptr[l]; q < sym.level_ptr[l + 1]; ++q) {
      const int sn = sym.level_snodes[q];
      snl.push_back(sn);
      const int n = sym.sn_col_begin[sn + 1] - sym.sn_col_begin[sn];
      const int mb = sym.sn_rows_ptr[sn + 1] - sym.sn_rows_ptr[sn] - n;
      const int chunks = std::max(1, (mb + 63) / 64);
      for (int c = 0; c < chunks; ++c) {
        dt.push_back(sn);
        dt.push_back(c);
      }
      const int schunks = std::max(1, (mb + kSolveRows - 1) / kSolveRows);
      for (int c = 0; c < schunks; ++c) {
        sol.push_back(sn);
        sol.push_back(c);
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
    lvl_solve_ptr_.push_back((int32_t)sol.size() / 2);
    lvl_trsm_ptr_.push_back((int32_t)dt.size() / 2);
    lvl_upd_ptr_.push_back((int32_t)upd.size() / 3);
  }
  lvl_snodes_.upload(snl, s);
  solve_items_.upload(sol, s);
  trsm_items_.upload(dt, s);
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

SupernodalLDLT::~SupernodalLDLT() {
  if (fact_graph_) (void)hipGraphExecDestroy(fact_graph_);
  for (auto& g : solve_graphs_) (void)hipGraphExecDestroy(g.exec);
}

namespace {
template <typename F>
hipGraphExec_t capture(hipStream_t s, F&& launches) {
  hipGraph_t g;
  hipGraphExec_t exec;
  HIP_OK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  launches();
  HIP_OK(hipStreamEndCapture(s, &g));
  HIP_OK(hipGraphInstantiate(&exec, g, nullptr, nullptr, 0));
  HIP_OK(hipGraphDestroy(g));
  return exec;
}
}  // namespace

void SupernodalLDLT::factorize_assembled(hipStream_t s) {
  if (sym_.n_levels == 0) return;
  if (!fact_graph_) fact_graph_ = capture(s, [&] { launch_factorize(s); });
  HIP_OK(hipGraphLaunch(fact_graph_, s));
}

void SupernodalLDLT::solve(const double* b, double* x, hipStream_t s) {
  if (sym_.n == 0) return;
  for (const auto& g : solve_graphs_)
    if (g.b == b && g.x == x) {
      HIP_OK(hipGraphLaunch(g.exec, s));
      return;
    }
  if (solve_graphs_.size() >= 8) {  // bound the cache; the oldest goes
    (void)hipGraphExecDestroy(solve_graphs_.front().exec);
    solve_graphs_.erase(solve_graphs_.begin());
  }
  solve_graphs_.push_back({b, x, capture(s, [&] { launch_solve(b, x, s); })});
  HIP_OK(hipGraphLaunch(solve_graphs_.back().exec, s));
}

void SupernodalLDLT::launch_factorize(hipStream_t s) {
  const SnDesc* desc = reinterpret_cast<const SnDesc*>(desc_.get());
  for (int l = 0; l < sym_.n_levels; ++l) {
    const int nt = lvl_trsm_ptr_[l + 1] - lvl_trsm_ptr_[l];
    if (nt)
      hipLaunchKernelGGL(diag_trsm_kernel, dim3(nt), dim3(256), 0, s,
                         trsm_items_.get() + 2 * lvl_trsm_ptr_[l], desc, panel_.get(), l11_.get(),
                         D_.get());
    const int nu = lvl_upd_ptr_[l + 1] - lvl_upd_ptr_[l];
    if (nu)
      hipLaunchKernelGGL(update_kernel, dim3(nu), dim3(256), 0, s,
                         upd_items_.get() + 3 * lvl_upd_ptr_[l], desc, panel_.get(), D_.get(),
                         scat_col_.get(), scat_rel_.get(), upd_rel_.get());
  }
  HIP_OK(hipGetLastError());
}

void SupernodalLDLT::launch_solve(const double* b, double* x, hipStream_t s) {
  const int n = sym_.n;
  const SnDesc* desc = reinterpret_cast<const SnDesc*>(desc_.get());
  hipLaunchKernelGGL(permute_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, n, perm_.get(), b,
                     y_.get(), part_.get());
  for (int l = 0; l < sym_.n_levels; ++l) {
    const int nb = lvl_solve_ptr_[l + 1] - lvl_solve_ptr_[l];
    hipLaunchKernelGGL(fwd_kernel, dim3(nb), dim3(256), 0, s,
                       solve_items_.get() + 2 * lvl_solve_ptr_[l], desc, rows_.get(),
                       panel_.get(), l11_.get(), y_.get(), yout_.get());
  }
  hipLaunchKernelGGL(dscale_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, n, D_.get(),
                     yout_.get());
  for (int l = sym_.n_levels - 1; l >= 0; --l) {
    const int nb = lvl_solve_ptr_[l + 1] - lvl_solve_ptr_[l];
    hipLaunchKernelGGL(bwd_kernel, dim3(nb), dim3(256), 0, s,
                       solve_items_.get() + 2 * lvl_solve_ptr_[l], desc, rows_.get(),
                       panel_.get(), l11_.get(), yout_.get(), part_.get(), counter_.get());
  }
  hipLaunchKernelGGL(unpermute_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, n, perm_.get(),
                     yout_.get(), x);
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
