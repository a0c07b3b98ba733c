// Device numeric supernodal LDL^T: assembly, per-level fused diagonal-block factorization + panel
// triangular solve, right-looking update (SYRK-like) scattered into ancestor panels, and the
// triangular solves. FP64, except the trailing updates of the mixed-precision form (FP32 matrix
// cores, set_f32_updates).
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
// Update tile (rows). 32-row tiles on every level, with shared memory sized for them (four
// blocks per CU at 33 KB, seven with the columns staged in two halves): a level's update time is the latency of its slowest tiles, and a 64-row tile's
// product (K = 64 on the FP64 matrix cores) and 4096-entry scatter take ~15 us where four 32-row
// tiles run side by side in ~6 (measured: 1000-camera system 464 -> 388 us of updates per
// factorization, 100 000-node pose graph 460 -> 205 us).
constexpr int kTile = 32;
constexpr int kCus = 256;  // compute units of an MI355X (block-count thresholds of the launches)

// l11_off: the factored diagonal block L11 (n x n, column-major, strictly lower part used) is
// kept apart from the panel, so the row-chunk blocks of one supernode, which all read the
// unfactored A11 from the panel, never see it half overwritten.
struct SnDesc {
  int64_t panel_off, l11_off;
  int32_t ncols, ld, col_begin, rows_off;
};

// One task of a level launch: its supernode's descriptor together with the task's own indices,
// fetched in one scalar-load round trip (an index list followed by a descriptor lookup costs two
// dependent round trips at the head of every latency-bound launch).
struct SnTask {
  SnDesc d;
  int32_t s, a, b, pad;  // supernode; chunk or tile row; tile column
};

__global__ void assemble_kernel(int64_t nnz, const int64_t* __restrict__ map,
                                const double* __restrict__ vals, double* __restrict__ panel) {
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < nnz;
       q += (int64_t)gridDim.x * blockDim.x)
    panel[map[q]] = vals[q];
}

// Diagonal task (s, chunk): one 512-thread block eliminates the 64 x 64 diagonal block of
// supernode s together with the chunk's 64 below rows in one right-looking sweep of kStep-column
// steps (the diagonal block is factored redundantly by every chunk of s; chunk 0 writes L11 to
// l11 and D), leaving L21 = A21 L11^-T D^-1 for its rows, written back into the panel in place.
//
// The 128 x 64 working matrix (rows 0..63 the diagonal block padded with identity, rows 64..127
// the chunk's below rows) lives in matrix-core accumulators: wave w holds rows 16w .. 16w + 15 as
// four 16 x 16 tiles. Step k0 = 0, kStep, .., 64 - kStep:
//  1. the lanes holding columns k0 .. k0 + kStep - 1 publish them to cb (LDS; for k0 > 0 right
//     after the previous step's products for those columns, before its other tiles');
//  2. row owners (thread t < 128 owns row t) factor the kStep x kStep pivot block (redundantly,
//     a chain of kStep reciprocals), form their row's w = c Lp^-T (c = the row's published
//     entries, Lp the pivot block's unit lower factor) and L = w D^-1, and write w to Wt and L
//     to Lc (the pivot rows get Lp and D instead);
//  3. every wave applies A(i, j) -= sum_u L(i, u) w(j, u) to its live tiles with
//     v_mfma_f64_16x16x4f64 (operands read straight from Lc and Wt).
// Each row's elimination is computed once, and an update needs a few LDS reads per lane per step
// (a register-patch FMA form read 16 per column, and was bound by them). The owner phase is a
// dependent FP64 chain that grows with the cube of kStep per step: 4-column steps beat 8-column
// ones (measured), the extra barriers notwithstanding. Tiles that are dead for the rest of the
// sweep are skipped (the step loop is unrolled; the row test is wave-uniform): columns below
// k0 + kStep and diagonal-block row tiles above it. Within live tiles nothing is masked: entries
// above the diagonal and finished columns pick up products of other such entries but are never
// read again (the final values are in Lc), and the entries still to be eliminated only ever
// combine lower-triangle values.
constexpr int kDiagThreads = 512;
constexpr int kStep = 4;                   // columns eliminated per sweep step (4 or 8)
constexpr int kDiagRows = 2 * kMaxCols;    // diagonal block + one chunk of below rows
// LDS layouts, padded so the owners' row-wise writes, the publishing lanes' writes and the
// matrix-core operand reads are free of bank conflicts (row-major arrays with 8-double rows cost
// 16-way conflicts on the owners' writes, four fifths of the owner phase)
constexpr int kCbLd = kDiagRows + 4;       // cb[u][row]: published columns
constexpr int kLcLd = kDiagRows + 2;       // Lc[col][row]: the working matrix, then L and D
constexpr int kDiagSmem = kStep * kCbLd + kStep * kDiagRows + kMaxCols * kLcLd;

// 1 / d by the hardware estimate and two Newton steps (the IEEE division sequence is twice as
// long, and these sit on the sweep's critical path).
__device__ __forceinline__ double refined_rcp(double d) {
  double r = __builtin_amdgcn_rcp(d);
  r = fma(r, fma(-d, r, 1.0), r);
  return fma(r, fma(-d, r, 1.0), r);
}

// Step 2 for row t: cb[u][i] = A(i, k0 + u) after the previous steps; writes Wt[u][t] = w_u and
// Lc[k0 + u][t] = L(t, k0 + u).
__device__ __forceinline__ void sweep_owner(int k0, int t, const double (*cb)[kCbLd],
                                            double (*Wt)[kDiagRows], double (*Lc)[kLcLd]) {
  double d[kStep], dinv[kStep], Lp[kStep][kStep];
  // pivot block, column by column: with wp_uv = L_uv d_v (row u's, formed when column u starts),
  // d_u = A_uu - sum_v wp_uv L_uv and L_ru = (A_ru - sum_v L_rv wp_uv) / d_u
#pragma unroll
  for (int u = 0; u < kStep; ++u) {
    double wp[kStep];
#pragma unroll
    for (int v = 0; v < u; ++v) wp[v] = Lp[u][v] * d[v];
    double dd = cb[u][k0 + u];
#pragma unroll
    for (int v = 0; v < u; ++v) dd -= wp[v] * Lp[u][v];
    d[u] = dd;
    dinv[u] = refined_rcp(dd);
#pragma unroll
    for (int r = u + 1; r < kStep; ++r) {
      double x = cb[u][k0 + r];
#pragma unroll
      for (int v = 0; v < u; ++v) x -= Lp[r][v] * wp[v];
      Lp[r][u] = x * dinv[u];
    }
  }
  double w[kStep], l[kStep];
#pragma unroll
  for (int u = 0; u < kStep; ++u) {
    double c = cb[u][t];
#pragma unroll
    for (int v = 0; v < u; ++v) c -= w[v] * Lp[u][v];
    w[u] = c;
    l[u] = c * dinv[u];
  }
  // pivot row k0 + r: Lp row left of the diagonal, D on it (w_r = d_r), zero right of it
  const int r = t - k0;
#pragma unroll
  for (int u = 0; u < kStep; ++u) {
    Wt[u][t] = w[u];
    Lc[k0 + u][t] = (r < 0 || r >= kStep || u < r) ? l[u] : (u == r ? w[u] : 0.0);
  }
}

// Value of v in lane k (k wave-uniform), through two scalar lane reads: a short dependency for
// the triangular recurrences (an LDS-based shuffle costs a full LDS round trip per step).
__device__ __forceinline__ double lane_value(double v, int k) {
  const int2 u = __builtin_bit_cast(int2, v);
  return __builtin_bit_cast(double, make_int2(__builtin_amdgcn_readlane(u.x, k),
                                              __builtin_amdgcn_readlane(u.y, k)));
}

// Forward solve fused into the diagonal task (kFwd): with y = P b plus the updates of every
// descendant (earlier levels), the block forms y_own = L11^-1 y_own (wave 0, lane r holding y_r)
// once L11 is in LDS, chunk 0 writes it to yout, and the chunk's below rows get
// y(row) -= L21(row, :) y_own atomically (siblings share ancestors). The same work as a separate
// forward launch per level, without the launch.
template <bool kFwd>
__device__ __forceinline__ void diag_body(const SnTask& tk, double* __restrict__ panel,
                                          double* __restrict__ l11,
                                          double* __restrict__ D, const int32_t* __restrict__ rows,
                                          double* __restrict__ y, double* __restrict__ yout,
                                          double* smem) {
  typedef double d4 __attribute__((ext_vector_type(4)));
  auto cb = reinterpret_cast<double (*)[kCbLd]>(smem);
  auto Wt = reinterpret_cast<double (*)[kDiagRows]>(smem + kStep * kCbLd);
  auto Lc = reinterpret_cast<double (*)[kLcLd]>(smem + kStep * kCbLd + kStep * kDiagRows);
  const SnDesc sd = tk.d;
  const int chunk = tk.a;
  const int n = sd.ncols, ld = sd.ld, mb = ld - n;
  double* P = panel + sd.panel_off;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lr = lane & 15, lk = lane >> 4;
  // forward-solve operands, fetched up front (clamped indices, selected at use)
  double y_in = 0.0;
  int32_t y_row = 0;
  if (kFwd && wave == 0) {
    y_in = y[sd.col_begin + min(lane, n - 1)];
    y_row = rows[sd.rows_off + (mb > 0 ? n + min(chunk * kMaxCols + lane, mb - 1) : 0)];
  }
  // Stage the working matrix in Lc with coalesced loads (consecutive threads take consecutive
  // rows of one column): one unconditional load per element from a clamped valid address, then
  // a select (a guarded load is compiled as a branch that waits for each load in turn).
  constexpr int kPer = kDiagRows * kMaxCols / kDiagThreads;
  double v[kPer];
#pragma unroll
  for (int q = 0; q < kPer; ++q) {  // all loads first
    const int idx = tid + kDiagThreads * q;
    const int i = idx & (kDiagRows - 1), j = idx / kDiagRows;
    const int jc = min(j, n - 1), ic = min(i, n - 1);
    const int r = chunk * kMaxCols + i - kMaxCols;
    const int row = (i >= kMaxCols) & (mb > 0) ? n + min(r, mb - 1) : ic;
    v[q] = P[(int64_t)jc * ld + row];
  }
#pragma unroll
  for (int q = 0; q < kPer; ++q) {
    const int idx = tid + kDiagThreads * q;
    const int i = idx & (kDiagRows - 1), j = idx / kDiagRows;
    const int r = chunk * kMaxCols + i - kMaxCols;
    const int below = i >= kMaxCols;
    // below rows: inside the chunk; diagonal block: lower triangle inside n x n, identity padding
    const int keep = (below & (r < mb) & (j < n)) | (!below & (i < n) & (j < n) & (i >= j));
    const double fill = (!below & (i == j)) ? 1.0 : 0.0;
    Lc[j][i] = keep ? v[q] : fill;
  }
  __syncthreads();
  // Wave w holds rows 16w .. 16w + 15 as four 16 x 16 accumulator tiles (columns 16c ..), in the
  // v_mfma_f64_16x16x4f64 layout: acc[c][r] = A(16w + lane / 16 + 4r, 16c + lane % 16).
  d4 acc[4];
#pragma unroll
  for (int c = 0; c < 4; ++c)
#pragma unroll
    for (int r = 0; r < 4; ++r) acc[c][r] = Lc[16 * c + lr][16 * wave + lk + 4 * r];
  // (the first barrier of the sweep orders these reads before the owners' Lc writes)
  const bool below_wave = wave >= kMaxCols / 16;
  // Steps past the supernode's last column are skipped (the padding stays the identity, which
  // its elimination would not change): a 6-column supernode takes 2 steps, not 16. (Skipping the
  // padding's MFMA tiles, or ending the solve chains at n, measured slower: the extra runtime
  // conditions cost the unrolled loops their schedule.)
  // publish columns k0 .. k0 + kStep - 1 (tile column k0 / 16, lanes k0 % 16 ..) of the rows
  // from k0 on
  auto publish = [&](int k0) {
    const int off = k0 & 15, cp = k0 >> 4;
    if ((below_wave || 16 * wave + 15 >= k0) && lr >= off && lr < off + kStep) {
#pragma unroll
      for (int r = 0; r < 4; ++r) cb[lr - off][16 * wave + lk + 4 * r] = acc[cp][r];
    }
  };
  publish(0);
#pragma unroll
  for (int st = 0; st < kMaxCols / kStep; ++st) {
    const int k0 = kStep * st;
    if (k0 >= n) break;
    __syncthreads();
    if (tid < kDiagRows) sweep_owner(k0, tid, cb, Wt, Lc);
    __syncthreads();
    const int kn = k0 + kStep;  // first column still to be eliminated
    if (kn < kMaxCols && (below_wave || 16 * wave + 15 >= kn)) {
      // A(i, j) -= sum_u L(i, k0 + u) w(j, u) on the matrix cores, kStep / 4 K = 4 pieces per tile:
      // lane l feeds A(l % 16, l / 16) = -L(16w + l % 16, k0 + 4h + l / 16) and
      // B(l / 16, l % 16) = w(16c + l % 16, 4h + l / 16)
      double av[kStep / 4];
#pragma unroll
      for (int h = 0; h < kStep / 4; ++h) av[h] = -Lc[k0 + 4 * h + lk][16 * wave + lr];
      // The tile of the next step's pivot columns first, published at once (the owners of this
      // step are done with cb); the other tiles' products then run behind the next step's
      // barrier and owner phase (measured: 16.4 -> 16.0 us per diagonal launch).
      const int cpn = kn >> 4;
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        if (c != cpn) continue;
#pragma unroll
        for (int h = 0; h < kStep / 4; ++h)
          acc[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[h], Wt[4 * h + lk][16 * c + lr], acc[c],
                                                         0, 0, 0);
      }
      if (kn < n) publish(kn);
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        if (c == cpn || 16 * c + 15 < kn) continue;  // done above / columns already eliminated
#pragma unroll
        for (int h = 0; h < kStep / 4; ++h)
          acc[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[h], Wt[4 * h + lk][16 * c + lr], acc[c],
                                                         0, 0, 0);
      }
    }
  }
  __syncthreads();
  if (kFwd && wave == 0) {
    const int r = lane;
    double yi = r < n ? y_in : 0.0;
#pragma unroll
    for (int k = 0; k < kMaxCols; ++k) yi -= (r > k ? Lc[k][r] : 0.0) * lane_value(yi, k);
    if (chunk == 0 && r < n) yout[sd.col_begin + r] = yi;
    double acc = 0.0;
#pragma unroll
    for (int k = 0; k < kMaxCols; ++k) acc += Lc[k][kMaxCols + r] * lane_value(yi, k);
    if (chunk * kMaxCols + r < mb) unsafeAtomicAdd(&y[y_row], -acc);
    return;
  }
  // Write back with coalesced stores (with the forward solve, wave 0 is busy with it and the
  // other waves store): L11 (strict lower, column-major n x n) and D from chunk 0, the chunk's
  // L21 rows into the panel. The thread index passes through an empty asm so the compiler
  // recomputes these addresses here instead of holding the load phase's across the sweep.
  constexpr int kStorers = kFwd ? kDiagThreads - 64 : kDiagThreads;
  int tid_w = kFwd ? tid - 64 : tid;
  asm volatile("" : "+v"(tid_w));
#pragma unroll 4
  for (int idx = tid_w; idx < kDiagRows * kMaxCols; idx += kStorers) {
    const int i = idx & (kDiagRows - 1), j = idx / kDiagRows;
    const double v = Lc[j][i];
    if (i < kMaxCols) {
      if (chunk == 0 && i < n && j < n) {
        if (i > j) l11[sd.l11_off + (int64_t)j * n + i] = v;
        if (i == j) D[sd.col_begin + i] = v;
      }
    } else {
      const int r = chunk * kMaxCols + i - kMaxCols;
      if (r < mb && j < n) P[(int64_t)j * ld + n + r] = v;
    }
  }
}

template <bool kFwd>
__global__ void __launch_bounds__(kDiagThreads) diag_trsm_kernel(
    const SnTask* __restrict__ tasks, double* __restrict__ panel, double* __restrict__ l11,
    double* __restrict__ D, const int32_t* __restrict__ rows, double* __restrict__ y,
    double* __restrict__ yout) {
  __shared__ double smem[kDiagSmem];
  diag_body<kFwd>(tasks[blockIdx.x], panel, l11, D, rows, y, yout, smem);
}

// Update task (s, ti, tj), one 256-thread block: U = L21 D L21^T tile (ti, tj) of KT x KT,
// scattered as -U into ancestor panels. For below-row j of the supernode, scat_col[j] is the
// panel offset of its column in the target supernode and scat_rel[j] + i indexes upd_rel for the
// target row position of below-row i >= j (i < mb). The launches use KT = kTile = 32 (the body
// also takes 64-row tiles).
//
// T = float is the mixed-precision form of the update: both row tiles are rounded to FP32 (the
// D scaling applied first, in FP64), the product runs on the FP32 matrix cores
// (v_mfma_f32_16x16x4f32, FP32 accumulation) and is scattered as FP64 like the FP64 form; the
// diagonal blocks and panels stay FP64 (diag_trsm_kernel), and the caller refines its solves
// against the FP64 matrix.
template <int KT, typename T>
__device__ __forceinline__ void update_body(const SnTask& tk, double* __restrict__ panel,
                                            const double* __restrict__ D,
                                            const int64_t* __restrict__ scat_col,
                                            const int32_t* __restrict__ scat_rel,
                                            const int32_t* __restrict__ upd_rel, double* smem) {
  static_assert(KT == 64 || KT == 32, "update tiles are 64 or 32 rows");
  constexpr bool kF64 = sizeof(T) == 8;
  constexpr int kKh = kMaxCols / 2;  // columns staged per half
  auto As = reinterpret_cast<T (*)[KT + 1]>(smem);
  auto Bs = reinterpret_cast<T (*)[KT + 1]>(reinterpret_cast<T*>(smem) + kKh * (KT + 1));
  const SnDesc sd = tk.d;
  const int ti = tk.a, tj = tk.b;
  const int n = sd.ncols, ld = sd.ld, mb = ld - n;
  const double* P = panel + sd.panel_off;
  const double* Ds = D + sd.col_begin;
  // Stage both row tiles, half the columns at a time (unconditional loads from clamped
  // addresses; an update item exists only when mb > 0). Two halves take half the shared memory
  // (17 KB: seven blocks per CU instead of four), and a supernode of at most 32 columns skips the
  // second, all-padding half (the 100 000-node pose graph's updates 206 -> 192 us per
  // factorization; the 1000-camera system's 63-column supernodes unchanged).
  auto stage = [&](int h) {
#pragma unroll
    for (int t = 0; t < kKh * KT / 256; ++t) {  // blockDim.x == 256
      const int idx = threadIdx.x + 256 * t;
      const int k = idx / KT, r = idx % KT;
      const int kk = kKh * h + k;
      const int ri = ti * KT + r, rj = tj * KT + r;
      const int kc = min(kk, n - 1);
      const double va = P[(int64_t)kc * ld + n + min(ri, mb - 1)];
      const double vb = P[(int64_t)kc * ld + n + min(rj, mb - 1)] * Ds[kc];
      As[k][r] = (kk < n && ri < mb) ? T(va) : T(0);
      Bs[k][r] = (kk < n && rj < mb) ? T(vb) : T(0);
    }
  };
  stage(0);
  // The product on the FP64 matrix cores, in 16 x 16 output blocks, K = 64 in two halves of 8
  // steps of v_mfma_f64_16x16x4f64: a 64-row tile gives wave w row block w and all four column blocks, a
  // 32-row tile gives it row block w / 2 and column block w % 2. Lane l feeds
  // A(i = l % 16, k = l / 16) = As[k][16 rb + i] and B(k, j = l % 16) = Bs[k][16 c + j] and holds
  // D(l / 16 + 4 r, l % 16) of block c in accumulator register r (FP64) or D(4 (l / 16) + r,
  // l % 16) (FP32).
  typedef T d4 __attribute__((ext_vector_type(4)));
  constexpr int NBW = KT == 64 ? 4 : 1;  // output blocks per wave
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // uniform: scalar metadata
  const int lr = lane & 15, lk = lane >> 4;
  const int rb = KT == 64 ? wave : wave >> 1, c0 = KT == 64 ? 0 : wave & 1;
  // Scatter targets, fetched while the tiles load and before the product loop (the row
  // positions depend on the metadata: two dependent loads, off the critical path here). The
  // scatter runs column by column: wave w takes tile columns w, w + 4, .. and each atomic
  // instruction covers KT consecutive rows of 64 / KT columns (a few cache lines) instead of 16
  // columns x 4 rows. Row indices are clamped into the column's update segment (rows j ..
  // mb - 1); clamped entries are masked at the scatter.
  constexpr int kCpi = 64 / KT;             // columns per scatter instruction
  constexpr int kNq = KT / 4 / kCpi;        // scatter instructions per lane
  const int srow = lane % KT, ssub = lane / KT;
  int64_t cbase[kNq];
  int32_t rel[kNq];
  {
    const int i = ti * KT + srow;
#pragma unroll
    for (int q = 0; q < kNq; ++q) {
      const int j = min(tj * KT + wave + 4 * (kCpi * q + ssub), mb - 1);
      cbase[q] = scat_col[sd.rows_off + n + j];
      const int32_t rbase = scat_rel[sd.rows_off + n + j];
      rel[q] = upd_rel[rbase + min(max(i, j), mb - 1)];
    }
  }
  __syncthreads();
  d4 acc[NBW] = {};
  for (int h = 0; h < 2; ++h) {
    if (h == 1) {
      if (n <= kKh) break;  // uniform: the second half is all padding
      __syncthreads();
      stage(1);
      __syncthreads();
    }
#pragma unroll 4
    for (int k0 = 0; k0 < kKh; k0 += 4) {
      const T av = As[k0 + lk][16 * rb + lr];
      T bv[NBW];
#pragma unroll
      for (int c = 0; c < NBW; ++c) bv[c] = Bs[k0 + lk][16 * (c0 + c) + lr];
#pragma unroll
      for (int c = 0; c < NBW; ++c) {
        if constexpr (kF64)
          acc[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv[c], acc[c], 0, 0, 0);
        else
          acc[c] = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bv[c], acc[c], 0, 0, 0);
      }
    }
  }
  // transpose through LDS (the tiles' space, as FP64): Tr[column][row]
  __syncthreads();
  auto Tr = reinterpret_cast<double (*)[KT + 1]>(smem);
#pragma unroll
  for (int c = 0; c < NBW; ++c)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int row = kF64 ? lk + 4 * r : 4 * lk + r;
      Tr[16 * (c0 + c) + lr][16 * rb + row] = double(acc[c][r]);
    }
  __syncthreads();
  const int i = ti * KT + srow;
#pragma unroll
  for (int q = 0; q < kNq; ++q) {
    const int jl = wave + 4 * (kCpi * q + ssub), j = tj * KT + jl;
    const double v = Tr[jl][srow];
    if (j < mb && i < mb && i >= j) unsafeAtomicAdd(panel + cbase[q] + rel[q], -v);
  }
}

template <int KT, typename T>
__global__ void __launch_bounds__(256) update_kernel(const SnTask* __restrict__ tasks,
                                                     double* __restrict__ panel,
                                                     const double* __restrict__ D,
                                                     const int64_t* __restrict__ scat_col,
                                                     const int32_t* __restrict__ scat_rel,
                                                     const int32_t* __restrict__ upd_rel) {
  __shared__ double smem[kMaxCols * (KT + 1)];
  update_body<KT, T>(tasks[blockIdx.x], panel, D, scat_col, scat_rel, upd_rel, smem);
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

// below-rows per solve block: a backward block sums them 64 at a time (chunks of 128 beat 256 and
// 512: more, shorter blocks per level); a forward block takes one per thread
constexpr int kSolveRows = 128;
static_assert(kSolveRows % 64 == 0 && kSolveRows <= 256, "solve chunk: whole waves, one block");

// Forward, block (s, chunk): y_own = L11^-1 y_own (one wavefront, lane r holds y_r, L11 from LDS;
// every chunk block of s solves redundantly, chunk 0 writes the result to yout) and, for the
// chunk's below-rows i, y(row i) -= L21(i, :) y_own with atomics into yacc (siblings at the same
// level share ancestors). A thread takes one below-row and all 64 (padded) columns; its 64 panel
// loads are issued before the triangular solve so they overlap it. Columns past n are clamped to
// column n - 1 and multiplied by ys = 0.
__global__ void __launch_bounds__(256) fwd_kernel(const SnTask* __restrict__ tasks,
                                                  const int32_t* __restrict__ rows,
                                                  const double* __restrict__ panel,
                                                  const double* __restrict__ l11,
                                                  double* __restrict__ yacc,
                                                  double* __restrict__ yout) {
  __shared__ double Ls[kMaxCols][kMaxCols + 1];
  __shared__ double ys[kMaxCols];
  const SnTask tk = tasks[blockIdx.x];
  const int chunk = tk.a;
  const SnDesc sd = tk.d;
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
  if (threadIdx.x < kSolveRows && i < mb) {  // a chunk may be narrower than the block
    double acc = 0;
#pragma unroll
    for (int k = 0; k < kMaxCols; ++k) acc += lrow[k] * ys[k];
    unsafeAtomicAdd(&yacc[rw[i]], -acc);
  }
}

// Backward, block (s, chunk): part(own) += L21(chunk rows, :)^T y(chunk rows) (wave w takes
// columns w, w + 4, ..., w + 60: 16 accumulators per lane, lanes over the rows, reduced across the
// wave), added atomically into part; the last chunk block of s to finish (a per-supernode arrival
// counter, reset by that block for the next solve) then forms y_own = L11^-T (D^-1 y_own - part)
// with one wavefront (the D^-1 of the solve folded in: own entries are read by their supernode's
// blocks only). part must be zero on entry (the permutation kernel clears it). A supernode of
// one chunk (at most kSolveRows below-rows: most of a pose graph's) skips part, the atomics and
// the arrival: its block's column sums are final and go through LDS.
//
// kLean (levels with more blocks than fit two to a CU, e.g. the thousands of small leaf
// supernodes of a pose graph): L11 is held in registers instead of LDS until the block knows it
// is the last, then written over the partials' space, so the block needs 33 KB of LDS instead of
// 66 KB and four blocks fit a CU instead of two (the 100 000-node pose graph's backward solve
// 0.43 -> 0.35 ms). On levels of a few blocks the extra barrier only adds latency.
template <bool kLean>
__global__ void __launch_bounds__(256) bwd_kernel(const SnTask* __restrict__ tasks,
                                                  const int32_t* __restrict__ rows,
                                                  const double* __restrict__ panel,
                                                  const double* __restrict__ l11,
                                                  const double* __restrict__ D,
                                                  double* __restrict__ yout,
                                                  double* __restrict__ part,
                                                  int32_t* __restrict__ counter) {
  // the waves' column partials and L11; with kLean, L11 takes the partials' space at the end
  constexpr int kRed = (kMaxCols / 4) * (64 + 1);
  constexpr int kLs = kMaxCols * (kMaxCols + 1);
  static_assert(4 * kRed >= kLs, "L11 fits the partials' space");
  __shared__ double buf[4 * kRed + (kLean ? 0 : kLs)];
  auto red_all = reinterpret_cast<double (*)[kRed]>(buf);
  auto Ls = reinterpret_cast<double (*)[kMaxCols + 1]>(kLean ? buf : buf + 4 * kRed);
  __shared__ double colsum[kMaxCols];  // a single-chunk supernode's column sums
  __shared__ int last;
  const SnTask tk = tasks[blockIdx.x];
  const int s = tk.s, chunk = tk.a;
  const SnDesc sd = tk.d;
  const int n = sd.ncols, ld = sd.ld, mb = ld - n;
  const int nchunks = max(1, (mb + kSolveRows - 1) / kSolveRows);
  const double* P = panel + sd.panel_off + n;
  const int32_t* rw = rows + sd.rows_off + n;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  // Every block loads L11 and y_own up front: only the last block of s uses them, but this way
  // their loads overlap the partial sums instead of following the arrival counter (y_own is
  // written by no other block of this launch).
  constexpr int kL11 = kLean ? kMaxCols * kMaxCols / 256 : 1;  // blockDim.x == 256
  double l11v[kL11];
  if constexpr (kLean) {
    const double* Lc = l11 + sd.l11_off;
#pragma unroll
    for (int t = 0; t < kL11; ++t) {
      const int idx = threadIdx.x + 256 * t;
      const int i = idx % kMaxCols, k = idx / kMaxCols;
      const double v = Lc[(int64_t)min(k, n - 1) * n + min(i, n - 1)];  // clamped, then masked
      l11v[t] = (i < n && k < n && i > k) ? v : 0.0;
    }
  } else {
    stage_l11(l11 + sd.l11_off, n, Ls);
  }
  double yown = 0;
  if (threadIdx.x < n) yown = yout[sd.col_begin + threadIdx.x] / D[sd.col_begin + threadIdx.x];
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
    // Each wave's 16 column sums across its 64 lanes, through LDS: the partials go in as
    // red[c][lane], lane l then adds 16 of column l / 4's (quarter l % 4) and two quad shuffles
    // finish. (Six rounds of 64-lane shuffles for 16 values took longer than the loads.)
    auto red = reinterpret_cast<double (*)[64 + 1]>(&red_all[wave][0]);
#pragma unroll
    for (int c = 0; c < kPer; ++c) red[c][lane] = acc[c];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const int c = lane >> 2, qtr = lane & 3;
    double v = 0;
#pragma unroll
    for (int t = 0; t < 16; ++t) v += red[c][16 * qtr + t];
    v += __shfl_xor(v, 1, 64);
    v += __shfl_xor(v, 2, 64);
    const int k = wave + 4 * c;
    if (qtr == 0 && k < n) {
      if (nchunks == 1) colsum[k] = v;  // the block holds every row: the sums are final
      else unsafeAtomicAdd(&part[sd.col_begin + k], v);
    }
  }
  if (nchunks == 1) {  // uniform: no other block of s, no arrival
    __syncthreads();
  } else {
    // Every wave waits for its own atomics into part before the block arrives: a workgroup
    // barrier does not wait for the memory operations of the other waves, and an agent fence
    // in the arriving thread would cover only its own wave (and write back and invalidate the
    // L2).
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    if (threadIdx.x == 0)
      last = __hip_atomic_fetch_add(&counter[s], 1, __ATOMIC_RELAXED,
                                    __HIP_MEMORY_SCOPE_AGENT) == nchunks - 1;
    __syncthreads();
    if (!__builtin_amdgcn_readfirstlane(last)) return;  // uniform across the block
  }
  if constexpr (kLean) {  // every wave is past its partials (the barriers above)
#pragma unroll
    for (int t = 0; t < kL11; ++t) {
      const int idx = threadIdx.x + 256 * t;
      Ls[idx % kMaxCols][idx / kMaxCols] = l11v[t];
    }
    __syncthreads();
  }
  // every chunk's sums are in part: read them at agent scope (past the per-XCD L2)
  if (threadIdx.x < 64) {
    const int r = threadIdx.x;
    double yi = 0;
    if (r < n) {
      if (nchunks == 1)
        yi = mb > 0 ? yown - colsum[r] : yown;
      else
        yi = yown - __hip_atomic_load(&part[sd.col_begin + r], __ATOMIC_RELAXED,
                                      __HIP_MEMORY_SCOPE_AGENT);
    }
#pragma unroll
    for (int k = kMaxCols - 1; k >= 0; --k) yi -= Ls[k][r] * lane_value(yi, k);
    if (r < n) yout[sd.col_begin + r] = yi;
    if (r == 0 && nchunks > 1) counter[s] = 0;
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
    }
    constexpr int kt = kTile;
    for (int q = sym.level_ptr[l]; q < sym.level_ptr[l + 1]; ++q) {
      const int sn = sym.level_snodes[q];
      const int n = sym.sn_col_begin[sn + 1] - sym.sn_col_begin[sn];
      const int mb = sym.sn_rows_ptr[sn + 1] - sym.sn_rows_ptr[sn] - n;
      const int T = (mb + kt - 1) / kt;
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
  auto upload_tasks = [&](const std::vector<int32_t>& items, int per, DBuf<uint8_t>& out) {
    std::vector<SnTask> t(std::max<size_t>(1, items.size() / per));
    for (size_t q = 0; q < items.size() / per; ++q) {
      const int32_t* it = &items[q * per];
      t[q] = SnTask{desc[it[0]], it[0], it[1], per > 2 ? it[2] : 0, 0};
    }
    out.resize(t.size() * sizeof(SnTask));
    HIP_OK(hipMemcpyAsync(out.get(), t.data(), t.size() * sizeof(SnTask), hipMemcpyHostToDevice, s));
  };
  upload_tasks(sol, 2, solve_tasks_);
  upload_tasks(dt, 2, trsm_tasks_);
  upload_tasks(upd, 3, upd_tasks_);
  HIP_OK(hipStreamSynchronize(s));
}

void SupernodalLDLT::factorize(const double* a_values, hipStream_t s, const double* b) {
  fill_zero(panel_.get(), sym_.panel_size(), s);
  const int64_t nnz = (int64_t)asm_map_.size();
  if (nnz) {
    const int grid = (int)std::min<int64_t>(grid_for(nnz, 256), 8192);
    hipLaunchKernelGGL(assemble_kernel, dim3(grid), dim3(256), 0, s, nnz, asm_map_.get(), a_values,
                       panel_.get());
  }
  factorize_assembled(s, b);
}

SupernodalLDLT::~SupernodalLDLT() {
  for (auto* cache : {&fact_graphs_, &solve_graphs_, &bwd_graphs_})
    for (auto& g : *cache) (void)hipGraphExecDestroy(g.exec);
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

void SupernodalLDLT::factorize_assembled(hipStream_t s, const double* b) {
  if (sym_.n_levels == 0) return;
  fused_ = b != nullptr;
  HIP_OK(hipGraphLaunch(cached_graph(fact_graphs_, b, nullptr, s,
                                     [&] { launch_factorize(b, s); }), s));
}

void SupernodalLDLT::solve(const double* b, double* x, hipStream_t s) {
  if (sym_.n == 0) return;
  HIP_OK(hipGraphLaunch(cached_graph(solve_graphs_, b, x, s, [&] { launch_solve(b, x, s); }), s));
}

void SupernodalLDLT::solve_fused(double* x, hipStream_t s) {
  if (sym_.n == 0) return;
  SYMHIP_CHECK(fused_, "solve_fused needs a factorization given the right-hand side");
  HIP_OK(hipGraphLaunch(cached_graph(bwd_graphs_, nullptr, x, s, [&] { launch_backward(x, s); }),
                        s));
}

template <typename F>
hipGraphExec_t SupernodalLDLT::cached_graph(std::vector<CachedGraph>& cache, const double* b,
                                            double* x, hipStream_t s, F&& launches) {
  for (const auto& g : cache)
    if (g.b == b && g.x == x && g.f32 == f32_updates_) return g.exec;
  if (cache.size() >= 8) {  // bound the cache; the oldest goes
    (void)hipGraphExecDestroy(cache.front().exec);
    cache.erase(cache.begin());
  }
  cache.push_back({b, x, f32_updates_, capture(s, launches)});
  return cache.back().exec;
}

void SupernodalLDLT::launch_factorize(const double* b, hipStream_t s) {
  const SnTask* trsm = reinterpret_cast<const SnTask*>(trsm_tasks_.get());
  const SnTask* upd = reinterpret_cast<const SnTask*>(upd_tasks_.get());
  if (b)
    hipLaunchKernelGGL(permute_kernel, dim3(grid_for(sym_.n, 256)), dim3(256), 0, s, sym_.n,
                       perm_.get(), b, y_.get(), part_.get());
  for (int l = 0; l < sym_.n_levels; ++l) {
    const int nt = lvl_trsm_ptr_[l + 1] - lvl_trsm_ptr_[l];
    if (nt) {
      if (b)
        hipLaunchKernelGGL(diag_trsm_kernel<true>, dim3(nt), dim3(kDiagThreads), 0, s,
                           trsm + lvl_trsm_ptr_[l], panel_.get(),
                           l11_.get(), D_.get(), rows_.get(), y_.get(), yout_.get());
      else
        hipLaunchKernelGGL(diag_trsm_kernel<false>, dim3(nt), dim3(kDiagThreads), 0, s,
                           trsm + lvl_trsm_ptr_[l], panel_.get(),
                           l11_.get(), D_.get(), nullptr, nullptr, nullptr);
    }
    const int nu = lvl_upd_ptr_[l + 1] - lvl_upd_ptr_[l];
    if (nu) {
      auto kern = f32_updates_ ? update_kernel<kTile, float> : update_kernel<kTile, double>;
      hipLaunchKernelGGL(kern, dim3(nu), dim3(256), 0, s, upd + lvl_upd_ptr_[l], panel_.get(),
                         D_.get(), scat_col_.get(), scat_rel_.get(), upd_rel_.get());
    }
  }
  HIP_OK(hipGetLastError());
}

void SupernodalLDLT::launch_solve(const double* b, double* x, hipStream_t s) {
  const int n = sym_.n;
  const SnTask* tasks = reinterpret_cast<const SnTask*>(solve_tasks_.get());
  hipLaunchKernelGGL(permute_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, n, perm_.get(), b,
                     y_.get(), part_.get());
  for (int l = 0; l < sym_.n_levels; ++l) {
    const int nb = lvl_solve_ptr_[l + 1] - lvl_solve_ptr_[l];
    hipLaunchKernelGGL(fwd_kernel, dim3(nb), dim3(256), 0, s,
                       tasks + lvl_solve_ptr_[l], rows_.get(),
                       panel_.get(), l11_.get(), y_.get(), yout_.get());
  }
  launch_backward(x, s);
}

// The backward levels (D^-1 folded in) and the inverse permutation, from the forward result in
// yout.
void SupernodalLDLT::launch_backward(double* x, hipStream_t s) {
  const int n = sym_.n;
  const SnTask* tasks = reinterpret_cast<const SnTask*>(solve_tasks_.get());
  for (int l = sym_.n_levels - 1; l >= 0; --l) {
    const int nb = lvl_solve_ptr_[l + 1] - lvl_solve_ptr_[l];
    auto kern = nb > 2 * kCus ? bwd_kernel<true> : bwd_kernel<false>;
    hipLaunchKernelGGL(kern, dim3(nb), dim3(256), 0, s, tasks + lvl_solve_ptr_[l], rows_.get(),
                       panel_.get(), l11_.get(), D_.get(), yout_.get(), part_.get(),
                       counter_.get());
  }
  hipLaunchKernelGGL(unpermute_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, n, perm_.get(),
                     yout_.get(), x);
  HIP_OK(hipGetLastError());
}

void SupernodalLDLT::factor_values(const std::vector<int32_t>& l_colptr,
                                   const std::vector<int32_t>& l_rowind, double* values,
                                   hipStream_t s) const {
  std::vector<double> panels(sym_.panel_size()), l11(l11_.size());
  panel_.download(panels.data(), panels.size(), s);
  l11_.download(l11.data(), l11.size(), s);
  HIP_OK(hipStreamSynchronize(s));
  int64_t l11_off = 0;
  for (int32_t k = 0; k < sym_.n_snodes; ++k) {
    const int32_t c0 = sym_.sn_col_begin[k], c1 = sym_.sn_col_begin[k + 1], n = c1 - c0;
    const int32_t r0 = sym_.sn_rows_ptr[k], nrows = sym_.sn_rows_ptr[k + 1] - r0;
    const int32_t* rows = sym_.sn_rows.data() + r0;
    for (int32_t j = c0; j < c1; ++j) {
      const int32_t jj = j - c0;
      for (int32_t q = l_colptr[j]; q < l_colptr[j + 1]; ++q) {
        const int32_t r = l_rowind[q];
        if (r < c1) {
          values[q] = l11[l11_off + (int64_t)jj * n + (r - c0)];
        } else {
          const int32_t* pos = std::lower_bound(rows + n, rows + nrows, r);
          SYMHIP_CHECK(pos != rows + nrows && *pos == r,
                       "factor pattern entry outside the supernodal panel");
          values[q] = panels[sym_.sn_panel_off[k] + (int64_t)jj * nrows + (pos - rows)];
        }
      }
    }
    l11_off += (int64_t)n * n;
  }
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
