// Batched LM for small robot-3D-localization problems: one persistent launch, one wavefront per
// problem running the whole Levenberg-Marquardt loop (the state-block bookkeeping of
// LevenbergMarquardtSolver::Iterate, levenberg_marquardt_solver.tcc:138-343, with the shared
// accept / lambda logic of lm_logic.h).
//
// Structure exploited: the optimized keys are the poses of one chain, odometry factors join
// consecutive poses and matching factors touch one pose, so H is block tridiagonal with 6 x 6
// blocks. Lane i linearizes pose i (its matching factors and the odometry factor to pose i + 1,
// whose pose-(i + 1) terms are handed to lane i + 1 by a shuffle); the wavefront then solves the
// damped system by block-tridiagonal Cholesky, six lanes holding the rows of each 6 x 6 block. Per state block and pose the linearization keeps
// 63 doubles: the lower diagonal block (21), H(i + 1, i) (36, row-major) and the rhs (6).
#include "./batch.h"

#include <algorithm>
#include <cmath>
#include <mutex>
#include <type_traits>

#include "./factor_traits.h"
#include "./geo.h"
#include "./lm_logic.h"
#include "./problem.h"
#include "./util.h"

namespace symhip {

namespace {

constexpr int kLin = 63;
constexpr int kMaxPoses = 64;

struct BatchView {
  int64_t total_poses;
  const int32_t* pose_begin;  // [n_problems + 1] global pose index
  const int32_t* pose_soff;   // [total_poses] offset of the initial pose value in cvals
  const int32_t* match_ptr;   // [total_poses + 1]
  const int32_t* match_args;  // 3 per match: landmark, measurement, sigma offsets in cvals
  const int32_t* odo_args;    // 3 per pose: measurement, sigmas, epsilon offsets (-1: last pose)
  const double* cvals;
  double* vals;  // [3][total_poses][7]
  double* lin;   // [3][total_poses][kLin]
  double* maxd;  // [total_poses][6]
  double* vout;  // storage copy; best poses written back
  symhip_iteration_stats* recs;
  int32_t rec_stride;
  symhip_optimization_stats* stats;
  int64_t* ticks;  // [n_problems][2]
  const int32_t* order;  // workgroup -> problem (null: the identity)
};

__device__ __forceinline__ double wave_total(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ __forceinline__ int tri(int r, int c) { return r * (r + 1) / 2 + c; }

// Relinearize state block b: lane i < n writes pose i's D, H(i + 1, i), rhs. Returns the error
// 0.5 |r|^2 of the problem (the same value in every lane). With S = float (the precision form)
// the Jacobians come from the FP32 factor bodies while the residual, and with it the cost and
// the gradient J^T r, comes from the FP64 bodies: the cost stays FP64-accurate, so LM's gain
// ratio and relative-reduction exit behave as in FP64 (FP32 residuals put a ~1e-7 floor under
// the cost, which kept LM from its early exit on some problems). The products J^T J are formed
// in FP64 from the FP32 Jacobians, so H stays positive semi-definite (FP32 sums of products
// made a near-singular H slightly indefinite, and the block Cholesky then failed).
template <typename S>
__device__ double linearize(const BatchView& v, int b, int64_t g, int lane, bool live) {
  double D[21] = {}, gv[6] = {}, Bv[36] = {}, Dn[21] = {}, gn[6] = {};
  double e = 0;
  int has_next = 0;
  if (live) {
    const double* pose = v.vals + (b * v.total_poses + g) * 7;
    for (int m = v.match_ptr[g]; m < v.match_ptr[g + 1]; ++m) {
      const double* a[4] = {pose, v.cvals + v.match_args[3 * m], v.cvals + v.match_args[3 * m + 1],
                            v.cvals + v.match_args[3 * m + 2]};
      S rs[3], J[3][6];
      factor_matching<S>(a, rs, J);
      double r[3], Jd[3][6];
      if constexpr (std::is_same_v<S, double>) {
#pragma unroll
        for (int q = 0; q < 3; ++q) r[q] = rs[q];
      } else {
        factor_matching<double>(a, r, Jd);
      }
#pragma unroll
      for (int i = 0; i < 6; ++i) {
#pragma unroll
        for (int q = 0; q < 3; ++q) gv[i] += (double)J[q][i] * r[q];
#pragma unroll
        for (int j = 0; j <= i; ++j) {
          double h = 0;
#pragma unroll
          for (int q = 0; q < 3; ++q) h += (double)J[q][i] * (double)J[q][j];
          D[tri(i, j)] += h;
        }
      }
#pragma unroll
      for (int q = 0; q < 3; ++q) e += r[q] * r[q];
    }
    const int32_t* oa = v.odo_args + 3 * g;
    if (oa[0] >= 0) {
      has_next = 1;
      const double* a[5] = {pose, pose + 7, v.cvals + oa[0], v.cvals + oa[1], v.cvals + oa[2]};
      S rs[6], J[6][12];
      factor_odometry<S>(a, rs, J);
      double r[6], Jd[6][12];
      if constexpr (std::is_same_v<S, double>) {
#pragma unroll
        for (int q = 0; q < 6; ++q) r[q] = rs[q];
      } else {
        factor_odometry<double>(a, r, Jd);
      }
#pragma unroll
      for (int i = 0; i < 6; ++i) {
        double ga = 0, gb = 0;
#pragma unroll
        for (int q = 0; q < 6; ++q) {
          ga += (double)J[q][i] * r[q];
          gb += (double)J[q][6 + i] * r[q];
        }
        gv[i] += ga;
        gn[i] = gb;
#pragma unroll
        for (int j = 0; j < 6; ++j) {
          double hb = 0;
#pragma unroll
          for (int q = 0; q < 6; ++q) hb += (double)J[q][6 + i] * (double)J[q][j];
          Bv[i * 6 + j] = hb;  // H(i + 1, i): rows of pose i + 1, cols of pose i
        }
#pragma unroll
        for (int j = 0; j <= i; ++j) {
          double ha = 0, hn = 0;
#pragma unroll
          for (int q = 0; q < 6; ++q) {
            ha += (double)J[q][i] * (double)J[q][j];
            hn += (double)J[q][6 + i] * (double)J[q][6 + j];
          }
          D[tri(i, j)] += ha;
          Dn[tri(i, j)] = hn;
        }
      }
#pragma unroll
      for (int q = 0; q < 6; ++q) e += r[q] * r[q];
    }
  }
  // the odometry factor's pose-(i + 1) terms belong to the next lane
  const int prev = __shfl_up(has_next, 1, 64);
  const bool take = lane >= 1 && prev != 0;
#pragma unroll
  for (int k = 0; k < 21; ++k) {
    const double t = __shfl_up(Dn[k], 1, 64);
    if (take) D[k] += t;
  }
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    const double t = __shfl_up(gn[k], 1, 64);
    if (take) gv[k] += t;
  }
  if (live) {
    double* L = v.lin + (b * v.total_poses + g) * kLin;
#pragma unroll
    for (int k = 0; k < 21; ++k) L[k] = D[k];
#pragma unroll
    for (int k = 0; k < 36; ++k) L[21 + k] = Bv[k];
#pragma unroll
    for (int k = 0; k < 6; ++k) L[57 + k] = gv[k];
  }
  return 0.5 * wave_total(e);
}

// Value of v in lane k (k wave-uniform) through two scalar lane reads.
__device__ __forceinline__ double lane_bcast(double v, int k) {
  const int2 u = __builtin_bit_cast(int2, v);
  return __builtin_bit_cast(double, make_int2(__builtin_amdgcn_readlane(u.x, k),
                                              __builtin_amdgcn_readlane(u.y, k)));
}

__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Block-tridiagonal Cholesky H = L L^T (L block lower bidiagonal: C_i on the diagonal,
// W_i = B_i C_i^-T below) and the two triangular solves, by the whole wavefront: lane r < 6
// holds row r of the 6 x 6 blocks of the current pose (the other lanes shadow row 5), and the
// values another row needs cross by scalar lane reads, so a pose costs a few dozen dependent
// steps instead of the ~6 000 cycles of one lane doing every entry. Each entry is formed by the
// same operations in the same order as a one-lane sweep would form it. A: packed lower diagonal
// blocks (overwritten by C_i), B: H(i + 1, i) row-major (overwritten by W_i), x: rhs in,
// H^-1 rhs out (LDS, written by the row lanes, read back after a wavefront fence).
__device__ void block_tridiagonal_solve(int n, double (*A)[21], double (*B)[36], double (*x)[6],
                                        int lane) {
  const int r = lane < 6 ? lane : 5;
  const bool row = lane < 6;
  double wprev[6] = {}, yprev[6] = {};
  for (int i = 0; i < n; ++i) {
    // M = A_i - W_{i-1} W_{i-1}^T, row r (entries c <= r)
    double m[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) m[c] = A[i][tri(r, c <= r ? c : r)];
    if (i > 0) {
#pragma unroll
      for (int c = 0; c < 6; ++c) {
        double sum = 0;
#pragma unroll
        for (int k = 0; k < 6; ++k) sum += wprev[k] * B[i - 1][c * 6 + k];
        m[c] -= sum;
      }
    }
    // Cholesky of M; every entry of C_i ends up in Cb (uniform) on the way
    // (one reciprocal per column; the triangular solves below multiply by it)
    double Cb[6][6], Ci[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      const double ckk = sqrt(lane_bcast(m[k], k));
      Cb[k][k] = ckk;
      Ci[k] = 1.0 / ckk;
      const double lrk = m[k] * Ci[k];  // row r's entry of column k (r > k)
#pragma unroll
      for (int c = k + 1; c < 6; ++c) Cb[c][k] = lane_bcast(lrk, c);
      if (r > k) m[k] = lrk;
      else if (r == k) m[k] = ckk;
#pragma unroll
      for (int c = k + 1; c < 6; ++c)
        if (c <= r && r > k) m[c] -= lrk * Cb[c][k];
    }
    if (row) {
#pragma unroll
      for (int c = 0; c < 6; ++c)
        if (c <= r) A[i][tri(r, c)] = m[c];
    }
    // W_i row r = B_i row r C_i^-T, and y_i = C_i^-1 (x_i - W_{i-1} y_{i-1})
    double w[6];
    if (i + 1 < n) {
#pragma unroll
      for (int c = 0; c < 6; ++c) {
        double v = B[i][r * 6 + c];
#pragma unroll
        for (int k = 0; k < c; ++k) v -= w[k] * Cb[c][k];
        w[c] = v * Ci[c];
      }
    }
    double b = x[i][r];
    if (i > 0) {
#pragma unroll
      for (int k = 0; k < 6; ++k) b -= wprev[k] * yprev[k];
    }
    double bu[6], y[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) bu[k] = lane_bcast(b, k);
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      double v = bu[q];
#pragma unroll
      for (int k = 0; k < q; ++k) v -= Cb[q][k] * y[k];
      y[q] = v * Ci[q];
    }
    if (row) {
      x[i][r] = y[r < 6 ? r : 5];
      if (i + 1 < n) {
#pragma unroll
        for (int c = 0; c < 6; ++c) B[i][r * 6 + c] = w[c];
      }
    }
    wave_lds_sync();  // W_i is read by all rows at pose i + 1
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      wprev[k] = w[k];
      yprev[k] = y[k];
    }
  }
  // x_i = C_i^-T (y_i - W_i^T x_{i+1})
  double xnext[6] = {};
  for (int i = n - 1; i >= 0; --i) {
    double b = x[i][r];
    if (i + 1 < n) {
#pragma unroll
      for (int k = 0; k < 6; ++k) b -= B[i][k * 6 + r] * xnext[k];
    }
    double bu[6], xs[6], di[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      bu[k] = lane_bcast(b, k);
      di[k] = 1.0 / A[i][tri(k, k)];  // off the substitution's dependent chain
    }
#pragma unroll
    for (int q = 5; q >= 0; --q) {
      double v = bu[q];
#pragma unroll
      for (int k = q + 1; k < 6; ++k) v -= A[i][tri(k, q)] * xs[k];
      xs[q] = v * di[q];
    }
    if (row) x[i][r] = xs[r < 6 ? r : 5];
#pragma unroll
    for (int k = 0; k < 6; ++k) xnext[k] = xs[k];
  }
  wave_lds_sync();
}

// kWaves: wavefronts per SIMD the register allocation must allow (1: no spills, the linearization
// holds ~350 VGPRs; 2: twice the resident problems at the cost of ~120 spilled registers); the
// batch autotuner picks per batch.
template <typename S, int kWaves>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(kWaves)))
batch_lm_kernel(BatchView v, symhip_optimizer_params p, double eps) {
  __shared__ double sA[kMaxPoses][21];
  __shared__ double sB[kMaxPoses][36];
  __shared__ double sx[kMaxPoses][6];
  const int prob = v.order ? v.order[blockIdx.x] : (int)blockIdx.x, lane = threadIdx.x;
  const int64_t t_start = wall_clock64();
  const int g0 = v.pose_begin[prob], n = v.pose_begin[prob + 1] - g0;
  const bool live = lane < n;
  const int64_t g = g0 + min(lane, n - 1);
  const int64_t TP = v.total_poses;
  // Reset: the start Values go into the New block (block 1)
  if (live)
    for (int k = 0; k < 7; ++k) v.vals[(TP + g) * 7 + k] = v.cvals[v.pose_soff[g] + k];
  // the first linearization reads the next lane's start pose (odometry), written just above
  __syncthreads();
  int init = 0, nw = 1, best = 0, fre = 2, iteration = -1, best_index = 0;
  int status = 0, reason = 0, linearized = 0, nrec = 0;
  bool have_max = false, done = false;
  double err[3] = {0, 0, 0};
  double lambda = p.initial_lambda, nu = p.dynamic_lambda_update_beta;
  auto record = [&](const symhip_iteration_stats& r) {
    if (lane == 0 && v.recs && nrec < v.rec_stride)
      v.recs[(int64_t)prob * v.rec_stride + nrec] = r;
    ++nrec;
  };
  auto relin = [&](int b) {
    err[b] = linearize<S>(v, b, g, lane, live);
    linearized |= 1 << b;
  };
  for (int it = 0; it < p.iterations && !done; ++it) {
    const int t = init;
    init = nw;
    nw = t;
    ++iteration;
    if (!((linearized >> init) & 1)) {
      relin(init);
      if (best != init) {
        if (best != nw) fre = best;
        best = init;
      }
    }
    if (iteration == 0) {
      symhip_iteration_stats r{};
      r.iteration = -1;
      r.new_error = err[init];
      r.current_lambda = lambda;
      record(r);
      if (!isfinite(err[init])) {
        status = SYMHIP_STATUS_FAILED;
        reason = SYMHIP_FAILURE_INITIAL_ERROR_NOT_FINITE;
        break;
      }
    }
    // DampHessian (levenberg_marquardt_solver.tcc:22-54) out of place, and the staged system
    const double* Li = v.lin + (init * TP + g) * kLin;
    double damp[6] = {}, u[6] = {};
    if (live) {
      for (int k = 0; k < 6; ++k) {
        const double dk = Li[tri(k, k)];
        double d = 0;
        if (p.use_diagonal_damping) {
          if (p.keep_max_diagonal_damping) {
            const double md = have_max ? fmax(v.maxd[g * 6 + k], dk) : fmax(dk, p.diagonal_damping_min);
            v.maxd[g * 6 + k] = md;
            d = md * lambda;
          } else {
            d = fmax(dk, p.diagonal_damping_min) * lambda;
          }
        }
        if (p.use_unit_damping) d += lambda;
        damp[k] = d;
      }
      for (int k = 0; k < 21; ++k) sA[lane][k] = Li[k];
      for (int k = 0; k < 6; ++k) sA[lane][tri(k, k)] += damp[k];
      for (int k = 0; k < 36; ++k) sB[lane][k] = Li[21 + k];
      for (int k = 0; k < 6; ++k) sx[lane][k] = Li[57 + k];
    }
    if (p.use_diagonal_damping && p.keep_max_diagonal_damping) have_max = true;
    __syncthreads();
    block_tridiagonal_solve(n, sA, sB, sx, lane);
    __syncthreads();
    double lde = 0;
    if (live) {
      for (int k = 0; k < 6; ++k) {
        u[k] = -sx[lane][k];
        lde += u[k] * (Li[57 + k] - damp[k] * u[k]);
      }
      // UpdateNewFromInit: New = Init retracted by the update
      const double* src = v.vals + (init * TP + g) * 7;
      double* dst = v.vals + (nw * TP + g) * 7;
      for (int k = 0; k < 7; ++k) dst[k] = src[k];
      rot3_retract_inplace(dst, u, eps);
      dst[4] += u[3];
      dst[5] += u[4];
      dst[6] += u[5];
    }
    __syncthreads();  // sA/sB/sx are rewritten next iteration
    relin(nw);
    lde = wave_total(lde);
    const LmDecision d = lm_decide(err[init], err[nw], err[best], lde, lambda, nu, iteration, p,
                                   0, 0.0, 0.0, 0.0, eps);
    record(d.rec);
    lambda = d.lambda;
    nu = d.nu;
    if (!d.accept) {
      const int s = nw;
      nw = init;
      init = s;
    } else {
      if (d.set_best) {
        if (best != nw) {
          if (best != init) fre = best;
          best = nw;
        }
        best_index = nrec - 1;
      }
      if (best == init) {
        init = fre;
        fre = best;
      }
    }
    status = d.status;
    reason = d.reason;
    done = status != 0;
  }
  if (!done && status == 0) status = SYMHIP_STATUS_HIT_ITERATION_LIMIT;
  if (live)
    for (int k = 0; k < 7; ++k) v.vout[v.pose_soff[g] + k] = v.vals[(best * TP + g) * 7 + k];
  if (lane == 0) {
    v.stats[prob] = symhip_optimization_stats{nrec, best_index, status, reason};
    v.ticks[2 * prob] = t_start;
    v.ticks[2 * prob + 1] = wall_clock64();
  }
}

}  // namespace

PackedBatch pack_batch(const symhip_problem_desc* descs, int32_t n_problems,
                       const symhip_optimizer_params& params) {
  SYMHIP_CHECK(descs && n_problems > 0, "batch: no problems");
  SYMHIP_CHECK(params.iterations > 0, "batch: iterations must be positive");
  SYMHIP_CHECK(!params.enable_bold_updates, "batch: bold updates are not supported");
  PackedBatch pk;
  pk.pose_begin.push_back(0);
  pk.match_ptr.push_back(0);
  for (int32_t pi = 0; pi < n_problems; ++pi) {
    HostProblem hp;
    build_host_problem(&descs[pi], hp);
    const int32_t n = (int32_t)hp.state_keys.size();
    SYMHIP_CHECK(n >= 1 && n <= kMaxPoses, "batch: each problem needs 1..64 optimized poses");
    std::vector<int32_t> pos_of(hp.n_keys, -1);
    for (int32_t i = 0; i < n; ++i) {
      SYMHIP_CHECK(hp.types[hp.state_keys[i]] == SYMHIP_TYPE_POSE3,
                   "batch: optimized keys must be Pose3");
      pos_of[hp.state_keys[i]] = i;
    }
    const int64_t base = (int64_t)pk.cvals.size();
    pk.cvals.insert(pk.cvals.end(), hp.values.begin(), hp.values.end());
    SYMHIP_CHECK(pk.cvals.size() < (size_t)INT32_MAX, "batch: storage too large");
    auto off = [&](int32_t key) { return (int32_t)(base + hp.storage_offset[key]); };
    auto constant = [&](int32_t key) {
      SYMHIP_CHECK(pos_of[key] < 0, "batch: factor constants must not be optimized");
      return off(key);
    };
    std::vector<std::vector<int32_t>> matches(n);
    std::vector<int32_t> odo(3 * (size_t)n, -1);
    for (int32_t f = 0; f < hp.n_factors; ++f) {
      const int32_t* a = hp.fargs.data() + hp.fargs_ptr[f];
      if (hp.ftype[f] == SYMHIP_FACTOR_MATCHING) {
        const int32_t i = pos_of[a[0]];
        SYMHIP_CHECK(i >= 0, "batch: matching factor on a constant pose");
        matches[i].insert(matches[i].end(), {constant(a[1]), constant(a[2]), constant(a[3])});
      } else if (hp.ftype[f] == SYMHIP_FACTOR_ODOMETRY) {
        const int32_t i = pos_of[a[0]], j = pos_of[a[1]];
        SYMHIP_CHECK(i >= 0 && j == i + 1,
                     "batch: odometry factors must join consecutive optimized poses (a, a + 1)");
        SYMHIP_CHECK(odo[3 * i] < 0, "batch: two odometry factors leave the same pose");
        odo[3 * i] = constant(a[2]);
        odo[3 * i + 1] = constant(a[3]);
        odo[3 * i + 2] = constant(a[4]);
      } else {
        throw Error("batch: only matching and odometry factors are supported");
      }
    }
    for (int32_t i = 0; i < n; ++i) {
      pk.pose_soff.push_back(off(hp.state_keys[i]));
      pk.match_args.insert(pk.match_args.end(), matches[i].begin(), matches[i].end());
      pk.match_ptr.push_back((int32_t)(pk.match_args.size() / 3));
      pk.odo_args.insert(pk.odo_args.end(), odo.begin() + 3 * i, odo.begin() + 3 * i + 3);
    }
    pk.pose_begin.push_back(pk.pose_begin.back() + n);
  }
  return pk;
}

BatchLocalization::BatchLocalization(const symhip_problem_desc* descs, int32_t n_problems,
                                     const symhip_optimizer_params& params, bool compute_f32,
                                     double epsilon)
    : n_problems_(n_problems), p_(params), f32_(compute_f32), eps_(epsilon) {
  PackedBatch pk = pack_batch(descs, n_problems, params);
  std::vector<int32_t>& pose_begin = pk.pose_begin;
  std::vector<double>& cvals = pk.cvals;
  total_poses_ = pose_begin.back();
  cvals_host_size_ = cvals.size();
  HIP_OK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
  HIP_OK(hipEventCreate(&ev0_));
  HIP_OK(hipEventCreate(&ev1_));
  pose_begin_.upload(pose_begin, stream_);
  pose_soff_.upload(pk.pose_soff, stream_);
  match_ptr_.upload(pk.match_ptr, stream_);
  match_args_.upload(pk.match_args.empty() ? std::vector<int32_t>{0} : pk.match_args, stream_);
  odo_args_.upload(pk.odo_args, stream_);
  cvals_.upload(cvals, stream_);
  vout_.upload(cvals, stream_);
  vals_.resize((size_t)3 * total_poses_ * 7);
  lin_.resize((size_t)3 * total_poses_ * kLin);
  maxd_.resize((size_t)total_poses_ * 6);
  recs_.resize((size_t)n_problems_ * (p_.iterations + 1));
  stats_.resize(n_problems_);
  ticks_.resize(2 * (size_t)n_problems_);
  {
    std::vector<int32_t> order(n_problems_);
    for (int32_t i = 0; i < n_problems_; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) {
      return pose_begin[a + 1] - pose_begin[a] > pose_begin[b + 1] - pose_begin[b];
    });
    order_.upload(order, stream_);
  }
  HIP_OK(hipStreamSynchronize(stream_));
  build_graph(config_);
}

BatchLocalization::~BatchLocalization() {
  for (auto& g : graphs_)
    if (g) (void)hipGraphExecDestroy(g);
  if (ev0_) (void)hipEventDestroy(ev0_);
  if (ev1_) (void)hipEventDestroy(ev1_);
  if (stream_) (void)hipStreamDestroy(stream_);
}

// Launch configuration c: bit 0 = two wavefronts per SIMD (else one), bit 1 = problems with
// the most poses first (else in the given order).
void BatchLocalization::build_graph(int c) {
  if (graphs_[c]) return;
  BatchView v{total_poses_,      pose_begin_.get(), pose_soff_.get(), match_ptr_.get(),
              match_args_.get(), odo_args_.get(),   cvals_.get(),     vals_.get(),
              lin_.get(),        maxd_.get(),       vout_.get(),      recs_.get(),
              p_.iterations + 1, stats_.get(),      ticks_.get(),
              (c & 2) ? order_.get() : nullptr};
  hipGraph_t graph;
  HIP_OK(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
  auto launch = [&](auto kernel) {
    hipLaunchKernelGGL(kernel, dim3(n_problems_), dim3(64), 0, stream_, v, p_, eps_);
  };
  if (f32_) {
    if (c & 1) launch(batch_lm_kernel<float, 2>); else launch(batch_lm_kernel<float, 1>);
  } else {
    if (c & 1) launch(batch_lm_kernel<double, 2>); else launch(batch_lm_kernel<double, 1>);
  }
  HIP_OK(hipStreamEndCapture(stream_, &graph));
  HIP_OK(hipGraphInstantiate(&graphs_[c], graph, nullptr, nullptr, 0));
  HIP_OK(hipGraphDestroy(graph));
}

// The autotuner: every configuration solves the batch `reps` times (the solves are complete and
// give the same results: the configurations change occupancy and order, not the arithmetic) and
// the one with the shortest launch is kept for later solves.
void BatchLocalization::tune(int reps, double* config_ms) {
  SYMHIP_CHECK(reps >= 1, "batch: tune needs reps >= 1");
  double best = 0;
  for (int c = 0; c < kConfigs; ++c) {
    build_graph(c);
    double ms = 0;
    for (int r = 0; r < reps; ++r) {
      HIP_OK(hipEventRecord(ev0_, stream_));
      HIP_OK(hipGraphLaunch(graphs_[c], stream_));
      HIP_OK(hipEventRecord(ev1_, stream_));
      HIP_OK(hipStreamSynchronize(stream_));
      float t = 0;
      HIP_OK(hipEventElapsedTime(&t, ev0_, ev1_));
      ms = r == 0 ? t : std::min<double>(ms, t);
    }
    if (config_ms) config_ms[c] = ms;
    if (c == 0 || ms < best) {
      best = ms;
      config_ = c;
    }
  }
}

void BatchLocalization::solve(double* values_out, symhip_optimization_stats* stats_out,
                              symhip_iteration_stats* iters_out, double* solve_ms,
                              double* total_ms) {
  HIP_OK(hipEventRecord(ev0_, stream_));
  HIP_OK(hipGraphLaunch(graphs_[config_], stream_));
  HIP_OK(hipEventRecord(ev1_, stream_));
  if (values_out) vout_.download(values_out, cvals_host_size_, stream_);
  if (stats_out) stats_.download(stats_out, n_problems_, stream_);
  if (iters_out) recs_.download(iters_out, recs_.size(), stream_);
  std::vector<int64_t> ticks;
  if (solve_ms) {
    ticks.resize(ticks_.size());
    ticks_.download(ticks.data(), ticks.size(), stream_);
  }
  HIP_OK(hipStreamSynchronize(stream_));
  if (total_ms) {
    float ms = 0;
    HIP_OK(hipEventElapsedTime(&ms, ev0_, ev1_));
    *total_ms = ms;
  }
  if (solve_ms) {
    int dev = 0, khz = 0;
    HIP_OK(hipGetDevice(&dev));
    HIP_OK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
    SYMHIP_CHECK(khz > 0, "batch: wall clock rate unavailable");
    for (int32_t i = 0; i < n_problems_; ++i)
      solve_ms[i] = (double)(ticks[2 * i + 1] - ticks[2 * i]) / (double)khz;
  }
}

// ---------------------------------------------------------------------------------------------
// Allocation-free solve: the kernel's working state (three state blocks of poses and their
// linearizations, the running maximum of the diagonal and the per-problem clock stamps) lives in
// a file-scope device array instead of buffers allocated per batch. Every call on a device shares
// that device's array, so calls that overlap on one device race. Nothing in it has to start at
// zero: each launch writes every entry it later reads (state block 1 from the inputs, the
// linearizations before they are read, the maximum diagonal on its first damping).
namespace {

constexpr int64_t kPerPoseScratch = 3 * 7 + 3 * kLin + 6;  // state blocks, linearizations, max diag
constexpr int64_t kScratchDoubles =
    SYMHIP_BATCH_NOALLOC_MAX_POSES * kPerPoseScratch + 2 * SYMHIP_BATCH_NOALLOC_MAX_POSES;

static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// The address of the current device's copy of `scratch`, looked up once per device.
double* device_scratch() {
  int dev = 0;
  HIP_OK(hipGetDevice(&dev));
  SYMHIP_CHECK(dev >= 0 && dev < kMaxDevices, "batch: device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    HIP_OK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    SYMHIP_CHECK(p != nullptr, "batch: scratch symbol has no device address");
    g_scratch_ptr[dev] = p;
  }
  return static_cast<double*>(g_scratch_ptr[dev]);
}

__global__ void copy_storage_kernel(int64_t n, const double* __restrict__ src,
                                    double* __restrict__ dst) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

}  // namespace

void batch_solve_noalloc(const symhip_batch_tables& t, const symhip_optimizer_params& p,
                         int32_t precision, double epsilon, double* values_out,
                         symhip_optimization_stats* stats_out, symhip_iteration_stats* iters_out,
                         hipStream_t stream) {
  SYMHIP_CHECK(t.n_problems > 0 && t.total_poses >= t.n_problems && t.storage_size > 0,
               "batch_noalloc: empty or inconsistent tables");
  SYMHIP_CHECK(t.total_poses <= SYMHIP_BATCH_NOALLOC_MAX_POSES,
               "batch_noalloc: more poses than the scratch array holds "
               "(SYMHIP_BATCH_NOALLOC_MAX_POSES)");
  SYMHIP_CHECK(t.pose_begin && t.pose_soff && t.match_ptr && t.odo_args && t.values &&
                   values_out && stats_out,
               "batch_noalloc: null table or output pointer");
  SYMHIP_CHECK(p.iterations > 0, "batch: iterations must be positive");
  SYMHIP_CHECK(!p.enable_bold_updates, "batch: bold updates are not supported");
  SYMHIP_CHECK(precision == SYMHIP_PRECISION_F64 || precision == SYMHIP_PRECISION_F32_JAC,
               "batch_noalloc: unknown precision");
  const int64_t TP = t.total_poses;
  double* base = device_scratch();
  double* vals = base;
  double* lin = vals + 3 * TP * 7;
  double* maxd = lin + 3 * TP * kLin;
  int64_t* ticks = reinterpret_cast<int64_t*>(maxd + TP * 6);
  // the output storage starts as the inputs; the kernel writes each problem's best poses into it
  hipLaunchKernelGGL(copy_storage_kernel, dim3((unsigned)((t.storage_size + 255) / 256)),
                     dim3(256), 0, stream, t.storage_size, t.values, values_out);
  BatchView v{TP,          t.pose_begin, t.pose_soff, t.match_ptr,        t.match_args,
              t.odo_args,  t.values,     vals,        lin,                maxd,
              values_out,  iters_out,    p.iterations + 1, stats_out,     ticks, nullptr};
  if (precision == SYMHIP_PRECISION_F32_JAC)
    hipLaunchKernelGGL((batch_lm_kernel<float, 1>), dim3(t.n_problems), dim3(64), 0, stream, v,
                       p, epsilon);
  else
    hipLaunchKernelGGL((batch_lm_kernel<double, 1>), dim3(t.n_problems), dim3(64), 0, stream, v,
                       p, epsilon);
  HIP_OK(hipGetLastError());
}

}  // namespace symhip
