// Preconditioned conjugate gradients on a block-sparse symmetric positive definite matrix with
// uniform ND x ND blocks (the reduced camera system of the Schur path), entirely on the device.
#pragma once

#include <vector>

#include "./common.h"

namespace symhip {

struct PcgResult {
  int iterations = 0;
  bool converged = false;
  bool breakdown = false;   // p.q <= 0 or a non-finite sum: S is not SPD to working precision
  double rel_residual = 0;  // ||b - S x|| / ||b|| by the CG recurrence
};

// Block-Jacobi PCG. Storage: full symmetric block-CSR (both triangles), each block row-major.
// The preconditioner is the inverse of each diagonal block, optionally corrected by a caller-
// supplied per-node ND x ND additive term (the learned correction).
class BlockPcg {
 public:
  // row_ptr (n + 1) / col: block pattern of the full matrix; diag[a] = block index of (a, a).
  // f32: the blocks are stored in FP32 (blocks32(); the sums stay FP64), else FP64 (blocks()).
  void setup(int nd, int32_t n_nodes, const std::vector<int32_t>& row_ptr,
             const std::vector<int32_t>& col, const std::vector<int32_t>& diag, hipStream_t s,
             bool f32 = false);
  double* blocks() const { return S_.get(); }
  float* blocks32() const { return S32_.get(); }
  double* precond() const { return Minv_.get(); }
  int64_t n_blocks() const { return n_blk_; }
  // Minv[a] = inverse(S_aa) (+ correction[a] if given, device, n * nd * nd).
  void build_preconditioner(const double* correction, hipStream_t s);
  // x = S^-1 b to ||r|| <= rel_tol ||b|| or max_iters. b, x: device, n * nd. Synchronizes the
  // stream every check_every iterations (and at the end) to read the convergence flag.
  // Iterations run in captured chunks of kChunk; the host reads the stop flag between chunks.
  PcgResult solve(const double* b, double* x, int max_iters, double rel_tol, hipStream_t s,
                  int check_every = kChunk);
  static constexpr int kChunk = 16;  // even: a chunk ends on slot 1 and the next starts on 0
  // y = S x for FP64 blocks S of this pattern (device vectors), for checks and step-error
  // measurement.
  void multiply_with(const double* S, const double* x, double* y, hipStream_t s);

 private:
  int nd_ = 0;
  int32_t n_ = 0;
  int64_t n_blk_ = 0;
  int rows_per_block_ = 0, n_upd_blocks_ = 0;
  DBuf<int32_t> row_ptr_, col_, diag_;
  bool f32_ = false;
  DBuf<double> S_, Minv_, r_, z_, p_[2], q_, partA_, partB_, scal_;
  DBuf<float> S32_;
  double* host_scal_ = nullptr;  // pinned
  hipGraphExec_t chunk_ = nullptr;  // kChunk iterations, captured for the output vector x
  double* chunk_x_ = nullptr;
  void capture_chunk(double* x, hipStream_t s);
 public:
  BlockPcg() = default;
  BlockPcg(const BlockPcg&) = delete;
  BlockPcg& operator=(const BlockPcg&) = delete;
  ~BlockPcg();
};

}  // namespace symhip
