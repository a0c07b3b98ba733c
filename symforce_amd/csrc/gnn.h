// Learned block-Jacobi correction for PCG on the reduced camera system: a small message-passing
// network over the camera covisibility graph predicts, per camera, a lower-triangular U_a with a
// positive (exponential) diagonal, entries bounded through 2 tanh(o / 20); the preconditioner block is M_a = L_a^-T U_a U_a^T L_a^-1
// with S_aa = L_a L_a^T, symmetric positive definite for any output, and U = I (zero output) is
// exactly block Jacobi. Training minimizes the scale-invariant probe loss
// min_c |c M S X - X|^2 / |X|^2 = 1 - <M S X, X>^2 / (|M S X|^2 |X|^2) (the spread of the
// eigenvalues of M S seen by random probes X) on Schur systems logged from the optimizer's own
// solves. All GEMMs are
// BF16 matrix-core products with FP32 accumulation; the per-camera algebra is FP64.
#pragma once

#include <vector>

#include "./common.h"

namespace symhip {

class LearnedPrecond {
 public:
  static constexpr int kNd = 9, kF = 16, kH = 64, kO = 48, kTri = 45;
  static constexpr int kParams = (2 * kF) * kH + kH + 2 * ((2 * kH) * kH + kH) + kH * kO + kO;

  // Pattern of the full symmetric block-CSR matrix (9 x 9 blocks).
  void setup(int32_t n_nodes, const std::vector<int32_t>& row_ptr, const std::vector<int32_t>& col,
             const std::vector<int32_t>& diag, hipStream_t s);
  bool ready() const { return n_ > 0; }
  void set_params(const float* host, hipStream_t s);
  void get_params(float* host, hipStream_t s) const;
  void init_params(uint64_t seed, hipStream_t s);

  // Minv (n x 81, FP64) for the system S (device blocks).
  void preconditioner(const double* S, double* Minv, hipStream_t s);
  // Loss and gradient (device) for probes X (n * 9 x P, column-major per probe: X[p * n9 + i])
  // and Y = S X. Returns the loss (synchronizes).
  float loss_grad(const double* S, const double* X, const double* Y, int P, hipStream_t s);
  const float* grad() const { return grad_.get(); }
  // Standard normal probes: out[i] for i < n from (seed, i) by a counter-based hash + Box-Muller.
  static void probes(double* out, int64_t n, uint64_t seed, hipStream_t s);
  // One Adam step on the current gradient (step counts from 1).
  void adam(float lr, int step, hipStream_t s);

 private:
  void forward(const double* S, hipStream_t s);
  int32_t n_ = 0;
  int64_t nblk_ = 0;
  DBuf<int32_t> row_ptr_, col_, diag_, mirror_;
  DBuf<float> w_, x_, z1_, p1_, h1_, z2_, p2_, h2_, z3_, p3_, h3_, o_, do_, dh_, dz_, dp_;
  DBuf<float> params_, grad_, m_, v_;
  DBuf<double> nrm_, part_, red_;
};

}  // namespace symhip
