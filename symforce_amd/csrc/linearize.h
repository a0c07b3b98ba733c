// Batched per-factor-type linearization into the reference's hessian_lower CSC layout.
#pragma once

#include <vector>

#include "./common.h"
#include "./problem.h"

namespace symhip {

struct FactorGroup {
  int type = 0;
  int32_t n = 0;
  DBuf<int32_t> arg_off;  // [NA][n] storage offsets of each argument (SoA over factors)
  DBuf<int32_t> res_off;  // [n]
  DBuf<int32_t> tan_off;  // [NL][n] tangent offsets of linearized args (-1: not optimized)
  DBuf<int32_t> cs_ptr;   // [n] first col-start of the factor
};

// Linearizer::Relinearize (opt/linearizer.cc:57-112) on the device: zero rhs/H, evaluate all
// factors of each type in one launch, scatter residual / rhs / H through the reference's storage
// maps with FP64 atomics.
class CscLinearizer {
 public:
  void setup(const HostProblem& p, hipStream_t s);
  // values: device Values storage. Any output may be null (e.g. residual-only evaluation).
  void run(const double* values, double* residual, double* hval, double* rhs, hipStream_t s) const;
  int64_t nnz = 0;
  int32_t N = 0, M = 0;

 private:
  std::vector<FactorGroup> groups_;
  DBuf<int32_t> col_starts_;
};

}  // namespace symhip
