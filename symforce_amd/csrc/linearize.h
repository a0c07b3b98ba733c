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
  int64_t contrib_off = 0;  // first entry of the group's local outputs in the contribution buffer
};

// Linearizer::Relinearize (opt/linearizer.cc:57-112) on the device, without atomics. Pass 1, one
// thread per factor and a launch per factor type: the residual into its rows and the factor's
// local rhs J^T r and Hessian blocks J^T J (lower diagonal blocks, full off-diagonal blocks) into
// a per-factor contribution record. Pass 2, one thread per rhs / hessian_lower entry: the sum of
// its contributions, gathered through a host-built list in factor order (the order in which the
// reference's UpdateFromLinearizedDenseFactorIntoSparse adds them, linearizer.cc:358-433).
class CscLinearizer {
 public:
  void setup(const HostProblem& p, hipStream_t s);
  // values: device Values storage. Any output may be null (e.g. residual-only evaluation).
  void run(const double* values, double* residual, double* hval, double* rhs, hipStream_t s) const;
  int64_t nnz = 0;
  int32_t N = 0, M = 0;

 private:
  std::vector<FactorGroup> groups_;
  DBuf<double> contrib_;
  DBuf<int32_t> hptr_, hidx_, rptr_, ridx_;  // gather lists: CSR over hval and rhs entries
};

}  // namespace symhip
