// Host-side problem model: keys, Values storage, state index, per-factor linearization helpers
// and the hessian_lower CSC pattern, all laid out as the reference Linearizer builds them.
#pragma once

#include <cstdint>
#include <vector>

#include "./common.h"

namespace symhip {

struct HostProblem {
  // keys in Values append order
  int32_t n_keys = 0;
  std::vector<int8_t> letters;
  std::vector<int64_t> subs, supers;
  std::vector<int32_t> types, storage_offset;
  int64_t storage_total = 0;
  std::vector<double> values;

  // factors in the caller's order
  int32_t n_factors = 0;
  std::vector<int32_t> ftype, fargs_ptr, fargs;

  // state index (Linearizer::BuildInitialLinearization, opt/linearizer.cc:150-158)
  std::vector<int32_t> state_keys;  // key indices in state order
  std::vector<int32_t> tan_off;     // per key: tangent offset, -1 when not optimized
  int32_t N = 0, M = 0;

  // per factor: residual offset, and the reference's hessian_storage_col_starts flattened
  // (opt/internal/linearizer_utils.h:98-143) with cs_ptr[f] the first entry of factor f.
  std::vector<int32_t> res_off, cs_ptr, col_starts;

  // hessian_lower pattern (compressed CSC with full diagonal, rows sorted)
  std::vector<int32_t> colptr, rowind;
  int64_t nnz() const { return (int64_t)rowind.size(); }

  // key-level adjacency (lower): for each state position k, coupled state positions > k
  std::vector<std::vector<int32_t>> adj;
};

// Builds everything above from the boundary description; throws symhip::Error on bad input.
void build_host_problem(const symhip_problem_desc* d, HostProblem& p);

// The sparse Jacobian (M x N, CSC, rows sorted) of the reference's linearizer under
// include_jacobians (linearizer.cc:475-538 triplets, one D x td block per factor and optimized
// key), and where each factor's columns land in it: for factor f and column t of its tangent
// layout (linearized arguments in order, T = FactorSpec::tdim columns), starts[sptr[f] + t] is
// the position of row res_off[f] of that column in the value array (its D rows follow), or -1
// when the argument is not optimized.
struct JacobianStructure {
  std::vector<int32_t> colptr, rowind, sptr;
  std::vector<int64_t> starts;
};
void build_jacobian_structure(const HostProblem& p, JacobianStructure& j);

}  // namespace symhip
