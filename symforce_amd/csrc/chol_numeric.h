// Device numeric supernodal LDL^T factorization and solves (see chol_symbolic.h for the layout).
#pragma once

#include <vector>

#include "./chol_symbolic.h"
#include "./common.h"

namespace symhip {

// Replaces SparseCholeskySolver::Factorize / Solve (opt/sparse_cholesky/
// sparse_cholesky_solver.tcc:108-259): A = P^T L D L^T P with unit lower L, computed by dense
// tile kernels over supernodal panels, level by level in the supernodal elimination tree.
class SupernodalLDLT {
 public:
  // colptr/rowind: lower CSC pattern (host) of the matrices that will be factorized.
  void setup(const CholSymbolic& sym, const int32_t* colptr, const int32_t* rowind,
             hipStream_t s);
  // Assembles A (device values in the CSC order given to setup) into panels and factorizes.
  // With b (device, original order) the forward solve L y = P b runs inside the factorization's
  // diagonal kernels, and solve_fused finishes x = A^-1 b.
  void factorize(const double* a_values, hipStream_t s, const double* b = nullptr);
  // Panels already assembled by the caller (e.g. the Schur path writes them directly). The
  // per-level launch sequence is captured in a hipGraph per b on the first call and replayed.
  void factorize_assembled(hipStream_t s, const double* b = nullptr);
  // x = A^-1 b for device vectors in the original (unpermuted) order; x may alias b.
  // Captured per (b, x) pair and replayed.
  void solve(const double* b, double* x, hipStream_t s);
  // x = A^-1 b for the b given to the last factorization (backward half only); x may alias b.
  void solve_fused(double* x, hipStream_t s);
  SupernodalLDLT() = default;
  SupernodalLDLT(const SupernodalLDLT&) = delete;
  SupernodalLDLT& operator=(const SupernodalLDLT&) = delete;
  ~SupernodalLDLT();

  // Mixed-precision form: the trailing supernode updates (update_kernel) run on the FP32 matrix
  // cores; diagonal blocks, panel solves, D and the triangular solves stay FP64. The factor is
  // then accurate to FP32 in the updates, and callers refine solves against the FP64 matrix.
  // May change between factorizations (each form's launch sequence is captured on first use).
  void set_f32_updates(bool on) { f32_updates_ = on; }
  bool f32_updates() const { return f32_updates_; }
  double* panels() const { return panel_.get(); }
  const CholSymbolic& symbolic() const { return sym_; }
  // Downloads the panels (L21 below the diagonal blocks; the diagonal blocks keep A11) and D.
  void download_factor(std::vector<double>& panels, std::vector<double>& d, hipStream_t s) const;
  // The factor's values on the reference's scalar pattern (scalar_factor_pattern of the same
  // permutation): strictly lower unit L, from the factored diagonal blocks and the panels.
  void factor_values(const std::vector<int32_t>& l_colptr, const std::vector<int32_t>& l_rowind,
                     double* values, hipStream_t s) const;
  int64_t assembly_nnz() const { return (int64_t)asm_map_.size(); }

 private:
  void launch_factorize(const double* b, hipStream_t s);
  void launch_solve(const double* b, double* x, hipStream_t s);
  void launch_backward(double* x, hipStream_t s);
  struct CachedGraph {
    const double* b;
    double* x;
    bool f32;  // factorization graphs: the update form captured
    hipGraphExec_t exec;
  };
  template <typename F>
  hipGraphExec_t cached_graph(std::vector<CachedGraph>& cache, const double* b, double* x,
                              hipStream_t s, F&& launches);
  // captured launch sequences, one per argument pair seen (a few state blocks)
  std::vector<CachedGraph> fact_graphs_, solve_graphs_, bwd_graphs_;
  bool fused_ = false;  // the last factorization ran the forward solve
  bool f32_updates_ = false;
  CholSymbolic sym_;
  DBuf<double> panel_, l11_, D_, y_, yout_, part_;
  DBuf<int32_t> counter_;  // per-supernode arrival counters of the backward pass (self-resetting)
  DBuf<int64_t> asm_map_, scat_col_;
  DBuf<int32_t> rows_, perm_, upd_rel_, scat_rel_;
  DBuf<uint8_t> trsm_tasks_, upd_tasks_, solve_tasks_;  // per-level task lists (chol_numeric.hip)
  // work lists (flattened over levels)
  std::vector<int32_t> lvl_snode_ptr_, lvl_trsm_ptr_, lvl_upd_ptr_, lvl_solve_ptr_;
  DBuf<int32_t> lvl_snodes_;
};

}  // namespace symhip
