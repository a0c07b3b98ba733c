// Oracle (test infrastructure only) -- CPU FP64 restatement of SymForce's optimizer hot path.
//
// Only the tests, __graft_entry__.smoke() and the cpu_baseline leg of bench.py load this library, and
// only as the checker / the timed CPU baseline. The shipped library (symforce_amd) never links it.
//
// Every function cites the reference file it restates (paths relative to the reference root).
// Parity status of each part (how it is pinned) is listed in tests/test_oracle_*.py.
#pragma once

#include <stdint.h>

#include "../include/symhip_types.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_problem oracle_problem;

enum oracle_solver_kind {
  ORACLE_SOLVER_CHOLESKY = 0, /* SparseCholeskySolver (sparse_cholesky_solver.tcc) */
  ORACLE_SOLVER_SCHUR = 1,    /* SparseSchurSolver (sparse_schur_solver.tcc) + adapter */
};

/* Creates a problem; returns NULL and sets the error string on failure. */
oracle_problem* oracle_problem_create(const symhip_problem_desc* desc, int include_jacobians);
void oracle_problem_destroy(oracle_problem* p);
const char* oracle_last_error(void);

/* Sizes after the first linearization: N (tangent), M (residual), nnz of H_lower, nnz of J. */
int oracle_dims(oracle_problem* p, int64_t* out4);
/* CSC structure of hessian_lower (colptr N+1, rowind nnz). */
int oracle_hessian_structure(oracle_problem* p, int32_t* colptr, int32_t* rowind);
int oracle_jacobian_structure(oracle_problem* p, int32_t* colptr, int32_t* rowind);
/* Storage offsets (Values::Data index) of each key in append order. */
int oracle_storage_offsets(oracle_problem* p, int32_t* offsets);
/* Optimized keys in state order (key indices) and their tangent offsets. */
int oracle_state_keys(oracle_problem* p, int32_t* key_index, int32_t* tangent_offset);
/* Per-factor hessian_storage_col_starts (reference linearization_dense_key_helper_t), flattened
   in factor order; returns count when out is NULL. */
int64_t oracle_hessian_col_starts(oracle_problem* p, int32_t* out);

/* Relinearize at `values` (full storage). Any output may be NULL. */
int oracle_linearize(oracle_problem* p, const double* values, double* residual,
                     double* hessian_values, double* rhs, double* jacobian_values);

/* Sparse Cholesky (LDL^T) on a lower CSC matrix. ordering: NULL = natural (Eigen NaturalOrdering
   returns an empty permutation), else the elimination order (new -> old), i.e. the
   `inv_permutation_` the reference's ordering functor fills. */
typedef struct oracle_cholesky oracle_cholesky;
oracle_cholesky* oracle_cholesky_create(int32_t n, const int32_t* colptr, const int32_t* rowind,
                                        const int32_t* ordering);
void oracle_cholesky_destroy(oracle_cholesky* c);
int oracle_cholesky_factorize(oracle_cholesky* c, const double* values);
int oracle_cholesky_solve(oracle_cholesky* c, const double* b, double* x, int32_t nrhs);
/* L (strictly lower, CSC) nnz, structure/values, and D. */
int64_t oracle_cholesky_nnz(oracle_cholesky* c);
int oracle_cholesky_factor(oracle_cholesky* c, int32_t* colptr, int32_t* rowind, double* lvals,
                           double* d, int32_t* parent);

/* Sparse Schur solver on lower CSC A with a block-diagonal bottom-right C of size c_dim. */
typedef struct oracle_schur oracle_schur;
oracle_schur* oracle_schur_create(int32_t n, const int32_t* colptr, const int32_t* rowind,
                                  int32_t c_dim, const int32_t* s_ordering);
void oracle_schur_destroy(oracle_schur* s);
int oracle_schur_factorize(oracle_schur* s, const double* values);
int oracle_schur_solve(oracle_schur* s, const double* b, double* x);
/* Dense lower S (b_dim x b_dim, column-major) of the last factorization. */
int oracle_schur_dense_s(oracle_schur* s, double* out);

/* Optimize in place (Optimizer::Optimize, optimizer.tcc:70-89).
   solver: ORACLE_SOLVER_*; ordering may be NULL (natural). For the Schur solver, c_dim is the
   trailing block-diagonal dimension and ordering applies to S.
   iters: capacity max_iters; best_values receives the best Values storage (may be NULL). */
int oracle_optimize(oracle_problem* p, const symhip_optimizer_params* params, int solver,
                    const int32_t* ordering, int32_t c_dim, double epsilon, int32_t num_iterations,
                    double* values_inout, symhip_iteration_stats* iters, int32_t max_iters,
                    symhip_optimization_stats* out_stats, double* updates_out);

/* As oracle_optimize, also writing the wall time of each Iterate() call (seconds) into
   iter_seconds (capacity num_iterations). Iteration 0 includes the one-time symbolic analysis. */
int oracle_optimize_timed(oracle_problem* p, const symhip_optimizer_params* params, int solver,
                          const int32_t* ordering, int32_t c_dim, double epsilon,
                          int32_t num_iterations, double* values_inout,
                          symhip_iteration_stats* iters, int32_t max_iters,
                          symhip_optimization_stats* out_stats, double* iter_seconds);

/* GncOptimizer<Optimizer>::Optimize (opt/gnc_optimizer.h:49-118): stages of oracle_optimize's
   LM (natural ordering) with the Scalar key mu_key (index into the problem's keys) raised from
   g->mu_initial by g->mu_step to g->mu_max after each stage ending in SUCCESS. */
int oracle_optimize_gnc(oracle_problem* p, const symhip_optimizer_params* params, int solver,
                        double epsilon, int32_t num_iterations, const symhip_gnc_params* g,
                        int32_t mu_key, double* values_inout, symhip_iteration_stats* iters,
                        int32_t max_iters, symhip_optimization_stats* out_stats);

/* Default parameters (optimizer.cc:8-56). */
void oracle_default_params(symhip_optimizer_params* p);

/* Factor-level linearization for a single factor (test hook): args storage pointers. */
int oracle_linearize_factor(int32_t type, const double* const* args, double* residual,
                            double* jacobian, double* hessian, double* rhs);

#ifdef __cplusplus
}
#endif
