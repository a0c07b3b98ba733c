/*
 * symforce_amd C ABI: the drop-in boundary between the host optimizer surface (sym::Optimizer,
 * sym::LevenbergMarquardtSolver, sym::Factor, sym::Values — see include/sym/) and the MI355X HIP
 * kernels. Plain pointers and sizes only; every call returns 0 on success and a negative code on
 * failure, with the message available from symhip_last_error(). Host buffers are caller-owned;
 * device buffers are owned by the handle. Handles are not thread safe (as the reference objects,
 * opt/optimizer.h:23): use one per thread; each handle runs on its own HIP stream.
 */
#ifndef SYMHIP_H_
#define SYMHIP_H_

#include <stdint.h>

#include "symhip_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Message of the last failed call on this thread. */
const char* symhip_last_error(void);
/* Number of visible devices (0 when no GPU / runtime unavailable). */
int symhip_device_count(void);

/* ---------------------------------------------------------------------------------------------
 * Problem: keys, Values, factors and the linearization structure.
 * Replaces the symbolic half of sym::Linearizer (opt/linearizer.cc:148-356): state index,
 * hessian_lower pattern and per-factor storage maps, built once.
 * ------------------------------------------------------------------------------------------- */
typedef struct symhip_problem* symhip_problem_t;

int symhip_problem_create(const symhip_problem_desc* desc, symhip_problem_t* out);
int symhip_problem_destroy(symhip_problem_t problem);
/* out[0]=N (tangent dim), out[1]=M (residual dim), out[2]=nnz(hessian_lower),
   out[3]=Values storage size, out[4]=number of optimized keys */
int symhip_problem_dims(symhip_problem_t problem, int64_t* out5);
/* CSC structure of hessian_lower (colptr N+1, rowind nnz). */
int symhip_problem_hessian_structure(symhip_problem_t problem, int32_t* colptr, int32_t* rowind);
/* Flattened hessian_storage_col_starts of every factor in factor order
   (linearization_dense_key_helper_t, lcmtypes/symforce.lcm:67-91). Returns the count; pass
   NULL to query it. */
int64_t symhip_problem_col_starts(symhip_problem_t problem, int32_t* out);
/* Optimized keys in state order and their tangent offsets (index_entry_t.offset). */
int symhip_problem_state_keys(symhip_problem_t problem, int32_t* key_index, int32_t* tangent_offset);

/* The registered factor type `code` (symhip_factor_type): out[0] = arguments, out[1] =
   linearized arguments (the first ones; the rest are constants), out[2] = residual dimension,
   out[3] = tangent dimension of the linearized arguments. */
int symhip_factor_info(int32_t code, int32_t* out4);
/* One registered factor body evaluated on the host (what the device kernels run): residual (D)
   and Jacobian (D x T, row-major, over the linearized arguments' tangents in argument order) at
   the argument storage pointers args (one per argument, as symhip_problem_desc lists them).
   precision SYMHIP_PRECISION_F64 runs the FP64 body, SYMHIP_PRECISION_F32_JAC the FP32 body of
   the precision forms (relative rotations, camera-frame points and measured-minus-predicted
   terms formed in FP64 and rounded once, the rest in FP32). No reference counterpart (the
   reference evaluates generated functions in its Scalar); for checks of the bodies. */
int symhip_factor_linearize_host(int32_t code, const double* const* args, int32_t precision,
                                 double* residual, double* jacobian);

/* Linearizer::Relinearize (opt/linearizer.cc:54-120) at the given Values storage. Host buffers in
   and out (copies over PCIe; the optimizer keeps everything resident instead). Outputs may be
   NULL. hessian_values are in the reference CSC order. */
int symhip_linearize(symhip_problem_t problem, const double* values, double* residual,
                     double* hessian_values, double* rhs);


/* ---------------------------------------------------------------------------------------------
 * Batched optimizer: many small independent problems in one persistent launch (one wavefront per
 * problem runs the whole LM loop; no reference counterpart, the reference runs one Optimizer per
 * thread, opt/optimizer.h:23). Supported problems: robot-3D-localization structure (optimized
 * keys are up to 64 Pose3, MatchingFactor on one pose, OdometryFactor between consecutive
 * optimized poses). The LM semantics are those of symhip_optimizer_optimize.
 * ------------------------------------------------------------------------------------------- */
typedef struct symhip_batch* symhip_batch_t;

/* precision: SYMHIP_PRECISION_F64, or SYMHIP_PRECISION_F32_JAC (factor bodies in FP32; Values,
   Hessian blocks, solve and cost in FP64). */
int symhip_batch_create(const symhip_problem_desc* descs, int32_t n_problems,
                        const symhip_optimizer_params* params, int32_t precision, double epsilon,
                        symhip_batch_t* out);
int symhip_batch_destroy(symhip_batch_t batch);
/* Total Values storage (doubles) over all problems, in problem order. */
int64_t symhip_batch_storage_size(symhip_batch_t batch);
/* Solves every problem from its initial Values. Outputs may be NULL: values_out (storage size)
   gets the best Values; stats_out one record per problem; iters_out (params.iterations + 1) per
   problem; solve_ms the per-problem time inside the launch; total_ms the launch time. */
int symhip_batch_solve(symhip_batch_t batch, double* values_out,
                       symhip_optimization_stats* stats_out, symhip_iteration_stats* iters_out,
                       double* solve_ms, double* total_ms);

/* Autotuner of the batched solve (no counterpart in the reference): solves the whole batch
   `reps` times in each of the 4 launch configurations (bit 0: the kernel built for two
   wavefronts per SIMD instead of one, i.e. twice the resident problems for ~120 spilled
   registers; bit 1: problems with the most poses launched first) and keeps the fastest for
   later symhip_batch_solve calls. config_ms (4 entries, may be NULL) receives each one's best
   launch time, *chosen (may be NULL) the configuration kept. The results of every
   configuration are the same. */
int symhip_batch_tune(symhip_batch_t batch, int32_t reps, double* config_ms, int32_t* chosen);
/* Selects launch configuration `config` (0-3, as above) for later solves, e.g. one a previous
   tuning run of the same batch shape chose. */
int symhip_batch_set_config(symhip_batch_t batch, int32_t config);

/* Allocation-free variant of symhip_batch_create + symhip_batch_solve. No counterpart in the
   reference: it is for callers that cannot allocate device memory (code under HIP graph capture,
   serving loops, harnesses that forbid hipMalloc). The caller packs the batch on the host with
   symhip_batch_pack_sizes / symhip_batch_pack (no device memory), copies the tables into device
   buffers it owns, and calls symhip_batch_solve_noalloc with device pointers for the tables and
   the outputs. The kernel's working state (3 state blocks of poses, their linearizations, the
   running maximum of the diagonal, clock stamps: 216 doubles per pose) lives in one file-scope
   device array per GPU sized for SYMHIP_BATCH_NOALLOC_MAX_POSES poses in total (28.6 MB, held
   by every GPU the library runs on until the process ends); larger batches are rejected with an
   error. The call allocates nothing and launches on `stream` (a hipStream_t; NULL is the null
   stream) without synchronizing. Calls that overlap on one GPU share that array and race:
   serialize them. */
#define SYMHIP_BATCH_NOALLOC_MAX_POSES 16384
/* out[0] = problems, out[1] = total optimized poses, out[2] = matching factors,
   out[3] = Values storage (doubles) over all problems. Validates like symhip_batch_create. */
int symhip_batch_pack_sizes(const symhip_problem_desc* descs, int32_t n_problems,
                            const symhip_optimizer_params* params, int64_t* out4);
/* Host tables, sized from symhip_batch_pack_sizes: pose_begin (problems + 1), pose_soff (poses),
   match_ptr (poses + 1), match_args (3 * matching factors), odo_args (3 * poses), values
   (storage: the initial Values of every problem, concatenated in problem order). */
int symhip_batch_pack(const symhip_problem_desc* descs, int32_t n_problems,
                      const symhip_optimizer_params* params, int32_t* pose_begin,
                      int32_t* pose_soff, int32_t* match_ptr, int32_t* match_args,
                      int32_t* odo_args, double* values);
typedef struct symhip_batch_tables {
  int32_t n_problems;
  int64_t total_poses;
  int64_t storage_size;
  /* device pointers to the packed tables (match_args may be NULL without matching factors) */
  const int32_t* pose_begin;
  const int32_t* pose_soff;
  const int32_t* match_ptr;
  const int32_t* match_args;
  const int32_t* odo_args;
  const double* values;
} symhip_batch_tables;
/* values_out (storage_size doubles), stats_out (n_problems records) and iters_out
   (n_problems * (params.iterations + 1) records, or NULL) are device pointers; the results are
   those of symhip_batch_solve on the same problems. */
int symhip_batch_solve_noalloc(const symhip_batch_tables* tables,
                               const symhip_optimizer_params* params, int32_t precision,
                               double epsilon, double* values_out,
                               symhip_optimization_stats* stats_out,
                               symhip_iteration_stats* iters_out, void* stream);

/* ---------------------------------------------------------------------------------------------
 * Optimizer: sym::Optimizer<double> + LevenbergMarquardtSolver with the linearization, linear
 * solve, damping, retraction and accept/lambda logic on the device (opt/optimizer.tcc:21-89,
 * opt/levenberg_marquardt_solver.tcc:138-343).
 * ------------------------------------------------------------------------------------------- */
enum symhip_linear_solver {
  SYMHIP_SOLVER_CHOLESKY = 0, /* SparseCholeskySolver replacement: supernodal LDL^T on H */
  SYMHIP_SOLVER_SCHUR = 1,    /* SparseSchurSolver replacement: landmark elimination + RCS */
};
enum symhip_ordering {
  SYMHIP_ORDERING_NATURAL = 0,
  SYMHIP_ORDERING_MIN_DEGREE = 1,
  SYMHIP_ORDERING_NESTED_DISSECTION = 2, /* recursive BFS-level separators (METIS's family) */
};
enum symhip_precision {
  SYMHIP_PRECISION_F64 = 0,      /* plain FP64 everywhere */
  SYMHIP_PRECISION_F32_JAC = 1,  /* FP32 Jacobians/Hessian blocks, FP64 residual/cost/decisions */
  SYMHIP_PRECISION_F32_UPDATES = 2, /* sparse Cholesky path: mixed-precision supernodal LDL^T */
};

typedef struct symhip_solver_options {
  int32_t linear_solver; /* symhip_linear_solver */
  int32_t ordering;      /* symhip_ordering (fill-reducing ordering of the factorized matrix) */
  int32_t precision;     /* symhip_precision. SYMHIP_PRECISION_F32_JAC on the optimizer: Schur
                            path only; the Hessian-side products (W = J_cam^T J_point, V) come
                            from FP32 factor bodies, W and Z = W L_p^-T are stored in FP32 and so
                            is S for PCG; residual, gradient, matrix-core sums, LDL^T, cost and
                            gain ratio stay FP64. Not combinable with the learned preconditioner
                            or system logging (both need FP64 systems).
                            SYMHIP_PRECISION_F32_UPDATES: sparse Cholesky path only; the
                            trailing supernode updates of the LDL^T run on the FP32 matrix cores
                            (diagonal blocks, panel solves, D and the triangular solves FP64) and
                            every solve is refined against the FP64 damped Hessian by
                            conjugate gradients preconditioned by that factor (FP64 products,
                            residuals and scalars) to a relative residual of 1e-13 or its
                            stagnation below 1e-10, at most 30 iterations; a solve that misses
                            it refactors with FP64 updates. */
  int32_t use_pcg;       /* Schur path: block-Jacobi preconditioned CG on the reduced camera
                            system. 0 off (direct LDL^T), 1 truncated (an iterate that missed
                            pcg_rel_tol within pcg_max_iters is used; LDL^T only when CG
                            breaks down), 2 strict (LDL^T whenever the tolerance is missed) */
  int32_t pcg_max_iters;
  double pcg_rel_tol;
  int32_t pcg_preconditioner; /* SYMHIP_PRECOND_BLOCK_JACOBI or SYMHIP_PRECOND_LEARNED */
  /* Truncated PCG only: once an accepted LM step's relative reduction falls below this value the
     remaining iterations of the run use the direct LDL^T (0: never). Until then the LM early
     exit on a small relative reduction is deferred: a truncated step that meets it switches the
     run to the direct LDL^T instead of ending it (a stalled inexact solve is not convergence).
     At the switch lambda is relaxed to min(lambda, initial_lambda) (and nu to its beta). */
  double pcg_exact_below_reduction;
  /* Truncated PCG only, inexact-Newton forcing (0: off): after an accepted LM step with relative
     reduction r the next solve uses tolerance min(pcg_rel_tol, pcg_forcing * sqrt(r)) (at least
     1e-10) and an iteration limit of pcg_max_iters * (1 + 2 log10(pcg_rel_tol / tolerance)), at
     most 20 * pcg_max_iters. */
  double pcg_forcing;
} symhip_solver_options;

enum symhip_pcg_preconditioner {
  SYMHIP_PRECOND_BLOCK_JACOBI = 0, /* inverse diagonal camera blocks */
  SYMHIP_PRECOND_LEARNED = 1,      /* block Jacobi corrected per camera by the trained network
                                      (symhip_optimizer_train_preconditioner); untrained it is
                                      exactly block Jacobi */
};

typedef struct symhip_optimizer* symhip_optimizer_t;

/* epsilon: the optimizer's epsilon (Optimizer ctor, default kDefaultEpsilon = 10 * DBL_EPSILON). */
int symhip_optimizer_create(const symhip_problem_desc* desc, const symhip_optimizer_params* params,
                            const symhip_solver_options* options, double epsilon,
                            symhip_optimizer_t* out);
int symhip_optimizer_destroy(symhip_optimizer_t opt);
/* Optimizer::Optimize(values, num_iterations, populate_best_linearization, stats): values is
   the full Values storage, updated in place with the best values. num_iterations < 0 uses
   params.iterations. iters receives up to max_iters records (iteration -1 first). */
int symhip_optimizer_optimize(symhip_optimizer_t opt, double* values, int32_t num_iterations,
                              symhip_iteration_stats* iters, int32_t max_iters,
                              symhip_optimization_stats* stats);
/* GncOptimizer<Optimizer>::Optimize (opt/gnc_optimizer.h:49-118): Levenberg-Marquardt stages
   to convergence, the Scalar key mu_key_index (an index into desc's keys; a constant read by the
   factors, e.g. SYMHIP_FACTOR_BARRON_VECTOR5) set to gnc->mu_initial and raised by mu_step up to
   mu_max after each stage that ends in SUCCESS. Stages before the last exit early at
   gnc_update_min_reduction; damping is relaxed to the initial lambda between stages. The
   records of all stages go to iters (one iteration -1 record, first); num_iterations (< 0:
   params.iterations) bounds the total. values is updated with the last stage's best Values. */
int symhip_optimizer_optimize_gnc(symhip_optimizer_t opt, double* values, int32_t num_iterations,
                                  const symhip_gnc_params* gnc, int32_t mu_key_index,
                                  symhip_iteration_stats* iters, int32_t max_iters,
                                  symhip_optimization_stats* stats);
/* Linearization of the best Values of the last optimize (stats.best_linearization). */
int symhip_optimizer_best_linearization(symhip_optimizer_t opt, double* residual,
                                        double* hessian_values, double* rhs);
int symhip_optimizer_update_params(symhip_optimizer_t opt, const symhip_optimizer_params* params);
/* Device-resident variant for throughput measurement: upload once, optimize without host
   transfers of Values, read the best Values back afterwards. */
int symhip_optimizer_set_values(symhip_optimizer_t opt, const double* values);
int symhip_optimizer_optimize_resident(symhip_optimizer_t opt, int32_t num_iterations,
                                       symhip_iteration_stats* iters, int32_t max_iters,
                                       symhip_optimization_stats* stats);
int symhip_optimizer_get_values(symhip_optimizer_t opt, double* values);
/* Host wall-clock time (ms since the start of the last optimize call) at which each iteration
   record of that call became available. Returns the number of entries written. */
int32_t symhip_optimizer_iteration_times(symhip_optimizer_t opt, double* ms, int32_t max);
/* Device phases timed with HIP events when timing is enabled (one phase = the kernels of one
   step of Iterate; the Schur phases are single kernels). */
enum symhip_phase {
  SYMHIP_PHASE_LINEARIZE = 0,    /* relinearization of the new state block (+ cost reduction) */
  SYMHIP_PHASE_DAMP = 1,         /* DampHessian */
  SYMHIP_PHASE_POINT_ELIM = 2,   /* Schur: landmark block factorization, Z = W L^-T */
  SYMHIP_PHASE_RCS_ASSEMBLE = 3, /* Schur: reduced camera system blocks into the LDL^T panels */
  SYMHIP_PHASE_REDUCED_RHS = 4,  /* Schur: reduced right-hand side */
  SYMHIP_PHASE_FACTOR = 5,       /* supernodal LDL^T factorization (incl. assembly) */
  SYMHIP_PHASE_SOLVE = 6,        /* triangular solves */
  SYMHIP_PHASE_BACK_SUB = 7,     /* Schur: landmark back-substitution */
  SYMHIP_PHASE_RETRACT = 8,      /* UpdateNewFromInit: copy + retraction */
  SYMHIP_PHASE_DECIDE = 9,       /* linear error reduction + accept / lambda decision */
  SYMHIP_PHASE_LIN_POINTS = 10,  /* Schur linearization: landmark-major kernel alone */
  SYMHIP_PHASE_LIN_CAMS = 11,    /* Schur linearization: camera-major kernel alone */
  SYMHIP_NUM_PHASES = 12,
};
int symhip_optimizer_enable_timing(symhip_optimizer_t opt, int32_t enable);
/* Accumulated milliseconds and occurrence counts per phase since timing was enabled. */
int symhip_optimizer_phase_times(symhip_optimizer_t opt, double* total_ms, int64_t* counts);

/* Symbolic facts of the linear solver: out[0]=factorized dim, out[1]=supernodes,
   out[2]=panel doubles, out[3]=levels, out[4]=factor flops (rounded), out[5]=reduced-camera-
   system blocks (Schur path, else 0), out[6]=landmark observation pairs, out[7]=observations,
   out[8]=flops of the diagonal blocks and panels (cols^3/3 + below-rows x cols^2 per
   supernode), out[9]=flops of the trailing updates (below-rows^2 x cols). */
int symhip_optimizer_solver_info(symhip_optimizer_t opt, int64_t* out10);
/* Iterative (PCG) linear solver statistics since creation: out[0]=solves, out[1]=total CG
   iterations, out[2]=fallbacks to the direct LDL^T, out[3]=most iterations in one solve.
   No reference counterpart (the reference solvers are direct). With
   SYMHIP_PRECISION_F32_UPDATES: out[0]=solves, out[1]=refinement (CG) iterations,
   out[2]=fallbacks to the FP64 factorization, out[3]=most iterations in one solve. */
int symhip_optimizer_linear_solver_stats(symhip_optimizer_t opt, int64_t* out4);
/* Mixed-precision sparse Cholesky (SYMHIP_PRECISION_F32_UPDATES) only; no reference counterpart.
   track >= 0 turns the recording of every solve's final relative residual |b - A x| / |b|
   (A the damped Hessian, FP64) on (1) or off (0) and clears the record; track < 0 only reads.
   Copies up to max recorded residuals into out (may be NULL); returns the number recorded, or
   -1 on error. */
int32_t symhip_optimizer_refinement_residuals(symhip_optimizer_t opt, int32_t track, double* out,
                                              int32_t max);

/* Learned PCG preconditioner (Schur path, use_pcg != 0, 9-dof cameras). No reference
   counterpart. The network has SYMHIP_LEARNED_PRECOND_PARAMS float parameters. */
#define SYMHIP_LEARNED_PRECOND_PARAMS 21744
/* Keep copies of the next max reduced camera systems and right-hand sides as the optimizer
   assembles them (training data); clears earlier copies. */
int symhip_optimizer_log_systems(symhip_optimizer_t opt, int32_t max);
int32_t symhip_optimizer_logged_systems(symhip_optimizer_t opt);
/* Block pattern of the reduced camera system (full symmetric block-CSR, 9x9 row-major blocks):
   returns the camera count n (< 0 on error); row_ptr (n + 1) and col (row_ptr[n]) may be NULL. */
int32_t symhip_optimizer_rcs_pattern(symhip_optimizer_t opt, int32_t* row_ptr, int32_t* col);
int symhip_optimizer_logged_system(symhip_optimizer_t opt, int32_t idx, double* blocks,
                                   double* rhs);
/* get / set may be NULL. */
int symhip_optimizer_preconditioner_params(symhip_optimizer_t opt, float* get, const float* set);
/* Loss and gradient on logged system idx for host probes X (P columns of n * 9, P <= 7). */
int symhip_optimizer_preconditioner_loss_grad(symhip_optimizer_t opt, int32_t idx,
                                              const double* X, int32_t P, float* loss,
                                              float* grad);
/* Adam training on logged systems 0 .. n_systems - 1 (all when n_systems <= 0), round robin,
   fresh probes each step; losses (steps) may be NULL. */
int symhip_optimizer_train_preconditioner(symhip_optimizer_t opt, int32_t steps, float lr,
                                          int32_t P, uint64_t seed, int32_t n_systems,
                                          float* losses);
/* PCG on logged system idx with block Jacobi (learned = 0) or the network (1): iterations to
   rel_tol (negative when max_iters was reached) and the relative error of the step against the
   supernodal LDL^T solve (step_err may be NULL). */
int symhip_optimizer_eval_preconditioner(symhip_optimizer_t opt, int32_t idx, double rel_tol,
                                         int32_t max_iters, int32_t learned, int32_t* iterations,
                                         double* step_err);

/* debug_stats (optimizer_params_t.debug_stats): the vectors PopulateIterationStats records for
   iteration record `index` of the last optimize call (levenberg_marquardt_solver.tcc:115-122,
   :165-170): update (N; record 0, the iteration -1 record, has none and leaves it untouched),
   the Values storage of the state block the record describes, and its residual (M). Any output
   may be NULL. Fails when the optimizer was created without debug_stats. */
int symhip_optimizer_debug_record(symhip_optimizer_t opt, int32_t index, double* update,
                                  double* values, double* residual);
/* The warnings of the last optimize call, each followed by the record separator 0x1e (a warning
   may span several lines), in the wording of the reference's
   spdlog::warn calls: zero on the damped Hessian diagonal (CheckHessianDiagonal,
   levenberg_marquardt_solver.tcc:56-90), non-finite initial or new error (:180-183, :257-262)
   and, with optimizer_params_t.debug_checks, the non-finite update (:225-227) and every factor
   whose linearization is not finite, with its keys, values and residual
   (opt/internal/linearizer_utils.h:307-367). Each is also written to stderr as it happens.
   Copies at most cap - 1 bytes and a NUL into buf (buf may be NULL when cap is 0); returns the
   full length in bytes, or a negative code. */
int64_t symhip_optimizer_warnings(symhip_optimizer_t opt, char* buf, int64_t cap);
/* include_jacobians (optimizer_params_t.include_jacobians): OptimizationStats.jacobian_sparsity
   (levenberg_marquardt_solver.tcc:172-175), the M x N Jacobian's CSC pattern (colptr N + 1,
   rowind nnz; rows sorted), and its values: of the best linearization of the last optimize
   (Linearization::jacobian of GetBestLinearization), or of iteration record `index` under
   debug_stats (optimization_iteration_t.jacobian_values, :115-122, :165-170). The sparsity call
   returns nnz (pass NULL outputs to query it) or a negative code; all fail when the optimizer
   was created without include_jacobians. check_derivatives requires include_jacobians, as in
   the reference (optimizer.tcc:39). */
int64_t symhip_optimizer_jacobian_sparsity(symhip_optimizer_t opt, int32_t* colptr,
                                           int32_t* rowind);
int symhip_optimizer_best_jacobian(symhip_optimizer_t opt, double* values);
int symhip_optimizer_debug_jacobian(symhip_optimizer_t opt, int32_t index, double* values);
/* OptimizationStats.linear_solver_ordering and .cholesky_factor_sparsity
   (levenberg_marquardt_solver.tcc:205-212): the sparse Cholesky path's Permutation().indices()
   (N entries, old -> new) and the strictly lower CSC pattern of its factor L (colptr N + 1,
   rowind nnz). Returns nnz(L) (pass NULL outputs to query it), or a negative code; the Schur
   path has neither (the reference's SparseSchurSolver exposes neither). */
int64_t symhip_optimizer_factor_sparsity(symhip_optimizer_t opt, int32_t* ordering,
                                         int32_t* colptr, int32_t* rowind);

/* LevenbergMarquardtSolver::Iterate's accept / reject / lambda step
   (levenberg_marquardt_solver.tcc:239-340), the same function the device optimizer runs, for
   host loops over a user linearization (sym::LevenbergMarquardtSolver in include/sym/).
   e_init / e_new / e_best: errors of the Init, New and Best state blocks; linear_delta:
   update . (rhs - damping * update); last_norm2, update_norm2, last_dot_update: the bold-update
   inputs |last|^2, |update|^2, last . update. */
typedef struct symhip_lm_decision {
  double lambda;
  double nu;
  symhip_iteration_stats record;
  int32_t status;         /* symhip_status, 0 while iterating */
  int32_t failure_reason; /* symhip_failure_reason */
  int32_t accept;
  int32_t set_best;
} symhip_lm_decision;
int symhip_lm_decide(double e_init, double e_new, double e_best, double linear_delta,
                     double lambda, double nu, int32_t iteration,
                     const symhip_optimizer_params* params, int32_t have_last_update,
                     double last_norm2, double update_norm2, double last_dot_update,
                     double epsilon, symhip_lm_decision* out);
/* Values::Retract over the given keys (opt/values.cc:294-327): values (host storage) is moved
   by delta (host tangent vector) at each key's storage and tangent offsets. */
int symhip_retract(int32_t n_keys, const int32_t* types, const int32_t* storage_offsets,
                   const int32_t* tangent_offsets, double* values, const double* delta,
                   double epsilon);

/* ---------------------------------------------------------------------------------------------
 * Linear solver adapter: SparseCholeskySolver (opt/sparse_cholesky/sparse_cholesky_solver.h:
 * 27-160) over a lower CSC matrix. node_ptr (n_nodes + 1 entries) groups consecutive scalars
 * into nodes that are ordered together; NULL means one node per scalar.
 * ordering: SYMHIP_ORDERING_*, or given_order (node elimination order) when non-NULL.
 * ------------------------------------------------------------------------------------------- */
typedef struct symhip_cholesky* symhip_cholesky_t;

int symhip_cholesky_create(int32_t n, const int32_t* colptr, const int32_t* rowind,
                           const int32_t* node_ptr, int32_t n_nodes, int32_t ordering,
                           const int32_t* given_order, symhip_cholesky_t* out);
int symhip_cholesky_destroy(symhip_cholesky_t c);
/* Factorize(A) with host values in the CSC order given at creation. */
int symhip_cholesky_factorize(symhip_cholesky_t c, const double* values);
/* Mixed-precision factorization (no reference counterpart): on != 0 runs the trailing supernode
   updates on the FP32 matrix cores (diagonal blocks, panels, D and the solves stay FP64). The
   factor is then FP32-accurate in the updates; Solve does not refine (the optimizer's
   SYMHIP_PRECISION_F32_UPDATES form does). Call before the first Factorize. */
int symhip_cholesky_set_f32_updates(symhip_cholesky_t c, int32_t on);
/* Solve(b) for nrhs column-major right-hand sides (host buffers). */
int symhip_cholesky_solve(symhip_cholesky_t c, const double* b, double* x, int32_t nrhs);
/* inv_permutation (new -> old), the reference's Permutation().inverse() convention. */
int symhip_cholesky_permutation(symhip_cholesky_t c, int32_t* inv_perm);
/* Reference elimination tree parent and L column counts for the chosen permutation
   (sparse_cholesky_solver.tcc:55-89). */
int symhip_cholesky_etree(symhip_cholesky_t c, int32_t* parent, int32_t* counts);
/* D of the last factorization (permuted order). */
int symhip_cholesky_d(symhip_cholesky_t c, double* d);
/* out[0]=n, out[1]=supernodes, out[2]=panel doubles, out[3]=levels, out[4]=factor flops. */
int symhip_cholesky_info(symhip_cholesky_t c, int64_t* out5);
/* SparseCholeskySolver::L() (sparse_cholesky_solver.h:146-148): the strictly lower unit factor of
   the permuted matrix as CSC on the reference's pattern for this permutation (colptr n + 1,
   rowind / values nnz; values need a factorization). Returns nnz; NULL outputs query it. */
int64_t symhip_cholesky_factor(symhip_cholesky_t c, int32_t* colptr, int32_t* rowind,
                               double* values);

#ifdef __cplusplus
}
#endif

#endif /* SYMHIP_H_ */
