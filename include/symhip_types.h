/*
 * Shared plain-data vocabulary of the symforce_amd C ABI: value types, factor types, the problem
 * description that crosses the boundary, optimizer parameters and per-iteration statistics.
 *
 * Numbering of the value types follows the reference's `type_t` enum
 * (lcmtypes/symforce_types.lcm:12-125) so that index entries can be exchanged unchanged.
 * Parameters mirror `optimizer_params_t` (lcmtypes/symforce.lcm:134-200) field for field, and the
 * statistics mirror `optimization_iteration_t` / `optimization_stats_t` (lcmtypes/symforce.lcm:229-336).
 */
#ifndef SYMHIP_TYPES_H_
#define SYMHIP_TYPES_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Value types (reference type_t numbering). */
enum symhip_value_type {
  SYMHIP_TYPE_SCALAR = 1,   /* storage 1, tangent 1 */
  SYMHIP_TYPE_ROT3 = 3,     /* storage 4 [x y z w], tangent 3 */
  SYMHIP_TYPE_POSE3 = 5,    /* storage 7 [qx qy qz qw tx ty tz], tangent 6 */
  SYMHIP_TYPE_VECTOR2 = 12, /* storage 2, tangent 2 */
  SYMHIP_TYPE_VECTOR3 = 13, /* storage 3, tangent 3 */
  SYMHIP_TYPE_VECTOR4 = 14, /* storage 4, tangent 4 */
  SYMHIP_TYPE_VECTOR5 = 15, /* storage 5, tangent 5 */
  SYMHIP_TYPE_VECTOR6 = 16, /* storage 6, tangent 6 */
  SYMHIP_TYPE_MATRIX33 = 37, /* storage 9 column-major, tangent 9 */
  SYMHIP_TYPE_MATRIX66 = 64, /* storage 36 column-major, tangent 36 */
};

/*
 * Factor types with a device implementation. Each lists its arguments in order; arguments marked
 * (*) are the ones the factor is linearized with respect to (its "optimized keys" in the
 * reference's Factor::Hessian form). All other arguments are constants read from the Values.
 */
enum symhip_factor_type {
  /* value(*) Pose3, prior Pose3, sqrt_info Matrix66, epsilon Scalar -> res 6
     gen/cpp/sym/factors/prior_factor_pose3.h:33 */
  SYMHIP_FACTOR_PRIOR_POSE3 = 1,
  /* a(*) Pose3, b(*) Pose3, a_T_b Pose3, sqrt_info Matrix66, epsilon -> res 6
     gen/cpp/sym/factors/between_factor_pose3.h:34 */
  SYMHIP_FACTOR_BETWEEN_POSE3 = 2,
  /* value(*) Rot3, prior Rot3, sqrt_info Matrix33, epsilon -> res 3 */
  SYMHIP_FACTOR_PRIOR_ROT3 = 3,
  /* a(*) Rot3, b(*) Rot3, a_T_b Rot3, sqrt_info Matrix33, epsilon -> res 3 */
  SYMHIP_FACTOR_BETWEEN_ROT3 = 4,
  /* cam_T_world(*) Pose3, intrinsics(*) V3 (f,k1,k2), point(*) V3, pixel V2, epsilon -> res 2
     symforce/examples/bundle_adjustment_in_the_large/gen/snavely_reprojection_factor.h:33-41 */
  SYMHIP_FACTOR_SNAVELY = 5,
  /* as SNAVELY plus huber_k Scalar before epsilon; whitened-norm Huber loss -> res 2 */
  SYMHIP_FACTOR_SNAVELY_HUBER = 6,
  /* world_T_cam(*) Pose3, point(*) V3, pixel V2, K V4 (fx fy cx cy), cauchy_c Scalar, epsilon
     -> res 2; linear camera, whitened-norm Cauchy loss */
  SYMHIP_FACTOR_REPROJ_CAUCHY = 7,
  /* world_T_body(*) Pose3, world_t_landmark V3, body_t_landmark V3, sigma Scalar -> res 3
     symforce/examples/robot_3d_localization/gen/matching_factor.h:27-34 */
  SYMHIP_FACTOR_MATCHING = 8,
  /* world_T_a(*) Pose3, world_T_b(*) Pose3, a_T_b Pose3, diagonal_sigmas V6, epsilon -> res 6
     symforce/examples/robot_3d_localization/gen/odometry_factor.h:28-35 */
  SYMHIP_FACTOR_ODOMETRY = 9,
  /* as BETWEEN_POSE3 plus cauchy_c Scalar before epsilon; whitened-norm Cauchy loss -> res 6 */
  SYMHIP_FACTOR_BETWEEN_POSE3_CAUCHY = 10,
  /* x(*) V5, y V5, mu Scalar, epsilon Scalar -> res 5: per element the Barron loss of x - y with
     alpha = 2 - 1/(1 - mu + epsilon), delta 1, unit information (the graduated non-convexity
     factor of test/symforce_gnc_codegen_test.py:20-29, symforce/opt/noise_models.py:318-436) */
  SYMHIP_FACTOR_BARRON_VECTOR5 = 11,
};

/* A problem as it crosses the boundary. Keys are given in Values append order; `values` is the
   flat storage of all keys in that order (reference Values::Data(), opt/values.h:30-324). */
/* optimizer_gnc_params_t (lcmtypes/symforce.lcm:203-227): graduated non-convexity schedule. */
typedef struct symhip_gnc_params {
  double gnc_update_min_reduction; /* early exit threshold of every stage but the last */
  double mu_initial;
  double mu_step;                  /* mu <= 0 or mu_initial >= mu_max: one plain stage */
  double mu_max;
} symhip_gnc_params;

typedef struct symhip_problem_desc {
  int32_t n_keys;
  const int8_t* key_letters;
  const int64_t* key_subs;
  const int64_t* key_supers;
  const int32_t* key_types;
  const double* values;
  int32_t n_factors;
  const int32_t* factor_types;
  const int32_t* factor_arg_ptr; /* n_factors + 1 offsets into factor_args */
  const int32_t* factor_args;    /* key indices */
  /* Optimized keys in state order. n_opt_keys == 0 selects ComputeKeysToOptimize():
     every key some factor linearizes, sorted by Key::LexicalLessThan (opt/factor.h:423-449). */
  int32_t n_opt_keys;
  const int32_t* opt_keys;
} symhip_problem_desc;

enum symhip_lambda_update_type {
  SYMHIP_LAMBDA_INVALID = 0,
  SYMHIP_LAMBDA_STATIC = 1,
  SYMHIP_LAMBDA_DYNAMIC = 2,
};

/* Mirrors optimizer_params_t (lcmtypes/symforce.lcm:134-200). */
typedef struct symhip_optimizer_params {
  int32_t verbose;
  int32_t debug_stats;
  int32_t check_derivatives;
  int32_t include_jacobians;
  int32_t debug_checks;
  double initial_lambda;
  double lambda_lower_bound;
  double lambda_upper_bound;
  int32_t lambda_update_type;
  double lambda_up_factor;
  double lambda_down_factor;
  double dynamic_lambda_update_beta;
  double dynamic_lambda_update_gamma;
  int32_t dynamic_lambda_update_p;
  int32_t use_diagonal_damping;
  int32_t use_unit_damping;
  int32_t keep_max_diagonal_damping;
  double diagonal_damping_min;
  int32_t iterations;
  double early_exit_min_reduction;
  double early_exit_min_absolute_error;
  int32_t enable_bold_updates;
} symhip_optimizer_params;

enum symhip_status {
  SYMHIP_STATUS_INVALID = 0,
  SYMHIP_STATUS_SUCCESS = 1,
  SYMHIP_STATUS_HIT_ITERATION_LIMIT = 2,
  SYMHIP_STATUS_FAILED = 3,
};

enum symhip_failure_reason {
  SYMHIP_FAILURE_INVALID = 0,
  SYMHIP_FAILURE_LAMBDA_OUT_OF_BOUNDS = 1,
  SYMHIP_FAILURE_INITIAL_ERROR_NOT_FINITE = 2,
};

/* Mirrors optimization_iteration_t without the debug vectors (lcmtypes/symforce.lcm:229-262). */
typedef struct symhip_iteration_stats {
  int32_t iteration;
  int32_t update_accepted;
  double current_lambda;
  double new_error_linear;
  double new_error;
  double relative_reduction;
  double update_angle_change;
} symhip_iteration_stats;

/* Mirrors the scalar fields of optimization_stats_t (lcmtypes/symforce.lcm:301-336). */
typedef struct symhip_optimization_stats {
  int32_t num_iterations; /* records written, including the iteration -1 record */
  int32_t best_index;
  int32_t status;
  int32_t failure_reason;
} symhip_optimization_stats;

#ifdef __cplusplus
}
#endif

#endif /* SYMHIP_TYPES_H_ */
