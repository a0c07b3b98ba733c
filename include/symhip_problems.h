/*
 * Seeded synthetic problem generators (host only, no GPU). They build the benchmark and test
 * problems as plain symhip_problem_desc data so the same bytes feed the HIP optimizer and any
 * checker. All generators draw from std::mt19937 with the given seed; distributions are
 * std::normal_distribution / std::uniform_real_distribution, as the reference tests use
 * (test/symforce_optimizer_test.cc:79-131, gen/cpp/sym/ops/matrix/storage_ops.h:211-218).
 */
#ifndef SYMHIP_PROBLEMS_H_
#define SYMHIP_PROBLEMS_H_

#include "symhip_types.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct symhip_generated symhip_generated;

/* Pose3 chain with priors at both ends and between factors: the reference's
   CreatePoseSmoothingProblem (test/symforce_optimizer_test.cc:79-131). */
symhip_generated* symhip_gen_pose_smoothing(int32_t num_keys, uint32_t seed);
/* Rot3 chain (test/symforce_optimizer_test.cc:183-264). */
symhip_generated* symhip_gen_rot_smoothing(int32_t num_keys, uint32_t seed);
/* Rot3 all-pairs between factors, key 0 frozen (test/symforce_optimizer_test.cc:276-348). */
symhip_generated* symhip_gen_rot_frozen(int32_t num_keys, uint32_t seed);

/* Bundle adjustment with a fixed linear camera on a circle, Cauchy loss, prior on camera 0
   (plus a translation-only prior on camera 1 that fixes the scale gauge). */
symhip_generated* symhip_gen_ba_circle(int32_t n_cams, int32_t n_points, uint32_t seed);

/* Snavely-camera bundle adjustment along a random-walk trajectory: Pose3 + (f, k1, k2) per
   camera, power-law track lengths in [2, 40] over consecutive cameras, Huber loss, prior on
   camera 0. Keys 'c' (pose), 'i' (intrinsics), 'p' (points) so points come last. */
symhip_generated* symhip_gen_ba_snavely(int32_t n_cams, int32_t n_points, uint32_t seed);

/* Pose graph: random walk with 1 m steps and N(0, 0.2 rad) turns, odometry between factors and
   loop closures between nodes within 2 m (1 % outliers, Cauchy loss), prior on node 0. */
symhip_generated* symhip_gen_pose_graph(int32_t n_nodes, int32_t n_loops, uint32_t seed);

/* One robot-3D-localization problem (matching + odometry factors, poses optimized, landmarks
   constant): n_poses on a smooth random trajectory, n_landmarks world points, each pose sees
   `visible_fraction` of the landmarks. */
symhip_generated* symhip_gen_localization(int32_t n_poses, int32_t n_landmarks,
                                          double visible_fraction, uint32_t seed);

/* Reads a Bundle Adjustment in the Large text file, as the reference example's ReadProblem
   (examples/bundle_adjustment_in_the_large/bundle_adjustment_in_the_large.cc:60-118): Pose3
   from angle-axis via Rot3::FromTangent, keys 'c', 'i', 'p', 'P', 'e'. huber_k > 0 uses the
   Huber-whitened Snavely factor. NULL on a malformed or missing file. */
symhip_generated* symhip_load_bal(const char* path, double huber_k);
/* Writes the Snavely observations/cameras/points of a problem as BAL text (0 on success). */
int symhip_save_bal(const symhip_problem_desc* desc, const char* path);

void symhip_gen_free(symhip_generated* g);
/* Points `out` at the generator's arrays (valid until symhip_gen_free). */
int symhip_gen_desc(const symhip_generated* g, symhip_problem_desc* out);
/* Named scalar facts about the problem: "c_dim" (trailing landmark tangent dim, 0 if none),
   "epsilon", "n_cams", "n_points", "n_obs". Returns NaN for unknown names. */
double symhip_gen_info(const symhip_generated* g, const char* name);
/* Values storage size in doubles. */
int64_t symhip_gen_storage_size(const symhip_generated* g);

#ifdef __cplusplus
}
#endif

#endif /* SYMHIP_PROBLEMS_H_ */
