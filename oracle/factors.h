// Oracle (test infrastructure only): factor residuals of the registered factor types, written once
// as scalar templates and differentiated by dual numbers. See geo.h for the semantics references.
#pragma once

#include <cstring>
#include <stdexcept>
#include <vector>

#include "../include/symhip_types.h"
#include "./geo.h"

namespace oracle {

inline int storage_dim(int type) {
  switch (type) {
    case SYMHIP_TYPE_SCALAR: return 1;
    case SYMHIP_TYPE_ROT3: return 4;
    case SYMHIP_TYPE_POSE3: return 7;
    case SYMHIP_TYPE_VECTOR2: return 2;
    case SYMHIP_TYPE_VECTOR3: return 3;
    case SYMHIP_TYPE_VECTOR4: return 4;
    case SYMHIP_TYPE_VECTOR5: return 5;
    case SYMHIP_TYPE_VECTOR6: return 6;
    case SYMHIP_TYPE_MATRIX33: return 9;
    case SYMHIP_TYPE_MATRIX66: return 36;
  }
  throw std::runtime_error("oracle: unknown value type");
}
inline int tangent_dim(int type) {
  switch (type) {
    case SYMHIP_TYPE_ROT3: return 3;
    case SYMHIP_TYPE_POSE3: return 6;
    default: return storage_dim(type);
  }
}

struct FactorSpec {
  int nargs;
  int arg_types[6];
  int lin_mask;  // bit k set: argument k is linearized
  int res_dim;
};

inline FactorSpec factor_spec(int type) {
  const int P = SYMHIP_TYPE_POSE3, R = SYMHIP_TYPE_ROT3, S = SYMHIP_TYPE_SCALAR,
            V2 = SYMHIP_TYPE_VECTOR2, V3 = SYMHIP_TYPE_VECTOR3, V4 = SYMHIP_TYPE_VECTOR4,
            V5 = SYMHIP_TYPE_VECTOR5, V6 = SYMHIP_TYPE_VECTOR6, M3 = SYMHIP_TYPE_MATRIX33,
            M6 = SYMHIP_TYPE_MATRIX66;
  switch (type) {
    case SYMHIP_FACTOR_PRIOR_POSE3: return {4, {P, P, M6, S}, 0b1, 6};
    case SYMHIP_FACTOR_BETWEEN_POSE3: return {5, {P, P, P, M6, S}, 0b11, 6};
    case SYMHIP_FACTOR_PRIOR_ROT3: return {4, {R, R, M3, S}, 0b1, 3};
    case SYMHIP_FACTOR_BETWEEN_ROT3: return {5, {R, R, R, M3, S}, 0b11, 3};
    case SYMHIP_FACTOR_SNAVELY: return {5, {P, V3, V3, V2, S}, 0b111, 2};
    case SYMHIP_FACTOR_SNAVELY_HUBER: return {6, {P, V3, V3, V2, S, S}, 0b111, 2};
    case SYMHIP_FACTOR_REPROJ_CAUCHY: return {6, {P, V3, V2, V4, S, S}, 0b11, 2};
    case SYMHIP_FACTOR_MATCHING: return {4, {P, V3, V3, S}, 0b1, 3};
    case SYMHIP_FACTOR_ODOMETRY: return {5, {P, P, P, V6, S}, 0b11, 6};
    case SYMHIP_FACTOR_BETWEEN_POSE3_CAUCHY: return {6, {P, P, P, M6, S, S}, 0b11, 6};
    case SYMHIP_FACTOR_BARRON_VECTOR5: return {4, {V5, V5, S, S}, 0b1, 5};
  }
  throw std::runtime_error("oracle: unknown factor type");
}

// ----- argument views --------------------------------------------------------------------------
template <typename T>
inline Pose3T<T> pose_const(const double* p) {
  return {{T(p[0]), T(p[1]), T(p[2]), T(p[3])}, {T(p[4]), T(p[5]), T(p[6])}};
}
template <typename T>
inline Quat<T> rot_const(const double* p) {
  return {T(p[0]), T(p[1]), T(p[2]), T(p[3])};
}
template <typename T>
inline std::array<T, 3> v3_const(const double* p) {
  return {T(p[0]), T(p[1]), T(p[2])};
}

// Retraction of a constant storage point by a zero-valued dual perturbation occupying lanes
// [lane, lane + tangent): first-order expansion of retract(x, delta) at delta = 0.
// For rotations d/dw [sin(t/2)/t * w, cos(t/2)] at w = 0 is [I/2; 0] (the epsilon term of the
// retraction contributes O(eps^2) and is dropped).
template <int N>
inline Quat<Dual<N>> rot_perturbed(const double* q, int lane) {
  Quat<Dual<N>> d{Dual<N>(0.0), Dual<N>(0.0), Dual<N>(0.0), Dual<N>(1.0)};
  d.x.d[lane + 0] = 0.5;
  d.y.d[lane + 1] = 0.5;
  d.z.d[lane + 2] = 0.5;
  return qmul(rot_const<Dual<N>>(q), d);
}
template <int N>
inline Pose3T<Dual<N>> pose_perturbed(const double* p, int lane) {
  Pose3T<Dual<N>> r;
  r.q = rot_perturbed<N>(p, lane);
  for (int i = 0; i < 3; ++i) {
    r.t[i] = Dual<N>(p[4 + i]);
    r.t[i].d[lane + 3 + i] = 1.0;
  }
  return r;
}
template <int N>
inline std::array<Dual<N>, 3> v3_perturbed(const double* p, int lane) {
  std::array<Dual<N>, 3> r;
  for (int i = 0; i < 3; ++i) {
    r[i] = Dual<N>(p[i]);
    r[i].d[lane + i] = 1.0;
  }
  return r;
}

// Robust whitening of a residual vector by its norm (symforce/opt/noise_models.py:68-90):
// w(x) = f(n)/n * x, n = sqrt(|x|^2 + eps), f = sqrt(2 rho(n)).
//   Huber:  rho(n) = n^2/2 (n <= k), k (n - k/2) (n > k)
//   Cauchy: rho(n) = c^2/2 log(1 + n^2/c^2)
template <typename T, int D>
inline void whiten_huber(std::array<T, D>& r, double k, double eps) {
  T sq = T(eps);
  for (int i = 0; i < D; ++i) sq = sq + r[i] * r[i];
  const T n = sqrt(sq);
  if (value(n) <= k) return;  // quadratic branch: scale 1
  const T s = sqrt(k * (2.0 * n - k)) / n;
  for (int i = 0; i < D; ++i) r[i] = s * r[i];
}
template <typename T, int D>
inline void whiten_cauchy(std::array<T, D>& r, double c, double eps) {
  T sq = T(eps);
  for (int i = 0; i < D; ++i) sq = sq + r[i] * r[i];
  const T n = sqrt(sq);
  const T s = c * sqrt(log(1.0 + sq / (c * c))) / n;
  for (int i = 0; i < D; ++i) r[i] = s * r[i];
}

// BarronNoiseModel (symforce/opt/noise_models.py:318-436) with delta 1, scalar_information 1,
// x_epsilon = alpha_epsilon = eps, alpha from mu by compute_alpha_from_mu (:382-396), applied to
// each element of x - y: barron_error (:398-416) and whiten_scalar (:418-436), the factor of
// test/symforce_gnc_codegen_test.py:20-29.
template <typename T, int D>
inline std::array<T, D> res_barron(const std::array<T, D>& x, const double* y, double mu,
                                   double eps) {
  const double alpha = 2.0 - 1.0 / (1.0 - mu + eps);
  const double b = std::fabs(alpha - 2.0) + eps;
  const double sign_no_zero = alpha >= 0.0 ? 1.0 : -1.0;
  const double d = alpha + sign_no_zero * eps;
  std::array<T, D> r;
  for (int i = 0; i < D; ++i) {
    const T e = x[i] - y[i];
    const T err = (b / d) * (pow((e * e) / b + 1.0, d / 2.0) - 1.0);
    r[i] = sqrt(2.0 * err + eps) - std::sqrt(eps);
  }
  return r;
}

template <typename T>
inline std::array<T, 6> apply_sqrt_info6(const double* S, const std::array<T, 6>& e) {
  std::array<T, 6> r;
  for (int i = 0; i < 6; ++i) {
    T acc = T(0.0);
    for (int j = 0; j < 6; ++j) acc = acc + S[j * 6 + i] * e[j];
    r[i] = acc;
  }
  return r;
}
template <typename T>
inline std::array<T, 3> apply_sqrt_info3(const double* S, const std::array<T, 3>& e) {
  std::array<T, 3> r;
  for (int i = 0; i < 3; ++i) {
    T acc = T(0.0);
    for (int j = 0; j < 3; ++j) acc = acc + S[j * 3 + i] * e[j];
    r[i] = acc;
  }
  return r;
}

// ----- residual bodies -------------------------------------------------------------------------
// Each takes the (possibly perturbed) linearized arguments and the constant argument storage.
template <typename T>
std::array<T, 6> res_prior_pose3(const Pose3T<T>& value, const double* prior, const double* S,
                                 double eps) {
  return apply_sqrt_info6(S, pose_local(pose_const<T>(prior), value, eps));
}
template <typename T>
std::array<T, 6> res_between_pose3(const Pose3T<T>& a, const Pose3T<T>& b, const double* a_T_b,
                                   const double* S, double eps) {
  return apply_sqrt_info6(
      S, pose_local(pose_const<T>(a_T_b), pose_compose(pose_inverse(a), b), eps));
}
template <typename T>
std::array<T, 3> res_prior_rot3(const Quat<T>& value, const double* prior, const double* S,
                                double eps) {
  return apply_sqrt_info3(S, qlog(qmul(qconj(rot_const<T>(prior)), value), eps));
}
template <typename T>
std::array<T, 3> res_between_rot3(const Quat<T>& a, const Quat<T>& b, const double* a_T_b,
                                  const double* S, double eps) {
  return apply_sqrt_info3(S, qlog(qmul(qconj(rot_const<T>(a_T_b)), qmul(qconj(a), b)), eps));
}
template <typename T>
std::array<T, 2> res_snavely(const Pose3T<T>& cam_T_world, const std::array<T, 3>& intr,
                             const std::array<T, 3>& point, const double* pixel) {
  const std::array<T, 3> pc = pose_act(cam_T_world, point);
  const T px = pc[0] / (-pc[2]);
  const T py = pc[1] / (-pc[2]);
  const T r2 = px * px + py * py;
  const T rad = 1.0 + intr[1] * r2 + intr[2] * r2 * r2;
  return {intr[0] * rad * px - pixel[0], intr[0] * rad * py - pixel[1]};
}
template <typename T>
std::array<T, 2> res_reproj_linear(const Pose3T<T>& world_T_cam, const std::array<T, 3>& point,
                                   const double* pixel, const double* K, double eps) {
  const std::array<T, 3> pc = pose_act(pose_inverse(world_T_cam), point);
  const T z = value(pc[2]) > eps ? pc[2] : T(eps);
  return {K[0] * pc[0] / z + K[2] - pixel[0], K[1] * pc[1] / z + K[3] - pixel[1]};
}
template <typename T>
std::array<T, 3> res_matching(const Pose3T<T>& world_T_body, const double* landmark,
                              const double* meas, double sigma) {
  const std::array<T, 3> p = pose_act(pose_inverse(world_T_body), v3_const<T>(landmark));
  return {(p[0] - meas[0]) / sigma, (p[1] - meas[1]) / sigma, (p[2] - meas[2]) / sigma};
}
template <typename T>
std::array<T, 6> res_odometry(const Pose3T<T>& a, const Pose3T<T>& b, const double* a_T_b,
                              const double* sigmas, double eps) {
  const std::array<T, 6> e =
      pose_local(pose_compose(pose_inverse(a), b), pose_const<T>(a_T_b), eps);
  std::array<T, 6> r;
  for (int i = 0; i < 6; ++i) r[i] = e[i] / sigmas[i];
  return r;
}

// ----- linearization ---------------------------------------------------------------------------
// Fills residual (res_dim), jacobian (res_dim x tdim, column-major), hessian (tdim x tdim, full
// symmetric = J^T J as the generated code's lower triangle mirrored), rhs (tdim) = J^T r.
// `args` are pointers to each argument's storage. tdim = sum of tangent dims of linearized args.
struct LinearizedFactor {
  int res_dim = 0, tdim = 0;
  std::vector<double> residual, jacobian, hessian, rhs;
};

template <int N, int D>
inline void unpack(const std::array<Dual<N>, D>& r, LinearizedFactor& out) {
  out.res_dim = D;
  out.tdim = N;
  out.residual.assign(D, 0.0);
  out.jacobian.assign(D * N, 0.0);
  for (int i = 0; i < D; ++i) {
    out.residual[i] = r[i].v;
    for (int j = 0; j < N; ++j) out.jacobian[j * D + i] = r[i].d[j];
  }
}

inline void finish_gauss_newton(LinearizedFactor& f) {
  const int D = f.res_dim, N = f.tdim;
  f.hessian.assign(N * N, 0.0);
  f.rhs.assign(N, 0.0);
  for (int a = 0; a < N; ++a) {
    for (int b = 0; b < N; ++b) {
      double acc = 0;
      for (int i = 0; i < D; ++i) acc += f.jacobian[a * D + i] * f.jacobian[b * D + i];
      f.hessian[b * N + a] = acc;
    }
    double acc = 0;
    for (int i = 0; i < D; ++i) acc += f.jacobian[a * D + i] * f.residual[i];
    f.rhs[a] = acc;
  }
}

inline void linearize_factor(int type, const std::vector<const double*>& a, LinearizedFactor& out) {
  switch (type) {
    case SYMHIP_FACTOR_PRIOR_POSE3: {
      auto r = res_prior_pose3(pose_perturbed<6>(a[0], 0), a[1], a[2], a[3][0]);
      unpack<6, 6>(r, out);
      break;
    }
    case SYMHIP_FACTOR_BETWEEN_POSE3: {
      auto r = res_between_pose3(pose_perturbed<12>(a[0], 0), pose_perturbed<12>(a[1], 6), a[2],
                                 a[3], a[4][0]);
      unpack<12, 6>(r, out);
      break;
    }
    case SYMHIP_FACTOR_BETWEEN_POSE3_CAUCHY: {
      auto r = res_between_pose3(pose_perturbed<12>(a[0], 0), pose_perturbed<12>(a[1], 6), a[2],
                                 a[3], a[5][0]);
      whiten_cauchy<Dual<12>, 6>(r, a[4][0], a[5][0]);
      unpack<12, 6>(r, out);
      break;
    }
    case SYMHIP_FACTOR_PRIOR_ROT3: {
      auto r = res_prior_rot3(rot_perturbed<3>(a[0], 0), a[1], a[2], a[3][0]);
      unpack<3, 3>(r, out);
      break;
    }
    case SYMHIP_FACTOR_BETWEEN_ROT3: {
      auto r = res_between_rot3(rot_perturbed<6>(a[0], 0), rot_perturbed<6>(a[1], 3), a[2], a[3],
                                a[4][0]);
      unpack<6, 3>(r, out);
      break;
    }
    case SYMHIP_FACTOR_SNAVELY: {
      auto r = res_snavely(pose_perturbed<12>(a[0], 0), v3_perturbed<12>(a[1], 6),
                           v3_perturbed<12>(a[2], 9), a[3]);
      unpack<12, 2>(r, out);
      break;
    }
    case SYMHIP_FACTOR_SNAVELY_HUBER: {
      auto r = res_snavely(pose_perturbed<12>(a[0], 0), v3_perturbed<12>(a[1], 6),
                           v3_perturbed<12>(a[2], 9), a[3]);
      whiten_huber<Dual<12>, 2>(r, a[4][0], a[5][0]);
      unpack<12, 2>(r, out);
      break;
    }
    case SYMHIP_FACTOR_REPROJ_CAUCHY: {
      auto r = res_reproj_linear(pose_perturbed<9>(a[0], 0), v3_perturbed<9>(a[1], 6), a[2], a[3],
                                 a[5][0]);
      whiten_cauchy<Dual<9>, 2>(r, a[4][0], a[5][0]);
      unpack<9, 2>(r, out);
      break;
    }
    case SYMHIP_FACTOR_MATCHING: {
      auto r = res_matching(pose_perturbed<6>(a[0], 0), a[1], a[2], a[3][0]);
      unpack<6, 3>(r, out);
      break;
    }
    case SYMHIP_FACTOR_ODOMETRY: {
      auto r = res_odometry(pose_perturbed<12>(a[0], 0), pose_perturbed<12>(a[1], 6), a[2], a[3],
                            a[4][0]);
      unpack<12, 6>(r, out);
      break;
    }
    case SYMHIP_FACTOR_BARRON_VECTOR5: {
      std::array<Dual<5>, 5> x;
      for (int i = 0; i < 5; ++i) {
        x[i] = Dual<5>(a[0][i]);
        x[i].d[i] = 1.0;
      }
      auto r = res_barron<Dual<5>, 5>(x, a[1], a[2][0], a[3][0]);
      unpack<5, 5>(r, out);
      break;
    }
    default:
      throw std::runtime_error("oracle: unknown factor type");
  }
  finish_gauss_newton(out);
}

// Residual only (double), used by finite-difference checks.
inline std::vector<double> evaluate_residual(int type, const std::vector<const double*>& a) {
  LinearizedFactor f;
  linearize_factor(type, a, f);
  return f.residual;
}

}  // namespace oracle
