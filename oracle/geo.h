// Oracle (test infrastructure only): CPU FP64 restatement of the geometry and factor residuals that
// the SymForce optimizer path evaluates. Nothing in the shipped library includes this file; only the
// tests, the smoke check and the CPU-baseline leg of the benchmark load the oracle.
//
// Jacobians are NOT hand-derived here: every residual is written once, templated on its scalar, and
// differentiated by forward-mode dual numbers with respect to the tangent-space perturbation that the
// retraction defines. That makes the oracle's derivatives independent of the hand-derived analytic
// Jacobians in the HIP kernels, so a sign/index mistake on either side shows up as a parity failure.
//
// Semantics followed (reference files):
//   Rot3 retract  q <- q * [sin(t/2)/t * w, cos(t/2)], t = sqrt(eps^2 + |w|^2)
//                 gen/cpp/sym/ops/rot3/lie_group_ops.cc:63-92
//   Pose3 retract rotation as Rot3, translation additive (product manifold, not SE(3))
//                 gen/cpp/sym/ops/pose3/lie_group_ops.cc:69-102
//   LocalCoordinates(a, b) = [log(a^-1 b) with min(1-eps,|w|) clamp, b.t - a.t]
//                 gen/cpp/sym/ops/pose3/lie_group_ops.cc:104-127
//   Pose3 composition / inverse / action: SE(3) group ops, gen/cpp/sym/pose3.h
//   Factor residuals: symforce/codegen/geo_factors_codegen.py:53-86 (between/prior),
//   symforce/examples/bundle_adjustment_in_the_large/bundle_adjustment_in_the_large.py:18-63,
//   symforce/examples/robot_3d_localization/robot_3d_localization.py:115-150,
//   robust whitening symforce/opt/noise_models.py:68-90 with norm(eps)=sqrt(|x|^2+eps)
//   (symforce/geo/matrix.py:610-614).
#pragma once

#include <algorithm>
#include <array>
#include <cmath>

namespace oracle {

// ---------------------------------------------------------------------------------------------
// Forward-mode dual number with N derivative lanes.
// ---------------------------------------------------------------------------------------------
template <int N>
struct Dual {
  double v = 0;
  std::array<double, N> d{};
  Dual() = default;
  Dual(double x) : v(x) {}  // NOLINT: implicit constants
};

template <int N>
inline Dual<N> operator+(const Dual<N>& a, const Dual<N>& b) {
  Dual<N> r(a.v + b.v);
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] + b.d[i];
  return r;
}
template <int N>
inline Dual<N> operator-(const Dual<N>& a, const Dual<N>& b) {
  Dual<N> r(a.v - b.v);
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] - b.d[i];
  return r;
}
template <int N>
inline Dual<N> operator-(const Dual<N>& a) {
  Dual<N> r(-a.v);
  for (int i = 0; i < N; ++i) r.d[i] = -a.d[i];
  return r;
}
template <int N>
inline Dual<N> operator*(const Dual<N>& a, const Dual<N>& b) {
  Dual<N> r(a.v * b.v);
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
  return r;
}
template <int N>
inline Dual<N> operator/(const Dual<N>& a, const Dual<N>& b) {
  Dual<N> r(a.v / b.v);
  const double inv = 1.0 / b.v;
  for (int i = 0; i < N; ++i) r.d[i] = (a.d[i] - r.v * b.d[i]) * inv;
  return r;
}
template <int N>
inline Dual<N> operator+(const Dual<N>& a, double b) { return a + Dual<N>(b); }
template <int N>
inline Dual<N> operator+(double a, const Dual<N>& b) { return Dual<N>(a) + b; }
template <int N>
inline Dual<N> operator-(const Dual<N>& a, double b) { return a - Dual<N>(b); }
template <int N>
inline Dual<N> operator-(double a, const Dual<N>& b) { return Dual<N>(a) - b; }
template <int N>
inline Dual<N> operator*(const Dual<N>& a, double b) {
  Dual<N> r(a.v * b);
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b;
  return r;
}
template <int N>
inline Dual<N> operator*(double a, const Dual<N>& b) { return b * a; }
template <int N>
inline Dual<N> operator/(const Dual<N>& a, double b) { return a * (1.0 / b); }
template <int N>
inline Dual<N> operator/(double a, const Dual<N>& b) { return Dual<N>(a) / b; }

template <int N>
inline Dual<N> apply1(const Dual<N>& a, double f, double df) {
  Dual<N> r(f);
  for (int i = 0; i < N; ++i) r.d[i] = df * a.d[i];
  return r;
}
template <int N>
inline Dual<N> sqrt(const Dual<N>& a) {
  const double s = std::sqrt(a.v);
  return apply1(a, s, 0.5 / s);
}
template <int N>
inline Dual<N> sin(const Dual<N>& a) { return apply1(a, std::sin(a.v), std::cos(a.v)); }
template <int N>
inline Dual<N> cos(const Dual<N>& a) { return apply1(a, std::cos(a.v), -std::sin(a.v)); }
template <int N>
inline Dual<N> acos(const Dual<N>& a) {
  return apply1(a, std::acos(a.v), -1.0 / std::sqrt(1.0 - a.v * a.v));
}
template <int N>
inline Dual<N> log(const Dual<N>& a) { return apply1(a, std::log(a.v), 1.0 / a.v); }
template <int N>
inline Dual<N> pow(const Dual<N>& a, double p) {
  return apply1(a, std::pow(a.v, p), p * std::pow(a.v, p - 1.0));
}
// d|x| = sign(x) dx, sign(0) = 0 (matches the generated code's sign() expansion)
template <int N>
inline Dual<N> fabs(const Dual<N>& a) {
  const double s = (a.v > 0) - (a.v < 0);
  return apply1(a, std::fabs(a.v), s);
}
// min(a, const): derivative of the selected branch
template <int N>
inline Dual<N> min_const(const Dual<N>& a, double c) { return a.v < c ? a : Dual<N>(c); }
inline double min_const(double a, double c) { return std::min(a, c); }
template <int N>
inline double value(const Dual<N>& a) { return a.v; }
inline double value(double a) { return a; }
template <int N>
inline double copysign1(const Dual<N>& a) { return std::copysign(1.0, a.v); }
inline double copysign1(double a) { return std::copysign(1.0, a); }

using std::acos;
using std::cos;
using std::fabs;
using std::log;
using std::sin;
using std::sqrt;

// ---------------------------------------------------------------------------------------------
// Quaternion [x, y, z, w] (SymForce storage order) and Pose3 [qx qy qz qw tx ty tz].
// ---------------------------------------------------------------------------------------------
template <typename T>
struct Quat {
  T x, y, z, w;
};

template <typename T>
inline Quat<T> qmul(const Quat<T>& a, const Quat<T>& b) {
  return {a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y,
          a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x,
          a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w,
          a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z};
}
template <typename T>
inline Quat<T> qconj(const Quat<T>& a) {
  return {-a.x, -a.y, -a.z, a.w};
}
// R(q) * v
template <typename T>
inline std::array<T, 3> qrot(const Quat<T>& q, const std::array<T, 3>& v) {
  const Quat<T> p{v[0], v[1], v[2], T(0.0)};
  const Quat<T> r = qmul(qmul(q, p), qconj(q));
  return {r.x, r.y, r.z};
}

// Rotation log with SymForce's clamp: 2 * sign(w) * acos(min(|w|, 1-eps)) / sqrt(1 - m^2) * xyz
template <typename T>
inline std::array<T, 3> qlog(const Quat<T>& q, double epsilon) {
  const T m = min_const(fabs(q.w), 1.0 - epsilon);
  const T f = 2.0 * copysign1(q.w) * acos(m) / sqrt(1.0 - m * m);
  return {f * q.x, f * q.y, f * q.z};
}

// Exp used by the retraction: [sin(t/2)/t * w, cos(t/2)], t = sqrt(eps^2 + |w|^2)
inline Quat<double> qexp(const double* w, double epsilon) {
  const double t = std::sqrt(epsilon * epsilon + w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
  const double s = std::sin(t / 2) / t;
  return {s * w[0], s * w[1], s * w[2], std::cos(t / 2)};
}

template <typename T>
struct Pose3T {
  Quat<T> q;
  std::array<T, 3> t;
};

template <typename T>
inline Pose3T<T> pose_inverse(const Pose3T<T>& a) {
  const Quat<T> qi = qconj(a.q);
  const std::array<T, 3> ti = qrot(qi, a.t);
  return {qi, {-ti[0], -ti[1], -ti[2]}};
}
template <typename T>
inline Pose3T<T> pose_compose(const Pose3T<T>& a, const Pose3T<T>& b) {
  const std::array<T, 3> rb = qrot(a.q, b.t);
  return {qmul(a.q, b.q), {a.t[0] + rb[0], a.t[1] + rb[1], a.t[2] + rb[2]}};
}
template <typename T>
inline std::array<T, 3> pose_act(const Pose3T<T>& a, const std::array<T, 3>& p) {
  const std::array<T, 3> r = qrot(a.q, p);
  return {r[0] + a.t[0], r[1] + a.t[1], r[2] + a.t[2]};
}
// LocalCoordinates(a, b) for the product manifold
template <typename T>
inline std::array<T, 6> pose_local(const Pose3T<T>& a, const Pose3T<T>& b, double epsilon) {
  const std::array<T, 3> w = qlog(qmul(qconj(a.q), b.q), epsilon);
  return {w[0], w[1], w[2], b.t[0] - a.t[0], b.t[1] - a.t[1], b.t[2] - a.t[2]};
}

// Plain double helpers on storage arrays
inline void rot3_retract(const double* q, const double* w, double epsilon, double* out) {
  const Quat<double> a{q[0], q[1], q[2], q[3]};
  const Quat<double> r = qmul(a, qexp(w, epsilon));
  out[0] = r.x;
  out[1] = r.y;
  out[2] = r.z;
  out[3] = r.w;
}
inline void pose3_retract(const double* p, const double* d, double epsilon, double* out) {
  rot3_retract(p, d, epsilon, out);
  out[4] = p[4] + d[3];
  out[5] = p[5] + d[4];
  out[6] = p[6] + d[5];
}

// Rot3::FromYawPitchRoll (gen/cpp/sym/rot3.h): q = Rz(yaw) * Ry(pitch) * Rx(roll)
inline Quat<double> quat_from_ypr(double yaw, double pitch, double roll) {
  const double cy = std::cos(yaw / 2), sy = std::sin(yaw / 2);
  const double cp = std::cos(pitch / 2), sp = std::sin(pitch / 2);
  const double cr = std::cos(roll / 2), sr = std::sin(roll / 2);
  const Quat<double> qz{0, 0, sy, cy}, qy{0, sp, 0, cp}, qx{sr, 0, 0, cr};
  return qmul(qmul(qz, qy), qx);
}

}  // namespace oracle
