// Device/host geometry and the registered factor bodies with hand-derived analytic Jacobians.
//
// Conventions (reference semantics):
//   Pose3 storage [qx qy qz qw tx ty tz]; tangent [w(3) t(3)]; retraction rotates on the right and
//   adds translation (gen/cpp/sym/ops/pose3/lie_group_ops.cc:69-102). Rot3 likewise with [w(3)].
//   LocalCoordinates(a, b) = [log(a^-1 b) with SymForce's min(|w|, 1-eps) clamp, b.t - a.t]
//   (gen/cpp/sym/ops/pose3/lie_group_ops.cc:104-127).
//   Jacobians are derivatives w.r.t. that tangent perturbation at zero; for a right perturbation
//   q <- q * [dw/2, 1] so d(q * b)/d(dw) = M(q)/2 with M(q) = [[w I + [v]x]; [-v^T]].
//   Hessian = J^T J (lower), rhs = J^T r, as the generated factors' `with_linearization`
//   (symforce/codegen/codegen.py:796-806).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

#include "../../include/symhip_types.h"

#define SYMHIP_HD __host__ __device__ __forceinline__

namespace symhip {

template <typename S>
struct V3 {
  S x, y, z;
};
template <typename S>
SYMHIP_HD V3<S> operator+(V3<S> a, V3<S> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <typename S>
SYMHIP_HD V3<S> operator-(V3<S> a, V3<S> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <typename S>
SYMHIP_HD V3<S> operator*(S s, V3<S> a) { return {s * a.x, s * a.y, s * a.z}; }
template <typename S>
SYMHIP_HD V3<S> cross(V3<S> a, V3<S> b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <typename S>
SYMHIP_HD S dot(V3<S> a, V3<S> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

template <typename S>
struct Quat {
  S x, y, z, w;
};
template <typename S>
SYMHIP_HD Quat<S> qmul(const Quat<S>& a, const Quat<S>& b) {
  return {a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y, a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x,
          a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w, a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z};
}
template <typename S>
SYMHIP_HD Quat<S> qconj(const Quat<S>& a) { return {-a.x, -a.y, -a.z, a.w}; }

// Row-major 3x3 rotation matrix of a unit quaternion.
template <typename S>
struct M3 {
  S m[9];
  SYMHIP_HD S operator()(int r, int c) const { return m[r * 3 + c]; }
};
template <typename S>
SYMHIP_HD M3<S> rotmat(const Quat<S>& q) {
  const S x = q.x, y = q.y, z = q.z, w = q.w;
  M3<S> R;
  R.m[0] = 1 - 2 * (y * y + z * z);
  R.m[1] = 2 * (x * y - z * w);
  R.m[2] = 2 * (x * z + y * w);
  R.m[3] = 2 * (x * y + z * w);
  R.m[4] = 1 - 2 * (x * x + z * z);
  R.m[5] = 2 * (y * z - x * w);
  R.m[6] = 2 * (x * z - y * w);
  R.m[7] = 2 * (y * z + x * w);
  R.m[8] = 1 - 2 * (x * x + y * y);
  return R;
}
template <typename S>
SYMHIP_HD V3<S> mul(const M3<S>& R, V3<S> v) {
  return {R.m[0] * v.x + R.m[1] * v.y + R.m[2] * v.z, R.m[3] * v.x + R.m[4] * v.y + R.m[5] * v.z,
          R.m[6] * v.x + R.m[7] * v.y + R.m[8] * v.z};
}
template <typename S>
SYMHIP_HD V3<S> mulT(const M3<S>& R, V3<S> v) {
  return {R.m[0] * v.x + R.m[3] * v.y + R.m[6] * v.z, R.m[1] * v.x + R.m[4] * v.y + R.m[7] * v.z,
          R.m[2] * v.x + R.m[5] * v.y + R.m[8] * v.z};
}
template <typename S>
SYMHIP_HD M3<S> matmul(const M3<S>& A, const M3<S>& B) {
  M3<S> C;
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) C.m[r * 3 + c] = A(r, 0) * B(0, c) + A(r, 1) * B(1, c) + A(r, 2) * B(2, c);
  return C;
}
template <typename S>
SYMHIP_HD M3<S> transpose(const M3<S>& A) {
  M3<S> C;
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) C.m[r * 3 + c] = A(c, r);
  return C;
}
// [v]x
template <typename S>
SYMHIP_HD M3<S> skew(V3<S> v) {
  M3<S> K;
  K.m[0] = 0;
  K.m[1] = -v.z;
  K.m[2] = v.y;
  K.m[3] = v.z;
  K.m[4] = 0;
  K.m[5] = -v.x;
  K.m[6] = -v.y;
  K.m[7] = v.x;
  K.m[8] = 0;
  return K;
}

// Rotation log with SymForce's clamp and its derivative w.r.t. a right perturbation of the
// logged quaternion: for q' = q * [G dth / 2, 1] (G a 3x3 matrix), d log(q') / d dth = D G.
// log(q) = 2 sgn(w) f(m) v with m = min(|w|, 1 - eps), f(m) = acos(m)/sqrt(1 - m^2).
// d log = 2 sgn(w) f I dv + 2 v f'(m) H dw, H = [|w| < 1 - eps]; with dq = M(q) dth/2:
//   dv = (w I + [v]x) dth/2, dw = -v^T dth/2.
template <typename S>
SYMHIP_HD void rot_log(const Quat<S>& q, S eps, V3<S>& out, M3<S>& D) {
  const S aw = q.w < 0 ? -q.w : q.w;
  // The clamp must be representable in S: an FP64-sized epsilon rounds 1 - eps to 1 in FP32,
  // and an exactly-identity relative rotation then gives acos(1) / sqrt(0) = 0 / 0.
  if constexpr (sizeof(S) == 4) eps = eps > S(8e-7) ? eps : S(8e-7);  // ~7 FP32 epsilons
  const S one_m_eps = S(1) - eps;
  const bool unclamped = aw < one_m_eps;
  const S m = unclamped ? aw : one_m_eps;
  // 1 - m^2 as (1 - m)(1 + m): 1 - m is exact near m = 1, where 1 - m * m loses ~8 digits to
  // cancellation at small angles (~1e-9 relative in f, 7% in f')
  const S sq = (S(1) - m) * (S(1) + m);
  const S f = acos(m) / sqrt(sq);
  const S sgn = q.w < 0 ? S(-1) : S(1);
  const V3<S> v{q.x, q.y, q.z};
  out = (S(2) * sgn * f) * v;
  const S fp = unclamped ? (m * f - S(1)) / sq : S(0);
  // D = sgn f (w I + [v]x) - fp v v^T   (the 1/2 of dq cancels the 2 of the log)
  const S a = sgn * f;
  D.m[0] = a * q.w - fp * v.x * v.x;
  D.m[1] = a * (-v.z) - fp * v.x * v.y;
  D.m[2] = a * (v.y) - fp * v.x * v.z;
  D.m[3] = a * (v.z) - fp * v.y * v.x;
  D.m[4] = a * q.w - fp * v.y * v.y;
  D.m[5] = a * (-v.x) - fp * v.y * v.z;
  D.m[6] = a * (-v.y) - fp * v.z * v.x;
  D.m[7] = a * (v.x) - fp * v.z * v.y;
  D.m[8] = a * q.w - fp * v.z * v.z;
}

template <typename S>
SYMHIP_HD Quat<S> load_quat(const double* p) {
  return {S(p[0]), S(p[1]), S(p[2]), S(p[3])};
}
template <typename S>
SYMHIP_HD V3<S> load_v3(const double* p) {
  return {S(p[0]), S(p[1]), S(p[2])};
}

// Relative Lie-group deltas for the FP32 bodies of the precision forms: the quantities that
// cancel -- a relative rotation q_a^-1 q_b, a measured-minus-predicted rotation or translation, a
// point in a camera frame -- are formed in FP64 from the FP64 Values and rounded once, so FP32
// rounding enters after the cancellation (each keeps FP32's relative accuracy instead of losing
// the ratio of absolute to relative magnitude, e.g. a 1 km pose and a 1 m step). With S =
// double these are the FP64 body's own expressions.
template <typename S>
SYMHIP_HD Quat<S> to_s(const Quat<double>& q) {
  return {S(q.x), S(q.y), S(q.z), S(q.w)};
}
template <typename S>
SYMHIP_HD V3<S> to_s(const V3<double>& v) {
  return {S(v.x), S(v.y), S(v.z)};
}
// q_a^-1 q_b in FP64
SYMHIP_HD Quat<double> rel_quat(const double* pa, const double* pb) {
  return qmul(qconj(load_quat<double>(pa)), load_quat<double>(pb));
}

// Retraction exponential [sin(t/2)/t w, cos(t/2)], t = sqrt(eps^2 + |w|^2).
SYMHIP_HD void rot3_retract_inplace(double* q, const double* w, double eps) {
  const double t = sqrt(eps * eps + w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
  const double s = sin(t / 2) / t;
  const Quat<double> d{s * w[0], s * w[1], s * w[2], cos(t / 2)};
  const Quat<double> r = qmul(Quat<double>{q[0], q[1], q[2], q[3]}, d);
  q[0] = r.x;
  q[1] = r.y;
  q[2] = r.z;
  q[3] = r.w;
}

// Values::Retract of one key (opt/values.cc:294-327, gen/cpp/sym/ops/pose3/lie_group_ops.cc:69-102):
// Pose3 rotates on the right and translates additively, Rot3 rotates on the right, every vector
// and matrix type adds its tangent (storage) entries.
SYMHIP_HD void retract_value(int type, double* v, const double* d, double eps) {
  if (type == SYMHIP_TYPE_POSE3) {
    rot3_retract_inplace(v, d, eps);
    v[4] += d[3];
    v[5] += d[4];
    v[6] += d[5];
  } else if (type == SYMHIP_TYPE_ROT3) {
    rot3_retract_inplace(v, d, eps);
  } else {
    const int td = type == SYMHIP_TYPE_SCALAR    ? 1
                   : type == SYMHIP_TYPE_VECTOR2  ? 2
                   : type == SYMHIP_TYPE_VECTOR3  ? 3
                   : type == SYMHIP_TYPE_VECTOR4  ? 4
                   : type == SYMHIP_TYPE_VECTOR5  ? 5
                   : type == SYMHIP_TYPE_VECTOR6  ? 6
                   : type == SYMHIP_TYPE_MATRIX33 ? 9
                                                  : 36;
    for (int k = 0; k < td; ++k) v[k] += d[k];
  }
}

// Robust whitening by the norm (symforce/opt/noise_models.py:68-90):
// w = s(n) r, n = sqrt(|r|^2 + eps); J_w = (s I + s'(n)/n r r^T) J.
// kind 1: Huber  s = 1 (n <= k), sqrt(k (2n - k))/n (n > k)
// kind 2: Cauchy s = c sqrt(log(1 + n^2/c^2))/n
template <typename S, int D, int T>
SYMHIP_HD void whiten_norm(int kind, S param, S eps, S* r, S (*J)[T]) {
  S sq = eps;
  for (int i = 0; i < D; ++i) sq += r[i] * r[i];
  const S n = sqrt(sq);
  S s, sp;
  if (kind == 1) {
    if (n <= param) return;
    const S g = sqrt(param * (S(2) * n - param));
    s = g / n;
    sp = param * (param - n) / (n * n * g);
  } else {
    const S c2 = param * param;
    const S L = log(S(1) + sq / c2);
    const S g = param * sqrt(L);
    s = g / n;
    const S gp = (n / (param * sqrt(L))) / (S(1) + sq / c2);
    sp = (gp * n - g) / (n * n);
  }
  const S k = sp / n;
  S Jw[D][T];
  for (int i = 0; i < D; ++i)
    for (int c = 0; c < T; ++c) {
      S acc = s * J[i][c];
      for (int j = 0; j < D; ++j) acc += k * r[i] * r[j] * J[j][c];
      Jw[i][c] = acc;
    }
  for (int i = 0; i < D; ++i) {
    r[i] = s * r[i];
    for (int c = 0; c < T; ++c) J[i][c] = Jw[i][c];
  }
}

// ----------------------------------------------------------------------------------------------
// Factor bodies. Each fills r[D] and J[D][T] (T = sum of linearized tangent dims, in argument
// order). `a` holds each argument's storage pointer.
// ----------------------------------------------------------------------------------------------

// local(prior, value) rotation part, perturbing `value` (q_rel = prior^-1 value).
template <typename S>
SYMHIP_HD void factor_prior_pose3(const double* const* a, S* r, S (*J)[6]) {
  const double* Sm = a[2];
  const S eps = S(a[3][0]);
  const Quat<S> qrel = to_s<S>(rel_quat(a[1], a[0]));
  V3<S> w;
  M3<S> D;
  rot_log(qrel, eps, w, D);
  const S e[6] = {w.x, w.y, w.z, S(a[0][4] - a[1][4]), S(a[0][5] - a[1][5]), S(a[0][6] - a[1][6])};
  S de[6][6] = {};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) de[i][j] = D(i, j);
  for (int i = 0; i < 3; ++i) de[3 + i][3 + i] = S(1);
  for (int i = 0; i < 6; ++i) {
    S acc = 0;
    for (int k = 0; k < 6; ++k) acc += S(Sm[k * 6 + i]) * e[k];
    r[i] = acc;
    for (int c = 0; c < 6; ++c) {
      S accj = 0;
      for (int k = 0; k < 6; ++k) accj += S(Sm[k * 6 + i]) * de[k][c];
      J[i][c] = accj;
    }
  }
}

// between(a, b) = a^-1 b, e = local(a_T_b, between); J over [dth_a dt_a dth_b dt_b].
template <typename S>
SYMHIP_HD void between_pose3_error(const double* pa, const double* pb, const double* pm, S eps,
                                   S* e, S (*de)[12]) {
  const Quat<double> qab_d = rel_quat(pa, pb);
  const V3<double> dt{pb[4] - pa[4], pb[5] - pa[5], pb[6] - pa[6]};
  const V3<double> tab_d = mulT(rotmat(load_quat<double>(pa)), dt);
  const Quat<S> qab = to_s<S>(qab_d);
  const M3<S> Ra = rotmat(load_quat<S>(pa)), Rab = rotmat(qab);
  const V3<S> tab = to_s<S>(tab_d);
  const Quat<S> qrel = to_s<S>(qmul(qconj(load_quat<double>(pm)), qab_d));
  V3<S> w;
  M3<S> D;
  rot_log(qrel, eps, w, D);
  e[0] = w.x;
  e[1] = w.y;
  e[2] = w.z;
  e[3] = S(tab_d.x - pm[4]);
  e[4] = S(tab_d.y - pm[5]);
  e[5] = S(tab_d.z - pm[6]);
  // rotation rows: d/dth_a = D (-Rab^T), d/dth_b = D
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      S acc = 0;
      for (int k = 0; k < 3; ++k) acc += D(i, k) * Rab(j, k);
      de[i][j] = -acc;
      de[i][3 + j] = 0;
      de[i][6 + j] = D(i, j);
      de[i][9 + j] = 0;
    }
  // translation rows: d/dth_a = [tab]x, d/dt_a = -Ra^T, d/dth_b = 0, d/dt_b = Ra^T
  const M3<S> K = skew(tab);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      de[3 + i][j] = K(i, j);
      de[3 + i][3 + j] = -Ra(j, i);
      de[3 + i][6 + j] = 0;
      de[3 + i][9 + j] = Ra(j, i);
    }
}

template <typename S>
SYMHIP_HD void factor_between_pose3(const double* const* a, S* r, S (*J)[12], int robust) {
  S e[6], de[6][12];
  const S eps = S(robust ? a[5][0] : a[4][0]);
  between_pose3_error<S>(a[0], a[1], a[2], eps, e, de);
  const double* Sm = a[3];
  for (int i = 0; i < 6; ++i) {
    S acc = 0;
    for (int k = 0; k < 6; ++k) acc += S(Sm[k * 6 + i]) * e[k];
    r[i] = acc;
    for (int c = 0; c < 12; ++c) {
      S accj = 0;
      for (int k = 0; k < 6; ++k) accj += S(Sm[k * 6 + i]) * de[k][c];
      J[i][c] = accj;
    }
  }
  if (robust) whiten_norm<S, 6, 12>(2, S(a[4][0]), eps, r, J);
}

template <typename S>
SYMHIP_HD void factor_prior_rot3(const double* const* a, S* r, S (*J)[3]) {
  const Quat<S> qrel = to_s<S>(rel_quat(a[1], a[0]));
  V3<S> w;
  M3<S> D;
  rot_log(qrel, S(a[3][0]), w, D);
  const S e[3] = {w.x, w.y, w.z};
  const double* Sm = a[2];
  for (int i = 0; i < 3; ++i) {
    S acc = 0;
    for (int k = 0; k < 3; ++k) acc += S(Sm[k * 3 + i]) * e[k];
    r[i] = acc;
    for (int c = 0; c < 3; ++c) {
      S accj = 0;
      for (int k = 0; k < 3; ++k) accj += S(Sm[k * 3 + i]) * D(k, c);
      J[i][c] = accj;
    }
  }
}

template <typename S>
SYMHIP_HD void factor_between_rot3(const double* const* a, S* r, S (*J)[6]) {
  const Quat<double> qab_d = rel_quat(a[0], a[1]);
  const M3<S> Rab = rotmat(to_s<S>(qab_d));
  V3<S> w;
  M3<S> D;
  rot_log(to_s<S>(qmul(qconj(load_quat<double>(a[2])), qab_d)), S(a[4][0]), w, D);
  const S e[3] = {w.x, w.y, w.z};
  S de[3][6];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      S acc = 0;
      for (int k = 0; k < 3; ++k) acc += D(i, k) * Rab(j, k);
      de[i][j] = -acc;
      de[i][3 + j] = D(i, j);
    }
  const double* Sm = a[3];
  for (int i = 0; i < 3; ++i) {
    S acc = 0;
    for (int k = 0; k < 3; ++k) acc += S(Sm[k * 3 + i]) * e[k];
    r[i] = acc;
    for (int c = 0; c < 6; ++c) {
      S accj = 0;
      for (int k = 0; k < 3; ++k) accj += S(Sm[k * 3 + i]) * de[k][c];
      J[i][c] = accj;
    }
  }
}

// BAL / Snavely projection: P = R X + t, p = -P_xy / P_z, u = f (1 + k1 r2 + k2 r2^2) p.
// J over [dth(3) dt(3) | f k1 k2 | X(3)].
template <typename S>
SYMHIP_HD void factor_snavely(const double* const* a, S* r, S (*J)[12], int robust) {
  const M3<S> R = rotmat(load_quat<S>(a[0]));
  const V3<S> X = load_v3<S>(a[2]);
  const S f = S(a[1][0]), k1 = S(a[1][1]), k2 = S(a[1][2]);
  // the point in the camera frame, formed in FP64
  const V3<S> P = to_s<S>(mul(rotmat(load_quat<double>(a[0])), load_v3<double>(a[2])) +
                          V3<double>{a[0][4], a[0][5], a[0][6]});
  const S iz = S(1) / P.z;
  const S px = -P.x * iz, py = -P.y * iz;
  const S r2 = px * px + py * py;
  const S rad = S(1) + k1 * r2 + k2 * r2 * r2;
  r[0] = f * rad * px - S(a[3][0]);
  r[1] = f * rad * py - S(a[3][1]);
  // du/dp = f (rad I + p drad^T), drad = 2 (k1 + 2 k2 r2) p
  const S g = S(2) * (k1 + S(2) * k2 * r2);
  const S dudp[2][2] = {{f * (rad + px * g * px), f * px * g * py},
                        {f * py * g * px, f * (rad + py * g * py)}};
  // dp/dP = [[-iz, 0, px iz], [0, -iz, py iz]]  (p = -P_xy/P_z => dp/dPz = P_xy/Pz^2 = -p iz)
  const S dpdP[2][3] = {{-iz, 0, -px * iz}, {0, -iz, -py * iz}};
  S dudP[2][3];
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 3; ++j) dudP[i][j] = dudp[i][0] * dpdP[0][j] + dudp[i][1] * dpdP[1][j];
  // dP/dth = -R [X]x ; dP/dt = I ; dP/dX = R
  const M3<S> RK = matmul(R, skew(X));
  for (int i = 0; i < 2; ++i) {
    for (int j = 0; j < 3; ++j) {
      J[i][j] = -(dudP[i][0] * RK(0, j) + dudP[i][1] * RK(1, j) + dudP[i][2] * RK(2, j));
      J[i][3 + j] = dudP[i][j];
      J[i][9 + j] = dudP[i][0] * R(0, j) + dudP[i][1] * R(1, j) + dudP[i][2] * R(2, j);
    }
  }
  const S p[2] = {px, py};
  for (int i = 0; i < 2; ++i) {
    J[i][6] = rad * p[i];
    J[i][7] = f * r2 * p[i];
    J[i][8] = f * r2 * r2 * p[i];
  }
  if (robust) whiten_norm<S, 2, 12>(1, S(a[4][0]), S(a[5][0]), r, J);
}

// Linear camera: P = R^T (X - t) with world_T_cam = (R, t); z = max(P_z, eps);
// u = fx P_x / z + cx, v = fy P_y / z + cy; Cauchy whitening. J over [dth dt | X].
template <typename S>
SYMHIP_HD void factor_reproj_cauchy(const double* const* a, S* r, S (*J)[9]) {
  const M3<S> R = rotmat(load_quat<S>(a[0]));
  const double* K = a[3];
  const S eps = S(a[5][0]);
  // the point in the camera frame, formed in FP64
  const V3<S> P = to_s<S>(mulT(rotmat(load_quat<double>(a[0])),
                               load_v3<double>(a[1]) - V3<double>{a[0][4], a[0][5], a[0][6]}));
  const bool zin = P.z > eps;
  const S z = zin ? P.z : eps;
  const S fx = S(K[0]), fy = S(K[1]);
  r[0] = fx * P.x / z + S(K[2]) - S(a[2][0]);
  r[1] = fy * P.y / z + S(K[3]) - S(a[2][1]);
  const S dudP[2][3] = {{fx / z, 0, zin ? -fx * P.x / (z * z) : S(0)},
                        {0, fy / z, zin ? -fy * P.y / (z * z) : S(0)}};
  // dP/dth = [P]x ; dP/dt = -R^T ; dP/dX = R^T
  const M3<S> PK = skew(P);
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 3; ++j) {
      J[i][j] = dudP[i][0] * PK(0, j) + dudP[i][1] * PK(1, j) + dudP[i][2] * PK(2, j);
      const S rt = dudP[i][0] * R(j, 0) + dudP[i][1] * R(j, 1) + dudP[i][2] * R(j, 2);
      J[i][3 + j] = -rt;
      J[i][6 + j] = rt;
    }
  whiten_norm<S, 2, 9>(2, S(a[4][0]), eps, r, J);
}

// Matching: pred = R^T (L - t); r = (pred - meas) / sigma. J over [dth dt].
template <typename S>
SYMHIP_HD void factor_matching(const double* const* a, S* r, S (*J)[6]) {
  const M3<S> R = rotmat(load_quat<S>(a[0]));
  // the landmark in the pose frame and its difference to the measurement, formed in FP64
  const V3<double> pred_d = mulT(rotmat(load_quat<double>(a[0])),
                                 load_v3<double>(a[1]) - V3<double>{a[0][4], a[0][5], a[0][6]});
  const V3<S> pred = to_s<S>(pred_d);
  const S inv = S(1) / S(a[3][0]);
  r[0] = S(pred_d.x - a[2][0]) * inv;
  r[1] = S(pred_d.y - a[2][1]) * inv;
  r[2] = S(pred_d.z - a[2][2]) * inv;
  const M3<S> PK = skew(pred);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      J[i][j] = PK(i, j) * inv;
      J[i][3 + j] = -R(j, i) * inv;
    }
}

// Odometry: pred = a^-1 b; e = local(pred, meas) = [log(pred^-1 meas), t_m - t_pred]; r = e/sigma.
template <typename S>
SYMHIP_HD void factor_odometry(const double* const* a, S* r, S (*J)[12]) {
  const Quat<double> qab_d = rel_quat(a[0], a[1]);
  const V3<double> dt{a[1][4] - a[0][4], a[1][5] - a[0][5], a[1][6] - a[0][6]};
  const V3<double> tab_d = mulT(rotmat(load_quat<double>(a[0])), dt);
  const M3<S> Ra = rotmat(load_quat<S>(a[0])), Rab = rotmat(to_s<S>(qab_d));
  const V3<S> tab = to_s<S>(tab_d);
  const Quat<S> qrel = to_s<S>(qmul(qconj(qab_d), load_quat<double>(a[2])));
  const M3<S> Rrel = rotmat(qrel);
  V3<S> w;
  M3<S> D;
  rot_log(qrel, S(a[4][0]), w, D);
  const double* sig = a[3];
  const S e[6] = {w.x, w.y, w.z, S(a[2][4] - tab_d.x), S(a[2][5] - tab_d.y),
                  S(a[2][6] - tab_d.z)};
  // rotation: d/dth_a = D Rrel^T Rab^T ; d/dth_b = -D Rrel^T
  const M3<S> RrT = transpose(Rrel);
  const M3<S> G = matmul(RrT, transpose(Rab));
  const M3<S> DA = matmul(D, G), DB = matmul(D, RrT);
  const M3<S> K = skew(tab);
  S de[6][12];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      de[i][j] = DA(i, j);
      de[i][3 + j] = 0;
      de[i][6 + j] = -DB(i, j);
      de[i][9 + j] = 0;
      de[3 + i][j] = -K(i, j);
      de[3 + i][3 + j] = Ra(j, i);
      de[3 + i][6 + j] = 0;
      de[3 + i][9 + j] = -Ra(j, i);
    }
  for (int i = 0; i < 6; ++i) {
    const S inv = S(1) / S(sig[i]);
    r[i] = e[i] * inv;
    for (int c = 0; c < 12; ++c) J[i][c] = de[i][c] * inv;
  }
}

// Graduated non-convexity factor: per element e = x_i - y_i, the Barron loss
//   rho(e) = (b / d) ((e^2 / b + 1)^(d / 2) - 1),  b = |alpha - 2| + eps,
//   d = alpha + eps (alpha >= 0) or alpha - eps,  alpha = 2 - 1 / (1 - mu + eps)
// (delta 1, unit information; symforce/opt/noise_models.py:382-436), whitened as
// r_i = sqrt(2 rho(e) + eps) - sqrt(eps). Its derivative in x_i:
// drho/de = e (e^2 / b + 1)^(d / 2 - 1), so dr_i/dx_i = drho/de / sqrt(2 rho + eps). (At
// alpha -> 0, mu -> 1/2, the form cancels: t^(d/2) - 1 is a few ulps.)
template <typename S>
SYMHIP_HD void factor_barron5(const double* const* a, S* r, S (*J)[5]) {
  const double mu = a[2][0], eps = a[3][0];
  const double alpha = 2.0 - 1.0 / (1.0 - mu + eps);
  const double b = fabs(alpha - 2.0) + eps;
  const double d = alpha >= 0.0 ? alpha + eps : alpha - eps;
  const double se = sqrt(eps);
  for (int i = 0; i < 5; ++i) {
    const double e = a[0][i] - a[1][i];
    const double t = e * e / b + 1.0;
    const double pd = pow(t, 0.5 * d);  // t^(d/2)
    const double rho = (b / d) * (pd - 1.0);
    const double w = sqrt(2.0 * rho + eps);
    r[i] = S(w - se);
    for (int c = 0; c < 5; ++c) J[i][c] = S(0);
    J[i][i] = S(e * (pd / t) / w);
  }
}

}  // namespace symhip
