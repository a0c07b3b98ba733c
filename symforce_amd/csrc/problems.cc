// Seeded synthetic problem generators (host only). See include/symhip_problems.h.
//
// Random draws use std::mt19937 with std::normal_distribution / std::uniform_*_distribution, the
// same engines the reference tests use, so problems built here from a reference test's recipe
// (test/symforce_optimizer_test.cc:79-131, :183-264, :276-348) reproduce its inputs exactly.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <limits>
#include <random>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/symhip_problems.h"

namespace {

constexpr int64_t kInvalid = std::numeric_limits<int64_t>::min();
constexpr double kDefaultEpsilon = 10.0 * std::numeric_limits<double>::epsilon();

struct Q {
  double x, y, z, w;
};
Q qmul(const Q& a, const Q& b) {
  return {a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y, a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x,
          a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w, a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z};
}
Q qconj(const Q& a) { return {-a.x, -a.y, -a.z, a.w}; }
void qrot(const Q& q, const double* v, double* out) {
  const Q p{v[0], v[1], v[2], 0.0};
  const Q r = qmul(qmul(q, p), qconj(q));
  out[0] = r.x;
  out[1] = r.y;
  out[2] = r.z;
}
// Rot3::FromTangent / retraction exponential: [sin(t/2)/t w, cos(t/2)], t = sqrt(eps^2+|w|^2)
Q qexp(const double* w, double eps) {
  const double t = std::sqrt(eps * eps + w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
  const double s = std::sin(t / 2) / t;
  return {s * w[0], s * w[1], s * w[2], std::cos(t / 2)};
}
Q ypr(double yaw, double pitch, double roll) {
  const Q qz{0, 0, std::sin(yaw / 2), std::cos(yaw / 2)};
  const Q qy{0, std::sin(pitch / 2), 0, std::cos(pitch / 2)};
  const Q qx{std::sin(roll / 2), 0, 0, std::cos(roll / 2)};
  return qmul(qmul(qz, qy), qx);
}
// Quaternion from a rotation matrix given by its columns (world_R_cam = [x y z]).
Q from_columns(const double* x, const double* y, const double* z) {
  const double m00 = x[0], m10 = x[1], m20 = x[2];
  const double m01 = y[0], m11 = y[1], m21 = y[2];
  const double m02 = z[0], m12 = z[1], m22 = z[2];
  const double tr = m00 + m11 + m22;
  Q q;
  if (tr > 0) {
    const double s = std::sqrt(tr + 1.0) * 2;
    q = {(m21 - m12) / s, (m02 - m20) / s, (m10 - m01) / s, 0.25 * s};
  } else if (m00 > m11 && m00 > m22) {
    const double s = std::sqrt(1.0 + m00 - m11 - m22) * 2;
    q = {0.25 * s, (m01 + m10) / s, (m02 + m20) / s, (m21 - m12) / s};
  } else if (m11 > m22) {
    const double s = std::sqrt(1.0 + m11 - m00 - m22) * 2;
    q = {(m01 + m10) / s, 0.25 * s, (m12 + m21) / s, (m02 - m20) / s};
  } else {
    const double s = std::sqrt(1.0 + m22 - m00 - m11) * 2;
    q = {(m02 + m20) / s, (m12 + m21) / s, 0.25 * s, (m10 - m01) / s};
  }
  const double n = std::sqrt(q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w);
  return {q.x / n, q.y / n, q.z / n, q.w / n};
}

struct Pose {
  Q q;
  double t[3];
};
Pose pinv(const Pose& a) {
  Pose r;
  r.q = qconj(a.q);
  double ti[3];
  qrot(r.q, a.t, ti);
  r.t[0] = -ti[0];
  r.t[1] = -ti[1];
  r.t[2] = -ti[2];
  return r;
}
Pose pcompose(const Pose& a, const Pose& b) {
  Pose r;
  r.q = qmul(a.q, b.q);
  double rb[3];
  qrot(a.q, b.t, rb);
  for (int i = 0; i < 3; ++i) r.t[i] = a.t[i] + rb[i];
  return r;
}
void pact(const Pose& a, const double* p, double* out) {
  qrot(a.q, p, out);
  for (int i = 0; i < 3; ++i) out[i] += a.t[i];
}
// Pose3 retraction (product manifold): rotation right-multiplied, translation added.
Pose pretract(const Pose& a, const double* d, double eps) {
  Pose r;
  r.q = qmul(a.q, qexp(d, eps));
  for (int i = 0; i < 3; ++i) r.t[i] = a.t[i] + d[3 + i];
  return r;
}

}  // namespace

struct symhip_generated {
  std::vector<int8_t> letters;
  std::vector<int64_t> subs, supers;
  std::vector<int32_t> types;
  std::vector<double> values;
  std::vector<int32_t> ftypes, fptr{0}, fargs;
  std::vector<int32_t> opt_keys;
  std::unordered_map<std::string, double> info;

  int32_t add_key(char letter, int64_t sub, int64_t super, int32_t type, const double* v, int n) {
    letters.push_back((int8_t)letter);
    subs.push_back(sub);
    supers.push_back(super);
    types.push_back(type);
    values.insert(values.end(), v, v + n);
    return (int32_t)letters.size() - 1;
  }
  int32_t add_pose(char l, int64_t sub, int64_t super, const Pose& p) {
    const double v[7] = {p.q.x, p.q.y, p.q.z, p.q.w, p.t[0], p.t[1], p.t[2]};
    return add_key(l, sub, super, SYMHIP_TYPE_POSE3, v, 7);
  }
  int32_t add_rot(char l, int64_t sub, int64_t super, const Q& q) {
    const double v[4] = {q.x, q.y, q.z, q.w};
    return add_key(l, sub, super, SYMHIP_TYPE_ROT3, v, 4);
  }
  int32_t add_scalar(char l, int64_t sub, int64_t super, double s) {
    return add_key(l, sub, super, SYMHIP_TYPE_SCALAR, &s, 1);
  }
  int32_t add_vec(char l, int64_t sub, int64_t super, int32_t type, std::initializer_list<double> v) {
    std::vector<double> tmp(v);
    return add_key(l, sub, super, type, tmp.data(), (int)tmp.size());
  }
  int32_t add_diag(char l, int64_t sub, int64_t super, int n, const double* d) {
    std::vector<double> m(n * n, 0.0);
    for (int i = 0; i < n; ++i) m[i * n + i] = d[i];
    return add_key(l, sub, super, n == 3 ? SYMHIP_TYPE_MATRIX33 : SYMHIP_TYPE_MATRIX66, m.data(),
                   n * n);
  }
  void add_factor(int32_t type, std::initializer_list<int32_t> args) {
    ftypes.push_back(type);
    fargs.insert(fargs.end(), args.begin(), args.end());
    fptr.push_back((int32_t)fargs.size());
  }
};

extern "C" {

symhip_generated* symhip_gen_pose_smoothing(int32_t num_keys, uint32_t seed) {
  auto* g = new symhip_generated();
  const double epsilon = 1e-10;
  Pose start{ypr(0, 0, 0), {0, 0, 0}};
  Pose last{ypr(M_PI, 0, 0), {5, 0, 0}};
  const double sp[6] = {10, 10, 10, 10, 10, 10};  // 1 / 0.1
  const double sb[6] = {2, 2, 2, 2, 2, 2};        // 1 / 0.5
  std::mt19937 gen(seed);
  std::vector<int32_t> keys;
  for (int i = 0; i < num_keys; ++i) {
    std::normal_distribution<double> dist{};
    double v[6];
    for (int k = 0; k < 6; ++k) v[k] = dist(gen);
    for (int k = 0; k < 6; ++k) v[k] *= 0.4;
    keys.push_back(g->add_pose('P', i, kInvalid, pretract(start, v, kDefaultEpsilon)));
  }
  const int32_t k_start = g->add_pose('a', 0, kInvalid, start);
  const int32_t k_last = g->add_pose('a', 1, kInvalid, last);
  const int32_t k_identity = g->add_pose('a', 2, kInvalid, Pose{{0, 0, 0, 1}, {0, 0, 0}});
  const int32_t k_sp = g->add_diag('s', 0, kInvalid, 6, sp);
  const int32_t k_sb = g->add_diag('s', 1, kInvalid, 6, sb);
  const int32_t k_eps = g->add_scalar('e', kInvalid, kInvalid, epsilon);
  g->add_factor(SYMHIP_FACTOR_PRIOR_POSE3, {keys[0], k_start, k_sp, k_eps});
  g->add_factor(SYMHIP_FACTOR_PRIOR_POSE3, {keys[num_keys - 1], k_last, k_sp, k_eps});
  for (int i = 0; i < num_keys - 1; ++i)
    g->add_factor(SYMHIP_FACTOR_BETWEEN_POSE3, {keys[i], keys[i + 1], k_identity, k_sb, k_eps});
  g->info["epsilon"] = kDefaultEpsilon;
  g->info["c_dim"] = 0;
  return g;
}

symhip_generated* symhip_gen_rot_smoothing(int32_t num_keys, uint32_t seed) {
  auto* g = new symhip_generated();
  const double epsilon = 1e-15;
  const Q start = ypr(0, 0, 0), last = ypr(M_PI, 0, 0);
  const double sp[3] = {10, 10, 10}, sb[3] = {2, 2, 2};
  std::mt19937 gen(seed);
  std::vector<int32_t> keys;
  for (int i = 0; i < num_keys; ++i) {
    std::normal_distribution<double> dist{};
    double v[3];
    for (int k = 0; k < 3; ++k) v[k] = dist(gen);
    for (int k = 0; k < 3; ++k) v[k] *= 0.4;
    keys.push_back(g->add_rot('R', i, kInvalid, qmul(start, qexp(v, kDefaultEpsilon))));
  }
  const int32_t k_start = g->add_rot('a', 0, kInvalid, start);
  const int32_t k_last = g->add_rot('a', 1, kInvalid, last);
  const int32_t k_identity = g->add_rot('a', 2, kInvalid, Q{0, 0, 0, 1});
  const int32_t k_sp = g->add_diag('s', 0, kInvalid, 3, sp);
  const int32_t k_sb = g->add_diag('s', 1, kInvalid, 3, sb);
  const int32_t k_eps = g->add_scalar('e', kInvalid, kInvalid, epsilon);
  g->add_factor(SYMHIP_FACTOR_PRIOR_ROT3, {keys[0], k_start, k_sp, k_eps});
  g->add_factor(SYMHIP_FACTOR_PRIOR_ROT3, {keys[num_keys - 1], k_last, k_sp, k_eps});
  for (int i = 0; i < num_keys - 1; ++i)
    g->add_factor(SYMHIP_FACTOR_BETWEEN_ROT3, {keys[i], keys[i + 1], k_identity, k_sb, k_eps});
  g->info["epsilon"] = kDefaultEpsilon;
  g->info["c_dim"] = 0;
  return g;
}

symhip_generated* symhip_gen_rot_frozen(int32_t num_keys, uint32_t seed) {
  auto* g = new symhip_generated();
  const double epsilon = 1e-15;
  const double sb[3] = {2, 2, 2};
  std::mt19937 gen(seed);
  std::vector<int32_t> keys;
  for (int i = 0; i < num_keys; ++i) {
    std::normal_distribution<double> dist{};
    double v[3];
    for (int k = 0; k < 3; ++k) v[k] = dist(gen);
    for (int k = 0; k < 3; ++k) v[k] *= 0.4;
    keys.push_back(g->add_rot('R', i, kInvalid, qexp(v, kDefaultEpsilon)));
  }
  const int32_t k_identity = g->add_rot('a', 2, kInvalid, Q{0, 0, 0, 1});
  const int32_t k_sb = g->add_diag('s', 1, kInvalid, 3, sb);
  const int32_t k_eps = g->add_scalar('e', kInvalid, kInvalid, epsilon);
  for (int i = 0; i < num_keys; ++i)
    for (int j = 0; j < num_keys; ++j)
      if (i != j)
        g->add_factor(SYMHIP_FACTOR_BETWEEN_ROT3, {keys[i], keys[j], k_identity, k_sb, k_eps});
  for (int i = 1; i < num_keys; ++i) g->opt_keys.push_back(keys[i]);
  g->info["epsilon"] = kDefaultEpsilon;
  g->info["c_dim"] = 0;
  return g;
}

symhip_generated* symhip_gen_ba_circle(int32_t n_cams, int32_t n_points, uint32_t seed) {
  auto* g = new symhip_generated();
  std::mt19937 gen(seed);
  std::normal_distribution<double> nrm(0.0, 1.0);
  std::uniform_real_distribution<double> uni(-2.0, 2.0);
  const double fx = 500, fy = 500, cx = 320, cy = 240;
  std::vector<Pose> truth(n_cams);
  for (int i = 0; i < n_cams; ++i) {
    const double th = 2 * M_PI * i / n_cams;
    const double pos[3] = {10 * std::cos(th), 10 * std::sin(th), 0};
    const double z[3] = {-std::cos(th), -std::sin(th), 0};
    const double y[3] = {0, 0, -1};
    const double x[3] = {y[1] * z[2] - y[2] * z[1], y[2] * z[0] - y[0] * z[2],
                         y[0] * z[1] - y[1] * z[0]};
    truth[i].q = from_columns(x, y, z);
    std::copy(pos, pos + 3, truth[i].t);
  }
  std::vector<std::array<double, 3>> pts(n_points);
  for (auto& p : pts)
    for (double& c : p) c = uni(gen);
  // observations (point-major)
  struct Obs {
    int cam, pt;
    double u, v;
  };
  std::vector<Obs> obs;
  for (int j = 0; j < n_points; ++j)
    for (int i = 0; i < n_cams; ++i) {
      double pc[3];
      pact(pinv(truth[i]), pts[j].data(), pc);
      if (pc[2] <= 0) continue;
      const double u = fx * pc[0] / pc[2] + cx, v = fy * pc[1] / pc[2] + cy;
      if (u < 0 || u >= 640 || v < 0 || v >= 480) continue;
      obs.push_back({i, j, u + nrm(gen), v + nrm(gen)});
    }
  std::vector<int32_t> kc(n_cams), kp(n_points);
  for (int i = 0; i < n_cams; ++i) {
    double d[6];
    for (int k = 0; k < 3; ++k) d[k] = 0.05 * nrm(gen);
    for (int k = 3; k < 6; ++k) d[k] = 0.1 * nrm(gen);
    kc[i] = g->add_pose('c', kInvalid, i, i == 0 ? truth[0] : pretract(truth[i], d, 0.0));
  }
  for (int j = 0; j < n_points; ++j) {
    const double v[3] = {pts[j][0] + 0.1 * nrm(gen), pts[j][1] + 0.1 * nrm(gen),
                         pts[j][2] + 0.1 * nrm(gen)};
    kp[j] = g->add_key('p', kInvalid, j, SYMHIP_TYPE_VECTOR3, v, 3);
  }
  const int32_t kK = g->add_vec('k', kInvalid, kInvalid, SYMHIP_TYPE_VECTOR4, {fx, fy, cx, cy});
  const int32_t kw = g->add_scalar('w', kInvalid, kInvalid, 1.0);
  const int32_t ke = g->add_scalar('e', kInvalid, kInvalid, kDefaultEpsilon);
  const int32_t kprior0 = g->add_pose('a', 0, kInvalid, truth[0]);
  const int32_t kprior1 = g->add_pose('a', 1, kInvalid, truth[1]);
  const double s0[6] = {1e3, 1e3, 1e3, 1e3, 1e3, 1e3};
  const double s1[6] = {0, 0, 0, 10, 10, 10};
  const int32_t ks0 = g->add_diag('s', 0, kInvalid, 6, s0);
  const int32_t ks1 = g->add_diag('s', 1, kInvalid, 6, s1);
  g->add_factor(SYMHIP_FACTOR_PRIOR_POSE3, {kc[0], kprior0, ks0, ke});
  if (n_cams > 1) g->add_factor(SYMHIP_FACTOR_PRIOR_POSE3, {kc[1], kprior1, ks1, ke});
  for (size_t o = 0; o < obs.size(); ++o) {
    const int32_t kx = g->add_vec('x', kInvalid, (int64_t)o, SYMHIP_TYPE_VECTOR2, {obs[o].u, obs[o].v});
    g->add_factor(SYMHIP_FACTOR_REPROJ_CAUCHY, {kc[obs[o].cam], kp[obs[o].pt], kx, kK, kw, ke});
  }
  g->info["epsilon"] = kDefaultEpsilon;
  g->info["c_dim"] = 3.0 * n_points;
  g->info["n_cams"] = n_cams;
  g->info["n_points"] = n_points;
  g->info["n_obs"] = (double)obs.size();
  return g;
}

symhip_generated* symhip_gen_ba_snavely(int32_t n_cams, int32_t n_points, uint32_t seed) {
  auto* g = new symhip_generated();
  std::mt19937 gen(seed);
  std::normal_distribution<double> nrm(0.0, 1.0);
  std::uniform_real_distribution<double> u01(0.0, 1.0);
  // Camera trajectory: planar random walk, 1 m steps, slowly turning heading; cameras look to
  // the left of the direction of travel. BAL convention: the camera looks along -z.
  std::vector<Pose> world_T_cam(n_cams);
  double pos[3] = {0, 0, 0}, heading = 0;
  for (int i = 0; i < n_cams; ++i) {
    if (i > 0) {
      heading += 0.03 * nrm(gen);
      pos[0] += std::cos(heading);
      pos[1] += std::sin(heading);
      pos[2] += 0.05 * nrm(gen);
    }
    const double x[3] = {std::cos(heading), std::sin(heading), 0};
    const double y[3] = {0, 0, 1};
    const double z[3] = {std::sin(heading), -std::cos(heading), 0};
    world_T_cam[i].q = from_columns(x, y, z);
    std::copy(pos, pos + 3, world_T_cam[i].t);
  }
  // Intrinsics truth (f, k1, k2)
  std::vector<std::array<double, 3>> intr(n_cams);
  for (auto& k : intr) k = {500 + 20 * nrm(gen), 1e-3 * nrm(gen), 1e-3 * nrm(gen)};
  // Track-length distribution: discrete power law on [2, 40] with exponent chosen for mean 5.
  const int tmin = 2, tmax = std::min(40, n_cams);
  auto mean_for = [&](double a) {
    double s = 0, m = 0;
    for (int t = tmin; t <= tmax; ++t) {
      s += std::pow(t, -a);
      m += t * std::pow(t, -a);
    }
    return m / s;
  };
  double lo = 0.0, hi = 6.0;
  for (int it = 0; it < 100; ++it) {
    const double mid = 0.5 * (lo + hi);
    if (mean_for(mid) > 5.0) lo = mid; else hi = mid;
  }
  const double alpha = 0.5 * (lo + hi);
  std::vector<double> cdf;
  double acc = 0;
  for (int t = tmin; t <= tmax; ++t) {
    acc += std::pow(t, -alpha);
    cdf.push_back(acc);
  }
  for (double& c : cdf) c /= acc;
  struct Obs {
    int cam, pt;
    double u, v;
  };
  std::vector<Obs> obs;
  std::vector<std::array<double, 3>> pts(n_points);
  std::vector<Pose> cam_T_world(n_cams);
  for (int i = 0; i < n_cams; ++i) cam_T_world[i] = pinv(world_T_cam[i]);
  for (int j = 0; j < n_points; ++j) {
    const double r = u01(gen);
    const int t = tmin + (int)(std::lower_bound(cdf.begin(), cdf.end(), r) - cdf.begin());
    const int s = (int)(u01(gen) * (n_cams - t + 1)) % (n_cams - t + 1);
    const int m = s + t / 2;
    const double depth = 6.0 + 6.0 * u01(gen) + 0.5 * t;
    const double along = (u01(gen) - 0.5) * 2.0;
    const double height = (u01(gen) - 0.5) * 4.0;
    // point in the middle camera's frame (looking along -z): x along track, y up, z = -depth
    const double pc[3] = {along, height, -depth};
    pact(world_T_cam[m], pc, pts[j].data());
    for (int i = s; i < s + t; ++i) {
      double q[3];
      pact(cam_T_world[i], pts[j].data(), q);
      if (q[2] > -3.0) continue;  // behind or too close: not observed
      const double px = -q[0] / q[2], py = -q[1] / q[2];
      if (std::fabs(px) > 1.2 || std::fabs(py) > 1.2) continue;  // outside the field of view
      const double r2 = px * px + py * py;
      const double rad = 1.0 + intr[i][1] * r2 + intr[i][2] * r2 * r2;
      obs.push_back({i, j, intr[i][0] * rad * px + nrm(gen), intr[i][0] * rad * py + nrm(gen)});
    }
  }
  std::vector<int32_t> kc(n_cams), ki(n_cams), kp(n_points);
  for (int i = 0; i < n_cams; ++i) {
    // perturb the camera's pose in the world (orientation about its own center, position), then
    // store the inverse as cam_T_world
    double d[6];
    for (int k = 0; k < 3; ++k) d[k] = 0.05 * nrm(gen);
    for (int k = 3; k < 6; ++k) d[k] = 0.1 * nrm(gen);
    kc[i] = g->add_pose('c', kInvalid, i,
                        i == 0 ? cam_T_world[0] : pinv(pretract(world_T_cam[i], d, 0.0)));
  }
  for (int i = 0; i < n_cams; ++i)
    ki[i] = g->add_key('i', kInvalid, i, SYMHIP_TYPE_VECTOR3, intr[i].data(), 3);
  for (int j = 0; j < n_points; ++j) {
    const double v[3] = {pts[j][0] + 0.1 * nrm(gen), pts[j][1] + 0.1 * nrm(gen),
                         pts[j][2] + 0.1 * nrm(gen)};
    kp[j] = g->add_key('p', kInvalid, j, SYMHIP_TYPE_VECTOR3, v, 3);
  }
  const int32_t kh = g->add_scalar('h', kInvalid, kInvalid, 2.0);
  const int32_t ke = g->add_scalar('e', kInvalid, kInvalid, kDefaultEpsilon);
  const int32_t kprior0 = g->add_pose('a', 0, kInvalid, cam_T_world[0]);
  const int32_t kprior1 = g->add_pose('a', 1, kInvalid, cam_T_world[std::min(1, n_cams - 1)]);
  const double s0[6] = {1e3, 1e3, 1e3, 1e3, 1e3, 1e3};
  const double s1[6] = {0, 0, 0, 10, 10, 10};
  const int32_t ks0 = g->add_diag('s', 0, kInvalid, 6, s0);
  const int32_t ks1 = g->add_diag('s', 1, kInvalid, 6, s1);
  g->add_factor(SYMHIP_FACTOR_PRIOR_POSE3, {kc[0], kprior0, ks0, ke});
  if (n_cams > 1) g->add_factor(SYMHIP_FACTOR_PRIOR_POSE3, {kc[1], kprior1, ks1, ke});
  for (size_t o = 0; o < obs.size(); ++o) {
    const int32_t kx = g->add_vec('x', kInvalid, (int64_t)o, SYMHIP_TYPE_VECTOR2, {obs[o].u, obs[o].v});
    g->add_factor(SYMHIP_FACTOR_SNAVELY_HUBER,
                  {kc[obs[o].cam], ki[obs[o].cam], kp[obs[o].pt], kx, kh, ke});
  }
  g->info["epsilon"] = kDefaultEpsilon;
  g->info["c_dim"] = 3.0 * n_points;
  g->info["n_cams"] = n_cams;
  g->info["n_points"] = n_points;
  g->info["n_obs"] = (double)obs.size();
  g->info["track_alpha"] = alpha;
  return g;
}

symhip_generated* symhip_gen_pose_graph(int32_t n_nodes, int32_t n_loops, uint32_t seed) {
  auto* g = new symhip_generated();
  std::mt19937 gen(seed);
  std::normal_distribution<double> nrm(0.0, 1.0);
  std::uniform_real_distribution<double> u01(0.0, 1.0);
  std::vector<Pose> truth(n_nodes);
  double heading = 0, pos[3] = {0, 0, 0};
  for (int i = 0; i < n_nodes; ++i) {
    if (i > 0) {
      heading += 0.2 * nrm(gen);
      pos[0] += std::cos(heading);
      pos[1] += std::sin(heading);
    }
    truth[i].q = ypr(heading, 0, 0);
    std::copy(pos, pos + 3, truth[i].t);
  }
  auto noisy = [&](const Pose& rel) {
    double d[6];
    for (int k = 0; k < 3; ++k) d[k] = 0.01 * nrm(gen);
    for (int k = 3; k < 6; ++k) d[k] = 0.05 * nrm(gen);
    return pretract(rel, d, 0.0);
  };
  std::vector<Pose> odo(n_nodes > 0 ? n_nodes - 1 : 0);
  for (int i = 0; i + 1 < n_nodes; ++i) odo[i] = noisy(pcompose(pinv(truth[i]), truth[i + 1]));
  // Loop closure candidates: nodes within 2 m, at least 20 steps apart (2 m grid hash).
  std::unordered_map<int64_t, std::vector<int>> grid;
  auto cell = [](double x) { return (int64_t)std::floor(x / 2.0); };
  std::vector<std::pair<int, int>> cand;
  for (int i = 0; i < n_nodes; ++i) {
    const int64_t cxi = cell(truth[i].t[0]), cyi = cell(truth[i].t[1]);
    for (int64_t dx = -1; dx <= 1; ++dx)
      for (int64_t dy = -1; dy <= 1; ++dy) {
        auto it = grid.find(((cxi + dx) << 32) ^ ((cyi + dy) & 0xffffffff));
        if (it == grid.end()) continue;
        for (int j : it->second) {
          if (i - j < 20) continue;
          const double ddx = truth[i].t[0] - truth[j].t[0], ddy = truth[i].t[1] - truth[j].t[1];
          if (ddx * ddx + ddy * ddy < 4.0) cand.push_back({j, i});
        }
      }
    grid[(cxi << 32) ^ (cyi & 0xffffffff)].push_back(i);
  }
  std::shuffle(cand.begin(), cand.end(), gen);
  if ((int)cand.size() > n_loops) cand.resize(n_loops);
  std::sort(cand.begin(), cand.end());
  // Initial guess: odometry integration from the true first node.
  std::vector<int32_t> kx(n_nodes);
  Pose cur = truth[0];
  for (int i = 0; i < n_nodes; ++i) {
    if (i > 0) cur = pcompose(cur, odo[i - 1]);
    kx[i] = g->add_pose('x', i, kInvalid, cur);
  }
  const double si[6] = {100, 100, 100, 20, 20, 20};
  const double sp[6] = {1e3, 1e3, 1e3, 1e3, 1e3, 1e3};
  const int32_t ksi = g->add_diag('s', 0, kInvalid, 6, si);
  const int32_t ksp = g->add_diag('s', 1, kInvalid, 6, sp);
  const int32_t ke = g->add_scalar('e', kInvalid, kInvalid, kDefaultEpsilon);
  const int32_t kc = g->add_scalar('w', kInvalid, kInvalid, 1.0);
  const int32_t kprior = g->add_pose('a', 0, kInvalid, truth[0]);
  g->add_factor(SYMHIP_FACTOR_PRIOR_POSE3, {kx[0], kprior, ksp, ke});
  for (int i = 0; i + 1 < n_nodes; ++i) {
    const int32_t km = g->add_pose('m', 0, i, odo[i]);
    g->add_factor(SYMHIP_FACTOR_BETWEEN_POSE3, {kx[i], kx[i + 1], km, ksi, ke});
  }
  int n_out = 0;
  for (size_t l = 0; l < cand.size(); ++l) {
    const auto [a, b] = cand[l];
    Pose meas;
    if (u01(gen) < 0.01) {
      meas.q = ypr(2 * M_PI * u01(gen), 0.5 * (u01(gen) - 0.5), 0.5 * (u01(gen) - 0.5));
      for (int k = 0; k < 3; ++k) meas.t[k] = 10 * (u01(gen) - 0.5);
      n_out++;
    } else {
      meas = noisy(pcompose(pinv(truth[a]), truth[b]));
    }
    const int32_t km = g->add_pose('m', 1, (int64_t)l, meas);
    g->add_factor(SYMHIP_FACTOR_BETWEEN_POSE3_CAUCHY, {kx[a], kx[b], km, ksi, kc, ke});
  }
  g->info["epsilon"] = kDefaultEpsilon;
  g->info["c_dim"] = 0;
  g->info["n_loops"] = (double)cand.size();
  g->info["n_outliers"] = n_out;
  return g;
}

symhip_generated* symhip_gen_localization(int32_t n_poses, int32_t n_landmarks,
                                          double visible_fraction, uint32_t seed) {
  auto* g = new symhip_generated();
  std::mt19937 gen(seed);
  std::normal_distribution<double> nrm(0.0, 1.0);
  std::uniform_real_distribution<double> u01(0.0, 1.0);
  std::vector<Pose> truth(n_poses);
  double heading = 2 * M_PI * u01(gen), pos[3] = {0, 0, 0};
  for (int i = 0; i < n_poses; ++i) {
    if (i > 0) {
      heading += 0.1 * nrm(gen);
      pos[0] += 0.5 * std::cos(heading);
      pos[1] += 0.5 * std::sin(heading);
      pos[2] += 0.02 * nrm(gen);
    }
    truth[i].q = ypr(heading, 0.02 * nrm(gen), 0.02 * nrm(gen));
    std::copy(pos, pos + 3, truth[i].t);
  }
  std::vector<std::array<double, 3>> lm(n_landmarks);
  for (auto& l : lm) {
    const int anchor = (int)(u01(gen) * n_poses) % n_poses;
    for (int k = 0; k < 3; ++k) l[k] = truth[anchor].t[k] + 10 * (u01(gen) - 0.5);
  }
  const double msig = 0.05;
  const double osig[6] = {0.01, 0.01, 0.01, 0.02, 0.02, 0.02};
  std::vector<int32_t> kw(n_poses), kl(n_landmarks);
  for (int i = 0; i < n_poses; ++i) {
    double d[6];
    for (int k = 0; k < 3; ++k) d[k] = 0.1 * nrm(gen);
    for (int k = 3; k < 6; ++k) d[k] = 0.3 * nrm(gen);
    kw[i] = g->add_pose('w', kInvalid, i, pretract(truth[i], d, 0.0));
  }
  for (int j = 0; j < n_landmarks; ++j)
    kl[j] = g->add_key('l', kInvalid, j, SYMHIP_TYPE_VECTOR3, lm[j].data(), 3);
  const int32_t kms = g->add_scalar('g', kInvalid, kInvalid, msig);
  const int32_t kos = g->add_key('s', kInvalid, kInvalid, SYMHIP_TYPE_VECTOR6, osig, 6);
  const int32_t ke = g->add_scalar('e', kInvalid, kInvalid, kDefaultEpsilon);
  for (int i = 0; i < n_poses; ++i) {
    bool any = false;
    for (int j = 0; j < n_landmarks; ++j) {
      const bool last_chance = (j == n_landmarks - 1) && !any;
      if (!(u01(gen) < visible_fraction) && !last_chance) continue;
      any = true;
      double b[3];
      pact(pinv(truth[i]), lm[j].data(), b);
      for (int k = 0; k < 3; ++k) b[k] += msig * nrm(gen);
      const int32_t km = g->add_key('m', i, j, SYMHIP_TYPE_VECTOR3, b, 3);
      g->add_factor(SYMHIP_FACTOR_MATCHING, {kw[i], kl[j], km, kms});
    }
  }
  for (int i = 0; i + 1 < n_poses; ++i) {
    Pose rel = pcompose(pinv(truth[i]), truth[i + 1]);
    double d[6];
    for (int k = 0; k < 6; ++k) d[k] = osig[k] * nrm(gen);
    const int32_t ko = g->add_pose('o', kInvalid, i, pretract(rel, d, 0.0));
    g->add_factor(SYMHIP_FACTOR_ODOMETRY, {kw[i], kw[i + 1], ko, kos, ke});
  }
  g->info["epsilon"] = kDefaultEpsilon;
  g->info["c_dim"] = 0;
  return g;
}

// BAL text format (bundle_adjustment_in_the_large.cc:60-118): header "cameras points
// observations", observations "camera point x y", cameras "rx ry rz tx ty tz f k1 k2" (angle-axis
// rotation of cam_T_world, converted with Rot3::FromTangent at the default epsilon), points
// "x y z". Keys as the reference example's gen/keys.h: 'c' pose, 'i' intrinsics, 'p' point,
// 'P' pixel (super = index), 'e' epsilon. huber_k > 0 selects the Huber-whitened factor with
// threshold key 'h', else the plain SnavelyReprojectionFactor. Returns NULL on a read error.
symhip_generated* symhip_load_bal(const char* path, double huber_k) {
  std::ifstream f(path);
  if (!f) return nullptr;
  int64_t nc = 0, np = 0, no = 0;
  if (!(f >> nc >> np >> no) || nc <= 0 || np <= 0 || no <= 0) return nullptr;
  struct O {
    int64_t cam, pt;
    double u, v;
  };
  std::vector<O> obs((size_t)no);
  for (auto& o : obs)
    if (!(f >> o.cam >> o.pt >> o.u >> o.v) || o.cam < 0 || o.cam >= nc || o.pt < 0 || o.pt >= np)
      return nullptr;
  auto* g = new symhip_generated();
  std::vector<int32_t> kc(nc), ki(nc), kp(np);
  std::vector<std::array<double, 3>> intr(nc);
  for (int64_t i = 0; i < nc; ++i) {
    double r[3], t[3];
    if (!(f >> r[0] >> r[1] >> r[2] >> t[0] >> t[1] >> t[2] >> intr[i][0] >> intr[i][1] >>
          intr[i][2])) {
      delete g;
      return nullptr;
    }
    Pose p;
    p.q = qexp(r, kDefaultEpsilon);
    std::copy(t, t + 3, p.t);
    kc[i] = g->add_pose('c', kInvalid, i, p);
  }
  for (int64_t i = 0; i < nc; ++i)
    ki[i] = g->add_key('i', kInvalid, i, SYMHIP_TYPE_VECTOR3, intr[i].data(), 3);
  for (int64_t j = 0; j < np; ++j) {
    double v[3];
    if (!(f >> v[0] >> v[1] >> v[2])) {
      delete g;
      return nullptr;
    }
    kp[j] = g->add_key('p', kInvalid, j, SYMHIP_TYPE_VECTOR3, v, 3);
  }
  const bool robust = huber_k > 0;
  const int32_t kh = robust ? g->add_scalar('h', kInvalid, kInvalid, huber_k) : -1;
  const int32_t ke = g->add_scalar('e', kInvalid, kInvalid, kDefaultEpsilon);
  for (int64_t o = 0; o < no; ++o) {
    const int32_t kx = g->add_vec('P', kInvalid, o, SYMHIP_TYPE_VECTOR2, {obs[o].u, obs[o].v});
    if (robust)
      g->add_factor(SYMHIP_FACTOR_SNAVELY_HUBER, {kc[obs[o].cam], ki[obs[o].cam], kp[obs[o].pt], kx,
                                                  kh, ke});
    else
      g->add_factor(SYMHIP_FACTOR_SNAVELY, {kc[obs[o].cam], ki[obs[o].cam], kp[obs[o].pt], kx, ke});
  }
  g->info["epsilon"] = kDefaultEpsilon;
  g->info["c_dim"] = 3.0 * (double)np;
  g->info["n_cams"] = (double)nc;
  g->info["n_points"] = (double)np;
  g->info["n_obs"] = (double)no;
  return g;
}

// Writes the Snavely observations, cameras and points of a problem in BAL text (rotation as the
// angle-axis logarithm of the cam_T_world quaternion, cameras and points numbered in key
// order). Factors other than the Snavely reprojection (e.g. gauge priors) have no BAL form
// and are skipped. Returns 0, or -1 when there is no Snavely factor or the file cannot be written.
int symhip_save_bal(const symhip_problem_desc* d, const char* path) {
  if (!d || !path) return -1;
  std::vector<int64_t> off((size_t)d->n_keys + 1, 0);
  for (int32_t k = 0; k < d->n_keys; ++k) {
    int w = 0;
    switch (d->key_types[k]) {
      case SYMHIP_TYPE_SCALAR: w = 1; break;
      case SYMHIP_TYPE_ROT3: w = 4; break;
      case SYMHIP_TYPE_POSE3: w = 7; break;
      case SYMHIP_TYPE_VECTOR2: w = 2; break;
      case SYMHIP_TYPE_VECTOR3: w = 3; break;
      case SYMHIP_TYPE_VECTOR4: w = 4; break;
      case SYMHIP_TYPE_VECTOR6: w = 6; break;
      case SYMHIP_TYPE_MATRIX33: w = 9; break;
      case SYMHIP_TYPE_MATRIX66: w = 36; break;
      default: return -1;
    }
    off[k + 1] = off[k] + w;
  }
  // cameras and points numbered in key order (the order they were added to the Values)
  std::vector<int32_t> cam_of_pose((size_t)d->n_keys, -1), pt_key;
  std::vector<int64_t> num((size_t)d->n_keys, -1);
  std::vector<char> is_cam((size_t)d->n_keys, 0), is_pt((size_t)d->n_keys, 0);
  for (int32_t fi = 0; fi < d->n_factors; ++fi) {
    const int t = d->factor_types[fi];
    if (t != SYMHIP_FACTOR_SNAVELY && t != SYMHIP_FACTOR_SNAVELY_HUBER) continue;
    const int32_t* a = d->factor_args + d->factor_arg_ptr[fi];
    is_cam[a[0]] = 1;
    cam_of_pose[a[0]] = a[1];
    is_pt[a[2]] = 1;
  }
  std::vector<int32_t> cams, pts;
  for (int32_t k = 0; k < d->n_keys; ++k) {
    if (is_cam[k]) {
      num[k] = (int64_t)cams.size() / 2;
      cams.push_back(k);
      cams.push_back(cam_of_pose[k]);
    } else if (is_pt[k]) {
      num[k] = (int64_t)pts.size();
      pts.push_back(k);
    }
  }
  struct O {
    int64_t cam, pt;
    int32_t px;
  };
  std::vector<O> obs;
  for (int32_t fi = 0; fi < d->n_factors; ++fi) {
    const int t = d->factor_types[fi];
    if (t != SYMHIP_FACTOR_SNAVELY && t != SYMHIP_FACTOR_SNAVELY_HUBER) continue;
    const int32_t* a = d->factor_args + d->factor_arg_ptr[fi];
    obs.push_back({num[a[0]], num[a[2]], a[3]});
  }
  if (obs.empty()) return -1;
  std::ofstream f(path);
  if (!f) return -1;
  f << std::setprecision(17);
  f << cams.size() / 2 << " " << pts.size() << " " << obs.size() << "\n";
  for (const O& o : obs) {
    const double* x = d->values + off[o.px];
    f << o.cam << " " << o.pt << " " << x[0] << " " << x[1] << "\n";
  }
  for (size_t c = 0; c < cams.size(); c += 2) {
    const double* p = d->values + off[cams[c]];
    Q q{p[0], p[1], p[2], p[3]};
    if (q.w < 0) q = {-q.x, -q.y, -q.z, -q.w};
    const double vn = std::sqrt(q.x * q.x + q.y * q.y + q.z * q.z);
    const double sc = vn > 0 ? 2.0 * std::atan2(vn, q.w) / vn : 2.0;
    const double* k = d->values + off[cams[c + 1]];
    f << sc * q.x << " " << sc * q.y << " " << sc * q.z << " " << p[4] << " " << p[5] << " "
      << p[6] << " " << k[0] << " " << k[1] << " " << k[2] << "\n";
  }
  for (int32_t k : pts) {
    const double* v = d->values + off[k];
    f << v[0] << " " << v[1] << " " << v[2] << "\n";
  }
  return f.good() ? 0 : -1;
}

void symhip_gen_free(symhip_generated* g) { delete g; }

int symhip_gen_desc(const symhip_generated* g, symhip_problem_desc* d) {
  d->n_keys = (int32_t)g->letters.size();
  d->key_letters = g->letters.data();
  d->key_subs = g->subs.data();
  d->key_supers = g->supers.data();
  d->key_types = g->types.data();
  d->values = g->values.data();
  d->n_factors = (int32_t)g->ftypes.size();
  d->factor_types = g->ftypes.data();
  d->factor_arg_ptr = g->fptr.data();
  d->factor_args = g->fargs.data();
  d->n_opt_keys = (int32_t)g->opt_keys.size();
  d->opt_keys = g->opt_keys.empty() ? nullptr : g->opt_keys.data();
  return 0;
}

double symhip_gen_info(const symhip_generated* g, const char* name) {
  auto it = g->info.find(name);
  return it == g->info.end() ? std::nan("") : it->second;
}

int64_t symhip_gen_storage_size(const symhip_generated* g) { return (int64_t)g->values.size(); }

}  // extern "C"
