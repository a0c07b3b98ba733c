"""Small problem builders for tests (pure Python, deterministic)."""
from __future__ import annotations

import json
from pathlib import Path

import numpy as np

from symforce_amd import _capi as capi

INVALID = np.iinfo(np.int64).min
GOLDEN = Path(__file__).resolve().parent / "golden"


class ProblemBuilder:
    def __init__(self):
        self.letters, self.subs, self.supers, self.types, self.values = [], [], [], [], []
        self.ftypes, self.fptr, self.fargs = [], [0], []
        self.opt_keys = None
        self.info = {"epsilon": 10 * np.finfo(np.float64).eps, "c_dim": 0}

    def key(self, letter, sub=INVALID, sup=INVALID, type_=capi.TYPE_SCALAR, value=0.0):
        v = np.atleast_1d(np.asarray(value, np.float64)).ravel(order="F")
        assert v.size == capi.STORAGE_DIM[type_]
        self.letters.append(ord(letter))
        self.subs.append(sub)
        self.supers.append(sup)
        self.types.append(type_)
        self.values.extend(v.tolist())
        return len(self.letters) - 1

    def factor(self, ftype, args):
        self.ftypes.append(ftype)
        self.fargs.extend(args)
        self.fptr.append(len(self.fargs))

    def build(self) -> capi.Problem:
        return capi.Problem(self.letters, self.subs, self.supers, self.types, self.values,
                            self.ftypes, self.fptr, self.fargs, self.opt_keys, self.info)


def robot3d_problem() -> tuple[capi.Problem, dict]:
    """The reference robot_3d_localization problem (run_dynamic_size.cc:25-105, common.h:31-62)
    from the extracted measurement fixture; poses start at identity."""
    d = json.loads((GOLDEN / "robot_3d_localization.json").read_text())
    b = ProblemBuilder()
    n, m = d["num_poses"], d["num_landmarks"]
    poses = [b.key("w", INVALID, i, capi.TYPE_POSE3, [0, 0, 0, 1, 0, 0, 0]) for i in range(n)]
    lms = [b.key("W", INVALID, j, capi.TYPE_VECTOR3, d["landmark_positions"][j]) for j in range(m)]
    sig = b.key("o", type_=capi.TYPE_VECTOR6, value=d["odometry_diagonal_sigmas"])
    odo = [b.key("O", INVALID, i, capi.TYPE_POSE3, d["odometry_relative_pose_measurements"][i])
           for i in range(n - 1)]
    msig = b.key("m", value=d["matching_sigma"])
    meas = {}
    for i in range(n):
        for j in range(m):
            meas[i, j] = b.key("b", i, j, capi.TYPE_VECTOR3,
                               d["body_t_landmark_measurements"][i][j])
    eps = b.key("e", value=10 * np.finfo(np.float64).eps)
    for i in range(n):
        for j in range(m):
            b.factor(capi.FACTOR_MATCHING, [poses[i], lms[j], meas[i, j], msig])
    for i in range(n - 1):
        b.factor(capi.FACTOR_ODOMETRY, [poses[i], poses[i + 1], odo[i], sig, eps])
    return b.build(), d


def gnc_problem(ys, x0=None) -> tuple[capi.Problem, int]:
    """The reference's GNC problem (test/symforce_gnc_test.cc:26-48): Barron factors on a
    Vector5 'x' against constant samples ('y', i), with the convexity parameter 'u' and the
    factor epsilon 'e'; keys in the order the reference's Values hold them ('u' is appended by
    GncOptimizer::Optimize). Returns the problem and the key index of 'u'."""
    b = ProblemBuilder()
    x = b.key("x", type_=capi.TYPE_VECTOR5, value=np.ones(5) if x0 is None else x0)
    e = b.key("e", value=10 * np.finfo(np.float64).eps)
    yk = [b.key("y", i, INVALID, capi.TYPE_VECTOR5, y) for i, y in enumerate(ys)]
    u = b.key("u", value=0.0)
    for k in yk:
        b.factor(capi.FACTOR_BARRON_VECTOR5, [x, k, u, e])
    b.info["epsilon"] = 1e-12  # the optimizer epsilon the reference test passes
    return b.build(), u


def gnc_reference_samples() -> np.ndarray:
    """The 20 samples of the reference's GNC test (tests/golden/gnc_problem.json)."""
    return np.array(json.loads((GOLDEN / "gnc_problem.json").read_text())["ys"])
