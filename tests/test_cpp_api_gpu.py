"""The C++ host surface (include/sym/symforce_amd.h) used the way the reference's own test uses
sym::Optimizer (test/symforce_optimizer_test.cc:142-176), run on the device optimizer."""
import json
import subprocess
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu
BIN = Path(__file__).resolve().parent / "cpp" / "_build" / "pose_smoothing"


def test_cpp_pose_smoothing_known_answer():
    assert BIN.exists(), "build with python -m symforce_amd.build"
    out = subprocess.run([str(BIN)], capture_output=True, text=True, timeout=120, check=True)
    r = json.loads(out.stdout.strip().splitlines()[-1])
    assert r["iteration"] == 12
    assert r["lambda"] == pytest.approx(0.0039, rel=1e-1)
    assert r["error"] == pytest.approx(7.801, rel=1e-3)
    assert r["status"] == 1
