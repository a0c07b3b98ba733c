"""In-tree build of the native libraries (no JIT cache, so the .so files travel with the tree).

    python -m symforce_amd.build            # everything
    python -m symforce_amd.build --only hip # just libsymhip.so

libsymhip.so        HIP kernels + host runtime + C ABI, for gfx950 only
libsymhip_problems.so  seeded problem generators (host only)
oracle/_build/liboracle.so  CPU FP64 checker (built by oracle/Makefile; test infrastructure)
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
REPO = PKG.parent
CSRC = PKG / "csrc"
OUT = PKG / "_build"
OBJ = OUT / "obj"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "-x", "hip", "--offload-arch=gfx950", "-O3", "-std=c++20", "-fPIC",
    # keep the compiler from emitting packed FP32 math (v_pk_*_f32)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-result",
    "-munsafe-fp-atomics",
]
HIP_SOURCES = sorted([p for p in CSRC.glob("*.hip")] +
                     [p for p in CSRC.glob("*.cc") if p.name != "problems.cc"])


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(map(str, cmd)) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {cmd[-1]}")
    return r


def build_problems():
    OUT.mkdir(parents=True, exist_ok=True)
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-o",
          str(OUT / "libsymhip_problems.so"), str(CSRC / "problems.cc")])


def build_hip(jobs: int = 8):
    OBJ.mkdir(parents=True, exist_ok=True)
    headers = list(CSRC.glob("*.h")) + list((REPO / "include").glob("*.h"))
    newest_header = max(h.stat().st_mtime for h in headers)
    srcs = [s for s in HIP_SOURCES if s.name != "problems.cc"]

    def compile_one(src: Path):
        obj = OBJ / (src.name + ".o")
        if obj.exists() and obj.stat().st_mtime > max(src.stat().st_mtime, newest_header):
            return obj
        _run([HIPCC, *HIP_FLAGS, "-c", str(src), "-o", str(obj)])
        return obj

    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(compile_one, srcs))
    tmp = OUT / "libsymhip.so.tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)])
    os.replace(tmp, OUT / "libsymhip.so")


def build_cpp_tests():
    """C++ programs written against the sym:: host surface (include/sym/), run by the GPU tests."""
    out = REPO / "tests" / "cpp" / "_build"
    out.mkdir(parents=True, exist_ok=True)
    for src in (REPO / "tests" / "cpp").glob("*.cc"):
        _run(["g++", "-O2", "-std=c++17", "-Wall", "-I", str(REPO / "include"), str(src), "-o",
              str(out / src.stem), "-L", str(OUT), "-lsymhip", "-lsymhip_problems",
              "-Wl,-rpath,$ORIGIN/../../../symforce_amd/_build"])


def build_oracle():
    _run(["make", "-s", "-C", str(REPO / "oracle")])


def build_all(only: str | None = None):
    if only in (None, "problems"):
        build_problems()
    if only in (None, "oracle"):
        build_oracle()
    if only in (None, "hip"):
        build_hip()
    if only in (None, "cpp"):
        build_cpp_tests()


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--only", choices=["hip", "problems", "oracle", "cpp"], default=None)
    a = ap.parse_args()
    build_all(a.only)
    print("built", sorted(p.name for p in OUT.glob("*.so")))
