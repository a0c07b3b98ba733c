"""The C ABI libraries load on a machine without a GPU and export every function their headers
declare (include/symhip.h for libsymhip.so, include/symhip_problems.h for the generators).
No compute call is made here."""
import ctypes as C
import re
from pathlib import Path

import pytest

from symforce_amd import _capi as capi

INCLUDE = Path(__file__).resolve().parent.parent / "include"
DECL = re.compile(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\**\s*(symhip_[a-z0-9_]+)\s*\(",
                  re.M)


def declared(header):
    text = (INCLUDE / header).read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)  # drop comments
    names = set(DECL.findall(text))
    assert names, f"no declarations parsed from {header}"
    return names


@pytest.mark.parametrize("header,loader", [
    ("symhip.h", capi.hip_lib),
    ("symhip_problems.h", capi.problems_lib),
])
def test_library_exports_every_declared_function(header, loader):
    lib = loader()
    missing = sorted(n for n in declared(header) if not hasattr(lib, n))
    assert not missing, f"{header}: not exported: {missing}"


def test_no_device_is_reported_without_a_gpu_runtime_call_failing():
    # symhip_device_count is the one call safe without a GPU: it reports 0 or the count
    assert capi.hip_lib().symhip_device_count() >= 0


def test_declarations_cover_the_entry_points_the_bindings_use():
    names = declared("symhip.h")
    for n in ("symhip_optimizer_create", "symhip_optimizer_optimize", "symhip_batch_solve",
              "symhip_batch_solve_noalloc", "symhip_cholesky_factorize", "symhip_linearize"):
        assert n in names
    # every function the Python binding sets a signature for is declared in the header
    bound = {name for name, _, _ in capi._EXTRA_SIGNATURES}
    assert bound <= names, sorted(bound - names)
    assert isinstance(C.sizeof(capi.BatchTables), int)
