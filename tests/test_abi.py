"""The C-ABI library: it loads, exports every entry point include/rt/rt.h declares, and
reports errors as status codes (no GPU needed for these calls)."""
import ctypes as C
import os
import re
import subprocess

import pytest

from rustracer_amd import rt

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "rt", "rt.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(rt_[a-z_0-9]+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_lists_the_entry_points():
    assert set(declared_functions()) == set(rt.EXPORTS)


def test_library_exports_every_declared_symbol():
    if not os.path.exists(rt.LIB_PATH):
        pytest.fail("librt_gfx950.so is not built (python -m rustracer_amd.build)")
    out = subprocess.check_output(["nm", "-D", "--defined-only", rt.LIB_PATH], text=True)
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    missing = [n for n in declared_functions() if n not in exported]
    assert not missing, missing


def test_library_loads_and_reports_errors_without_a_device():
    lib = rt.load_library()
    assert lib.rt_version().decode().startswith("rt-gfx950")
    h = C.c_void_p()
    assert lib.rt_create(None, 0, None) == -1  # RT_ERR_INVALID_ARGUMENT
    assert b"NULL" in lib.rt_last_error(None)
    assert lib.rt_sync(None) == -1
    assert lib.rt_build_blas_batch(None, None, 0, None) == -1
    assert lib.rt_trace_rays(None, 0, None, None, None, 0, 0, 0, None) == -1
    lib.rt_destroy(None)  # no-op
    del h


def test_kernels_are_built_for_gfx950():
    data = open(rt.LIB_PATH, "rb").read()
    assert b"gfx950" in data
