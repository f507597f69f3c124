"""CPU-only checks of the drop-in boundary: the C-ABI library loads and exports
every entry point include/fdmgpu.h declares; no compute call is made."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(ROOT, "include", "fdmgpu.h")).read()
    return sorted(set(re.findall(r"\b(fdmgpu_[a-z0-9_]+)\s*\(", src)))


def test_header_matches_python_exports():
    import fdm_amd
    assert declared() == sorted(fdm_amd.EXPORTS)


def test_library_exports_every_symbol():
    import fdm_amd
    L = fdm_amd.lib()
    for name in declared():
        assert hasattr(L, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", fdm_amd.LIBPATH], capture_output=True, text=True).stdout
    for name in declared():
        assert re.search(rf"\bT {name}\b", out), name


def test_strerror_and_version():
    import fdm_amd
    L = fdm_amd.lib()
    assert L.fdmgpu_version() >= 100
    assert b"power of two" in L.fdmgpu_strerror(fdm_amd.ERR_INVALID_SIZE)


def test_argument_validation_without_gpu():
    import ctypes as C

    import fdm_amd
    L = fdm_amd.lib()
    h = C.c_void_p()
    # bad precision / mask / null plan are rejected before any device call
    assert L.fdmgpu_plan_cube(C.byref(h), 9, 0, 1, 1, 1, 1, 1, 1, 7, 7, 7) == fdm_amd.ERR_INVALID_ARG
    assert L.fdmgpu_plan_cube(C.byref(h), 1, 3, 1, 1, 1, 1, 1, 1, 7, 7, 7) == fdm_amd.ERR_INVALID_ARG
    assert L.fdmgpu_plan_cube(None, 1, 0, 1, 1, 1, 1, 1, 1, 7, 7, 7) == fdm_amd.ERR_INVALID_ARG
    assert L.fdmgpu_destroy(None) == fdm_amd.ERR_INVALID_ARG


def test_cpp_shim_compiles():
    # the header-only drop-in (include/fdm/*.h) must compile against the C ABI
    src = os.path.join(ROOT, "tests", "cpp", "ut_lapl_cube.cpp")
    if not os.path.exists(src):
        pytest.skip("no C++ shim test")
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-I", os.path.join(ROOT, "include"), src],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
