"""The C++ drop-in (include/fdm/*.h) used like the reference's unit tests
(ut/ut_lapl_cube.cpp, ut/ut_lapl_rect.cpp) on the GPU."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "ut_lapl_cube")


@pytest.mark.gpu
def test_cpp_dropin_reference_scenarios():
    print("[case] C++ drop-in ut_lapl_cube", flush=True)
    if not os.path.exists(BIN):
        subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "tests", "cpp")])
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    assert "OK (0 failures)" in r.stdout
