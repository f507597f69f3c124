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


NS_BIN = os.path.join(ROOT, "tests", "cpp", "ut_ns_cube")


@pytest.mark.gpu
def test_cpp_nscube_class_verbose(tmp_path):
    """fdm::NSCube<double,false>(Config) (include/fdm/ns_cube.h) with ns:verbose=1:
    the reference's per-step line (src/ns_cube.cpp:52-56: dt*time_index, then
    maxabs of p, u, v, w, x, RHS, F, G, H, each "%.1e") against the oracle's
    cavity stepped alike, and the destructor's "Step times" line
    (src/ns_cube.cpp:15-24).  The program itself checks the timed eager steps
    bitwise against graph-replayed ones and maxabs() against host maxima."""
    import math

    import numpy as np

    import oracle as O
    n, steps = 15, 6
    print(f"[case] C++ NSCube class {n}^3 verbose, {steps} steps", flush=True)
    if not os.path.exists(NS_BIN):
        subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "tests", "cpp")])
    ini = tmp_path / "ns_cube.ini"
    ini.write_text(f"[ns]\nnx = {n}\nnz = {n}\nRe = 250\ndt = 0.01\nverbose = 1\n")
    r = subprocess.run([NS_BIN, str(ini), f"--ns:steps={steps}", "--solver:gpu_precision=f64"],
                       capture_output=True, text=True, timeout=120, cwd=tmp_path)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    lines = r.stdout.splitlines()
    assert lines[-1] == "OK (0 failures)"
    step_lines = [l for l in lines if l.endswith(" ") and len(l.split()) == 10]
    assert len(step_lines) == steps
    times = [l for l in lines if l.startswith("Step times (ms): init_bound=")]
    assert len(times) == 1
    vals = [float(t.split("=")[1]) for t in times[0].split(": ")[1].split(", ")]
    assert len(vals) == 4 and all(math.isfinite(v) and v > 0 for v in vals)
    o = O.NsCube(n, n, n, u0=1.0, Re=250.0, dt=0.01)
    order = ("p", "u", "v", "w", "x", "RHS", "F", "G", "H")
    for t, line in enumerate(step_lines, 1):
        o.step(1)
        got = line.split()
        assert got[0] == "%.1e" % (0.01 * t)
        for name, s in zip(order, got[1:]):
            ref = float(np.abs(o.get(name)).max())
            if ref == 0.0:
                assert float(s) == 0.0, (t, name, s)
            else:   # the printed value is the 2-digit rounding of one within 1e-12 of the oracle's
                half_ulp = 0.5 * 10.0 ** (math.floor(math.log10(ref)) - 1)
                assert abs(float(s) - ref) <= half_ulp * (1 + 1e-9) + 1e-12 * ref, (t, name, s, ref)
