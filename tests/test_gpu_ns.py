"""GPU NSCube step (fdmgpu_ns_*) against the CPU oracle's NSCube.

Bars: with the FP64 pressure solve the velocity and pressure fields match the
oracle to 1e-11 relative (max-norm over each field) after tens of steps; with
the FP32+refine pressure solve (residual 1e-10 per step) to 1e-8 after 100
steps (the cavity's bar: FP64 CPU agreement to 1e-8 relative at step 100).
Steps after the second are
graph replays, and the comparisons are made between step batches (host
copies in between).
"""
import sys

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def relmax(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.fixture(scope="module")
def fdm():
    import fdm_amd
    return fdm_amd


@pytest.mark.parametrize("shape,steps", [((31, 31, 31), (2, 3, 15)), ((31, 15, 15), (5, 10))])
def test_ns_f64_vs_oracle(fdm, shape, steps):
    nx, ny, nz = shape
    say("ns f64", shape, steps)
    kw = dict(u0=1.0, Re=250.0, dt=0.01)
    g = fdm.NSCube(nx, ny, nz, solver="f64", **kw)
    o = O.NsCube(nx, ny, nz, **kw)
    for s in steps:
        g.step(s)
        o.step(s)
        for f in ("u", "v", "w", "p"):
            assert relmax(g.get(f), o.get(f)) <= 1e-11, (f, g.time_index)


def test_ns_refine_vs_oracle_100_steps(fdm):
    n = 63
    say("ns refine", n)
    kw = dict(u0=1.0, Re=250.0, dt=0.01)
    g = fdm.NSCube(n, n, n, solver="refine", **kw)
    o = O.NsCube(n, n, n, **kw)
    g.step(100)
    o.step(100)
    st = g.solver_stats()
    assert 0 <= st.rel_residual <= 1e-10 and 1 <= st.iterations <= 8
    for f in ("u", "v", "w", "p"):
        assert relmax(g.get(f), o.get(f)) <= 1e-8, f


def test_ns_refine_field_upload(fdm):
    """A field uploaded between steps (graph replay already running) restarts
    the pressure refinement from zero; the run keeps matching the oracle fed
    the same field."""
    n = 31
    say("ns refine with a field upload", n)
    kw = dict(u0=1.0, Re=250.0, dt=0.01)
    g = fdm.NSCube(n, n, n, solver="refine", **kw)
    o = O.NsCube(n, n, n, **kw)
    g.step(6)
    o.step(6)
    u = o.get("u") * 0.5
    g.set("u", u)
    o.set("u", u)
    g.step(1)
    o.step(1)
    st = g.solver_stats()
    assert 0 <= st.rel_residual <= 1e-10
    g.step(10)
    o.step(10)
    st = g.solver_stats()
    assert 0 <= st.rel_residual <= 1e-10
    for f in ("u", "v", "w", "p"):
        assert relmax(g.get(f), o.get(f)) <= 1e-8, f


def test_ns_refine_not_converged_is_reported(fdm):
    """A projection that misses its tolerance (1e-30 in two inner solves, steps
    replayed from the graph included) is reported by the next field copy as
    ERR_NOT_CONVERGED; the same run with the default settings reports none."""
    n = 15
    say("ns refine not converged flag", n)
    kw = dict(u0=1.0, Re=250.0, dt=0.01)
    ok = fdm.NSCube(n, n, n, solver="refine", **kw)
    ok.step(4)
    assert np.isfinite(ok.get("u")).all()
    bad = fdm.NSCube(n, n, n, solver="refine", **kw)
    bad.set_solver_refine(1e-30, 3)
    bad.step(4)
    with pytest.raises(fdm.FdmGpuError) as e:
        bad.get("u")
    assert e.value.code == fdm.ERR_NOT_CONVERGED
    ok.close()
    bad.close()


def test_ns_errors(fdm):
    say("ns errors")
    with pytest.raises(fdm.FdmGpuError) as e:      # lid loop bound nz+1 beyond u's x range
        fdm.NSCube(15, 15, 31)
    assert e.value.code == fdm.ERR_UNSUPPORTED
    with pytest.raises(fdm.FdmGpuError) as e:      # Poisson transform length 13
        fdm.NSCube(12, 12, 12, solver="f64")
    assert e.value.code == fdm.ERR_INVALID_SIZE
