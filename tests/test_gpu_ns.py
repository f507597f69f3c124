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


@pytest.mark.parametrize("solver", ["f64", "refine"])
def test_ns_timed_steps_and_maxabs(fdm, solver):
    """fdmgpu_ns_step_timed: the same fields, bitwise, as fdmgpu_ns_step (eager
    first steps and graph replays alike), four positive phase times;
    fdmgpu_ns_maxabs: max |value| per whole array, NaN skipped
    (tensor::maxabs, reference src/tensor.h:269-273)."""
    n = 31
    say("ns timed steps + maxabs", solver)
    kw = dict(u0=1.0, Re=250.0, dt=0.01, solver=solver)
    a = fdm.NSCube(n, n, 15, **kw)
    b = fdm.NSCube(n, n, 15, **kw)
    for _ in range(5):
        ms = a.step_timed()
        assert len(ms) == 4 and all(t > 0 for t in ms), ms
    b.step(5)
    assert a.time_index == b.time_index == 5
    m = a.maxabs()
    for name in fdm.NS_FIELDS:
        fa = a.get(name)
        assert np.array_equal(fa, b.get(name)), name
        assert m[name] == float(np.abs(fa).max()), name
    # NaN and -inf planted in a ghost face: maxabs skips the NaN, takes |-inf|
    u = a.get("u")
    u[0, 0, 0] = np.nan
    a.set("u", u)
    assert a.maxabs()["u"] == float(np.nanmax(np.abs(u)))
    u[0, 0, 1] = -np.inf
    a.set("u", u)
    assert a.maxabs()["u"] == np.inf
    a.close()
    b.close()
