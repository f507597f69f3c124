"""fdmgpu_solve_batch_noalloc (no device allocation, static scratch) through the
C ABI against fdmgpu_solve_batch_device: bitwise the same answers on two calls
in a row (the second with a smaller batch of smaller problems, so it runs on
the first call's leftover scratch), for an in-place refine solve and for a
plain FP64 plan; the refine results also meet the refinement bar (relative
residual <= 1e-10 on the host, error <= 1e-9 vs the FP64 oracle).  A batch
whose scratch exceeds the static array is rejected with ERR_TOO_LARGE and
leaves the answer untouched."""
import sys

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def fdm():
    import fdm_amd
    return fdm_amd


@pytest.fixture(scope="module")
def hip():
    from fdm_amd import hip as h
    return h


def geom(n):
    h = 2 * np.pi / n
    return (h, h, h, (n + 1) * h, (n + 1) * h, (n + 1) * h)


def batch_pair(fdm, hip, n, count, prec, seed, inplace=False):
    """(noalloc answer, solve_batch answer, rhs) for `count` n^3 problems."""
    g = geom(n)
    F = np.stack([O.field(seed + i, (n, n, n)) for i in range(count)])
    s = fdm.LaplCube(*g, n, n, n, precision=prec)
    out = []
    for entry in ("solve_batch_noalloc", "solve_batch"):
        d_f = hip.DeviceArray.from_host(F)
        d_u = d_f if inplace else hip.DeviceArray(F.shape, np.float64)
        if not inplace:
            d_u.upload(np.full(F.shape, 7.0))
        getattr(s, entry)(d_u, d_f)
        hip.synchronize()
        out.append(d_u.download())
        if prec == "refine":
            st = s.stats()
            assert 0 <= st.max_rel_residual <= 1e-10, (entry, st.max_rel_residual)
    return out[0], out[1], F, s


def test_noalloc_matches_batch_entry_two_calls(fdm, hip):
    say("noalloc refine 63^3 x 4")
    a, b, F, _ = batch_pair(fdm, hip, 63, 4, "refine", 50)
    assert np.array_equal(a, b)
    g = geom(63)
    assert O.rel_residual(F[3], a[3], g[0], g[1], g[2]) <= 1e-10
    ref = O.lapl_cube(F[3], *g)
    assert np.linalg.norm(a[3] - ref) / np.linalg.norm(ref) <= 1e-9
    say("noalloc refine 31^3 x 2 (second call, smaller)")
    a, b, F, _ = batch_pair(fdm, hip, 31, 2, "refine", 60)
    assert np.array_equal(a, b)
    g = geom(31)
    assert O.rel_residual(F[1], a[1], g[0], g[1], g[2]) <= 1e-10


def test_noalloc_inplace_and_plain(fdm, hip):
    say("noalloc refine 31^3 x 3 in place")
    a, b, F, _ = batch_pair(fdm, hip, 31, 3, "refine", 70, inplace=True)
    assert np.array_equal(a, b)
    say("noalloc f64 63^3 x 3")
    a, b, F, _ = batch_pair(fdm, hip, 63, 3, "f64", 80)
    assert np.array_equal(a, b)
    g = geom(63)
    ref = O.lapl_cube(F[2], *g)
    assert np.linalg.norm(a[2] - ref) / np.linalg.norm(ref) <= 1e-12


def test_noalloc_rejects_beyond_scratch(fdm, hip):
    n, count = 127, 5          # 5 x ~16.5 MB of FP32+refine scratch > 64 MiB
    say("noalloc too large", n, count)
    s = fdm.LaplCube(*geom(n), n, n, n, precision="refine")
    d_f = hip.DeviceArray((count, n, n, n), np.float64)
    d_u = hip.DeviceArray((count, n, n, n), np.float64)
    d_u.upload(np.full(d_u.shape, 7.0))
    with pytest.raises(fdm.FdmGpuError) as e:
        s.solve_batch_noalloc(d_u, d_f)
    assert e.value.code == fdm.ERR_TOO_LARGE
    hip.synchronize()
    assert (d_u.download() == 7.0).all()
