"""GPU LaplCyl3FFT2 (fdmgpu_plan_cyl3_fft2) against the CPU oracle.

Tolerances: FP64 ||u_gpu - u_cpu||_2 / ||u_cpu||_2 <= 1e-12 against the
oracle's cyclic-reduction and LU (gttrs-like) instantiations; FP32 <= 2e-5
against the FP64 oracle.  Every GPU step prints its case first.
"""
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


def rel2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def geom(nr, nz, zper, dr=0.05, dz=0.1, r0=0.025):
    zpoints = nz if zper else nz + 1
    return dr, dz, r0, (nr + 1) * dr, zpoints * dz


# (nphi, nz, nr, zper): nphi and nz+1 (nz periodic) powers of two
CASES = [(8, 7, 5, False), (32, 31, 40, False), (64, 63, 100, False), (128, 31, 17, False), (16, 16, 33, True),
         (64, 128, 64, True), (2, 1, 1, False), (256, 15, 8, False)]


@pytest.mark.parametrize("nphi,nz,nr,zper", CASES)
def test_cyl_vs_oracle(fdm, nphi, nz, nr, zper):
    say("cyl", nphi, nz, nr, zper)
    dr, dz, r0, lr, lz = geom(nr, nz, zper)
    rhs = O.field(51, (nphi, nz, nr))
    ref_cr = O.lapl_cyl3_fft2(rhs, dr, dz, r0, lz, zperiodic=zper, cyclic_reduction=True)
    ref_lu = O.lapl_cyl3_fft2(rhs, dr, dz, r0, lz, zperiodic=zper, cyclic_reduction=False)
    for prec, dt in (("f64", np.float64), ("f32", np.float32)):
        s = fdm.LaplCyl3FFT2(dr, dz, r0, lr, lz, nr, nz, nphi, precision=prec, zperiodic=zper)
        ans = np.zeros((nphi, nz, nr), dt)
        s.solve(ans, rhs.astype(dt))
        if prec == "f64":
            assert rel2(ans, ref_cr) <= 1e-12
            assert rel2(ans, ref_lu) <= 1e-12
        else:
            assert rel2(ans, ref_cr) <= 2e-5


def test_cyl_batch_bitwise(fdm):
    from fdm_amd import hip
    nphi, nz, nr = 32, 15, 24
    say("cyl batch")
    dr, dz, r0, lr, lz = geom(nr, nz, False)
    B = 4
    F = np.stack([O.field(60 + i, (nphi, nz, nr)) for i in range(B)])
    s = fdm.LaplCyl3FFT2(dr, dz, r0, lr, lz, nr, nz, nphi)
    d_f = hip.DeviceArray.from_host(F)
    d_u = hip.DeviceArray(F.shape, np.float64)
    s.solve_batch(d_u, d_f)
    hip.synchronize()
    got = d_u.download()
    for i in range(B):
        one = np.zeros((nphi, nz, nr))
        s.solve(one, F[i])
        assert np.array_equal(one, got[i])


def test_cyl_errors(fdm):
    say("cyl errors")
    with pytest.raises(fdm.FdmGpuError) as e:      # nphi not a power of two
        fdm.LaplCyl3FFT2(0.1, 0.1, 0.05, 1.0, 1.6, 9, 15, 12)
    assert e.value.code == fdm.ERR_INVALID_SIZE
    with pytest.raises(fdm.FdmGpuError) as e:      # no FP32+refine for the cylindrical operator
        fdm.LaplCyl3FFT2(0.1, 0.1, 0.05, 1.0, 1.6, 9, 15, 16, precision="refine")
    assert e.value.code == fdm.ERR_UNSUPPORTED
