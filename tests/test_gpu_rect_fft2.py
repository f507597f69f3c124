"""GPU LaplRectFFT2 (fdmgpu_plan_rect_fft2) against the CPU oracle.

Tolerances: FP64 plans ||u_gpu - u_cpu||_2 / ||u_cpu||_2 <= 1e-12; FP32 plans
<= 5e-6 against the FP64 oracle (float transforms along two axes); FP32+refine
(Dirichlet axes): relative residual <= 1e-10 and error vs the FP64 oracle
<= 1e-9.  Every GPU step prints its case first.
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


@pytest.fixture(scope="module")
def hip():
    from fdm_amd import hip
    return hip


def rel2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def geom(nx, ny, yper, xper, dx=0.3, dy=0.23):
    xpoints, ypoints = (nx if xper else nx + 1), (ny if yper else ny + 1)
    return dx, dy, xpoints * dx, ypoints * dy


CASES = [((15, 31), False, False), ((31, 31), False, False), ((127, 255), False, False), ((255, 511), False, False),
         ((16, 63), True, False), ((128, 127), True, False), ((15, 64), False, True), ((32, 64), True, True),
         ((256, 128), True, True), ((4, 8), True, True), ((1, 1), False, False)]


@pytest.mark.parametrize("shape,yper,xper", CASES)
def test_fft2_f64_f32_vs_oracle(fdm, shape, yper, xper):
    ny, nx = shape
    say("fft2", shape, yper, xper)
    dx, dy, lx, ly = geom(nx, ny, yper, xper)
    rhs = O.field(31, (ny, nx))
    ref = O.lapl_rect_fft2(rhs, dx, dy, lx, ly, yperiodic=yper, xperiodic=xper)
    for prec, dt, tol in (("f64", np.float64, 1e-12), ("f32", np.float32, 5e-6)):
        s = fdm.LaplRectFFT2(dx, dy, lx, ly, nx, ny, precision=prec, yperiodic=yper, xperiodic=xper)
        ans = np.zeros((ny, nx), dt)
        s.solve(ans, rhs.astype(dt))
        assert rel2(ans, ref) <= tol, prec


def test_fft2_scales_vs_oracle(fdm):
    ny, nx = 31, 63
    say("fft2 scales")
    dx, dy, lx, ly = geom(nx, ny, False, False)
    scale = np.r_[0.0, 1.0 / (1.0 + 0.05 * np.arange(1, nx + 1)) ** 2]
    rhs = O.field(32, (ny, nx))
    ref = O.lapl_rect_fft2(rhs, dx, dy, lx, ly, lm_y_scale=scale)
    s = fdm.LaplRectFFT2(dx, dy, lx, ly, nx, ny)
    s.set_scales(lm_y_scale=scale)
    ans = np.zeros_like(rhs)
    s.solve(ans, rhs)
    assert rel2(ans, ref) <= 1e-12
    r = fdm.LaplRectFFT2(dx, dy, lx, ly, nx, ny, precision="refine")
    with pytest.raises(fdm.FdmGpuError) as e:      # no real-space stencil for a spectral scale
        r.set_scales(lm_y_scale=scale)
    assert e.value.code == fdm.ERR_UNSUPPORTED


@pytest.mark.parametrize("shape", [(31, 63), (127, 127), (255, 127)])
def test_fft2_refine(fdm, shape):
    ny, nx = shape
    say("fft2 refine", shape)
    h = 2 * np.pi / nx
    dx, dy = h, 2 * np.pi / ny
    lx, ly = (nx + 1) * dx, (ny + 1) * dy
    rhs = O.field(33, (ny, nx))
    ref = O.lapl_rect_fft2(rhs, dx, dy, lx, ly)
    s = fdm.LaplRectFFT2(dx, dy, lx, ly, nx, ny, precision="refine")
    ans = np.zeros_like(rhs)
    s.solve(ans, rhs)
    st = s.stats()
    hx = dy if nx == ny else dx           # lm_x aliases lm_y when the point counts match
    assert 0 <= st.rel_residual <= 1e-10
    assert O.rel_residual(rhs, ans, hx, dy) <= 1e-10
    assert rel2(ans, ref) <= 1e-9


def test_fft2_batch_and_table_bitwise(fdm, hip):
    ny, nx = 64, 127
    say("fft2 batch")
    dx, dy, lx, ly = geom(nx, ny, True, False)
    B = 5
    F = np.stack([O.field(40 + i, (ny, nx)) for i in range(B)])
    s = fdm.LaplRectFFT2(dx, dy, lx, ly, nx, ny, yperiodic=True)
    d_f = hip.DeviceArray.from_host(F)
    d_u = hip.DeviceArray(F.shape, np.float64)
    s.solve_batch(d_u, d_f)
    hip.synchronize()
    got = d_u.download()
    for i in range(B):
        one = np.zeros((ny, nx))
        s.solve(one, F[i])
        assert np.array_equal(one, got[i])


def test_fft2_errors(fdm):
    say("fft2 errors")
    with pytest.raises(fdm.FdmGpuError) as e:
        fdm.LaplRectFFT2(0.1, 0.1, 1.0, 1.0, 30, 31)          # nx+1 = 31
    assert e.value.code == fdm.ERR_INVALID_SIZE
    with pytest.raises(fdm.FdmGpuError) as e:
        fdm.LaplRectFFT2(0.1, 0.1, 1.6, 1.6, 15, 16, precision="refine", yperiodic=True)
    assert e.value.code == fdm.ERR_UNSUPPORTED
