"""GPU parity of the periodic solves (LaplCube<..., periodic^3>, LaplRect<...,
periodic> in y) through the C ABI, against the CPU oracle, the reference's
golden vectors and the mathematics of the zero mode.

Reference semantics (src/lapl_cube.cpp:9-142, src/fft.cpp:108-212): real
half-complex DFT (pFFT_1) forward along z, y, x with factor h*sqrt(2/l),
divide by -(lambda_x + lambda_y + lambda_z) with lambda_k = 4/h^2 sin^2(k pi/n),
zero mode := 0, inverse series (pFFT) with factor sqrt(2/l).

Tolerances: FP64 relative L2 error vs the oracle <= 1e-12; FP32 <= 2e-6 vs the
FP64 oracle (2e-5 vs the reference's float golden vectors).
"""
import os
import sys

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "ref_vectors.npz"))


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
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def periodic_geom(nx, ny, nz, aniso=False):
    # [-pi, pi) with n points per axis: h = 2 pi / n, l = n h (the n-body setup)
    if aniso:
        dx, dy, dz = 0.11, 0.07, 0.13
    else:
        dx, dy, dz = 2 * np.pi / nx, 2 * np.pi / ny, 2 * np.pi / nz
    return (dx, dy, dz, nx * dx, ny * dy, nz * dz)


def lapl_periodic(u, hx, hy, hz):
    out = (np.roll(u, 1, 2) - 2 * u + np.roll(u, -1, 2)) / hx ** 2
    out += (np.roll(u, 1, 1) - 2 * u + np.roll(u, -1, 1)) / hy ** 2
    out += (np.roll(u, 1, 0) - 2 * u + np.roll(u, -1, 0)) / hz ** 2
    return out


@pytest.mark.parametrize("shape", [(2, 2, 2), (4, 4, 4), (8, 8, 8), (16, 16, 16), (32, 32, 32), (64, 64, 64),
                                   (128, 128, 128), (16, 32, 8), (8, 8, 256), (256, 16, 8), (32, 512, 4)])
def test_cube_periodic_f64_vs_oracle(fdm, shape):
    nx, ny, nz = shape
    say("periodic cube f64", shape)
    g = periodic_geom(nx, ny, nz, aniso=(nx != ny or ny != nz))
    rhs = O.field(11, (nz, ny, nx))
    ref = O.lapl_cube(rhs, *g, periodic=True)
    s = fdm.LaplCube(*g, nx, ny, nz, precision="f64", periodic=True)
    ans = np.zeros_like(rhs)
    s.solve(ans, rhs)
    assert rel2(ans, ref) <= 1e-12


@pytest.mark.parametrize("shape", [(16, 16, 16), (64, 64, 64), (128, 128, 128), (32, 64, 128)])
def test_cube_periodic_f32_vs_oracle(fdm, shape):
    nx, ny, nz = shape
    say("periodic cube f32", shape)
    g = periodic_geom(nx, ny, nz)
    rhs = O.field(12, (nz, ny, nx)).astype(np.float32)
    ref64 = O.lapl_cube(rhs.astype(np.float64), *g, periodic=True)
    ref = O.lapl_cube(rhs, *g, periodic=True)
    s = fdm.LaplCube(*g, nx, ny, nz, precision="f32", periodic=True)
    ans = np.zeros_like(rhs)
    s.solve(ans, rhs)
    assert rel2(ans, ref64) <= 2e-6
    assert rel2(ans, ref) <= 2e-6


@pytest.mark.parametrize("tag", ["d", "f"])
def test_cube_periodic_vs_reference_golden(fdm, tag):
    say("periodic cube golden", tag)
    rhs = GOLD[f"cubep_{tag}_16_rhs"]
    g = tuple(float(v) for v in GOLD[f"cubep_{tag}_16_geom"])
    want = GOLD[f"cubep_{tag}_16_ans"]
    s = fdm.LaplCube(*g, 16, 16, 16, precision="f64" if tag == "d" else "f32", periodic=True)
    ans = np.zeros_like(rhs)
    s.solve(ans, rhs)
    assert rel2(ans, want) <= (1e-12 if tag == "d" else 2e-5)


@pytest.mark.parametrize("n", [32, 128])
def test_cube_periodic_zero_mode(fdm, n):
    """The reference zeroes the (0,0,0) coefficient: the answer is the zero-mean
    solution of A u = f - mean(f) (A = periodic 7-point Laplacian)."""
    say("periodic zero mode", n)
    g = periodic_geom(n, n, n)
    f = O.field(13, (n, n, n)) + 0.75      # non-zero mean on purpose
    s = fdm.LaplCube(*g, n, n, n, precision="f64", periodic=True)
    u = np.zeros_like(f)
    s.solve(u, f)
    h = g[0]
    assert abs(u.mean()) <= 1e-12 * np.abs(u).max()
    r = lapl_periodic(u, h, h, h) - (f - f.mean())
    assert np.linalg.norm(r) / np.linalg.norm(f - f.mean()) <= 1e-11


def test_cube_periodic_device_batch(fdm):
    from fdm_amd import hip
    n, B = 64, 3
    say("periodic batch", n, B)
    g = periodic_geom(n, n, n)
    F = np.stack([O.field(20 + i, (n, n, n)) for i in range(B)])
    s = fdm.LaplCube(*g, n, n, n, precision="f64", periodic=True)
    d_f = hip.DeviceArray.from_host(F)
    d_u = hip.DeviceArray(F.shape, np.float64)
    st = hip.Stream()
    s.solve_batch(d_u, d_f, st)
    st.synchronize()
    U = d_u.download()
    for i in range(B):
        assert rel2(U[i], O.lapl_cube(F[i], *g, periodic=True)) <= 1e-12


@pytest.mark.parametrize("shape", [(5, 8), (31, 32), (100, 64), (7, 128), (511, 256)])
def test_rect_yperiodic_vs_oracle(fdm, shape):
    nx, ny = shape
    say("rect y-periodic", shape)
    dx, dy = 2 * np.pi / (nx + 1), 2 * np.pi / ny
    lx, ly = (nx + 1) * dx, ny * dy
    rhs = O.field(14, (ny, nx))
    ref = O.lapl_rect(rhs, dx, dy, lx, ly, yperiodic=True)
    s = fdm.LaplRect(dx, dy, lx, ly, nx, ny, precision="f64", yperiodic=True)
    ans = np.zeros_like(rhs)
    s.solve(ans, rhs)
    assert rel2(ans, ref) <= 1e-12
    s32 = fdm.LaplRect(dx, dy, lx, ly, nx, ny, precision="f32", yperiodic=True)
    r32 = rhs.astype(np.float32)
    a32 = np.zeros_like(r32)
    s32.solve(a32, r32)
    assert rel2(a32, O.lapl_rect(rhs, dx, dy, lx, ly, yperiodic=True)) <= 2e-6


def test_periodic_errors(fdm):
    say("periodic error paths")
    with pytest.raises(fdm.FdmGpuError) as e:
        fdm.LaplCube(0.1, 0.1, 0.1, 1, 1, 1, 6, 8, 8, periodic=True)     # n = 6 is not a power of two
    assert e.value.code == fdm.ERR_INVALID_SIZE
    with pytest.raises(fdm.FdmGpuError) as e:      # refinement needs l = n h (the solve's scale 1)
        fdm.LaplCube(0.1, 0.1, 0.1, 1.7, 1.6, 1.6, 16, 16, 16, precision="refine", periodic=True)
    assert e.value.code == fdm.ERR_UNSUPPORTED
    with pytest.raises(fdm.FdmGpuError) as e:      # periodic LaplRect: no projected refinement
        fdm.LaplRect(0.1, 0.1, 1.7, 1.6, 16, 16, precision="refine", yperiodic=True)
    assert e.value.code == fdm.ERR_UNSUPPORTED


@pytest.mark.parametrize("shape", [(8, 8, 8), (16, 16, 16), (64, 64, 64), (128, 128, 128), (16, 32, 8),
                                   (8, 8, 256), (256, 16, 8)])
def test_cube_periodic_refine(fdm, shape):
    """FP32+refine on periodic cubes: the zero-mean solution of A u = f - mean f
    (the reference's periodic solve, src/lapl_cube.cpp:82-84) with f of non-zero
    mean.  Bars: device-reported projected residual and the host's
    ||(f - mean f) - A u|| / ||f - mean f|| (FP64, numpy) <= 1e-10; within 1e-9
    (relative L2) of the FP64 oracle; mean(u) ~ 0."""
    nx, ny, nz = shape
    say("periodic cube refine", shape)
    g = periodic_geom(nx, ny, nz, aniso=(nx != ny or ny != nz))
    f = O.field(15, (nz, ny, nx)) + 0.75
    ref = O.lapl_cube(f, *g, periodic=True)
    s = fdm.LaplCube(*g, nx, ny, nz, precision="refine", periodic=True)
    u = np.zeros_like(f)
    s.solve(u, f)
    st = s.stats()
    assert 1 <= st.iterations <= 4 and 0 <= st.rel_residual <= 1e-10, (st.iterations, st.rel_residual)
    fp = f - f.mean()
    r = fp - lapl_periodic(u, s_hx(g, shape), g[1], s_hz(g, shape))
    assert np.linalg.norm(r) / np.linalg.norm(fp) <= 1e-10
    assert rel2(u, ref) <= 1e-9
    assert abs(u.mean()) <= 1e-9 * np.abs(u).max()


def s_hx(g, shape):
    # the operator the solve inverts takes lambda_x from y when nx == ny
    # (reference init_lm eigenvalue aliasing, src/lapl_cube.cpp:162)
    return g[1] if shape[0] == shape[1] else g[0]


def s_hz(g, shape):
    return g[1] if shape[2] == shape[1] else g[2]


def test_cube_periodic_refine_batch_in_place(fdm):
    """Three problems back to back, solved in place (ans == rhs): each equals
    its solo refine solve bitwise and sits within 1e-9 of the FP64 oracle."""
    from fdm_amd import hip
    n, B = 32, 3
    say("periodic refine batch in place", n, B)
    g = periodic_geom(n, n, n)
    F = np.stack([O.field(30 + i, (n, n, n)) + 0.5 * i for i in range(B)])
    s = fdm.LaplCube(*g, n, n, n, precision="refine", periodic=True)
    d = hip.DeviceArray.from_host(F)
    st = hip.Stream()
    s.solve_batch(d, d, st)
    st.synchronize()
    U = d.download()
    for i in range(B):
        solo = np.zeros_like(F[i])
        s.solve(solo, F[i])
        assert np.array_equal(U[i], solo), i
        assert rel2(U[i], O.lapl_cube(F[i], *g, periodic=True)) <= 1e-9
