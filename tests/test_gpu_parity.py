"""GPU parity: the HIP path (through the C ABI) against the CPU oracle and the
reference's golden vectors.

Tolerances (stated per test):
  * FP64 plans: ||u_gpu - u_cpu||_2 / ||u_cpu||_2 <= 1e-12 (plain GPU FP64
    must match the CPU path to <= 1e-12 relative);
  * FP32 plans: relative L2 error vs the FP64 and FP32 oracles <= 2e-6 (float
    rounding; the reference's own float path is ~4e-7 from the FP64 answer),
    and <= 2e-5 vs the reference's float golden vectors;
  * FP32+refine plans: relative residual ||f - A u|| / ||f|| <= 1e-10 computed
    in FP64 on the host, and error vs the FP64 oracle <= 1e-9.
Every GPU step prints its case first (flushed) so a failing box names it.
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


def cube_setup(nx, ny, nz, seed=1, aniso=False):
    if aniso:
        dx, dy, dz = 0.11, 0.07, 0.13
    else:
        dx = dy = dz = 2 * np.pi / max(nx, ny, nz)
    g = (dx, dy, dz, (nx + 1) * dx, (ny + 1) * dy, (nz + 1) * dz)
    rhs = O.field(seed, (nz, ny, nx))
    return g, rhs


@pytest.mark.parametrize("shape", [(1, 1, 1), (3, 3, 3), (7, 7, 7), (15, 15, 15), (31, 31, 31), (63, 63, 63),
                                   (127, 127, 127), (31, 15, 7), (7, 31, 63), (63, 127, 31), (255, 15, 31)])
def test_cube_f64_vs_oracle(fdm, shape):
    nx, ny, nz = shape
    say("cube f64", shape)
    g, rhs = cube_setup(nx, ny, nz, seed=3, aniso=(nx != ny))
    ref = O.lapl_cube(rhs, *g)
    s = fdm.LaplCube(*g, nx, ny, nz, precision="f64")
    ans = np.zeros_like(rhs)
    s.solve(ans, rhs)
    assert rel2(ans, ref) <= 1e-12


@pytest.mark.parametrize("shape", [(7, 7, 7), (31, 31, 31), (127, 127, 127), (63, 31, 15), (255, 7, 7)])
def test_cube_f32_vs_oracle(fdm, shape):
    nx, ny, nz = shape
    say("cube f32", shape)
    g, rhs = cube_setup(nx, ny, nz, seed=4)
    rhs32 = rhs.astype(np.float32)
    ref = O.lapl_cube(rhs32, *g)
    ref64 = O.lapl_cube(rhs32.astype(np.float64), *g)
    s = fdm.LaplCube(*g, nx, ny, nz, precision="f32")
    ans = np.zeros_like(rhs32)
    s.solve(ans, rhs32)
    # FP32 tolerance: the reference's float path is ~4e-7 away from the FP64
    # answer at 127^3; the GPU (float data, double tridiagonal elimination) ~6e-7
    assert rel2(ans, ref64) <= 2e-6
    assert rel2(ans, ref) <= 2e-6


@pytest.mark.parametrize("n", [15, 31])
@pytest.mark.parametrize("tag", ["d", "f"])
def test_cube_vs_reference_golden(fdm, n, tag):
    say("cube golden", n, tag)
    rhs = GOLD[f"cube_{tag}_{n}_rhs"]
    g = GOLD[f"cube_{tag}_{n}_geom"]
    s = fdm.LaplCube(*g, n, n, n, precision="f64" if tag == "d" else "f32")
    ans = np.zeros_like(rhs)
    s.solve(ans, rhs)
    assert rel2(ans, GOLD[f"cube_{tag}_{n}_ans"]) <= (1e-12 if tag == "d" else 2e-5)


def test_cube_aniso_quirk_golden(fdm):
    say("cube aniso golden")
    rhs = GOLD["cube_aniso_d_rhs"]
    g = GOLD["cube_aniso_d_geom"]
    nz, ny, nx = rhs.shape
    s = fdm.LaplCube(*g, nx, ny, nz, precision="f64")
    ans = np.zeros_like(rhs)
    s.solve(ans, rhs)
    assert rel2(ans, GOLD["cube_aniso_d_ans"]) <= 1e-12


@pytest.mark.parametrize("n", [31, 127])
def test_cube_manufactured_f64(fdm, n):
    say("cube manufactured f64", n)
    u, f, h, l = O.manufactured_cube(n, seed=1)
    s = fdm.LaplCube(h, h, h, l, l, l, n, n, n, precision="f64")
    ans = np.zeros_like(f)
    s.solve(ans, f)
    assert rel2(ans, u) <= 1e-12
    assert O.rel_residual(f, ans, h, h, h) <= 1e-13


def test_cube_inplace(fdm):
    say("cube in place f64 31")
    g, rhs = cube_setup(31, 31, 31, seed=5)
    ref = O.lapl_cube(rhs, *g)
    s = fdm.LaplCube(*g, 31, 31, 31)
    buf = rhs.copy()
    s.solve(buf, buf)
    assert rel2(buf, ref) <= 1e-12


@pytest.mark.parametrize("n", [31, 127])
def test_cube_refine(fdm, n):
    say("cube refine", n)
    u, f, h, l = O.manufactured_cube(n, seed=2)
    ref = O.lapl_cube(f, h, h, h, l, l, l)
    s = fdm.LaplCube(h, h, h, l, l, l, n, n, n, precision="refine")
    ans = np.zeros_like(f)
    s.solve(ans, f)
    st = s.stats()
    res = O.rel_residual(f, ans, h, h, h)
    print(f"   iterations={st.iterations} device_residual={st.rel_residual:.3e} host_residual={res:.3e} "
          f"err_vs_fp64={rel2(ans, ref):.3e}", flush=True)
    assert res <= 1e-10
    assert st.rel_residual <= 1e-10
    assert abs(st.rel_residual - res) <= 1e-12 + 0.05 * res
    assert rel2(ans, ref) <= 1e-9


# Shapes for the FP64 residual's z march (loads two planes ahead, the loop
# unrolled by three): z lengths 1, 3, 7, 15, 31 (tails of 1 and 2 planes),
# x tiles cut by the field edge (255 = 128 + 127, 100-wide rows do not occur:
# Dirichlet lengths are 2^k - 1), single problems whose columns are split into
# z chunks, and the axis-tile path (non-square planes).
@pytest.mark.parametrize("shape", [(1, 1, 1), (3, 3, 3), (7, 7, 1), (15, 7, 3), (31, 127, 7), (255, 15, 31),
                                   (63, 63, 15)])
def test_cube_refine_shapes(fdm, shape):
    nx, ny, nz = shape
    say("cube refine shape", shape)
    g, rhs = cube_setup(nx, ny, nz, seed=9)
    ref = O.lapl_cube(rhs, *g)
    s = fdm.LaplCube(*g, nx, ny, nz, precision="refine")
    ans = np.zeros_like(rhs)
    s.solve(ans, rhs)
    st = s.stats()
    res = O.rel_residual(rhs, ans, g[0], g[1], g[2])
    print(f"   iterations={st.iterations} device_residual={st.rel_residual:.3e} host_residual={res:.3e} "
          f"err_vs_fp64={rel2(ans, ref):.3e}", flush=True)
    assert res <= 1e-10
    assert st.rel_residual <= 1e-10
    assert abs(st.rel_residual - res) <= 1e-12 + 0.05 * res
    assert rel2(ans, ref) <= 1e-9


def test_cube_tile_batch_matches_single(fdm):
    from fdm_amd import hip
    say("cube batch of 5 on device, axis-tile path, f64 and refine, 31x15x7")
    nx, ny, nz = 31, 15, 7
    g, _ = cube_setup(nx, ny, nz)
    fs = np.stack([cube_setup(nx, ny, nz, seed=20 + k)[1] for k in range(5)])
    for prec in ("f64", "refine"):
        s = fdm.LaplCube(*g, nx, ny, nz, precision=prec)
        rhs = hip.DeviceArray.from_host(fs)
        out = hip.DeviceArray(fs.shape, np.float64)
        s.solve_batch(out, rhs)
        hip.synchronize()
        batch = out.download()
        for i in (0, 3, 4):
            one_r = hip.DeviceArray.from_host(fs[i])
            one = hip.DeviceArray(fs[i].shape, np.float64)
            s.solve(one, one_r)
            hip.synchronize()
            assert np.array_equal(one.download(), batch[i])
        ref = O.lapl_cube(fs[4], *g)
        assert rel2(batch[4], ref) <= (1e-12 if prec == "f64" else 1e-9)


def test_cube_device_batch_matches_single(fdm):
    from fdm_amd import hip
    say("cube batch of 3 on device, f64 and refine, 63^3")
    n = 63
    h, l = O.cube_geometry(n)
    fs = np.stack([O.manufactured_cube(n, seed=s)[1] for s in (11, 12, 13)])
    for prec in ("f64", "refine"):
        s = fdm.LaplCube(h, h, h, l, l, l, n, n, n, precision=prec)
        rhs = hip.DeviceArray.from_host(fs)
        out = hip.DeviceArray(fs.shape, np.float64)
        s.solve_batch(out, rhs)
        hip.synchronize()
        batch = out.download()
        for i in range(3):
            one_r = hip.DeviceArray.from_host(fs[i])
            one = hip.DeviceArray(fs[i].shape, np.float64)
            s.solve(one, one_r)
            hip.synchronize()
            # a problem's result does not depend on its batch neighbours: bitwise
            assert np.array_equal(one.download(), batch[i])
        ref = O.lapl_cube(fs[1], h, h, h, l, l, l)
        assert rel2(batch[1], ref) <= (1e-12 if prec == "f64" else 1e-9)


def test_timing_kernel_mask(fdm):
    """fdmgpu_set_timing_kernels: events around the named classes' launches
    only; the timed launches match the all-classes count for that class, and
    the results are the same with timing on or off (bitwise)."""
    from fdm_amd import hip
    say("timing mask, refine 31^3 batch of 2")
    n = 31
    h, l = O.cube_geometry(n)
    fs = np.stack([O.manufactured_cube(n, seed=s)[1] for s in (21, 22)])
    s = fdm.LaplCube(h, h, h, l, l, l, n, n, n, precision="refine")
    s.set_refine(1e-10, 2)
    rhs = hip.DeviceArray.from_host(fs)
    out = hip.DeviceArray(fs.shape, np.float64)
    s.solve_batch(out, rhs)
    hip.synchronize()
    plain = out.download()
    s.set_timing(True)
    s.solve_batch(out, rhs)
    full = s.timing()
    assert np.array_equal(out.download(), plain)
    assert "residual" in full and len(full) >= 3
    s.set_timing(True, kernels=["residual"])
    s.solve_batch(out, rhs)
    only = s.timing()
    assert list(only) == ["residual"] and only["residual"][1] == full["residual"][1]
    assert np.array_equal(out.download(), plain)
    s.set_timing(False)
    s.solve_batch(out, rhs)
    assert s.timing() == {}


def test_torch_interop_subprocess():
    # torch brings its own HIP runtime; with torch initialised first, libfdmgpu.so binds
    # to it and accepts torch tensors.  Run in a fresh process.
    import subprocess
    say("torch interop (subprocess)")
    code = r"""
import torch, numpy as np
assert torch.cuda.is_available()
torch.zeros(1, device='cuda')
import fdm_amd, oracle as O
n = 31
u, f, h, l = O.manufactured_cube(n, 1)
s = fdm_amd.LaplCube(h, h, h, l, l, l, n, n, n, precision='f64')
rhs = torch.from_numpy(f).cuda(); out = torch.zeros_like(rhs)
s.solve(out, rhs); torch.cuda.synchronize()
err = np.linalg.norm(out.cpu().numpy() - u) / np.linalg.norm(u)
print('torch-interop err', err); assert err < 1e-12
"""
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, "-c", code], cwd=root, capture_output=True, text=True, timeout=180)
    print(r.stdout, r.stderr[-2000:])
    assert r.returncode == 0


@pytest.mark.parametrize("shape", [(31, 31), (100, 63), (7, 127), (511, 255)])
def test_rect_f64_vs_oracle(fdm, shape):
    nx, ny = shape
    say("rect f64", shape)
    dx, dy = 1.0 / nx, 1.0 / ny
    rhs = O.field(8, (ny, nx))
    ref = O.lapl_rect(rhs, dx, dy, 1 + dx, 1 + dy)
    s = fdm.LaplRect(dx, dy, 1 + dx, 1 + dy, nx, ny)
    ans = np.zeros_like(rhs)
    s.solve(ans, rhs)
    assert rel2(ans, ref) <= 1e-12


def test_rect_scales_vs_oracle(fdm):
    say("rect scales f64 (cylindrical-like)")
    nx, ny = 40, 31
    dr, dy = 0.05, 0.1
    r = 1.0 + dr * np.arange(0, nx + 1)
    lm = np.r_[0, 1.0 / r[1:] ** 2]
    Ls = np.r_[0, (r[1:] - 0.5 * dr) / r[1:]]
    Us = np.r_[0, (r[1:] + 0.5 * dr) / r[1:]]
    rhs = O.field(9, (ny, nx))
    ref = O.lapl_rect(rhs, dr, dy, (nx + 1) * dr, (ny + 1) * dy, lm_y_scale=lm, L_scale=Ls, U_scale=Us)
    for prec, tol in (("f64", 1e-12), ("refine", 1e-9)):
        s = fdm.LaplRect(dr, dy, (nx + 1) * dr, (ny + 1) * dy, nx, ny, precision=prec)
        s.set_scales(lm, Ls, Us)
        ans = np.zeros_like(rhs)
        s.solve(ans, rhs)
        assert rel2(ans, ref) <= tol


def test_rect_biquadratic_exact(fdm):
    # ut/ut_lapl_rect.cpp:300-374: exact for a biquadratic, < 1e-14 (double) / 1e-5 (float)
    say("rect biquadratic")
    n = 31
    h = 1.0 / n
    K, J = np.meshgrid(np.arange(1, n + 1), np.arange(1, n + 1), indexing="ij")

    def a0(k, j):
        x, y = h * j - h / 2, h * k - h / 2
        return x * (x - 1) * y * (y - 1)

    x, y = h * J - h / 2, h * K - h / 2
    rhs = 2 * y * (y - 1) + 2 * x * (x - 1)
    rhs[0] -= a0(K[0] - 1, J[0]) / h / h
    rhs[:, 0] -= a0(K[:, 0], J[:, 0] - 1) / h / h
    rhs[:, -1] -= a0(K[:, -1], J[:, -1] + 1) / h / h
    rhs[-1] -= a0(K[-1] + 1, J[-1]) / h / h
    exact = a0(K, J)
    for prec, dt, tol in (("f64", np.float64, 1e-14), ("f32", np.float32, 1e-5)):
        s = fdm.LaplRect(h, h, 1 + h, 1 + h, n, n, precision=prec)
        ans = np.zeros((n, n), dt)
        s.solve(ans, rhs.astype(dt))
        assert np.max(np.abs(ans - exact)) / np.max(np.abs(exact)) < tol


def test_errors(fdm):
    say("error paths")
    with pytest.raises(fdm.FdmGpuError) as e:
        fdm.LaplCube(0.1, 0.1, 0.1, 1, 1, 1, 6, 7, 7)       # nx+1 = 7 is not a power of two
    assert e.value.code == fdm.ERR_INVALID_SIZE
    with pytest.raises(fdm.FdmGpuError) as e:
        fdm.LaplRect(0.1, 0.1, 1, 1, 10, 10)                  # ny+1 = 11
    assert e.value.code == fdm.ERR_INVALID_SIZE
    s = fdm.LaplRect(0.1, 0.1, 0.3, 3.2, 2, 31)
    with pytest.raises(fdm.FdmGpuError) as e:                 # every row [[d, u], [l, d]] with d*d = u*l
        s.set_scales(np.zeros(3), np.full(3, 2.0), np.full(3, 2.0))
    assert e.value.code == fdm.ERR_SINGULAR


@pytest.mark.parametrize("prec", ["f64", "f32", "refine"])
def test_rect_rows_needing_pivoting(fdm, prec):
    """Scales that break diagonal dominance, so that gtsv interchanges rows in
    29 of the 31 row systems (the reference's gtsv pivots, src/lapl_rect.cpp:90):
    the GPU's pivoting elimination against the oracle's LaplRect (dgtsv
    restated).  The row systems' condition numbers reach 372, which scales the
    bars: FP64 within 1e-12, FP32 within 1e-4 of the FP64 oracle (372 x float
    rounding), FP32+refine to a residual of 1e-10 and within 1e-9."""
    nx, ny = 100, 31
    say("rect pivoting", prec, nx, ny)
    h = 0.1
    g = (h, h, (nx + 1) * h, (ny + 1) * h)
    j = np.arange(nx + 1)
    lm = 0.01 + 0.08 * np.sin(0.37 * j)         # |d| = 2/h^2 + lam * s below |L| + |U| = 3.8/h^2
    Ls = 1.91 + 0.085 * np.cos(0.21 * j)
    Us = 1.92 + 0.085 * np.sin(0.13 * j)
    rhs = O.field(77, (ny, nx))
    ref = O.lapl_rect(rhs, *g, lm_y_scale=lm, L_scale=Ls, U_scale=Us)
    s = fdm.LaplRect(*g, nx, ny, precision=prec)
    s.set_scales(lm, Ls, Us)
    if prec == "f32":
        ans = np.zeros((ny, nx), np.float32)
        s.solve(ans, rhs.astype(np.float32))
        assert rel2(ans, ref) <= 1e-4
    else:
        ans = np.zeros_like(rhs)
        s.solve(ans, rhs)
        assert rel2(ans, ref) <= (1e-12 if prec == "f64" else 1e-9)
        if prec == "refine":
            assert 0 <= s.stats().rel_residual <= 1e-10


@pytest.mark.parametrize("periodic", [False, True])
def test_f64_streaming_planes_vs_plane_kernel(fdm, periodic):
    """The streaming FP64 plane kernel (default for planes of N = 128) against the
    one-plane-per-block kernel and the oracle: a batch of 3 (blocks walk two
    planes each), contiguous and through a pointer table; both within 1e-12 of
    the oracle."""
    from fdm_amd import hip
    n = 128 if periodic else 127
    say("f64 streaming planes", n, "periodic" if periodic else "dirichlet")
    h = 2 * np.pi / n
    l = n * h if periodic else (n + 1) * h
    F = np.stack([O.field(90 + k, (n, n, n)) for k in range(3)])
    if periodic:
        F -= F.mean(axis=(1, 2, 3), keepdims=True)
    s = fdm.LaplCube(h, h, h, l, l, l, n, n, n, precision="f64", periodic=periodic)
    assert s.variant[0] == 2   # the default for FP64 planes of N = 128
    d_f = hip.DeviceArray.from_host(F)
    out = {}
    for v in ((2, 0, 0, 0, 0), next(x for x in s.variants() if x[0] == 1)):
        if v[0] == 2:
            s.variant = s.variants()[0]
        else:
            s.variant = v
        d_u = hip.DeviceArray(F.shape, np.float64)
        s.solve_batch(d_u, d_f)
        hip.synchronize()
        out[v[0]] = d_u.download()
    # the same problems through a pointer table, reversed
    fs = [hip.DeviceArray.from_host(F[i]) for i in range(3)]
    us = [hip.DeviceArray(F[i].shape, np.float64) for i in range(3)]
    s.variant = s.variants()[0]
    tab_u = hip.DeviceArray.from_host(np.array([us[i].ptr.value for i in (2, 1, 0)], np.uint64))
    tab_f = hip.DeviceArray.from_host(np.array([fs[i].ptr.value for i in (2, 1, 0)], np.uint64))
    s.solve_table(tab_u, tab_f, 3)
    hip.synchronize()
    for i in range(3):
        assert np.array_equal(us[i].download(), out[2][i]), i
    for i in (0, 2):
        ref = O.lapl_cube(F[i], h, h, h, l, l, l, periodic=periodic)
        assert rel2(out[2][i], ref) <= 1e-12
        assert rel2(out[1][i], ref) <= 1e-12


@pytest.mark.parametrize("periodic", [False, True])
def test_f32_streaming_planes_vs_plane_kernel(fdm, periodic):
    """The streaming FP32 plane kernel (variant plane 2 of FP32 plans with planes
    of N = 128) against the one-plane-per-block kernel on a batch of 5 with 255
    or 256 planes each: with two blocks per CU every block walks three planes
    (256 CUs), of different problems.  Both within 2e-6 of the FP64 oracle and
    1e-6 of each other (the y pass differs in operation order: float
    rounding); the same problems through a reversed pointer table bitwise
    those of the batch."""
    from fdm_amd import hip
    n, nz = (128, 256) if periodic else (127, 255)
    B = 5
    say("f32 streaming planes", n, nz, B, "periodic" if periodic else "dirichlet")
    h = 2 * np.pi / n
    l = n * h if periodic else (n + 1) * h
    lz = nz * h if periodic else (nz + 1) * h
    F = np.stack([O.field(70 + k, (nz, n, n)) for k in range(B)])
    if periodic:
        F -= F.mean(axis=(1, 2, 3), keepdims=True)
    F32 = F.astype(np.float32)
    s = fdm.LaplCube(h, h, h, l, l, lz, n, n, nz, precision="f32", periodic=periodic)
    plane, stream = (1, 0, 0, 0, 0), (2, 0, 0, 0, 0)   # the default z solve
    assert plane in s.variants() and stream in s.variants()
    d_f = hip.DeviceArray.from_host(F32)
    out = {}
    for v in (stream, plane):
        s.variant = v
        d_u = hip.DeviceArray(F32.shape, np.float32)
        s.solve_batch(d_u, d_f)
        hip.synchronize()
        out[v[0]] = d_u.download()
    for i in range(B):
        assert rel2(out[2][i], out[1][i]) <= 1e-6, i
    fs = [hip.DeviceArray.from_host(F32[i]) for i in range(B)]
    us = [hip.DeviceArray(F32[i].shape, np.float32) for i in range(B)]
    s.variant = stream
    rev = list(range(B))[::-1]
    tab_u = hip.DeviceArray.from_host(np.array([us[i].ptr.value for i in rev], np.uint64))
    tab_f = hip.DeviceArray.from_host(np.array([fs[i].ptr.value for i in rev], np.uint64))
    s.solve_table(tab_u, tab_f, B)
    hip.synchronize()
    for i in range(B):
        assert np.array_equal(us[i].download(), out[2][i]), i
    ref = O.lapl_cube(F32[1].astype(np.float64), h, h, h, l, l, lz, periodic=periodic)
    assert rel2(out[2][1], ref) <= 2e-6
    assert rel2(out[1][1], ref) <= 2e-6


@pytest.mark.parametrize("periodic", [False, True])
def test_refine_streaming_planes_with_done_flags(fdm, periodic):
    """FP32+refine through the streaming FP32 plane kernel and through the plane
    kernel: a batch of 5 with 255 or 256 planes each (every block walks three
    planes, of different problems) where one problem is certified after 2
    inner solves and the others take a third, whose transforms skip the
    certified problem's planes wherever they fall in a block's sequence.  Per
    variant, every answer and iteration count in the batch is bitwise that of
    the problem solved alone, and within 1e-9 of the FP64 oracle."""
    from fdm_amd import hip
    n, nz = (128, 256) if periodic else (127, 255)
    B = 5
    say("refine streaming planes, done flags", n, nz, B, "periodic" if periodic else "dirichlet")
    h = 2 * np.pi / n
    l = n * h if periodic else (n + 1) * h
    lz = nz * h if periodic else (nz + 1) * h
    F = np.stack([O.field(50 + k, (nz, n, n)) for k in range(B)])
    if periodic:
        F -= F.mean(axis=(1, 2, 3), keepdims=True)
    s = fdm.LaplCube(h, h, h, l, l, lz, n, n, nz, precision="refine", periodic=periodic)
    plane, stream = (1, 0, 0, 0, 0), (2, 0, 0, 0, 0)   # separate residual pass, the default z solve
    assert plane in s.variants() and stream in s.variants()
    d_f = hip.DeviceArray.from_host(F)
    ref = O.lapl_cube(F[2], h, h, h, l, l, lz, periodic=periodic)
    for v in (plane, stream):
        s.variant = v
        say("  variant", v)
        # the residual of each problem after 2 inner solves
        s.set_refine(1e-30, 2)
        r2 = []
        for i in range(B):
            with pytest.raises(fdm.FdmGpuError):
                s.solve(np.zeros_like(F[i]), F[i])
            r2.append(s.stats().rel_residual)
        order = np.argsort(r2)
        lo, mid = r2[order[0]], r2[order[1]]
        assert mid > lo * 1.01, r2
        s.set_refine(float(np.sqrt(lo * mid)), 4)   # problem order[0] stops after 2 solves, the others after 3
        say("    certified after 2 solves:", int(order[0]), "r2", r2)
        alone = []
        for i in range(B):
            ans = np.zeros_like(F[i])
            s.solve(ans, F[i])
            assert s.stats().iterations == (2 if i == order[0] else 3), (i, r2)
            alone.append(ans)
        d_u = hip.DeviceArray(F.shape, np.float64)
        s.solve_batch(d_u, d_f)
        hip.synchronize()
        ans = d_u.download()
        for i in range(B):
            assert np.array_equal(ans[i], alone[i]), (v, i)
        assert rel2(ans[2], ref) <= 1e-9


@pytest.mark.parametrize("shape", [(127, 63, 31), (63, 127, 15), (31, 127, 63), (127, 31, 63), (63, 31, 127),
                                   (31, 63, 31), (15, 127, 7)])
def test_cube_rect_planes(fdm, shape):
    """Cubes whose planes are not square run through the rectangular plane
    kernel by default (three passes per solve); FP64 within 1e-12 of the
    oracle, FP32 within 2e-6 of the FP64 oracle, FP32+refine to a residual of
    1e-10 and within 1e-9; the axis-tile variant agrees to the same bars."""
    nx, ny, nz = shape
    say("cube rect planes", shape)
    g, rhs = cube_setup(nx, ny, nz, seed=8, aniso=True)
    ref = O.lapl_cube(rhs, *g)
    for prec in ("f64", "f32", "refine"):
        s = fdm.LaplCube(*g, nx, ny, nz, precision=prec)
        assert s.variant[0] == 1, prec
        tile = next(v for v in s.variants() if v[0] == 0)
        for v in (s.variant, tile):
            s.variant = v
            say("  ", prec, v)
            if prec == "f32":
                ans = np.zeros(rhs.shape, np.float32)
                s.solve(ans, rhs.astype(np.float32))
                assert rel2(ans, O.lapl_cube(rhs.astype(np.float32).astype(np.float64), *g)) <= 2e-6, v
            else:
                ans = np.zeros_like(rhs)
                s.solve(ans, rhs)
                assert rel2(ans, ref) <= (1e-12 if prec == "f64" else 1e-9), (prec, v)
                if prec == "refine":
                    assert 0 <= s.stats().rel_residual <= 1e-10
