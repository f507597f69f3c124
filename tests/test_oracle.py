"""Pins the CPU oracle (oracle/fdm_oracle.cpp) before anything trusts it.

Each oracle function is checked against at least one of:
  * the reference's own code (oracle/_ref, built from the reference sources;
    skipped where it is absent),
  * committed golden vectors written by tests/golden/make_golden.py,
  * closed forms / library routines (direct sums, scipy.fft.dst, scipy's
    LAPACK dgtsv, dense numpy solves, manufactured solutions),
chosen so that a dropped term, a wrong sign or index, or a transposed operand
fails.  These tests never touch a GPU.
"""
import os

import numpy as np
import pytest
import scipy.fft
import scipy.linalg.lapack as lapack

import oracle as O

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "ref_vectors.npz"))
needs_ref = pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built (no reference checkout)")


def rel(a, b):
    return float(np.max(np.abs(np.asarray(a, np.float64) - b)) / max(np.max(np.abs(b)), 1e-300))


# ---------------------------------------------------------------------------
# 1-D transforms
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("N", [2, 4, 8, 16, 64, 256, 1024])
def test_dst_vs_direct_sum(N):
    rng = np.random.default_rng(N)
    x = np.zeros(N)
    x[1:] = rng.uniform(-1, 1, N - 1)
    X = O.dst1(x, N, 0.5)
    D = O.sft_direct(x, N, 0.5)
    assert rel(X[1:], D[1:]) < 1e-14


@pytest.mark.parametrize("N", [4, 32, 128])
def test_dst_vs_scipy(N):
    rng = np.random.default_rng(N + 1)
    x = np.zeros(N)
    x[1:] = rng.uniform(-1, 1, N - 1)
    # scipy type-1 DST is 2 * sum x_j sin(pi (j+1)(k+1)/N) over the N-1 points
    ref = scipy.fft.dst(x[1:], type=1) / 2
    assert rel(O.dst1(x, N)[1:], ref) < 1e-14


def test_dst_self_inverse():
    # DST-I o DST-I = N/2 I (the identity behind sFFT(sFFT(s), 2/N) = s, ut/ut_fft.cpp:190-228)
    N = 512
    x = np.zeros(N)
    x[1:] = np.random.default_rng(3).uniform(-1, 1, N - 1)
    y = O.dst1(O.dst1(x, N, 1.0), N, 2.0 / N)
    assert rel(y[1:], x[1:]) < 1e-14


def test_dst_float_matches_double():
    N = 128
    x = np.zeros(N)
    x[1:] = np.random.default_rng(4).uniform(-1, 1, N - 1)
    assert rel(O.dst1(x.astype(np.float32), N)[1:], O.dst1(x, N)[1:]) < 2e-6


@pytest.mark.parametrize("N", [8, 32, 128])
@pytest.mark.parametrize("tag", ["d", "f"])
def test_transforms_vs_golden(N, tag):
    tol = 1e-13 if tag == "d" else 2e-5
    x = GOLD[f"sfft_{tag}_N{N}_in"]
    assert rel(O.dst1(x[:N].copy(), N, 0.37)[1:N], GOLD[f"sfft_{tag}_N{N}_out"][1:N].astype(np.float64)) < tol
    if tag == "d":
        xd = x.astype(np.float64)
        assert rel(O.dct1(xd, N, 0.37), GOLD[f"cfft_d_N{N}_out"]) < tol
        assert rel(O.rdft_fwd(xd[:N], N, 0.37), GOLD[f"pfft1_d_N{N}_out"][:N]) < tol
        assert rel(O.rdft_inv(xd[:N], N, 0.37), GOLD[f"pfft_d_N{N}_out"][:N]) < tol


@pytest.mark.parametrize("N", [4, 8, 64])
def test_rdft_vs_direct_and_roundtrip(N):
    rng = np.random.default_rng(N + 7)
    x = rng.uniform(-1, 1, N)
    X = O.rdft_fwd(x, N, 1.0)
    j = np.arange(N)
    assert abs(X[0] - x.sum()) < 1e-13
    assert abs(X[N // 2] - np.sum(x * (-1.0) ** j)) < 1e-13
    for k in range(1, N // 2):
        assert abs(X[k] - np.sum(x * np.cos(2 * np.pi * j * k / N))) < 1e-13
        assert abs(X[N - k] - np.sum(x * np.sin(2 * np.pi * j * k / N))) < 1e-13
    # pFFT(pFFT_1(x, 1), 2/N) = x (ut/ut_fft.cpp:73-77)
    assert rel(O.rdft_inv(X, N, 2.0 / N), x) < 1e-14


def test_dct_vs_direct():
    N = 16
    x = np.random.default_rng(11).uniform(-1, 1, N + 1)
    X = O.dct1(x, N, 1.0)
    j = np.arange(1, N)
    for k in range(N + 1):
        ref = 0.5 * x[0] + np.sum(x[1:N] * np.cos(np.pi * j * k / N)) + 0.5 * (-1) ** k * x[N]
        assert abs(X[k] - ref) < 1e-13


@needs_ref
@pytest.mark.parametrize("N", [4, 16, 256, 1024])
def test_transforms_vs_ref(N):
    R = O.ref()
    rng = np.random.default_rng(N + 5)
    x = rng.uniform(-1, 1, N + 1)
    S = np.zeros(N + 1)
    R.ref_fft_d(S, x.copy(), N, 1.0, 0)
    # the reference's Samarskii recursion itself carries ~N*eps error (test_dst_vs_direct_sum
    # bounds the oracle against a long-double direct sum)
    assert rel(O.dst1(x[:N].copy(), N)[1:N], S[1:N]) < 2e-13
    S = np.zeros(N + 1)
    R.ref_fft_d(S, x.copy(), N, 1.0, 2)
    assert rel(O.rdft_fwd(x[:N], N), S[:N]) < 1e-13
    S = np.zeros(N + 1)
    R.ref_fft_d(S, x.copy(), N, 1.0, 3)
    assert rel(O.rdft_inv(x[:N], N), S[:N]) < 1e-13


# ---------------------------------------------------------------------------
# LaplCube
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("n", [15, 31])
@pytest.mark.parametrize("tag", ["d", "f"])
def test_cube_vs_golden(n, tag):
    rhs = GOLD[f"cube_{tag}_{n}_rhs"]
    g = GOLD[f"cube_{tag}_{n}_geom"]
    ans = O.lapl_cube(rhs, *g)
    assert rel(ans, GOLD[f"cube_{tag}_{n}_ans"].astype(np.float64)) < (1e-13 if tag == "d" else 2e-5)


def test_cube_aniso_quirk_vs_golden():
    # nx == ny with dx != dy: lm_x aliases lm_y (src/lapl_cube.cpp:162), so the reference
    # inverts the operator with spacing dy on the x axis.
    rhs = GOLD["cube_aniso_d_rhs"]
    g = GOLD["cube_aniso_d_geom"]
    ans = O.lapl_cube(rhs, *g)
    assert rel(ans, GOLD["cube_aniso_d_ans"]) < 1e-13
    dx, dy, dz = g[:3]
    # the aliased operator is exactly what was solved: residual with hx := dy is tiny ...
    assert O.rel_residual(rhs, ans, dy, dy, dz) < 1e-13
    # ... and with the true dx it is not (the quirk is real and reproduced)
    assert O.rel_residual(rhs, ans, dx, dy, dz) > 1e-3


@pytest.mark.parametrize("tag", ["d", "f"])
def test_cube_periodic_vs_golden(tag):
    rhs = GOLD[f"cubep_{tag}_16_rhs"]
    g = GOLD[f"cubep_{tag}_16_geom"]
    ans = O.lapl_cube(rhs, *g, periodic=True)
    assert rel(ans, GOLD[f"cubep_{tag}_16_ans"].astype(np.float64)) < (1e-13 if tag == "d" else 2e-5)


@pytest.mark.parametrize("n", [2, 4])
def test_cube_periodic_dense_pseudoinverse(n):
    """Periodic LaplCube = zero-mean least-squares solution of the singular
    periodic 7-point system (the zeroed zero mode, src/lapl_cube.cpp:82-84)."""
    h = 2 * np.pi / n
    f = O.field(5, (n, n, n)) + 0.3
    u = O.lapl_cube(f, h, h, h, n * h, n * h, n * h, periodic=True)
    N = n ** 3
    A = np.zeros((N, N))
    for c in range(N):
        e = np.zeros(N)
        e[c] = 1
        v = e.reshape(n, n, n)
        A[:, c] = sum((np.roll(v, 1, a) - 2 * v + np.roll(v, -1, a)) / h ** 2 for a in range(3)).ravel()
    ref = np.linalg.lstsq(A, f.ravel(), rcond=None)[0]
    ref -= ref.mean()
    assert rel(u.ravel(), ref) < 1e-13


@pytest.mark.parametrize("n", [7, 31, 63])
def test_cube_manufactured(n):
    u, f, h, l = O.manufactured_cube(n, seed=1)
    ans = O.lapl_cube(f, h, h, h, l, l, l)
    assert rel(ans, u) < 1e-12
    assert O.rel_residual(f, ans, h, h, h) < 1e-14


def test_cube_dense_bruteforce():
    # distinct point counts per axis (no eigenvalue aliasing) and distinct spacings:
    # solve the assembled 7-point matrix densely.
    for (nx, ny, nz), hz_used in (((7, 3, 1), 0.5), ((7, 3, 3), 0.2)):
        dx, dy, dz = 0.3, 0.2, 0.5
        rhs = O.field(9, (nz, ny, nx))
        ans = O.lapl_cube(rhs, dx, dy, dz, (nx + 1) * dx, (ny + 1) * dy, (nz + 1) * dz)
        n = nx * ny * nz
        A = np.zeros((n, n))
        for c in range(n):
            e = np.zeros(n)
            e[c] = 1
            # with ny == nz the reference's lm_z aliases lm_y (src/lapl_cube.cpp:171): hz := dy
            A[:, c] = O.apply_lapl(e.reshape(nz, ny, nx), dx, dy, hz_used).ravel()
        ref = np.linalg.solve(A, rhs.ravel()).reshape(nz, ny, nx)
        assert rel(ans, ref) < 1e-13


def test_cube_analytic_ut():
    # ut/ut_lapl_cube.cpp:39-125: u = sin^2 x + cos^2 y + sin^2 z on a cell-centred
    # grid over [0,1]^3, boundary values moved to the rhs, 31^3, error < 1e-4.
    n = 31
    h = 1.0 / n

    def ans(i, k, j):
        x, y, z = h * j - h / 2, h * k - h / 2, h * i - h / 2
        return np.sin(x) ** 2 + np.cos(y) ** 2 + np.sin(z) ** 2

    def rp(i, k, j):
        x, y, z = h * j - h / 2, h * k - h / 2, h * i - h / 2
        return (2 * np.sin(y) ** 2 - 2 * np.cos(y) ** 2 - 2 * np.sin(x) ** 2 + 2 * np.cos(x) ** 2
                - 2 * np.sin(z) ** 2 + 2 * np.cos(z) ** 2)

    I, K, J = np.meshgrid(np.arange(1, n + 1), np.arange(1, n + 1), np.arange(1, n + 1), indexing="ij")
    rhs = rp(I, K, J)
    rhs[0] -= ans(I[0] - 1, K[0], J[0]) / h / h
    rhs[:, 0] -= ans(I[:, 0], K[:, 0] - 1, J[:, 0]) / h / h
    rhs[:, :, 0] -= ans(I[:, :, 0], K[:, :, 0], J[:, :, 0] - 1) / h / h
    rhs[:, :, -1] -= ans(I[:, :, -1], K[:, :, -1], J[:, :, -1] + 1) / h / h
    rhs[:, -1] -= ans(I[:, -1], K[:, -1] + 1, J[:, -1]) / h / h
    rhs[-1] -= ans(I[-1] + 1, K[-1], J[-1]) / h / h
    for dt in (np.float64, np.float32):
        sol = O.lapl_cube(rhs.astype(dt), h, h, h, 1 + h, 1 + h, 1 + h)
        exact = ans(I, K, J)
        assert np.max(np.abs(sol - exact)) / np.max(np.abs(exact)) < 1e-4


@needs_ref
@pytest.mark.parametrize("n,periodic", [(63, False), (127, False), (64, True)])
def test_cube_vs_ref(n, periodic):
    if periodic:
        h, l = 2 * np.pi / n, 2 * np.pi
    else:
        h, l = O.cube_geometry(n)
    rhs = O.field(2, (n, n, n))
    a = O.lapl_cube(rhs, h, h, h, l, l, l, periodic=periodic)
    r = O.ref_lapl_cube(rhs, h, h, h, l, l, l, periodic=periodic)
    assert rel(a, r) < 1e-13


def test_cube_rejects_non_pow2():
    with pytest.raises(ValueError):
        O.lapl_cube(np.zeros((5, 5, 5)), 1, 1, 1, 6, 6, 6)


# ---------------------------------------------------------------------------
# LaplRect / LaplRectFFT2 / gtsv
# ---------------------------------------------------------------------------
def test_gtsv_vs_scipy():
    rng = np.random.default_rng(5)
    for n in (2, 3, 17, 255):
        dl, d, du, b = rng.uniform(-1, 1, n - 1), rng.uniform(-1, 1, n), rng.uniform(-1, 1, n - 1), rng.uniform(-1, 1, n)
        # pad dl/du to n (the oracle takes length-n arrays like LAPACK callers usually allocate)
        x, info = O.gtsv(np.r_[dl, 0], d, np.r_[du, 0], b)
        _, _, _, xs, infos = lapack.dgtsv(dl, d, du, b)
        assert info == infos == 0
        assert rel(x, xs) < 1e-12   # pivoting path included (random, not dominant)


def test_gtsv_singular_info():
    x, info = O.gtsv(np.zeros(3), np.zeros(3), np.zeros(3), np.ones(3))
    assert info == 1


def _rect_matrix(nx, ny, dx, dy, lm_scale=None, Ls=None, Us=None):
    n = nx * ny
    A = np.zeros((n, n))
    for c in range(n):
        e = np.zeros(n)
        e[c] = 1
        A[:, c] = O.apply_lapl(e.reshape(ny, nx), dx, dy).ravel()
    return A


def test_rect_dense_bruteforce():
    nx, ny = 9, 7
    dx, dy = 0.3, 0.2
    rhs = O.field(3, (ny, nx))
    ans = O.lapl_rect(rhs, dx, dy, (nx + 1) * dx, (ny + 1) * dy)
    ref = np.linalg.solve(_rect_matrix(nx, ny, dx, dy), rhs.ravel()).reshape(ny, nx)
    assert rel(ans, ref) < 1e-13


def test_rect_scales_dense():
    # lm_y_scale / L_scale / U_scale (src/lapl_rect.h:57-59): row k solves
    #   L_j/dx^2 u_{j-1} + (-2/dx^2 - lm_y[k] s_j) u_j + U_j/dx^2 u_{j+1} = f_hat
    nx, ny = 6, 7
    dx, dy = 0.25, 0.2
    rng = np.random.default_rng(8)
    s = np.r_[0, rng.uniform(0.5, 1.5, nx)]
    Ls = np.r_[0, rng.uniform(0.5, 1.0, nx)]
    Us = np.r_[0, rng.uniform(0.5, 1.0, nx)]
    rhs = O.field(4, (ny, nx))
    ans = O.lapl_rect(rhs, dx, dy, (nx + 1) * dx, (ny + 1) * dy, lm_y_scale=s, L_scale=Ls, U_scale=Us)
    # brute force in the DST-y basis: y-eigenvectors are exact, so assemble the full operator
    N = ny + 1
    k = np.arange(1, N)
    lam = 4 / dy ** 2 * np.sin(k * np.pi / (2 * N)) ** 2
    Q = np.sin(np.pi * np.outer(k, k) / N)          # symmetric, Q Q = N/2 I
    n = nx * ny
    A = np.zeros((n, n))
    for kk in range(ny):
        for j in range(nx):
            r = kk * nx + j
            A[r, r] = -2 / dx ** 2 - lam[kk] * s[j + 1]
            if j > 0:
                A[r, r - 1] = Ls[j + 1] / dx ** 2
            if j < nx - 1:
                A[r, r + 1] = Us[j + 1] / dx ** 2
    fhat = (Q @ rhs)
    uhat = np.linalg.solve(A, fhat.ravel()).reshape(ny, nx)
    ref = (2.0 / N) * (Q @ uhat)
    assert rel(ans, ref) < 1e-12


def test_rect_biquadratic_exact():
    # ut/ut_lapl_rect.cpp:300-374: the 5-point scheme is exact for a biquadratic.
    n = 31
    h = 1.0 / n
    x1 = y1 = 0.0
    x2 = y2 = 1.0

    def ans0(k, j):
        x, y = x1 + h * j - h / 2, y1 + h * k - h / 2
        return (x - x1) * (x - x2) * (y - y1) * (y - y2)

    def rp0(k, j):
        x, y = x1 + h * j - h / 2, y1 + h * k - h / 2
        return 2 * (y - y1) * (y - y2) + 2 * (x - x1) * (x - x2)

    K, J = np.meshgrid(np.arange(1, n + 1), np.arange(1, n + 1), indexing="ij")
    rhs = rp0(K, J)
    rhs[0] -= ans0(K[0] - 1, J[0]) / h / h
    rhs[:, 0] -= ans0(K[:, 0], J[:, 0] - 1) / h / h
    rhs[:, -1] -= ans0(K[:, -1], J[:, -1] + 1) / h / h
    rhs[-1] -= ans0(K[-1] + 1, J[-1]) / h / h
    exact = ans0(K, J)
    for dt, tol in ((np.float64, 1e-14), (np.float32, 1e-5)):
        sol = O.lapl_rect(rhs.astype(dt), h, h, 1 + h, 1 + h)
        assert np.max(np.abs(sol - exact)) / np.max(np.abs(exact)) < tol


def test_rect_vs_fft2():
    # ut/ut_lapl_rect.cpp:384-456: LaplRect (gtsv) == LaplRectFFT2 at 511^2 to 1e-15
    # (absolute, assert_float_equal); analytic rhs of the ut file.
    n = 511
    h = 1.0 / n
    K, J = np.meshgrid(np.arange(1, n + 1), np.arange(1, n + 1), indexing="ij")
    x, y = h * J - h / 2, h * K - h / 2
    rhs = 2 * np.sin(y) ** 2 - 2 * np.cos(y) ** 2 - 2 * np.sin(x) ** 2 + 2 * np.cos(x) ** 2
    a1 = O.lapl_rect(rhs, h, h, 1 + h, 1 + h)
    a2 = O.lapl_rect_fft2(rhs, h, h, 1 + h, 1 + h)
    assert np.max(np.abs(a1 - a2)) < 1e-13


def _fft2_matrices(points, periodic, h, l):
    """Forward / inverse 1-D transforms of LaplRectFFT2 as dense matrices,
    from the series the reference's sFFT / pFFT_1 / pFFT compute
    (src/fft.cpp:108-212, 293-365): DST-I on the N-1 interior points of a
    Dirichlet axis, the half-complex real DFT on the N points of a periodic one;
    factor h*sqrt(2/l) forward, sqrt(2/l) inverse."""
    N = points
    sl = np.sqrt(2.0 / l)
    if not periodic:
        k = np.arange(1, N)
        S = np.sin(np.pi * np.outer(k, k) / N)
        return h * sl * S, sl * S
    j = np.arange(N)
    R = np.zeros((N, N))
    Q = np.zeros((N, N))
    R[0] = 1.0
    R[N // 2] = (-1.0) ** j
    Q[:, 0] = 0.5
    Q[:, N // 2] = 0.5 * (-1.0) ** j
    for k in range(1, N // 2):
        R[k] = np.cos(2 * np.pi * k * j / N)
        R[N - k] = np.sin(2 * np.pi * k * j / N)
        Q[:, k] = np.cos(2 * np.pi * k * j / N)
        Q[:, N - k] = np.sin(2 * np.pi * k * j / N)
    return h * sl * R, sl * Q


@pytest.mark.parametrize("ny,nx,yper,xper", [(7, 15, False, False), (15, 15, False, False), (8, 7, True, False),
                                             (7, 16, False, True), (8, 16, True, True)])
def test_fft2_dense(ny, nx, yper, xper):
    # LaplRectFFT2::solve (src/lapl_rect.cpp:113-206) as matrix products, with
    # init_lm (:25-41): lm_x aliases lm_y (spacing dy) in the all-Dirichlet
    # instantiation when the point counts match; a spectral lm_y_scale; the
    # double-periodic zero-mode coefficient 1 (:169-172).
    dx, dy = 0.3, 0.23
    ypoints, xpoints = (ny if yper else ny + 1), (nx if xper else nx + 1)
    lx, ly = xpoints * dx, ypoints * dy
    rhs = O.field(21, (ny, nx))
    scale = np.r_[1.3, 0.7 + 0.05 * np.arange(nx)]     # nx+1 entries, x1..xn read
    Tyf, Tyi = _fft2_matrices(ypoints, yper, dy, ly)
    Txf, Txi = _fft2_matrices(xpoints, xper, dx, lx)
    ky = np.arange(0, ny) if yper else np.arange(1, ny + 1)
    kx = np.arange(0, nx) if xper else np.arange(1, nx + 1)
    lmy = 4 / dy ** 2 * np.sin(ky * np.pi / ny if yper else ky * np.pi * 0.5 / (ny + 1)) ** 2
    lmx = 4 / dx ** 2 * np.sin(kx * np.pi / nx if xper else kx * np.pi * 0.5 / (nx + 1)) ** 2
    if not yper and not xper and xpoints == ypoints:
        lmx = lmy
    sc = scale[(0 if xper else 1):][:nx]
    W = Tyf @ rhs @ Txf.T
    with np.errstate(divide="ignore", invalid="ignore"):
        W = W / (-(lmy[:, None] * sc[None, :] + lmx[None, :]))
    if yper and xper:
        W[0, 0] = 1.0
    want = Tyi @ W @ Txi.T
    got = O.lapl_rect_fft2(rhs, dx, dy, lx, ly, yperiodic=yper, xperiodic=xper, lm_y_scale=scale)
    assert rel(got, want) < 1e-13


@pytest.mark.parametrize("yper,xper", [(True, False), (False, True), (True, True)])
def test_fft2_periodic_operator(yper, xper):
    # Physical check with default scales: the 5-point Laplacian (periodic wrap
    # on periodic axes) of the answer reproduces the rhs; double-periodic: the
    # rhs with its mean removed, and the answer's mean is the constant
    # slx*sly/4 that the zero-mode coefficient 1 adds.
    ny, nx = (16 if yper else 15), (16 if xper else 15)
    dx, dy = 0.2, 0.3
    ypoints, xpoints = (ny if yper else ny + 1), (nx if xper else nx + 1)
    lx, ly = xpoints * dx, ypoints * dy
    rhs = O.field(22, (ny, nx))
    u = O.lapl_rect_fft2(rhs, dx, dy, lx, ly, yperiodic=yper, xperiodic=xper)

    def lap(u):
        uy = np.roll(u, 1, 0) + np.roll(u, -1, 0) if yper else np.pad(u, ((1, 1), (0, 0)))[:-2] + np.pad(u, ((1, 1), (0, 0)))[2:]
        ux = np.roll(u, 1, 1) + np.roll(u, -1, 1) if xper else np.pad(u, ((0, 0), (1, 1)))[:, :-2] + np.pad(u, ((0, 0), (1, 1)))[:, 2:]
        return (uy - 2 * u) / dy ** 2 + (ux - 2 * u) / dx ** 2
    if yper and xper:
        assert rel(lap(u), rhs - rhs.mean()) < 1e-12
        assert abs(u.mean() - np.sqrt(2 / lx) * np.sqrt(2 / ly) / 4) < 1e-12
    else:
        assert rel(lap(u), rhs) < 1e-12


def _cyl_geometry(nr, nz, nphi, zper):
    dr, dz, r0 = 0.11, 0.2, 0.35
    zpoints = nz if zper else nz + 1
    return dr, dz, r0, zpoints * dz


def _cyl_apply(u, dr, dz, r0, zper):
    """The discrete operator LaplCyl3FFT2 inverts (src/lapl_cyl.cpp:55-73,
    130-170): flux-form r second difference, periodic phi second difference
    over r^2, z second difference (periodic wrap or zero Dirichlet)."""
    nphi, nz, nr = u.shape
    dphi = 2 * np.pi / nphi
    r = r0 + dr * np.arange(1, nr + 1)
    up = np.pad(u, ((0, 0), (0, 0), (1, 1)))
    Ar = ((r + 0.5 * dr) * (up[:, :, 2:] - u) - (r - 0.5 * dr) * (u - up[:, :, :-2])) / (dr ** 2 * r)
    Aphi = (np.roll(u, 1, 0) - 2 * u + np.roll(u, -1, 0)) / (dphi ** 2 * r ** 2)
    if zper:
        Az = (np.roll(u, 1, 1) - 2 * u + np.roll(u, -1, 1)) / dz ** 2
    else:
        uz = np.pad(u, ((0, 0), (1, 1), (0, 0)))
        Az = (uz[:, 2:] - 2 * u + uz[:, :-2]) / dz ** 2
    return Ar + Aphi + Az


@pytest.mark.parametrize("nphi,nz,nr,zper", [(8, 7, 5, False), (16, 15, 9, False), (8, 8, 6, True), (32, 16, 12, True)])
def test_cyl3_operator(nphi, nz, nr, zper):
    # the cylindrical 7-point operator of the answer reproduces the rhs, for
    # the LU (gttrs) and cyclic-reduction instantiations, double and float
    dr, dz, r0, lz = _cyl_geometry(nr, nz, nphi, zper)
    rhs = O.field(41, (nphi, nz, nr))
    for cr in (False, True):
        u = O.lapl_cyl3_fft2(rhs, dr, dz, r0, lz, zperiodic=zper, cyclic_reduction=cr)
        assert rel(_cyl_apply(u, dr, dz, r0, zper), rhs) < 1e-12
        uf = O.lapl_cyl3_fft2(rhs.astype(np.float32), dr, dz, r0, lz, zperiodic=zper, cyclic_reduction=cr)
        assert rel(uf, u) < 2e-5


def cyl_known_answer(nr, nz, nphi, solve):
    """ut/ut_lapl_cyl.cpp:320-406: u = (z-h0)(z-h1)(r-r0)(r-R)(sin phi + cos phi)
    on cell-centred r, z nodes, rhs = its continuous Laplacian
    P[Z (4r - r0 - R)/r + 2 Rr - Rr Z / r^2] with the boundary values moved to
    the rhs; returns max|u_h - u| / max|u|."""
    r0, R, h0, h1 = np.pi / 2, np.pi, 0.0, 10.0
    dr, dz, dphi = (R - r0) / nr, (h1 - h0) / nz, 2 * np.pi / nphi
    i = np.arange(nphi)[:, None, None]
    k = np.arange(0, nz + 2)[None, :, None]
    j = np.arange(0, nr + 2)[None, None, :]
    r, z, phi = r0 + dr * j - dr / 2, h0 + dz * k - dz / 2, dphi * (i + 1) - dphi / 2
    Z, Rr, P = (z - h0) * (z - h1), (r - r0) * (r - R), np.sin(phi) + np.cos(phi)
    u = Z * Rr * P
    f = P * (Z * (4 * r - r0 - R) / r + 2 * Rr - Rr * Z / r ** 2)
    rhs = f[:, 1:nz + 1, 1:nr + 1].copy()
    rr = r[:, :, 1:nr + 1]
    rhs[:, :, 0] -= ((rr - dr / 2) / rr)[:, :, 0] * u[:, 1:nz + 1, 0] / dr ** 2
    rhs[:, :, -1] -= ((rr + dr / 2) / rr)[:, :, -1] * u[:, 1:nz + 1, nr + 1] / dr ** 2
    rhs[:, 0, :] -= u[:, 0, 1:nr + 1] / dz ** 2
    rhs[:, -1, :] -= u[:, nz + 1, 1:nr + 1] / dz ** 2
    exact = u[:, 1:nz + 1, 1:nr + 1]
    got = solve(rhs, dr, dz, r0 - dr / 2, R - r0 + dr, h1 - h0 + dz, nr, nz, nphi)
    return np.abs(got - exact).max() / np.abs(exact).max()


def test_cyl3_known_answer_second_order():
    # the reference asserts err(16,15,16) / err(32,31,32) > 3.7
    def solve(rhs, dr, dz, r0, lr, lz, nr, nz, nphi):
        return O.lapl_cyl3_fft2(rhs, dr, dz, r0, lz)
    e1 = cyl_known_answer(16, 15, 16, solve)
    e2 = cyl_known_answer(32, 31, 32, solve)
    assert e1 / e2 > 3.7
    assert e2 < 1e-2


def test_cyl3_dense():
    # brute force: the dense operator solved directly (nonsingular: r > 0)
    nphi, nz, nr = 8, 3, 4
    for zper, nzz in ((False, 3), (True, 4)):
        dr, dz, r0, lz = _cyl_geometry(nr, nzz, nphi, zper)
        n = nphi * nzz * nr
        A = np.zeros((n, n))
        for c in range(n):
            e = np.zeros(n)
            e[c] = 1
            A[:, c] = _cyl_apply(e.reshape(nphi, nzz, nr), dr, dz, r0, zper).ravel()
        rhs = O.field(42, (nphi, nzz, nr))
        want = np.linalg.solve(A, rhs.ravel()).reshape(rhs.shape)
        assert rel(O.lapl_cyl3_fft2(rhs, dr, dz, r0, lz, zperiodic=zper), want) < 1e-12


def test_rect_yperiodic_dense():
    nx, ny = 5, 8
    dx, dy = 0.3, 0.25
    rhs = O.field(6, (ny, nx))
    ans = O.lapl_rect(rhs, dx, dy, (nx + 1) * dx, ny * dy, yperiodic=True)
    n = nx * ny
    A = np.zeros((n, n))
    for c in range(n):
        e = np.zeros(n)
        e[c] = 1
        u = e.reshape(ny, nx)
        Au = (np.roll(u, 1, 0) - 2 * u + np.roll(u, -1, 0)) / dy ** 2
        up = np.pad(u, ((0, 0), (1, 1)))
        Au = Au + (up[:, :-2] - 2 * u + up[:, 2:]) / dx ** 2
        A[:, c] = Au.ravel()
    ref = np.linalg.solve(A, rhs.ravel()).reshape(ny, nx)
    assert rel(ans, ref) < 1e-12


def test_rect_singular_reports_info():
    # zero scales make D = -2/dx^2 ... keep dominance broken: L=U=huge, lm scale = 0
    nx, ny = 4, 3
    rhs = np.ones((ny, nx))
    bad = np.r_[0, np.zeros(nx)]
    with pytest.raises(ArithmeticError):
        O.lapl_rect(rhs, 1.0, 1.0, 5.0, 4.0, lm_y_scale=bad, L_scale=np.r_[0, np.full(nx, 2.0)],
                    U_scale=np.r_[0, np.full(nx, 2.0)] * 0 + np.r_[0, [2.0, 0.0, 2.0, 0.0]])


# ---------------------------------------------------------------------------
# Thomas / cyclic reduction
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("n", [31, 127, 20, 100])
def test_tdiag_vs_golden_and_dense(n):
    k = f"tdiag_{n}"
    sub, diag, sup, rhs = GOLD[k + "_sub"], GOLD[k + "_diag"], GOLD[k + "_sup"], GOLD[k + "_rhs"]
    general = int(GOLD[k + "_general"][0])
    A = np.diag(diag) + np.diag(sup[:-1], 1) + np.diag(sub[:-1], -1)
    dense = np.linalg.solve(A, rhs)
    th = O.thomas(rhs, sub, diag, sup)
    e = np.r_[0.0, sub[:-1]]
    cr = O.cyclic_reduction(diag, e, sup, rhs, general=bool(general))
    assert rel(th, dense) < 1e-13
    assert rel(cr, dense) < 1e-13
    assert rel(th, GOLD[k + "_thomas"]) < 1e-14
    assert rel(cr, GOLD[k + "_cr"]) < 1e-15
    assert rel(cr, GOLD[k + "_crclass"]) < 1e-13


def test_cr_constant_coefficient_float():
    n = 255
    rng = np.random.default_rng(12)
    d = np.full(n, -4.3)
    e = np.r_[0.0, np.full(n - 1, 1.0)]
    f = np.r_[np.full(n - 1, 1.0), 0.0]
    b = rng.uniform(-1, 1, n)
    A = np.diag(d) + np.diag(f[:-1], 1) + np.diag(e[1:], -1)
    x = O.cyclic_reduction(d, e, f, b, dtype=np.float32)
    assert rel(x, np.linalg.solve(A, b)) < 1e-5


# ---------------------------------------------------------------------------
# NSCube::poisson RHS assembly
# ---------------------------------------------------------------------------
def test_ns_rhs_divergence_of_linear_field():
    # F = a x_face, G = b y_face, H = c z_face: the discrete divergence is a+b+c
    # exactly; boundary p terms subtract p/h^2 on each face (src/ns_cube.cpp:213-232).
    nx, ny, nz = 5, 4, 3
    dx, dy, dz, dt = 0.5, 0.25, 0.2, 0.1
    a, b, c = 1.5, -0.5, 2.0
    F = np.fromfunction(lambda i, k, j: a * j * dx, (nz, ny, nx + 1))
    G = np.fromfunction(lambda i, k, j: b * k * dy, (nz, ny + 1, nx))
    H = np.fromfunction(lambda i, k, j: c * i * dz, (nz + 1, ny, nx))
    p = np.zeros((nz + 2, ny + 2, nx + 2))
    rhs = O.ns_poisson_rhs(F, G, H, p, dx, dy, dz, dt)
    assert np.allclose(rhs, (a + b + c) / dt, rtol=0, atol=1e-12)
    p[0] = 3.0      # bottom face pressure
    p[:, :, -1] = 5.0  # x = nx+1 face
    rhs2 = O.ns_poisson_rhs(F, G, H, p, dx, dy, dz, dt)
    d = rhs2 - rhs
    assert np.allclose(d[0, :, :-1], -3.0 / dz ** 2)
    assert np.allclose(d[1:, :, -1], -5.0 / dx ** 2)
    assert np.allclose(d[0, :, -1], -3.0 / dz ** 2 - 5.0 / dx ** 2)
    assert np.allclose(d[1:, :, :-1], 0.0)


def test_field_is_uniform_and_deterministic():
    u = O.field(2, (64, 64, 64))
    assert u.min() >= -1 and u.max() < 1
    assert abs(u.mean()) < 0.01 and abs(u.var() - 1 / 3) < 0.01
    assert np.array_equal(u, O.field(2, (64, 64, 64)))
    assert not np.array_equal(u, O.field(3, (64, 64, 64)))
