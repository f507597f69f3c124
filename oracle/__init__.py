"""ctypes bindings to the CPU oracle (oracle/liboracle.so) and, when built,
to the reference's own code (oracle/_ref/libfdmref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py.  The product library (fdm_amd) never imports
this package.

Functions and the reference code they restate are documented in
oracle/fdm_oracle.cpp; each wrapper here is numpy-in / numpy-out.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_REFLIB = os.path.join(_HERE, "_ref", "libfdmref.so")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_d = C.c_double
_i = C.c_int


def build(force: bool = False) -> None:
    """Compile the oracle (and oracle/_ref when the reference checkout exists)."""
    if force or not os.path.exists(_LIB):
        subprocess.check_call(["make", "-s", "-C", _HERE, "all"])


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = C.CDLL(_LIB)
        L.orc_dst1_d.argtypes = [_dp, _dp, _i, _d]
        L.orc_dst1_f.argtypes = [_fp, _fp, _i, C.c_float]
        L.orc_dct1_d.argtypes = [_dp, _dp, _i, _d]
        L.orc_rdft_fwd_d.argtypes = [_dp, _dp, _i, _d]
        L.orc_rdft_inv_d.argtypes = [_dp, _dp, _i, _d]
        L.orc_sft_direct.argtypes = [_dp, _dp, _i, _d]
        cube = [_i, _d, _d, _d, _d, _d, _d, _i, _i, _i]
        L.orc_lapl_cube_d.argtypes = [_dp, _dp] + cube
        L.orc_lapl_cube_f.argtypes = [_fp, _fp] + cube
        L.orc_lapl_cube_d.restype = _i
        L.orc_lapl_cube_f.restype = _i
        rect = [_i, _d, _d, _d, _d, _i, _i]
        L.orc_lapl_rect_d.argtypes = [_dp, _dp] + rect + [C.c_void_p] * 3
        L.orc_lapl_rect_f.argtypes = [_fp, _fp] + rect + [C.c_void_p] * 3
        L.orc_lapl_rect_d.restype = _i
        L.orc_lapl_rect_f.restype = _i
        L.orc_lapl_rect_fft2_d.argtypes = [_dp, _dp, _i, _i, _d, _d, _d, _d, _i, _i, C.c_void_p]
        L.orc_lapl_rect_fft2_d.restype = _i
        cyl = [_i, _d, _d, _d, _d, _i, _i, _i, _i]
        L.orc_lapl_cyl3_fft2_d.argtypes = [_dp, _dp] + cyl
        L.orc_lapl_cyl3_fft2_f.argtypes = [_fp, _fp] + cyl
        L.orc_lapl_cyl3_fft2_d.restype = _i
        L.orc_lapl_cyl3_fft2_f.restype = _i
        L.orc_gtsv_d.argtypes = [_i, _dp, _dp, _dp, _dp]
        L.orc_gtsv_d.restype = _i
        L.orc_gtsv_f.argtypes = [_i, _fp, _fp, _fp, _fp]
        L.orc_gtsv_f.restype = _i
        L.orc_thomas_d.argtypes = [_dp, _dp, _dp, _dp, _i]
        L.orc_cr_d.argtypes = [_dp, _dp, _dp, _dp, _i, _i, _i]
        L.orc_cr_f.argtypes = [_fp, _fp, _fp, _fp, _i, _i, _i]
        L.orc_ns_poisson_rhs_d.argtypes = [_dp] * 5 + [_i, _i, _i, _d, _d, _d, _d]
        L.orc_field.argtypes = [_dp, C.c_uint64, C.c_long]
        L.orc_apply_lapl.argtypes = [_dp, _dp, _i, _i, _i, _d, _d, _d]
        L.orc_rel_residual.argtypes = [_dp, _dp, _i, _i, _i, _d, _d, _d]
        L.orc_rel_residual.restype = _d
        L.orc_num_threads.restype = _i
        L.orc_ns_create.argtypes = [_i, _i, _i] + [_d] * 9
        L.orc_ns_create.restype = C.c_void_p
        L.orc_ns_step.argtypes = [C.c_void_p, _i]
        L.orc_ns_step.restype = _i
        L.orc_ns_get.argtypes = [C.c_void_p, _i, C.c_void_p]
        L.orc_ns_get.restype = C.c_long
        L.orc_ns_set.argtypes = [C.c_void_p, _i, _dp]
        L.orc_ns_set.restype = C.c_long
        L.orc_ns_destroy.argtypes = [C.c_void_p]
        _lib = L
    return _lib


def ref_available() -> bool:
    return os.path.exists(_REFLIB)


def ref():
    """The reference's own code compiled from its sources (oracle/_ref)."""
    global _ref
    if _ref is None:
        if not ref_available():
            raise FileNotFoundError(_REFLIB)
        R = C.CDLL(_REFLIB)
        R.ref_fft_d.argtypes = [_dp, _dp, _i, _d, _i]
        R.ref_fft_f.argtypes = [_fp, _fp, _i, C.c_float, _i]
        R.ref_sft_d.argtypes = [_dp, _dp, _i, _d]
        R.ref_cft_d.argtypes = [_dp, _dp, _i, _d]
        cube = [_i, _d, _d, _d, _d, _d, _d, _i, _i, _i]
        R.ref_lapl_cube_d.argtypes = [_dp, _dp] + cube
        R.ref_lapl_cube_f.argtypes = [_fp, _fp] + cube
        R.ref_cube_create_d.argtypes = [_d] * 6 + [_i] * 3
        R.ref_cube_create_d.restype = C.c_void_p
        R.ref_cube_solve_d.argtypes = [C.c_void_p, _dp, _dp]
        R.ref_cube_destroy.argtypes = [C.c_void_p]
        R.ref_thomas_d.argtypes = [_dp, _dp, _dp, _dp, _i]
        R.ref_cr_d.argtypes = [_dp, _dp, _dp, _dp, _i, _i, _i]
        R.ref_cr_class_d.argtypes = [_dp, _dp, _dp, _dp, _i]
        _ref = R
    return _ref


# ---------------------------------------------------------------------------
# numpy-level helpers
# ---------------------------------------------------------------------------
def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def dst1(x: np.ndarray, N: int, scale: float = 1.0) -> np.ndarray:
    """DST-I of x[1..N-1] (length-N arrays, entry 0 ignored), like FFT<T>::sFFT."""
    dt = np.float32 if x.dtype == np.float32 else np.float64
    x = _c(x, dt)
    X = np.zeros(N, dt)
    (lib().orc_dst1_f if dt == np.float32 else lib().orc_dst1_d)(X, x, N, scale)
    return X


def sft_direct(x: np.ndarray, N: int, scale: float = 1.0) -> np.ndarray:
    X = np.zeros(N)
    lib().orc_sft_direct(X, _c(x, np.float64), N, scale)
    return X


def dct1(x: np.ndarray, N: int, scale: float = 1.0) -> np.ndarray:
    X = np.zeros(N + 1)
    lib().orc_dct1_d(X, _c(x, np.float64), N, scale)
    return X


def rdft_fwd(x: np.ndarray, N: int, scale: float = 1.0) -> np.ndarray:
    X = np.zeros(N)
    lib().orc_rdft_fwd_d(X, _c(x, np.float64), N, scale)
    return X


def rdft_inv(X: np.ndarray, N: int, scale: float = 1.0) -> np.ndarray:
    x = np.zeros(N)
    lib().orc_rdft_inv_d(x, _c(X, np.float64), N, scale)
    return x


def cube_geometry(n, x1=-np.pi, x2=np.pi, periodic=False):
    """(h, l) per axis as NSCube builds them (src/ns_cube.h:63,77): h=(x2-x1)/n,
    l = x2-x1+h (Dirichlet); the periodic tests use l = x2-x1 (ut/ut_lapl_cube.cpp:176-181)."""
    h = (x2 - x1) / n
    return h, ((x2 - x1) if periodic else (x2 - x1 + h))


def lapl_cube(rhs: np.ndarray, dx, dy, dz, lx, ly, lz, periodic=False) -> np.ndarray:
    """LaplCube<T>::solve on rhs[nz,ny,nx] (T from rhs.dtype)."""
    nz, ny, nx = rhs.shape
    dt = np.float32 if rhs.dtype == np.float32 else np.float64
    r = _c(rhs, dt)
    ans = np.zeros_like(r)
    f = lib().orc_lapl_cube_f if dt == np.float32 else lib().orc_lapl_cube_d
    rc = f(ans, r, int(periodic), dx, dy, dz, lx, ly, lz, nx, ny, nz)
    if rc:
        raise ValueError("LaplCube: transform length must be a power of two")
    return ans


def ref_lapl_cube(rhs: np.ndarray, dx, dy, dz, lx, ly, lz, periodic=False) -> np.ndarray:
    nz, ny, nx = rhs.shape
    dt = np.float32 if rhs.dtype == np.float32 else np.float64
    r = _c(rhs, dt)
    ans = np.zeros_like(r)
    f = ref().ref_lapl_cube_f if dt == np.float32 else ref().ref_lapl_cube_d
    f(ans, r, int(periodic), dx, dy, dz, lx, ly, lz, nx, ny, nz)
    return ans


def _scales_ptr(a, dt, keep):
    if a is None:
        return None
    a = _c(a, dt)
    keep.append(a)
    return a.ctypes.data_as(C.c_void_p)


def lapl_rect(rhs: np.ndarray, dx, dy, lx, ly, yperiodic=False, lm_y_scale=None, L_scale=None,
              U_scale=None) -> np.ndarray:
    """LaplRect<T>::solve on rhs[ny,nx]; raises on a gtsv failure (info != 0)."""
    ny, nx = rhs.shape
    dt = np.float32 if rhs.dtype == np.float32 else np.float64
    r = _c(rhs, dt)
    ans = np.zeros_like(r)
    keep = []
    f = lib().orc_lapl_rect_f if dt == np.float32 else lib().orc_lapl_rect_d
    info = f(ans, r, int(yperiodic), dx, dy, lx, ly, nx, ny, _scales_ptr(lm_y_scale, dt, keep),
             _scales_ptr(L_scale, dt, keep), _scales_ptr(U_scale, dt, keep))
    if info:
        raise ArithmeticError(f"gtsv info={info}")
    return ans


def lapl_rect_fft2(rhs: np.ndarray, dx, dy, lx, ly, yperiodic=False, xperiodic=False,
                   lm_y_scale=None) -> np.ndarray:
    ny, nx = rhs.shape
    r = _c(rhs, np.float64)
    ans = np.zeros_like(r)
    keep = []
    rc = lib().orc_lapl_rect_fft2_d(ans, r, int(yperiodic), int(xperiodic), dx, dy, lx, ly, nx, ny,
                                   _scales_ptr(lm_y_scale, np.float64, keep))
    if rc:
        raise ValueError("LaplRectFFT2: transform lengths must be powers of two")
    return ans


def lapl_cyl3_fft2(rhs: np.ndarray, dr, dz, r0, lz, zperiodic=False, cyclic_reduction=False) -> np.ndarray:
    """LaplCyl3FFT2 (src/lapl_cyl.cpp:11-128): rhs [nphi][nz][nr], r fastest."""
    nphi, nz, nr = rhs.shape
    dt = np.float32 if rhs.dtype == np.float32 else np.float64
    r = _c(rhs, dt)
    ans = np.zeros_like(r)
    f = lib().orc_lapl_cyl3_fft2_f if dt == np.float32 else lib().orc_lapl_cyl3_fft2_d
    rc = f(ans, r, int(zperiodic), dr, dz, r0, lz, nr, nz, nphi, int(cyclic_reduction))
    if rc:
        raise ValueError(f"LaplCyl3FFT2: invalid sizes or singular r system (code {rc})")
    return ans


def gtsv(dl, d, du, b, dtype=np.float64):
    """LAPACK ?gtsv restatement; returns (x, info)."""
    dl, d, du, b = (_c(v, dtype).copy() for v in (dl, d, du, b))
    f = lib().orc_gtsv_f if dtype == np.float32 else lib().orc_gtsv_d
    info = f(len(d), dl, d, du, b)
    return b, info


def thomas(B, A1, A2, A3):
    B, A1, A2, A3 = (_c(v, np.float64).copy() for v in (B, A1, A2, A3))
    lib().orc_thomas_d(B, A1, A2, A3, len(B))
    return B


def cyclic_reduction(d, e, f, b, general=False, dtype=np.float64):
    d, e, f, b = (_c(v, dtype).copy() for v in (d, e, f, b))
    n = len(b)
    q = int(np.ceil(np.log2(n + 1)))
    (lib().orc_cr_f if dtype == np.float32 else lib().orc_cr_d)(d, e, f, b, q, n, int(general))
    return b


def field(seed: int, shape) -> np.ndarray:
    """Seeded U[-1,1) field (counter-based splitmix64)."""
    n = int(np.prod(shape))
    u = np.empty(n)
    lib().orc_field(u, seed, n)
    return u.reshape(shape)


def apply_lapl(u: np.ndarray, hx, hy, hz=None) -> np.ndarray:
    """7-point (3-D u) / 5-point (2-D u) discrete Laplacian with zero Dirichlet boundary."""
    u3 = _c(u, np.float64)
    shape = u3.shape
    if u3.ndim == 2:
        u3 = u3.reshape(1, *shape)
        hz = 0.0
    nz, ny, nx = u3.shape
    f = np.empty_like(u3)
    lib().orc_apply_lapl(f, u3, nx, ny, nz, hx, hy, hz)
    return f.reshape(shape)


def rel_residual(f: np.ndarray, u: np.ndarray, hx, hy, hz=None) -> float:
    """||f - A u||_2 / ||f||_2 (FP64, long-double sums)."""
    f3 = _c(f, np.float64)
    u3 = _c(u, np.float64)
    if f3.ndim == 2:
        hz = 0.0
        f3 = f3.reshape(1, *f3.shape)
        u3 = u3.reshape(1, *u3.shape)
    nz, ny, nx = f3.shape
    return lib().orc_rel_residual(f3, u3, nx, ny, nz, hx, hy, hz)


def ns_poisson_rhs(F, G, H, p, dx, dy, dz, dt):
    """NSCube::poisson RHS assembly; F,G,H,p in NSCube's offset layouts."""
    nz, ny, nx1 = F.shape
    nx = nx1 - 1
    out = np.zeros((nz, ny, nx))
    lib().orc_ns_poisson_rhs_d(out, _c(F, np.float64), _c(G, np.float64), _c(H, np.float64),
                               _c(p, np.float64), nx, ny, nz, dx, dy, dz, dt)
    return out


def manufactured_cube(n: int, seed: int, x1=-np.pi, x2=np.pi):
    """The workload's seeded problem: u ~ U[-1,1) interior, f = A u (7-point)."""
    h, l = cube_geometry(n, x1, x2)
    u = field(seed, (n, n, n))
    f = apply_lapl(u, h, h, h)
    return u, f, h, l


# ---------------------------------------------------------------------------
# NSCube<double> (oracle restatement of src/ns_cube.cpp; see fdm_oracle.cpp)
# ---------------------------------------------------------------------------
NS_FIELDS = ("u", "v", "w", "p", "x", "F", "G", "H", "RHS")


def ns_shapes(nx, ny, nz):
    """Array shapes (z, y, x) with the reference's index ranges (src/ns_cube.h:66-75)."""
    return {"u": (nz + 2, ny + 2, nx + 3), "v": (nz + 2, ny + 3, nx + 2), "w": (nz + 3, ny + 2, nx + 2),
            "p": (nz + 2, ny + 2, nx + 2), "x": (nz, ny, nx), "F": (nz, ny, nx + 1), "G": (nz, ny + 1, nx),
            "H": (nz + 1, ny, nx), "RHS": (nz, ny, nx)}


class NsCube:
    """CPU oracle of NSCube<double,check>: NSCube(Config) parameters, step()."""

    def __init__(self, nx, ny, nz, x1=-np.pi, x2=np.pi, y1=-np.pi, y2=np.pi, z1=-np.pi, z2=np.pi, u0=1.0,
                 Re=1.0, dt=0.001):
        self.h = lib().orc_ns_create(nx, ny, nz, x1, x2, y1, y2, z1, z2, u0, Re, dt)
        if not self.h:
            raise ValueError("unsupported NSCube size (needs n >= 2 and nz <= nx)")
        self.nx, self.ny, self.nz = nx, ny, nz
        self.dx, self.dy, self.dz = (x2 - x1) / nx, (y2 - y1) / ny, (z2 - z1) / nz
        self.dt, self.Re, self.u0 = dt, Re, u0
        self.shapes = ns_shapes(nx, ny, nz)

    def step(self, count=1):
        rc = lib().orc_ns_step(self.h, count)
        if rc:
            raise RuntimeError(f"oracle NSCube step failed ({rc})")

    def get(self, name):
        out = np.empty(self.shapes[name], np.float64)
        lib().orc_ns_get(self.h, NS_FIELDS.index(name), out.ctypes.data_as(C.c_void_p))
        return out

    def set(self, name, a):
        a = _c(a, np.float64)
        assert a.shape == self.shapes[name]
        lib().orc_ns_set(self.h, NS_FIELDS.index(name), a)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_ns_destroy(self.h)
            self.h = None
