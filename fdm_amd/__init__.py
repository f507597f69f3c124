"""fdm_amd: MI355X (gfx950) fast Poisson solve behind fdm's LaplCube / LaplRect API.

Python mirror of the reference's solver classes (reference src/lapl_cube.h:9-106,
src/lapl_rect.h:11-69) over the C ABI in include/fdmgpu.h, implemented by the
HIP library fdm_amd/libfdmgpu.so.  There is no CPU fallback: if the library is
missing or no GPU is present, constructing a solver raises.

    from fdm_amd import LaplCube
    s = LaplCube(dx, dy, dz, lx, ly, lz, nx, ny, nz, precision="f64")
    s.solve(ans, rhs)     # numpy arrays (synchronous) or torch CUDA tensors (stream-ordered)
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, "libfdmgpu.so")

F32, F64, F32_REFINE64 = 0, 1, 2
PERIODIC_X, PERIODIC_Y, PERIODIC_Z = 1, 2, 4
_PREC = {"f32": F32, "float": F32, "f64": F64, "double": F64, "refine": F32_REFINE64,
         "f32_refine64": F32_REFINE64}

# error codes (include/fdmgpu.h)
OK, ERR_INVALID_SIZE, ERR_INVALID_ARG, ERR_UNSUPPORTED = 0, 1, 2, 3
ERR_OUT_OF_MEMORY, ERR_DEVICE, ERR_SINGULAR, ERR_NOT_CONVERGED, ERR_ADMISSION, ERR_TOO_LARGE = 4, 5, 6, 7, 8, 9

# every entry point include/fdmgpu.h declares
EXPORTS = ("fdmgpu_plan_cube", "fdmgpu_plan_rect", "fdmgpu_plan_rect_fft2", "fdmgpu_plan_cyl3_fft2",
           "fdmgpu_rect_set_scales", "fdmgpu_set_refine",
           "fdmgpu_solve_host", "fdmgpu_solve_device", "fdmgpu_solve_batch_device", "fdmgpu_solve_table_device",
           "fdmgpu_plan_variants", "fdmgpu_plan_set_variant", "fdmgpu_plan_get_variant", "fdmgpu_get_stats",
           "fdmgpu_plan_footprint", "fdmgpu_set_timing", "fdmgpu_get_timing", "fdmgpu_set_timing_kernels", "fdmgpu_group_create",
           "fdmgpu_group_solve", "fdmgpu_group_result", "fdmgpu_group_destroy", "fdmgpu_server_create",
           "fdmgpu_server_submit", "fdmgpu_server_set_refine", "fdmgpu_server_poll", "fdmgpu_server_wait", "fdmgpu_server_drain",
           "fdmgpu_server_forget", "fdmgpu_server_destroy", "fdmgpu_tune", "fdmgpu_tune_set", "fdmgpu_ns_create", "fdmgpu_ns_step", "fdmgpu_ns_get", "fdmgpu_ns_set", "fdmgpu_ns_info",
           "fdmgpu_ns_step_timed", "fdmgpu_ns_maxabs", "fdmgpu_ns_solver", "fdmgpu_ns_destroy", "fdmgpu_destroy",
           "fdmgpu_strerror",
           "fdmgpu_version", "fdmgpu_solve_batch_noalloc", "fdmgpu_group_set_refine")

# kernel classes for fdmgpu_get_timing (include/fdmgpu.h)
K_XY_FWD, K_XY_INV, K_X_TILE, K_Y_TILE, K_ZCR, K_ROWCR, K_RESIDUAL, K_OTHER = range(8)
KERNEL_NAMES = ("xy_fwd", "xy_inv", "x_tile", "y_tile", "zcr", "rowcr", "residual", "other", "xy_fwd_resid",
                "xy_inv_acc", "residual_rhs", "xy_fwd_r32")


class FdmGpuError(RuntimeError):
    """A non-zero fdmgpu return code (the reference aborts via verify(), src/verify.h)."""

    def __init__(self, code: int, where: str):
        self.code = code
        super().__init__(f"{where}: {_lib_handle().fdmgpu_strerror(code).decode()} (code {code})")


class Stats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("rel_residual", C.c_double), ("max_rel_residual", C.c_double)]


class Variant(C.Structure):
    """fdmgpu_variant: kernel variant (tile shapes) of a plan."""
    _fields_ = [("plane", C.c_int), ("tile_x", C.c_int), ("tile_y", C.c_int), ("zcr_gmul", C.c_int),
                ("fuse", C.c_int)]

    def astuple(self):
        return (self.plane, self.tile_x, self.tile_y, self.zcr_gmul, self.fuse)


class Problem(C.Structure):
    """fdmgpu_problem: one solve of a heterogeneous batch or serving request."""
    _fields_ = [("kind", C.c_int), ("precision", C.c_int), ("periodic", C.c_int),
                ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
                ("lx", C.c_double), ("ly", C.c_double), ("lz", C.c_double),
                ("d_ans", C.c_void_p), ("d_rhs", C.c_void_p)]


class Result(C.Structure):
    """fdmgpu_result: per-problem outcome (state 0 queued, 1 running, 2 done)."""
    _fields_ = [("state", C.c_int), ("status", C.c_int), ("iterations", C.c_int),
                ("rel_residual", C.c_double), ("queue_ms", C.c_double), ("latency_ms", C.c_double)]


def build(force: bool = True) -> None:
    """Compile libfdmgpu.so for gfx950 with hipcc (make -C fdm_amd)."""
    args = ["make", "-s", "-j8", "-C", _HERE]
    if force:
        args.append("-B")
    subprocess.check_call(args)


_lib = None


def _lib_handle():
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise ImportError(f"{LIBPATH} is not built: run `make -C fdm_amd` (HIP library required, "
                              "there is no CPU fallback)")
        L = C.CDLL(LIBPATH)
        vp, d, i = C.c_void_p, C.c_double, C.c_int
        L.fdmgpu_plan_cube.argtypes = [C.POINTER(vp), i, i, d, d, d, d, d, d, i, i, i]
        L.fdmgpu_plan_rect.argtypes = [C.POINTER(vp), i, i, d, d, d, d, i, i]
        L.fdmgpu_plan_rect_fft2.argtypes = [C.POINTER(vp), i, i, d, d, d, d, i, i]
        L.fdmgpu_plan_cyl3_fft2.argtypes = [C.POINTER(vp), i, i, d, d, d, d, d, i, i, i]
        L.fdmgpu_rect_set_scales.argtypes = [vp, vp, vp, vp]
        L.fdmgpu_set_refine.argtypes = [vp, d, i]
        L.fdmgpu_solve_host.argtypes = [vp, vp, vp]
        L.fdmgpu_solve_device.argtypes = [vp, vp, vp, vp]
        L.fdmgpu_solve_batch_device.argtypes = [vp, i, vp, vp, vp]
        L.fdmgpu_solve_batch_noalloc.argtypes = [vp, i, vp, vp, vp]
        L.fdmgpu_get_stats.argtypes = [vp, C.POINTER(Stats)]
        L.fdmgpu_plan_footprint.argtypes = [vp, i, C.POINTER(C.c_ulonglong), C.POINTER(C.c_ulonglong)]
        L.fdmgpu_set_timing.argtypes = [vp, i]
        L.fdmgpu_set_timing_kernels.argtypes = [vp, C.c_uint]
        L.fdmgpu_get_timing.argtypes = [vp, i, C.POINTER(d), C.POINTER(i)]
        ull = C.c_ulonglong
        L.fdmgpu_solve_table_device.argtypes = [vp, i, vp, vp, vp]
        L.fdmgpu_plan_variants.argtypes = [vp, i, C.POINTER(Variant), C.POINTER(i)]
        L.fdmgpu_plan_set_variant.argtypes = [vp, C.POINTER(Variant)]
        L.fdmgpu_plan_get_variant.argtypes = [vp, C.POINTER(Variant)]
        ull = C.c_ulonglong
        L.fdmgpu_group_create.argtypes = [C.POINTER(vp), i, C.POINTER(Problem), i, i]
        L.fdmgpu_group_solve.argtypes = [vp, vp]
        L.fdmgpu_group_set_refine.argtypes = [vp, d, i]
        L.fdmgpu_group_result.argtypes = [vp, i, C.POINTER(Result)]
        L.fdmgpu_group_destroy.argtypes = [vp]
        L.fdmgpu_server_create.argtypes = [C.POINTER(vp), ull, i, i]
        L.fdmgpu_server_submit.argtypes = [vp, C.POINTER(Problem), C.POINTER(ull)]
        L.fdmgpu_server_set_refine.argtypes = [vp, d, i]
        L.fdmgpu_server_poll.argtypes = [vp, ull, C.POINTER(Result)]
        L.fdmgpu_server_wait.argtypes = [vp, ull, C.POINTER(Result)]
        L.fdmgpu_server_drain.argtypes = [vp]
        L.fdmgpu_server_forget.argtypes = [vp, ull]
        L.fdmgpu_server_destroy.argtypes = [vp]
        L.fdmgpu_tune.argtypes = [i, C.POINTER(Problem), i, C.c_char_p]
        L.fdmgpu_tune_set.argtypes = [C.POINTER(Problem), C.POINTER(Variant)]
        L.fdmgpu_ns_create.argtypes = [C.POINTER(vp), i, i, i, i] + [d] * 9
        L.fdmgpu_ns_step.argtypes = [vp, i, vp]
        L.fdmgpu_ns_get.argtypes = [vp, i, vp, ull]
        L.fdmgpu_ns_set.argtypes = [vp, i, vp, ull]
        L.fdmgpu_ns_info.argtypes = [vp, i, C.POINTER(ull), C.POINTER(vp)]
        L.fdmgpu_ns_solver.argtypes = [vp, C.POINTER(vp)]
        L.fdmgpu_ns_step_timed.argtypes = [vp, C.POINTER(C.c_double), vp]
        L.fdmgpu_ns_maxabs.argtypes = [vp, C.POINTER(C.c_double)]
        L.fdmgpu_ns_destroy.argtypes = [vp]
        L.fdmgpu_destroy.argtypes = [vp]
        L.fdmgpu_strerror.argtypes = [i]
        L.fdmgpu_strerror.restype = C.c_char_p
        L.fdmgpu_version.restype = i
        _lib = L
    return _lib


def lib():
    """The loaded libfdmgpu.so (ctypes)."""
    return _lib_handle()


def _check(rc: int, where: str):
    if rc != OK:
        raise FdmGpuError(rc, where)


def _is_torch(a) -> bool:
    return type(a).__module__.startswith("torch")


def _is_dev(a) -> bool:
    return type(a).__name__ == "DeviceArray"


def _stream_handle(stream):
    if stream is None:
        return C.c_void_p()
    return getattr(stream, "h", stream)


class _Plan:
    _kind = ""

    def __init__(self, handle: C.c_void_p, precision: int, shape):
        self._h = handle
        self.precision = precision
        self.shape = tuple(shape)
        self.dtype = np.float32 if precision == F32 else np.float64

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _lib_handle().fdmgpu_destroy(h)
            except Exception:
                pass
            self._h = None

    # -- solve --------------------------------------------------------------
    def _check_array(self, a, name):
        if tuple(a.shape) != self.shape:
            raise ValueError(f"{name} has shape {tuple(a.shape)}, expected {self.shape}")

    def solve(self, ans, rhs, stream=None):
        """Reference semantics: ans = A^{-1} rhs; ans fully overwritten; ans may be rhs.

        numpy arrays: synchronous host solve (fdmgpu_solve_host).
        fdm_amd.hip.DeviceArray: enqueued on `stream` (fdmgpu_solve_device).
        torch CUDA tensors: enqueued on torch's current stream.  torch ships its
        own HIP runtime: import torch and initialise its device before this
        package loads libfdmgpu.so so that both share one runtime.
        """
        self._check_array(ans, "ans")
        self._check_array(rhs, "rhs")
        L = _lib_handle()
        if _is_dev(ans):
            for a, nm in ((ans, "ans"), (rhs, "rhs")):
                if a.dtype != self.dtype:
                    raise TypeError(f"{nm} must be {np.dtype(self.dtype).name}")
            _check(L.fdmgpu_solve_device(self._h, ans.ptr, rhs.ptr, _stream_handle(stream)), f"{self._kind}.solve")
            return ans
        if _is_torch(ans):
            import torch
            want = torch.float32 if self.precision == F32 else torch.float64
            for t, nm in ((ans, "ans"), (rhs, "rhs")):
                if t.dtype != want or not t.is_cuda or not t.is_contiguous():
                    raise TypeError(f"{nm} must be a contiguous CUDA {want} tensor")
            stream = torch.cuda.current_stream(ans.device).cuda_stream
            _check(L.fdmgpu_solve_device(self._h, C.c_void_p(ans.data_ptr()), C.c_void_p(rhs.data_ptr()),
                                         C.c_void_p(stream)), f"{self._kind}.solve")
            return ans
        for a, nm in ((ans, "ans"), (rhs, "rhs")):
            if a.dtype != self.dtype or not a.flags["C_CONTIGUOUS"]:
                raise TypeError(f"{nm} must be a C-contiguous {np.dtype(self.dtype).name} array")
        if not ans.flags["WRITEABLE"]:
            raise TypeError("ans must be writeable")
        _check(L.fdmgpu_solve_host(self._h, ans.ctypes.data_as(C.c_void_p), rhs.ctypes.data_as(C.c_void_p)),
               f"{self._kind}.solve")
        return ans

    def solve_batch(self, ans, rhs, stream=None):
        """count problems stacked on a leading axis (DeviceArray or torch CUDA tensors),
        solved in one launch sequence."""
        return self._batch(ans, rhs, stream, "fdmgpu_solve_batch_device", "solve_batch")

    def solve_batch_noalloc(self, ans, rhs, stream=None):
        """solve_batch with no device allocation (fdmgpu_solve_batch_noalloc): the
        refinement scratch comes from the library's static 64 MiB array, shared by
        every call on the GPU (calls must not overlap); ERR_TOO_LARGE beyond it."""
        return self._batch(ans, rhs, stream, "fdmgpu_solve_batch_noalloc", "solve_batch_noalloc")

    def _batch(self, ans, rhs, stream, entry, what):
        count = ans.shape[0]
        if tuple(ans.shape[1:]) != self.shape or tuple(rhs.shape) != tuple(ans.shape):
            raise ValueError("batch shape mismatch")
        if _is_dev(ans) and _is_dev(rhs):
            for a, nm in ((ans, "ans"), (rhs, "rhs")):
                if a.dtype != self.dtype:
                    raise TypeError(f"{nm} must be {np.dtype(self.dtype).name}")
            pa, pr, st = ans.ptr, rhs.ptr, _stream_handle(stream)
        elif _is_torch(ans) and _is_torch(rhs):
            import torch
            want = torch.float32 if self.precision == F32 else torch.float64
            for t, nm in ((ans, "ans"), (rhs, "rhs")):
                if t.dtype != want or not t.is_cuda or not t.is_contiguous():
                    raise TypeError(f"{nm} must be a contiguous CUDA {want} tensor")
            pa, pr = C.c_void_p(ans.data_ptr()), C.c_void_p(rhs.data_ptr())
            st = C.c_void_p(torch.cuda.current_stream(ans.device).cuda_stream)
        else:
            raise TypeError(f"{what} takes two fdm_amd.hip.DeviceArray or two torch CUDA tensors")
        _check(getattr(_lib_handle(), entry)(self._h, count, pa, pr, st), f"{self._kind}.{what}")
        return ans

    def set_timing(self, enable: bool = True, kernels=None):
        """Enable (and reset) per-kernel HIP-event timing: of every kernel
        class, or only of the classes named in `kernels` (KERNEL_NAMES)."""
        if kernels is None:
            _check(_lib_handle().fdmgpu_set_timing(self._h, int(enable)), "set_timing")
            return
        mask = 0
        for name in kernels:
            mask |= 1 << KERNEL_NAMES.index(name)
        _check(_lib_handle().fdmgpu_set_timing_kernels(self._h, mask if enable else 0), "set_timing")

    def timing(self):
        """{kernel name: (total ms, launches)} since the last set_timing (synchronises)."""
        out = {}
        for k, name in enumerate(KERNEL_NAMES):
            ms, n = C.c_double(), C.c_int()
            _check(_lib_handle().fdmgpu_get_timing(self._h, k, C.byref(ms), C.byref(n)), "get_timing")
            if n.value:
                out[name] = (ms.value, n.value)
        return out

    def variants(self):
        """Kernel variants valid for this plan, the default first: tuples (plane, tile_x,
        tile_y, zcr_gmul, fuse), plane 0 = x and y tile kernels, 1 = plane kernel,
        2 = streaming plane kernel (planes of N = 128), as fdmgpu_variant in fdmgpu.h."""
        L = _lib_handle()
        n = C.c_int()
        _check(L.fdmgpu_plan_variants(self._h, 0, None, C.byref(n)), "plan_variants")
        arr = (Variant * max(1, n.value))()
        _check(L.fdmgpu_plan_variants(self._h, n.value, arr, C.byref(n)), "plan_variants")
        return [arr[k].astuple() for k in range(n.value)]

    @property
    def variant(self):
        v = Variant()
        _check(_lib_handle().fdmgpu_plan_get_variant(self._h, C.byref(v)), "plan_get_variant")
        return v.astuple()

    @variant.setter
    def variant(self, t):
        v = Variant(*t)
        _check(_lib_handle().fdmgpu_plan_set_variant(self._h, C.byref(v)), "plan_set_variant")

    def solve_table(self, ans_tab, rhs_tab, count: int, stream=None):
        """Batch of `count` problems at device pointers ans_tab[p] / rhs_tab[p]
        (device arrays of uint64 pointers, e.g. hip.DeviceArray).  Every pointer
        must address a device array of the plan's shape and element type."""
        if count < 1:
            raise ValueError("count must be positive")
        for t, nm in ((ans_tab, "ans_tab"), (rhs_tab, "rhs_tab")):
            if not _is_dev(t):
                raise TypeError(f"{nm} must be an fdm_amd.hip.DeviceArray of device pointers")
            if np.dtype(t.dtype).itemsize != 8 or np.dtype(t.dtype).kind not in "ui":
                raise TypeError(f"{nm} must hold 64-bit pointers (uint64)")
            if int(np.prod(t.shape)) < count:
                raise ValueError(f"{nm} holds {int(np.prod(t.shape))} pointers, fewer than count={count}")
        _check(_lib_handle().fdmgpu_solve_table_device(self._h, count, ans_tab.ptr, rhs_tab.ptr,
                                                       _stream_handle(stream)), f"{self._kind}.solve_table")

    def set_refine(self, rel_tol: float = 1e-10, max_iter: int = 4):
        _check(_lib_handle().fdmgpu_set_refine(self._h, rel_tol, max_iter), "set_refine")

    def stats(self) -> Stats:
        st = Stats()
        _check(_lib_handle().fdmgpu_get_stats(self._h, C.byref(st)), "get_stats")
        return st

    def footprint(self, count: int = 1):
        r, p = C.c_ulonglong(), C.c_ulonglong()
        _check(_lib_handle().fdmgpu_plan_footprint(self._h, count, C.byref(r), C.byref(p)), "footprint")
        return r.value, p.value


def _prec(precision) -> int:
    if isinstance(precision, int):
        return precision
    if precision in (np.float32, np.dtype(np.float32)):
        return F32
    if precision in (np.float64, np.dtype(np.float64)):
        return F64
    return _PREC[str(precision).lower()]


class LaplCube(_Plan):
    """LaplCube<T,check,F>(dx,dy,dz,lx,ly,lz,nx,ny,nz) (reference src/lapl_cube.h:58-100).

    Arrays are [nz][ny][nx] interior points, x fastest.  precision: "f32"
    (LaplCube<float>), "f64" (LaplCube<double>) or "refine" (double in/out,
    FP32 solves with FP64 iterative refinement).  Transform lengths nx+1,
    ny+1, nz+1 must be powers of two (the reference aborts otherwise,
    src/fft.cpp:67).
    """
    _kind = "LaplCube"

    def __init__(self, dx, dy, dz, lx, ly, lz, nx, ny, nz, precision="f64", periodic=False):
        h = C.c_void_p()
        p = _prec(precision)
        _check(_lib_handle().fdmgpu_plan_cube(C.byref(h), p, 7 if periodic else 0, dx, dy, dz, lx, ly, lz,
                                              nx, ny, nz), "LaplCube")
        super().__init__(h, p, (nz, ny, nx))
        self.dx, self.dy, self.dz, self.lx, self.ly, self.lz = dx, dy, dz, lx, ly, lz
        self.nx, self.ny, self.nz = nx, ny, nz


class LaplRect(_Plan):
    """LaplRect<T,check,F>(dx,dy,lx,ly,nx,ny) (reference src/lapl_rect.h:40-63).

    Arrays are [ny][nx], x fastest.  lm_y_scale / L_scale / U_scale (length
    nx+1, index 0 unused, default ones) via set_scales (src/lapl_rect.h:57-59).
    """
    _kind = "LaplRect"

    def __init__(self, dx, dy, lx, ly, nx, ny, precision="f64", yperiodic=False):
        h = C.c_void_p()
        p = _prec(precision)
        _check(_lib_handle().fdmgpu_plan_rect(C.byref(h), p, PERIODIC_Y if yperiodic else 0, dx, dy, lx, ly, nx,
                                              ny), "LaplRect")
        super().__init__(h, p, (ny, nx))
        self.dx, self.dy, self.lx, self.ly, self.nx, self.ny = dx, dy, lx, ly, nx, ny

    def set_scales(self, lm_y_scale=None, L_scale=None, U_scale=None):
        keep = []

        def ptr(a):
            if a is None:
                return None
            a = np.ascontiguousarray(a, np.float64)
            if a.shape != (self.nx + 1,):
                raise ValueError("scales have nx+1 entries (index 0 unused)")
            keep.append(a)
            return a.ctypes.data_as(C.c_void_p)

        _check(_lib_handle().fdmgpu_rect_set_scales(self._h, ptr(lm_y_scale), ptr(L_scale), ptr(U_scale)),
               "LaplRect.set_scales")


class LaplRectFFT2(LaplRect):
    """LaplRectFFT2<T,check,F>(dx,dy,lx,ly,nx,ny) (reference src/lapl_rect.h:76-109):
    transforms along both axes.  yperiodic = tensor_flags<periodic>; both
    = tensor_flags<periodic, periodic> (zero mode coefficient 1).  Arrays are
    [ny][nx] over the reference's index ranges (all n points on a periodic
    axis).  set_scales(lm_y_scale=...) as the reference's public member."""
    _kind = "LaplRectFFT2"

    def __init__(self, dx, dy, lx, ly, nx, ny, precision="f64", yperiodic=False, xperiodic=False):
        h = C.c_void_p()
        p = _prec(precision)
        mask = (PERIODIC_Y if yperiodic else 0) | (PERIODIC_X if xperiodic else 0)
        _check(_lib_handle().fdmgpu_plan_rect_fft2(C.byref(h), p, mask, dx, dy, lx, ly, nx, ny), "LaplRectFFT2")
        _Plan.__init__(self, h, p, (ny, nx))
        self.dx, self.dy, self.lx, self.ly, self.nx, self.ny = dx, dy, lx, ly, nx, ny


class LaplCyl3FFT2(_Plan):
    """LaplCyl3FFT2<T,check,zflag,true>(dr,dz,r0,lr,lz,nr,nz,nphi) (reference
    src/lapl_cyl.h:173-249): arrays [nphi][nz][nr], r fastest; zperiodic =
    zflag periodic (z indices 0..nz-1).  The r systems are solved by cyclic
    reduction (the use_cyclic_reduction instantiation)."""
    _kind = "LaplCyl3FFT2"

    def __init__(self, dr, dz, r0, lr, lz, nr, nz, nphi, precision="f64", zperiodic=False):
        h = C.c_void_p()
        p = _prec(precision)
        _check(_lib_handle().fdmgpu_plan_cyl3_fft2(C.byref(h), p, PERIODIC_Z if zperiodic else 0, dr, dz, r0, lr, lz,
                                                   nr, nz, nphi), "LaplCyl3FFT2")
        super().__init__(h, p, (nphi, nz, nr))
        self.dr, self.dz, self.r0, self.lr, self.lz = dr, dz, r0, lr, lz
        self.nr, self.nz, self.nphi = nr, nz, nphi


NS_FIELDS = ("u", "v", "w", "p", "x", "F", "G", "H", "RHS")


def ns_shapes(nx, ny, nz):
    """Field shapes (z, y, x) with the reference's index ranges (src/ns_cube.h:66-75)."""
    return {"u": (nz + 2, ny + 2, nx + 3), "v": (nz + 2, ny + 3, nx + 2), "w": (nz + 3, ny + 2, nx + 2),
            "p": (nz + 2, ny + 2, nx + 2), "x": (nz, ny, nx), "F": (nz, ny, nx + 1), "G": (nz, ny + 1, nx),
            "H": (nz + 1, ny, nx), "RHS": (nz, ny, nx)}


class NSCube:
    """NSCube<double,check>(Config) (reference src/ns_cube.h:12-78) on the GPU: the
    lid-driven cavity whose pressure projection is the fast Poisson solve.
    solver: "refine" (FP32 transforms + FP64 refinement) or "f64"."""

    def __init__(self, nx=32, ny=32, nz=32, x1=-np.pi, x2=np.pi, y1=-np.pi, y2=np.pi, z1=-np.pi, z2=np.pi,
                 u0=1.0, Re=1.0, dt=0.001, solver="refine"):
        h = C.c_void_p()
        _check(_lib_handle().fdmgpu_ns_create(C.byref(h), _prec(solver), nx, ny, nz, x1, x2, y1, y2, z1, z2,
                                              u0, Re, dt), "NSCube")
        self._h = h
        self.nx, self.ny, self.nz = nx, ny, nz
        self.dx, self.dy, self.dz = (x2 - x1) / nx, (y2 - y1) / ny, (z2 - z1) / nz
        self.shapes = ns_shapes(nx, ny, nz)
        self.time_index = 0

    def step(self, count: int = 1, stream=None):
        _check(_lib_handle().fdmgpu_ns_step(self._h, count, _stream_handle(stream)), "NSCube.step")
        self.time_index += count

    def step_timed(self, stream=None) -> tuple:
        """One step, eagerly, with GPU events at its phases: (init_bound, FGH,
        poisson, update_uvwp) in ms (reference src/ns_cube.cpp:27-50)."""
        ms = (C.c_double * 4)()
        _check(_lib_handle().fdmgpu_ns_step_timed(self._h, ms, _stream_handle(stream)), "NSCube.step_timed")
        self.time_index += 1
        return tuple(ms)

    def maxabs(self) -> dict:
        """Largest |value| of every field, ghost layers included, reduced on the
        GPU (tensor::maxabs, reference src/tensor.h:269-273)."""
        out = (C.c_double * 9)()
        _check(_lib_handle().fdmgpu_ns_maxabs(self._h, out), "NSCube.maxabs")
        return dict(zip(NS_FIELDS, out))

    def get(self, name: str) -> np.ndarray:
        """A copy of field `name`.  Raises FdmGpuError (ERR_NOT_CONVERGED) when
        a FP32+refine projection of this run ended above its tolerance."""
        out = np.empty(self.shapes[name], np.float64)
        _check(_lib_handle().fdmgpu_ns_get(self._h, NS_FIELDS.index(name), out.ctypes.data_as(C.c_void_p),
                                           out.size), "NSCube.get")
        return out

    def set(self, name: str, a: np.ndarray):
        a = np.ascontiguousarray(a, np.float64)
        if a.shape != self.shapes[name]:
            raise ValueError(f"{name} has shape {self.shapes[name]}")
        _check(_lib_handle().fdmgpu_ns_set(self._h, NS_FIELDS.index(name), a.ctypes.data_as(C.c_void_p), a.size),
               "NSCube.set")

    def set_solver_refine(self, rel_tol: float = 1e-10, max_iter: int = 4):
        """Refinement settings of the FP32+refine pressure solver (fdmgpu_set_refine
        on the plan of fdmgpu_ns_solver); steps enqueued afterwards use them."""
        plan = C.c_void_p()
        _check(_lib_handle().fdmgpu_ns_solver(self._h, C.byref(plan)), "NSCube.solver")
        _check(_lib_handle().fdmgpu_set_refine(plan, rel_tol, max_iter), "NSCube.set_solver_refine")

    def solver_stats(self) -> Stats:
        plan = C.c_void_p()
        _check(_lib_handle().fdmgpu_ns_solver(self._h, C.byref(plan)), "NSCube.solver")
        st = Stats()
        _check(_lib_handle().fdmgpu_get_stats(plan, C.byref(st)), "get_stats")
        return st

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib_handle().fdmgpu_ns_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def version() -> int:
    return _lib_handle().fdmgpu_version()
