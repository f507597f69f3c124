"""Heterogeneous batches, serving and the kernel-variant autotuner (ctypes over
include/fdmgpu.h: fdmgpu_group_*, fdmgpu_server_*, fdmgpu_tune).

No reference counterpart: the reference solves one problem per
LaplCube::solve / LaplRect::solve call.  Problems live in device memory
(fdm_amd.hip.DeviceArray or any object with a `.ptr` device address).
"""
from __future__ import annotations

import ctypes as C

from . import F32_REFINE64, FdmGpuError, Problem, Result, Variant, _check, _lib_handle, _stream_handle

KIND_RECT, KIND_CUBE = 2, 3
GROUP_GRAPH = 1


def _addr(a):
    p = getattr(a, "ptr", a)
    return p.value if isinstance(p, C.c_void_p) else int(p)


def cube_problem(dx, dy, dz, lx, ly, lz, nx, ny, nz, ans, rhs, precision=F32_REFINE64, periodic=0) -> Problem:
    """A LaplCube solve (constructor arguments of LaplCube, src/lapl_cube.h:58-100)."""
    return Problem(KIND_CUBE, precision, periodic, nx, ny, nz, dx, dy, dz, lx, ly, lz, _addr(ans), _addr(rhs))


def rect_problem(dx, dy, lx, ly, nx, ny, ans, rhs, precision=F32_REFINE64, periodic=0) -> Problem:
    """A LaplRect solve (constructor arguments of LaplRect, src/lapl_rect.h:40-63)."""
    return Problem(KIND_RECT, precision, periodic, nx, ny, 1, dx, dy, 0.0, lx, ly, 0.0, _addr(ans), _addr(rhs))


class Group:
    """A fixed set of problems solved together (one launch per stage and solver,
    batches on `nstreams` streams; with graph=True later solves replay a hipGraph)."""

    def __init__(self, problems, nstreams: int = 0, graph: bool = True):
        arr = (Problem * len(problems))(*problems)
        h = C.c_void_p()
        _check(_lib_handle().fdmgpu_group_create(C.byref(h), len(problems), arr, nstreams,
                                                 GROUP_GRAPH if graph else 0), "group_create")
        self._h = h
        self.count = len(problems)

    def solve(self, stream=None):
        _check(_lib_handle().fdmgpu_group_solve(self._h, _stream_handle(stream)), "group_solve")

    def result(self, i: int) -> Result:
        r = Result()
        _check(_lib_handle().fdmgpu_group_result(self._h, i, C.byref(r)), "group_result")
        return r

    def set_refine(self, rel_tol: float = 1e-10, max_iter: int = 4):
        """Refinement settings of the group's FP32+refine solvers (recaptures the graph)."""
        _check(_lib_handle().fdmgpu_group_set_refine(self._h, rel_tol, max_iter), "group_set_refine")

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib_handle().fdmgpu_group_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Request loop: submit() returns a ticket; results via poll()/wait()."""

    def __init__(self, hbm_budget: int, max_batch: int = 64, nstreams: int = 0):
        h = C.c_void_p()
        _check(_lib_handle().fdmgpu_server_create(C.byref(h), int(hbm_budget), max_batch, nstreams), "server_create")
        self._h = h

    def submit(self, problem: Problem) -> int:
        t = C.c_ulonglong()
        _check(_lib_handle().fdmgpu_server_submit(self._h, C.byref(problem), C.byref(t)), "server_submit")
        return t.value

    def set_refine(self, rel_tol: float = 1e-10, max_iter: int = 4):
        _check(_lib_handle().fdmgpu_server_set_refine(self._h, rel_tol, max_iter), "server_set_refine")

    def poll(self, ticket: int) -> Result:
        r = Result()
        _check(_lib_handle().fdmgpu_server_poll(self._h, ticket, C.byref(r)), "server_poll")
        return r

    def wait(self, ticket: int) -> Result:
        r = Result()
        _check(_lib_handle().fdmgpu_server_wait(self._h, ticket, C.byref(r)), "server_wait")
        return r

    def drain(self):
        _check(_lib_handle().fdmgpu_server_drain(self._h), "server_drain")

    def forget(self, ticket: int):
        _check(_lib_handle().fdmgpu_server_forget(self._h, ticket), "server_forget")

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib_handle().fdmgpu_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def tune(problems, reps: int = 5, cache_path: str | None = None):
    """Time every kernel variant of each shape among `problems` (on device copies
    of their right-hand sides; the problems' own arrays are only read) and keep
    the fastest (process-wide; JSON file if given)."""
    arr = (Problem * len(problems))(*problems)
    _check(_lib_handle().fdmgpu_tune(len(problems), arr, reps, cache_path.encode() if cache_path else None), "tune")


__all__ = ["Group", "Server", "tune", "cube_problem", "rect_problem", "KIND_CUBE", "KIND_RECT", "FdmGpuError"]


def tune_set(problem: Problem, variant) -> None:
    """Pins `variant` (plane, tile_x, tile_y, zcr_gmul, fuse) for problems of
    `problem`'s shape in the process-wide table (fdmgpu_tune_set)."""
    v = Variant(*variant)
    _check(_lib_handle().fdmgpu_tune_set(C.byref(problem), C.byref(v)), "tune_set")
