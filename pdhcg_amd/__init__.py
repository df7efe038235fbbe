"""MI355X implementation of the PDHCG solve path behind the pdhcg.h C interface.

The product is the shared library ``libpdhcg_amd.so`` built from ``csrc/``
(HIP kernels for gfx950 plus the C interface).  This package locates and loads
it with ctypes and exposes thin helpers; there is no Python or CPU fallback:
``lib()`` raises if the library is missing.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

from . import capi
from .problem import QPData, build, default_params, solve as _solve

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpdhcg_amd.so")
_lib: Optional[C.CDLL] = None

PRECISION_FP64 = 0
PRECISION_MIXED = 1


def lib() -> C.CDLL:
    """Loads libpdhcg_amd.so (built by ``__graft_entry__.build()``)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`")
        L = C.CDLL(LIB_PATH)
        capi.declare_api(L, "")
        P = C.POINTER
        L.pdhcg_solver_create.restype = C.c_void_p
        L.pdhcg_solver_create.argtypes = [P(capi.qp_problem_t), P(capi.pdhg_parameters_t), C.c_int]
        L.pdhcg_solver_run.restype = C.c_int
        L.pdhcg_solver_run.argtypes = [C.c_void_p, C.c_int]
        L.pdhcg_solver_counts.restype = C.c_int
        L.pdhcg_solver_counts.argtypes = [C.c_void_p, P(C.c_int), P(C.c_int)]
        L.pdhcg_solver_synchronize.restype = C.c_int
        L.pdhcg_solver_synchronize.argtypes = [C.c_void_p]
        L.pdhcg_solver_result.restype = P(capi.pdhcg_result_t)
        L.pdhcg_solver_result.argtypes = [C.c_void_p]
        L.pdhcg_solver_free.restype = None
        L.pdhcg_solver_free.argtypes = [C.c_void_p]
        L.pdhcg_set_precision.restype = None
        L.pdhcg_set_precision.argtypes = [C.c_int]
        L.pdhcg_get_precision.restype = C.c_int
        L.pdhcg_request_cancel.restype = None
        L.pdhcg_request_cancel.argtypes = [C.c_int]
        L.pdhcg_solver_timing.restype = C.c_int
        L.pdhcg_solver_timing.argtypes = [C.c_void_p, C.c_int]
        L.pdhcg_solver_kernel_time.restype = C.c_int
        L.pdhcg_solver_kernel_time.argtypes = [C.c_void_p, C.c_int, P(C.c_double), P(C.c_int)]
        L.pdhcg_set_graph_window.restype = None
        L.pdhcg_set_graph_window.argtypes = [C.c_int]
        L.pdhcg_get_graph_window.restype = C.c_int
        L.pdhcg_solver_graph_stats.restype = C.c_int
        L.pdhcg_solver_graph_stats.argtypes = [C.c_void_p, P(C.c_int), P(C.c_int)]
        L.pdhcg_solve_batch.restype = C.c_int
        L.pdhcg_solve_batch.argtypes = [P(P(capi.qp_problem_t)), C.c_int, P(capi.pdhg_parameters_t),
                                        P(P(capi.pdhcg_result_t))]
        L.pdhcg_project_cones.restype = C.c_int
        L.pdhcg_project_cones.argtypes = [P(capi.cone_blocks_t), capi.c_double_p, capi.c_double_p, C.c_int]
        L.pdhcg_project_cones_seq.restype = C.c_int
        L.pdhcg_project_cones_seq.argtypes = [P(capi.cone_blocks_t), capi.c_double_p, capi.c_double_p, C.c_int,
                                              C.c_int]
        L.pdhcg_cone_dual_residual.restype = C.c_int
        L.pdhcg_cone_dual_residual.argtypes = [P(capi.cone_blocks_t), capi.c_double_p, capi.c_double_p,
                                               capi.c_double_p, C.c_int]
        _lib = L
    return _lib


def solve(p: QPData, precision: int = PRECISION_FP64, **params) -> Optional[capi.Solution]:
    """Solves through ``solve_qp_problem`` of the MI355X library."""
    L = lib()
    L.pdhcg_set_precision(precision)
    return _solve(L, "", p, default_params(L, "", **params))


def set_graph_window(enable: bool) -> None:
    """Graph replay of iteration windows for solvers created afterwards."""
    lib().pdhcg_set_graph_window(1 if enable else 0)


def solve_batch(problems, precision: int = PRECISION_FP64, **params) -> list:
    """Solves the problems together through pdhcg_solve_batch (one device
    workgroup per problem); returns one Solution per problem, or None for a
    problem the batch rejected.  PRECISION_MIXED runs FP32 iterates first and
    evaluates and terminates every problem in FP64."""
    L = lib()
    L.pdhcg_set_precision(precision)
    if len(problems) >= 64:
        # the library's create_qp_problem copies run outside the GIL, so a
        # pool of threads overlaps them across problems
        from concurrent.futures import ThreadPoolExecutor
        with ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 1)) as ex:
            handles = list(ex.map(lambda p: build(L, "", p), problems))
    else:
        handles = [build(L, "", p) for p in problems]
    try:
        count = len(problems)
        ptrs = (C.POINTER(capi.qp_problem_t) * max(count, 1))()
        for i, h in enumerate(handles):
            if h is not None:
                ptrs[i] = h.ptr
        res = (C.POINTER(capi.pdhcg_result_t) * max(count, 1))()
        prm = default_params(L, "", **params)
        if L.pdhcg_solve_batch(ptrs, count, C.byref(prm), res) != 0:
            raise RuntimeError("pdhcg_solve_batch failed")
        out = []
        for i, h in enumerate(handles):
            if h is None or not res[i]:
                out.append(None)
                continue
            n, m = h.contents.num_variables, h.contents.num_constraints
            out.append(capi.result_to_solution(res[i], n, m if m > 0 else 0))
            L.pdhcg_result_free(res[i])
        return out
    finally:
        for h in handles:
            if h is not None:
                h.close()


def project_cones(specs, x, d=None, fixed=None) -> np.ndarray:
    """Device projection of x onto diag(d) K for the blocks in ``specs``
    ((type, start, v_dim[, alpha]) tuples); returns a new array.  ``fixed``
    (per-variable 0/1) marks slots that keep their value: those blocks are
    projected onto their fixed section."""
    L = lib()
    out = capi.carray(x).copy()
    dd = None if d is None else capi.carray(d)
    b, _keep = capi.cone_blocks(specs, fixed)
    if L.pdhcg_project_cones(C.byref(b), capi.as_ptr(dd), capi.as_ptr(out), len(out)) != 0:
        raise RuntimeError("pdhcg_project_cones failed")
    return out


def project_cones_seq(specs, xs, d=None) -> np.ndarray:
    """Device projections of the rows of ``xs`` (count x n) in order through
    one cone set (pdhcg_project_cones_seq); returns a new array."""
    L = lib()
    out = np.ascontiguousarray(np.array(xs, dtype=np.float64, copy=True))
    count, n = out.shape
    dd = None if d is None else capi.carray(d)
    b, _keep = capi.cone_blocks(specs)
    if L.pdhcg_project_cones_seq(C.byref(b), capi.as_ptr(dd), capi.as_ptr(out), n, count) != 0:
        raise RuntimeError("pdhcg_project_cones_seq failed")
    return out


def cone_dual_residual(specs, rc, d=None) -> np.ndarray:
    """Device dual-cone residual (rc - P_{K*}(rc)) d on the cone slots (zero elsewhere)."""
    L = lib()
    r = capi.carray(rc)
    out = np.zeros_like(r)
    dd = None if d is None else capi.carray(d)
    b, _keep = capi.cone_blocks(specs)
    if L.pdhcg_cone_dual_residual(C.byref(b), capi.as_ptr(dd), capi.as_ptr(r), capi.as_ptr(out), len(r)) != 0:
        raise RuntimeError("pdhcg_cone_dual_residual failed")
    return out


class Stepper:
    """Owns a pdhcg_solver_t: create once, run the main loop in chunks."""

    def __init__(self, p: QPData, precision: int = PRECISION_FP64, **params):
        L = lib()
        self.L = L
        self.handle = build(L, "", p)
        if self.handle is None:
            raise ValueError("create_qp_problem rejected the problem")
        self.params = default_params(L, "", **params)
        self.s = L.pdhcg_solver_create(self.handle.ptr, C.byref(self.params), precision)
        if not self.s:
            raise RuntimeError("pdhcg_solver_create failed")

    def run(self, iterations: int) -> int:
        rc = self.L.pdhcg_solver_run(self.s, iterations)
        if rc < 0:
            raise RuntimeError("device iteration failed")
        return rc

    def synchronize(self):
        if self.L.pdhcg_solver_synchronize(self.s) != 0:
            raise RuntimeError("device synchronisation failed")

    def counts(self):
        t, i = C.c_int(), C.c_int()
        self.L.pdhcg_solver_counts(self.s, C.byref(t), C.byref(i))
        return t.value, i.value

    def graph_stats(self):
        """(windows, iterations) replayed from recorded graphs so far."""
        w, i = C.c_int(), C.c_int()
        self.L.pdhcg_solver_graph_stats(self.s, C.byref(w), C.byref(i))
        return w.value, i.value

    def result(self) -> capi.Solution:
        r = self.L.pdhcg_solver_result(self.s)
        if not r:
            raise RuntimeError("result extraction failed")
        h = self.handle.contents
        sol = capi.result_to_solution(r, h.num_variables, h.num_constraints)
        self.L.pdhcg_result_free(r)
        return sol

    def close(self):
        if self.s:
            self.L.pdhcg_solver_free(self.s)
            self.s = None
        if self.handle is not None:
            self.handle.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
