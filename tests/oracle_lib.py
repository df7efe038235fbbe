"""Loader for the CPU oracle (oracle/liboracle.so).

Test infrastructure only: the oracle is the checker the HIP library is
compared with, never part of the product path.
"""
from __future__ import annotations

import ctypes as C
import os
from functools import lru_cache

from pdhcg_amd import capi
from pdhcg_amd.problem import QPData, build, default_params, solve

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_PATH = os.path.join(ROOT, "oracle", "liboracle.so")


@lru_cache(maxsize=1)
def oracle() -> C.CDLL:
    if not os.path.exists(ORACLE_PATH):
        raise RuntimeError(f"{ORACLE_PATH} missing; run __graft_entry__.build()")
    L = capi.declare_api(C.CDLL(ORACLE_PATH), "orc_")
    P = C.POINTER
    L.orc_solver_create.restype = C.c_void_p
    L.orc_solver_create.argtypes = [P(capi.qp_problem_t), P(capi.pdhg_parameters_t)]
    L.orc_solver_run.restype = C.c_int
    L.orc_solver_run.argtypes = [C.c_void_p, C.c_int]
    L.orc_solver_total_count.restype = C.c_int
    L.orc_solver_total_count.argtypes = [C.c_void_p]
    L.orc_solver_inner_count.restype = C.c_int
    L.orc_solver_inner_count.argtypes = [C.c_void_p]
    L.orc_solver_free.restype = None
    L.orc_solver_free.argtypes = [C.c_void_p]
    L.orc_set_threads.restype = None
    L.orc_set_threads.argtypes = [C.c_int]
    L.orc_get_threads.restype = C.c_int
    L.orc_solver_scalings.restype = None
    dp = P(C.c_double)
    L.orc_solver_scalings.argtypes = [C.c_void_p, dp, dp, dp, dp, dp, dp]
    L.orc_csr_transpose.restype = None
    ip = P(C.c_int)
    L.orc_csr_transpose.argtypes = [C.c_int, C.c_int, ip, ip, dp, ip, ip, dp]
    return L


def oracle_params(**kw) -> capi.pdhg_parameters_t:
    return default_params(oracle(), "orc_", **kw)


def oracle_solve(p: QPData, **kw):
    return solve(oracle(), "orc_", p, oracle_params(**kw))


def oracle_build(p: QPData):
    return build(oracle(), "orc_", p)
