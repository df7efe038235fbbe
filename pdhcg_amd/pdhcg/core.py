"""The three functions of the reference's compiled Python module
(python_bindings/_core_bindings.cpp:253-310, 631-905, 1092-1120) over
libpdhcg_amd.so: get_default_params, solve_once and read_problem_file.

solve_once keeps the reference's argument list, parameter dictionary keys,
errors and result dictionary.  The solve runs in the library with the GIL
released (ctypes releases it for the foreign call); Ctrl-C during a solve
asks the library to stop at its next evaluation (USER_INTERRUPT) instead of
killing the interpreter, as the reference's SIGINT handler does.
"""
from __future__ import annotations

import ctypes as C
import signal
import threading
from typing import Any, Optional

import numpy as np
import scipy.sparse as sp

from .. import capi, lib as _lib
from ..problem import QPData, build
from .cones import ConeSpec, block_lengths

_STATUS = {0: ("UNSPECIFIED", -1), 1: ("OPTIMAL", 0), 2: ("PRIMAL_INFEASIBLE", 1), 3: ("DUAL_INFEASIBLE", 2),
           4: ("INFEASIBLE_OR_UNBOUNDED", 5), 5: ("TIME_LIMIT", 3), 6: ("ITERATION_LIMIT", 4),
           7: ("USER_INTERRUPT", -1), 8: ("FEAS_POLISH_SUCCESS", -1)}


def _param_fields(p: capi.pdhg_parameters_t):
    """(key, struct, field) for every dictionary key of the interface."""
    tc, rp, ip = p.termination_criteria, p.restart_params, p.inner_solver_parameters
    return [("verbose", p, "verbose"), ("termination_evaluation_frequency", p, "termination_evaluation_frequency"),
            ("eps_optimal_relative", tc, "eps_optimal_relative"),
            ("eps_feasible_relative", tc, "eps_feasible_relative"), ("time_sec_limit", tc, "time_sec_limit"),
            ("iteration_limit", tc, "iteration_limit"), ("curtis_reid_iterations", p, "curtis_reid_iterations"),
            ("l_inf_ruiz_iterations", p, "l_inf_ruiz_iterations"),
            ("has_pock_chambolle_alpha", p, "has_pock_chambolle_alpha"),
            ("pock_chambolle_alpha", p, "pock_chambolle_alpha"),
            ("bound_objective_rescaling", p, "bound_objective_rescaling"),
            ("use_cone_preserving_scaling", p, "use_cone_preserving_scaling"),
            ("artificial_restart_threshold", rp, "artificial_restart_threshold"),
            ("sufficient_reduction_for_restart", rp, "sufficient_reduction_for_restart"),
            ("necessary_reduction_for_restart", rp, "necessary_reduction_for_restart"), ("k_p", rp, "k_p"),
            ("reflection_coefficient", p, "reflection_coefficient"),
            ("feasibility_polishing", p, "feasibility_polishing"),
            ("eps_feas_polish_relative", tc, "eps_feas_polish_relative"), ("presolve", p, "presolve"),
            ("sv_max_iter", p, "sv_max_iter"), ("sv_tol", p, "sv_tol"),
            ("inner_iter_limit", ip, "iteration_limit"), ("inner_init_tol", ip, "initial_tolerance"),
            ("inner_min_tol", ip, "min_tolerance"), ("diag_jacobi_precond", p, "diag_jacobi_precond")]


def _defaults() -> capi.pdhg_parameters_t:
    p = capi.pdhg_parameters_t()
    _lib().set_default_parameters(C.byref(p))
    return p


def get_default_params() -> dict:
    """The solver's default parameters as a dictionary (reference keys)."""
    p = _defaults()
    d = {k: getattr(s, f) for k, s, f in _param_fields(p)}
    d["optimality_norm"] = "linf" if p.optimality_norm == 1 else "l2"
    return d


def params_struct(params: Optional[dict]) -> capi.pdhg_parameters_t:
    """Defaults overridden by the dictionary's known keys (others ignored)."""
    p = _defaults()
    if params is None:
        return p
    for k, s, f in _param_fields(p):
        if k in params:
            t = dict(s._fields_)[f]
            setattr(s, f, bool(params[k]) if t is C.c_bool else (int(params[k]) if t is C.c_int
                                                                   else float(params[k])))
    if "optimality_norm" in params:
        v = params["optimality_norm"]
        if not isinstance(v, str):
            raise ValueError("optimality_norm must be a string ('l2'/'linf')")
        if v.lower() not in ("l2", "linf"):
            raise ValueError(f"unknown optimality_norm '{v}' (expected 'l2' or 'linf')")
        p.optimality_norm = 1 if v.lower() == "linf" else 0
    return p


def _shape_check(mat, name):
    if not hasattr(mat, "shape") or len(mat.shape) != 2:
        raise ValueError(f"matrix {name} must be numpy.ndarray or scipy.sparse matrix (2D)")


def _vector(v, name, count):
    if v is None:
        return None
    a = np.ascontiguousarray(np.asarray(v, dtype=np.float64).ravel())
    if a.size != count:
        raise ValueError(f"{name} length {a.size} != {count}")
    return a


def _cone_tuples(spec: Optional[ConeSpec], ambient: int, affine: bool):
    if spec is None:
        return []
    if not isinstance(spec, ConeSpec):
        raise ValueError("cones must be a pdhcg.ConeSpec (columnar cone metadata)")
    spec.validate_ambient(ambient, allow_fixed=not affine, require_cover=affine)
    lengths = block_lengths(spec.types, spec.v_dims)
    out = []
    for (t, s, v, a), ln in zip(spec.as_tuples(), lengths):
        mask = None
        if spec.fixed_mask is not None:
            mask = spec.fixed_mask[s:s + int(ln)]
            if not mask.any():
                mask = None
        out.append((t, s, v, a, mask))
    return out


def _with_zero_tolerance(mat, tol):
    if mat is None or tol <= 0.0:
        return mat
    m = sp.csr_matrix(mat, dtype=np.float64, copy=True)
    m.data[np.abs(m.data) <= tol] = 0.0
    m.eliminate_zeros()
    return m


class _Interrupt:
    """Routes Ctrl-C to pdhcg_request_cancel while a solve runs."""

    def __init__(self, L):
        self.L, self.old = L, None

    def __enter__(self):
        L = self.L
        L.pdhcg_request_cancel(0)
        if threading.current_thread() is threading.main_thread():
            self.old = signal.signal(signal.SIGINT, lambda *_: L.pdhcg_request_cancel(1))
        return self

    def __exit__(self, *exc):
        if self.old is not None:
            signal.signal(signal.SIGINT, self.old)
        return False


def solve_once(Q, R, A, objective_vector, objective_constant=None, variable_lower_bound=None,
               variable_upper_bound=None, constraint_lower_bound=None, constraint_upper_bound=None,
               zero_tolerance: float = 0.0, params: Optional[dict] = None, primal_start=None, dual_start=None,
               D=None, cones: Optional[ConeSpec] = None, affine_F=None, affine_g=None,
               affine_cones: Optional[ConeSpec] = None) -> dict:
    """Builds the problem, solves it once on the device and returns the
    result dictionary (X, Pi, objectives, status, counters, residuals, rays)."""
    for mat, name in ((A, "A"), (Q, "Q"), (R, "R"), (affine_F, "affine_F")):
        if mat is not None:
            _shape_check(mat, name)
    n = 0
    for mat in (A, Q, R, affine_F):
        if mat is not None and mat.shape[1] > 0:
            n = int(mat.shape[1])
            break
    if n == 0 and objective_vector is not None:
        n = int(np.asarray(objective_vector).size)
    m = int(A.shape[0]) if A is not None else 0
    if affine_F is not None and affine_F.shape[1] != n:
        raise ValueError("affine_F column count must match the number of variables")
    if affine_F is None and (affine_g is not None or affine_cones is not None):
        raise ValueError("affine_F is required when affine_g or affine_cones is provided")
    if D is not None:
        if R is None:
            raise ValueError("D was provided but R is None; D is only meaningful with a low-rank R.")
        rank = int(R.shape[0])
        dd = np.asarray(D) if not sp.issparse(D) else D
        if not sp.issparse(dd) and dd.ndim == 1:
            if dd.size != rank:
                raise ValueError(f"D (diag) length {dd.size} must equal rank {rank}")
        elif dd.shape != (rank, rank):
            raise ValueError(f"D shape {dd.shape} must be ({rank}, {rank})")
    p_aff = int(affine_F.shape[0]) if affine_F is not None else 0
    acones = _cone_tuples(affine_cones, p_aff, True)
    if affine_F is not None and not acones:
        raise ValueError("affine_cones must describe every row of affine_F")
    data = QPData(
        c=_vector(objective_vector, "objective_vector", n) if objective_vector is not None else np.zeros(n),
        A=_with_zero_tolerance(A, zero_tolerance), Q=_with_zero_tolerance(Q, zero_tolerance),
        R=_with_zero_tolerance(R, zero_tolerance), D=D,
        con_lb=_vector(constraint_lower_bound, "constraint_lower_bound", m),
        con_ub=_vector(constraint_upper_bound, "constraint_upper_bound", m),
        var_lb=_vector(variable_lower_bound, "variable_lower_bound", n),
        var_ub=_vector(variable_upper_bound, "variable_upper_bound", n),
        objective_constant=float(objective_constant) if objective_constant is not None else 0.0,
        cones=_cone_tuples(cones, n, False), F=_with_zero_tolerance(affine_F, zero_tolerance),
        g=_vector(affine_g, "affine_g", p_aff), affine_cones=acones,
        primal_start=_vector(primal_start, "primal_start", n),
        dual_start=_vector(dual_start, "dual_start", m + p_aff))
    fmt = {}
    for key, mat in (("A", A), ("Q", Q), ("R", R), ("F", affine_F)):
        if mat is not None and not sp.issparse(mat):
            fmt[key] = "dense"
    data.fmt = fmt
    L = _lib()
    prm = params_struct(params)
    h = build(L, "", data)
    if h is None:
        raise RuntimeError("create_qp_problem failed.")
    try:
        with _Interrupt(L):
            res = L.solve_qp_problem(h.ptr, C.byref(prm))
    finally:
        h.close()
    if not res:
        raise RuntimeError("solve_qp_problem returned NULL.")
    try:
        r = res.contents
        x = np.ctypeslib.as_array(r.primal_solution, shape=(r.num_variables,)).copy() if r.num_variables else \
            np.zeros(0)
        y = np.ctypeslib.as_array(r.dual_solution, shape=(r.num_constraints,)).copy() if r.num_constraints else \
            np.zeros(0)
        status, code = _STATUS.get(int(r.termination_reason), ("UNKNOWN", -1))
        return {"X": x, "Pi": y, "PrimalObj": r.primal_objective_value, "DualObj": r.dual_objective_value,
                "ObjectiveGap": r.objective_gap, "RelativeObjectiveGap": r.relative_objective_gap,
                "Status": status, "StatusCode": code, "Iterations": r.total_count,
                "RescalingTimeSec": r.rescaling_time_sec, "RuntimeSec": r.cumulative_time_sec,
                "RelativePrimalResidual": r.relative_primal_residual,
                "RelativeDualResidual": r.relative_dual_residual,
                "MaxPrimalRayInfeas": r.max_primal_ray_infeasibility,
                "MaxDualRayInfeas": r.max_dual_ray_infeasibility,
                "PrimalRayLinObj": r.primal_ray_linear_objective, "DualRayObj": r.dual_ray_objective}
    finally:
        L.pdhcg_result_free(res)


def _csr_payload(comp, rows: int, cols: int, nnz: int):
    if not comp or rows <= 0:
        return None
    c = comp.contents
    indptr = np.ctypeslib.as_array(c.row_ptr, shape=(rows + 1,)).astype(np.int64)
    if nnz <= 0:
        return {"data": np.zeros(0), "indices": np.zeros(0, np.int32), "indptr": indptr, "shape": (rows, cols)}
    return {"data": np.ctypeslib.as_array(c.val, shape=(nnz,)).copy(),
            "indices": np.ctypeslib.as_array(c.col_ind, shape=(nnz,)).astype(np.int32),
            "indptr": indptr, "shape": (rows, cols)}


def read_problem_file(path: str) -> dict:
    """Reads an .mps / .qps / .cbf file (optionally .gz) with the library's
    parsers and returns its data as arrays (reference
    python_bindings/_core_bindings.cpp:967-1090): c, Q, the scalar rows A with
    their bounds (less the row constants), variable bounds, objective
    constant, variable cones, and the affine cone rows F, g with their cones
    (starts renumbered over F's rows); quadratic constraints are lifted into
    rotated second-order cones first."""
    L = _lib()
    stem = path[:-3] if path.endswith(".gz") else path
    fn = L.read_cbf_file if stem.endswith(".cbf") else L.read_mps_file
    fn.restype = C.POINTER(capi.qp_problem_t)
    fn.argtypes = [C.c_char_p]
    ptr = fn(path.encode())
    if not ptr:
        raise RuntimeError(f"failed to read problem file: {path}")
    if ptr.contents.num_quadratic_constraints > 0:
        L.qcqp_to_socp_qp.restype = C.POINTER(capi.qp_problem_t)
        L.qcqp_to_socp_qp.argtypes = [C.POINTER(capi.qp_problem_t), C.c_int]
        lifted = L.qcqp_to_socp_qp(ptr, capi.CONE_ROTATED_SOC)
        L.qp_problem_free(ptr)
        if not lifted:
            raise RuntimeError(f"QCQP -> SOCP transform failed for: {path}")
        ptr = lifted
    try:
        q = ptr.contents
        n, mt = q.num_variables, q.num_constraints
        arr = lambda p_, k: np.ctypeslib.as_array(p_, shape=(k,)).copy() if (p_ and k > 0) else np.zeros(k)

        def columnar(b, starts=None):
            k = b.num_cones
            types = np.array([b.type[i] for i in range(k)], np.int32)
            vd = np.array([b.v_dim[i] for i in range(k)], np.int32)
            return types, vd, {"types": types,
                               "starts": np.array(starts if starts is not None else
                                                  [b.start_idx[i] for i in range(k)], np.int32),
                               "v_dims": vd,
                               "power_alphas": np.array([b.power_alpha[i] if b.power_alpha else 0.0
                                                         for i in range(k)])}

        is_cone_row = np.zeros(mt, bool)
        affine_rows = []
        ac = q.affine_cones
        if ac.num_cones > 0:
            t_, v_, _ = columnar(ac)
            for st, ln in zip([ac.start_idx[i] for i in range(ac.num_cones)], block_lengths(t_, v_)):
                is_cone_row[st:st + int(ln)] = True
                affine_rows.extend(range(st, st + int(ln)))
        scalar_rows = np.nonzero(~is_cone_row)[0]
        full = _csr_payload(q.constraint_matrix, mt, n, q.constraint_matrix_num_nonzeros)
        M = sp.csr_matrix((full["data"], full["indices"], full["indptr"]), shape=(mt, n)) if full else \
            sp.csr_matrix((mt, n))
        const = arr(q.affine_cone_offset, mt) if q.affine_cone_offset else np.zeros(mt)
        cl, cu = arr(q.constraint_lower_bound, mt), arr(q.constraint_upper_bound, mt)
        pay = lambda S: {"data": S.data, "indices": S.indices, "indptr": S.indptr, "shape": S.shape}
        out = {"c": arr(q.objective_vector, n), "obj_const": q.objective_constant,
               "Q": _csr_payload(q.objective_sparse_matrix, n, n, q.objective_sparse_matrix_num_nonzeros),
               "A": pay(M[scalar_rows].tocsr()),
               "constr_lb": cl[scalar_rows] - const[scalar_rows], "constr_ub": cu[scalar_rows] - const[scalar_rows],
               "var_lb": arr(q.variable_lower_bound, n), "var_ub": arr(q.variable_upper_bound, n),
               "cones": None, "affine_F": None, "affine_g": None, "affine_cones": None}
        vc = q.cones
        if vc.num_cones > 0:
            _, _, col = columnar(vc)
            addr = C.c_void_p.from_buffer(vc, capi.cone_blocks_t.is_fixed.offset).value
            col["fixed_mask"] = np.array((C.c_uint8 * n).from_address(addr), dtype=np.uint8) if addr else None
            out["cones"] = col
        if affine_rows:
            rows = np.array(affine_rows)
            out["affine_F"] = pay(M[rows].tocsr())
            out["affine_g"] = const[rows]
            t_, v_, _ = columnar(ac)
            starts = np.concatenate([[0], np.cumsum(block_lengths(t_, v_))[:-1]]).astype(np.int32)
            out["affine_cones"] = columnar(ac, starts)[2]
        if q.primal_start:
            out["primal_start"] = arr(q.primal_start, n)
        return out
    finally:
        L.qp_problem_free(ptr)
