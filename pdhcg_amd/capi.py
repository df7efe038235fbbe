"""ctypes mirror of the PDHCG C interface types (include/pdhcg_types.h).

The structures here have the member order and C types of the reference
interface (reference include/pdhcg_types.h:28-300), so any library exporting
that interface can be driven from Python: the MI355X library
libpdhcg_amd.so, and in the tests also the CPU oracle.

Helpers build ``matrix_desc_t`` values from numpy / scipy matrices and keep
the backing arrays alive for as long as the descriptor is used.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Any, List, Optional

import numpy as np

c_int_p = C.POINTER(C.c_int)
c_double_p = C.POINTER(C.c_double)

# termination_reason_t
TERMINATION_REASONS = [
    "UNSPECIFIED", "OPTIMAL", "PRIMAL_INFEASIBLE", "DUAL_INFEASIBLE", "INFEASIBLE_OR_UNBOUNDED",
    "TIME_LIMIT", "ITERATION_LIMIT", "USER_INTERRUPT", "FEAS_POLISH_SUCCESS",
]
OPTIMAL = 1
ITERATION_LIMIT = 6
NORM_TYPE_L2, NORM_TYPE_L_INF = 0, 1
PDHCG_SPARSE_Q, PDHCG_DIAG_Q, PDHCG_LOW_RANK_PLUS_SPARSE_Q, PDHCG_LOW_RANK_Q, PDHCG_NON_Q = range(5)
CONE_ROTATED_SOC, CONE_STANDARD_SOC, CONE_EXPONENTIAL, CONE_POWER, CONE_PSD = range(5)
MATRIX_DENSE, MATRIX_CSR, MATRIX_CSC, MATRIX_COO = range(4)


class CsrComponent(C.Structure):
    _fields_ = [("row_ptr", c_int_p), ("col_ind", c_int_p), ("val", c_double_p)]


class cone_blocks_t(C.Structure):
    _fields_ = [
        ("num_cones", C.c_int), ("start_idx", c_int_p), ("v_dim", c_int_p), ("type", c_int_p),
        ("power_alpha", c_double_p), ("fixed_mask_size", C.c_int), ("is_fixed", C.c_char_p),
    ]


class cone_spec_t(C.Structure):
    _fields_ = [
        ("type", C.c_int), ("start_idx", C.c_int), ("v_dim", C.c_int), ("power_alpha", C.c_double),
        ("is_fixed", C.c_void_p),
    ]


class qp_problem_t(C.Structure):
    _fields_ = [
        ("num_variables", C.c_int), ("num_constraints", C.c_int), ("num_rank_lowrank_obj", C.c_int),
        ("variable_lower_bound", c_double_p), ("variable_upper_bound", c_double_p),
        ("objective_vector", c_double_p), ("objective_constant", C.c_double),
        ("constraint_matrix", C.POINTER(CsrComponent)), ("constraint_matrix_num_nonzeros", C.c_int),
        ("objective_sparse_matrix", C.POINTER(CsrComponent)), ("objective_sparse_matrix_num_nonzeros", C.c_int),
        ("objective_lowrank_matrix", C.POINTER(CsrComponent)), ("objective_lowrank_matrix_num_nonzeros", C.c_int),
        ("objective_lowrank_middle_matrix", C.POINTER(CsrComponent)),
        ("objective_lowrank_middle_matrix_num_nonzeros", C.c_int),
        ("constraint_lower_bound", c_double_p), ("constraint_upper_bound", c_double_p),
        ("affine_cone_offset", c_double_p), ("affine_cones", cone_blocks_t),
        ("num_quadratic_constraints", C.c_int), ("quadratic_constraint_row_indices", c_int_p),
        ("quadratic_constraint_matrices", C.POINTER(C.POINTER(CsrComponent))),
        ("quadratic_constraint_matrix_num_nonzeros", c_int_p),
        ("cones", cone_blocks_t), ("num_original_variables", C.c_int),
        ("primal_start", c_double_p), ("dual_start", c_double_p),
    ]


class restart_parameters_t(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "artificial_restart_threshold", "sufficient_reduction_for_restart", "necessary_reduction_for_restart",
        "k_p", "k_i", "k_d", "i_smooth")]


class termination_criteria_t(C.Structure):
    _fields_ = [
        ("eps_optimal_relative", C.c_double), ("eps_feasible_relative", C.c_double),
        ("eps_feas_polish_relative", C.c_double), ("eps_infeasible", C.c_double),
        ("time_sec_limit", C.c_double), ("iteration_limit", C.c_int),
    ]


class inner_solver_parameters_t(C.Structure):
    _fields_ = [("iteration_limit", C.c_int), ("initial_tolerance", C.c_double), ("min_tolerance", C.c_double)]


class grid_size_t(C.Structure):
    _fields_ = [("row_dims", C.c_int), ("col_dims", C.c_int), ("decided", C.c_bool)]


class pdhg_parameters_t(C.Structure):
    _fields_ = [
        ("curtis_reid_iterations", C.c_int), ("l_inf_ruiz_iterations", C.c_int),
        ("has_pock_chambolle_alpha", C.c_bool), ("pock_chambolle_alpha", C.c_double),
        ("bound_objective_rescaling", C.c_bool), ("use_cone_preserving_scaling", C.c_bool),
        ("verbose", C.c_int), ("termination_evaluation_frequency", C.c_int),
        ("sv_max_iter", C.c_int), ("sv_tol", C.c_double),
        ("termination_criteria", termination_criteria_t), ("restart_params", restart_parameters_t),
        ("reflection_coefficient", C.c_double), ("feasibility_polishing", C.c_bool),
        ("optimality_norm", C.c_int), ("inner_solver_parameters", inner_solver_parameters_t),
        ("presolve", C.c_bool), ("diag_jacobi_precond", C.c_bool), ("default_cone_type", C.c_int),
        ("partition_method", C.c_int), ("permute_method", C.c_int), ("grid_size", grid_size_t),
        ("permute_block_size", C.c_int),
    ]


class pdhcg_result_t(C.Structure):
    _fields_ = [
        ("num_variables", C.c_int), ("num_constraints", C.c_int), ("num_nonzeros", C.c_int),
        ("num_reduced_variables", C.c_int), ("num_reduced_constraints", C.c_int), ("num_reduced_nonzeros", C.c_int),
        ("primal_solution", c_double_p), ("dual_solution", c_double_p), ("reduced_cost", c_double_p),
        ("total_count", C.c_int), ("total_inner_count", C.c_int),
        ("rescaling_time_sec", C.c_double), ("cumulative_time_sec", C.c_double),
        ("absolute_primal_residual", C.c_double), ("relative_primal_residual", C.c_double),
        ("absolute_dual_residual", C.c_double), ("relative_dual_residual", C.c_double),
        ("primal_objective_value", C.c_double), ("dual_objective_value", C.c_double),
        ("objective_gap", C.c_double), ("relative_objective_gap", C.c_double),
        ("max_primal_ray_infeasibility", C.c_double), ("max_dual_ray_infeasibility", C.c_double),
        ("primal_ray_linear_objective", C.c_double), ("dual_ray_objective", C.c_double),
        ("termination_reason", C.c_int), ("feasibility_polishing_time", C.c_double),
        ("feasibility_iteration", C.c_int), ("presolve_time", C.c_double), ("presolve_status", C.c_int),
    ]


class _Dense(C.Structure):
    _fields_ = [("A", c_double_p)]


class _Csr(C.Structure):
    _fields_ = [("nnz", C.c_int), ("row_ptr", c_int_p), ("col_ind", c_int_p), ("vals", c_double_p)]


class _Csc(C.Structure):
    _fields_ = [("nnz", C.c_int), ("col_ptr", c_int_p), ("row_ind", c_int_p), ("vals", c_double_p)]


class _Coo(C.Structure):
    _fields_ = [("nnz", C.c_int), ("row_ind", c_int_p), ("col_ind", c_int_p), ("vals", c_double_p)]


class _MatrixData(C.Union):
    _fields_ = [("dense", _Dense), ("csr", _Csr), ("csc", _Csc), ("coo", _Coo)]


class matrix_desc_t(C.Structure):
    _fields_ = [("m", C.c_int), ("n", C.c_int), ("fmt", C.c_int), ("zero_tolerance", C.c_double),
                ("data", _MatrixData)]


def _dp(a: np.ndarray):
    return a.ctypes.data_as(c_double_p)


def _ip(a: np.ndarray):
    return a.ctypes.data_as(c_int_p)


@dataclass
class Desc:
    """A matrix_desc_t plus the arrays it points into."""
    desc: matrix_desc_t
    keep: List[Any] = field(default_factory=list)

    @property
    def ptr(self):
        return C.byref(self.desc)


def desc_diag(dvec) -> Desc:
    """CSR descriptor of diag(dvec) built directly with numpy, with the
    entries scipy's diags(dvec).tocsr() keeps (the nonzero ones)."""
    v = np.asarray(dvec, dtype=np.float64)
    keepm = v != 0.0
    rp = np.zeros(v.size + 1, dtype=np.int32)
    np.cumsum(keepm, out=rp[1:])
    ci = np.ascontiguousarray(np.nonzero(keepm)[0], dtype=np.int32)
    vals = np.ascontiguousarray(v[keepm])
    d = matrix_desc_t()
    d.m = d.n = int(v.size)
    d.fmt = MATRIX_CSR
    d.data.csr.nnz = int(vals.size)
    d.data.csr.row_ptr, d.data.csr.col_ind, d.data.csr.vals = _ip(rp), _ip(ci), _dp(vals)
    d.zero_tolerance = 0.0
    return Desc(d, [rp, ci, vals])


def desc_from(mat, fmt: str = "csr", zero_tolerance: float = 0.0) -> Optional[Desc]:
    """Descriptor from a numpy array (dense) or a scipy sparse matrix.

    fmt: 'dense', 'csr', 'csc' or 'coo' selects the input format handed to the
    library; the library converts to CSR itself."""
    if mat is None:
        return None
    import scipy.sparse as sp
    d = matrix_desc_t()
    keep: List[Any] = []
    if fmt == "dense":
        a = np.ascontiguousarray(mat.toarray() if sp.issparse(mat) else np.asarray(mat), dtype=np.float64)
        if a.ndim == 1:
            a = a.reshape(1, -1)
        d.m, d.n = a.shape
        d.fmt = MATRIX_DENSE
        d.data.dense.A = _dp(a)
        keep.append(a)
    else:
        s = sp.csr_matrix(mat) if not sp.issparse(mat) else mat
        d.m, d.n = s.shape
        if fmt == "csr":
            s = sp.csr_matrix(s)
            rp = np.ascontiguousarray(s.indptr, dtype=np.int32)
            ci = np.ascontiguousarray(s.indices, dtype=np.int32)
            v = np.ascontiguousarray(s.data, dtype=np.float64)
            d.fmt = MATRIX_CSR
            d.data.csr.nnz = int(v.size)
            d.data.csr.row_ptr, d.data.csr.col_ind, d.data.csr.vals = _ip(rp), _ip(ci), _dp(v)
            keep += [rp, ci, v]
        elif fmt == "csc":
            s = sp.csc_matrix(s)
            cp = np.ascontiguousarray(s.indptr, dtype=np.int32)
            ri = np.ascontiguousarray(s.indices, dtype=np.int32)
            v = np.ascontiguousarray(s.data, dtype=np.float64)
            d.fmt = MATRIX_CSC
            d.data.csc.nnz = int(v.size)
            d.data.csc.col_ptr, d.data.csc.row_ind, d.data.csc.vals = _ip(cp), _ip(ri), _dp(v)
            keep += [cp, ri, v]
        elif fmt == "coo":
            s = sp.coo_matrix(s)
            ri = np.ascontiguousarray(s.row, dtype=np.int32)
            ci = np.ascontiguousarray(s.col, dtype=np.int32)
            v = np.ascontiguousarray(s.data, dtype=np.float64)
            d.fmt = MATRIX_COO
            d.data.coo.nnz = int(v.size)
            d.data.coo.row_ind, d.data.coo.col_ind, d.data.coo.vals = _ip(ri), _ip(ci), _dp(v)
            keep += [ri, ci, v]
        else:
            raise ValueError(f"unknown matrix format {fmt!r}")
    d.zero_tolerance = zero_tolerance
    return Desc(d, keep)


def declare_api(lib: C.CDLL, prefix: str = "") -> C.CDLL:
    """Declares argtypes/restypes of the pdhcg.h entry points on ``lib``.

    ``prefix`` selects a renamed copy of the same interface (the oracle
    exports it as orc_*)."""
    P = C.POINTER
    f = getattr(lib, prefix + "create_qp_problem")
    f.restype = P(qp_problem_t)
    f.argtypes = [c_double_p, P(matrix_desc_t), P(matrix_desc_t), P(matrix_desc_t), P(matrix_desc_t),
                  c_double_p, c_double_p, c_double_p, c_double_p, c_double_p, C.c_int, P(cone_spec_t),
                  P(matrix_desc_t), c_double_p, C.c_int, P(cone_spec_t)]
    f = getattr(lib, prefix + "set_start_values")
    f.restype = None
    f.argtypes = [P(qp_problem_t), c_double_p, c_double_p]
    f = getattr(lib, prefix + "set_cone_fixed")
    f.restype = C.c_int
    f.argtypes = [P(qp_problem_t), C.c_int, C.c_int, C.c_double]
    f = getattr(lib, prefix + "solve_qp_problem")
    f.restype = P(pdhcg_result_t)
    f.argtypes = [P(qp_problem_t), P(pdhg_parameters_t)]
    f = getattr(lib, prefix + "set_default_parameters")
    f.restype = None
    f.argtypes = [P(pdhg_parameters_t)]
    f = getattr(lib, prefix + ("result_free" if prefix else "pdhcg_result_free"))
    f.restype = None
    f.argtypes = [P(pdhcg_result_t)]
    f = getattr(lib, prefix + "qp_problem_free")
    f.restype = None
    f.argtypes = [P(qp_problem_t)]
    return lib


def as_ptr(a: Optional[np.ndarray]):
    if a is None:
        return None
    return _dp(a)


def carray(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


def cone_specs(specs) -> Any:
    """Builds a cone_spec_t array from (type, start, v_dim[, alpha[, is_fixed]])
    tuples; is_fixed is a per-slot 0/1 sequence (or None)."""
    arr = (cone_spec_t * max(1, len(specs)))()
    arr.keep = []
    for i, s in enumerate(specs):
        t, start, vd = s[0], s[1], s[2]
        arr[i].type, arr[i].start_idx, arr[i].v_dim = int(t), int(start), int(vd)
        arr[i].power_alpha = float(s[3]) if len(s) > 3 else 0.0
        arr[i].is_fixed = None
        if len(s) > 4 and s[4] is not None:
            mask = (C.c_char * len(s[4]))(*[bytes([1 if v else 0]) for v in s[4]])
            arr.keep.append(mask)
            arr[i].is_fixed = C.cast(mask, C.c_void_p)
    return arr


def cone_blocks(specs, fixed=None):
    """Builds a cone_blocks_t from (type, start, v_dim[, alpha]) tuples;
    fixed, if given, is the per-variable 0/1 fixed-slot mask.

    Returns (struct, keep) where keep holds the arrays the struct points to."""
    k = len(specs)
    start = (C.c_int * max(1, k))(*[int(s[1]) for s in specs])
    vdim = (C.c_int * max(1, k))(*[int(s[2]) for s in specs])
    typ = (C.c_int * max(1, k))(*[int(s[0]) for s in specs])
    alpha = (C.c_double * max(1, k))(*[float(s[3]) if len(s) > 3 else 0.0 for s in specs])
    b = cone_blocks_t()
    b.num_cones = k
    b.start_idx = C.cast(start, c_int_p)
    b.v_dim = C.cast(vdim, c_int_p)
    b.type = C.cast(typ, c_int_p)
    b.power_alpha = C.cast(alpha, c_double_p)
    b.fixed_mask_size = 0
    b.is_fixed = None
    mask = None
    if fixed is not None:
        mask = C.create_string_buffer(bytes([1 if v else 0 for v in fixed]), len(fixed))
        b.fixed_mask_size = len(fixed)
        b.is_fixed = C.cast(mask, C.c_char_p)
    return b, (start, vdim, typ, alpha, mask)


@dataclass
class Solution:
    status: str
    x: np.ndarray
    y: np.ndarray
    reduced_cost: np.ndarray
    pobj: float
    dobj: float
    rel_primal: float
    rel_dual: float
    rel_gap: float
    abs_primal: float
    abs_dual: float
    iterations: int
    inner_iterations: int
    solve_time: float
    raw: dict


def result_to_solution(res_ptr, n: int, m: int) -> Solution:
    r = res_ptr.contents
    x = np.ctypeslib.as_array(r.primal_solution, shape=(n,)).copy() if n > 0 else np.zeros(0)
    y = np.ctypeslib.as_array(r.dual_solution, shape=(m,)).copy() if m > 0 else np.zeros(0)
    rc = np.ctypeslib.as_array(r.reduced_cost, shape=(n,)).copy() if n > 0 else np.zeros(0)
    raw = {name: getattr(r, name) for name, t in r._fields_ if t in (C.c_int, C.c_double)}
    return Solution(
        status=TERMINATION_REASONS[r.termination_reason], x=x, y=y, reduced_cost=rc,
        pobj=r.primal_objective_value, dobj=r.dual_objective_value,
        rel_primal=r.relative_primal_residual, rel_dual=r.relative_dual_residual,
        rel_gap=r.relative_objective_gap, abs_primal=r.absolute_primal_residual,
        abs_dual=r.absolute_dual_residual, iterations=r.total_count,
        inner_iterations=r.total_inner_count, solve_time=r.cumulative_time_sec, raw=raw)
