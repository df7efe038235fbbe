"""Problem data in Python, handed to any library exporting the pdhcg.h interface.

``QPData`` holds the inputs of ``create_qp_problem`` (reference
include/pdhcg.h:31-46) as numpy / scipy objects.  ``build`` calls the
library's create function and returns an owning handle; ``solve`` runs the
library's ``solve_qp_problem`` and converts the result.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Any, List, Optional, Sequence

import numpy as np

from . import capi


@dataclass
class QPData:
    c: np.ndarray
    A: Any = None                 # m x n (scipy sparse or dense ndarray)
    con_lb: Optional[np.ndarray] = None
    con_ub: Optional[np.ndarray] = None
    var_lb: Optional[np.ndarray] = None
    var_ub: Optional[np.ndarray] = None
    Q: Any = None                 # n x n
    R: Any = None                 # k x n
    D: Any = None                 # k x k, or a length-k vector for a diagonal D
    objective_constant: float = 0.0
    cones: Sequence[tuple] = ()   # (type, start, v_dim[, alpha])
    F: Any = None                 # affine cone rows
    g: Optional[np.ndarray] = None
    affine_cones: Sequence[tuple] = ()
    fmt: dict = field(default_factory=dict)  # per-matrix input format, default 'csr'
    primal_start: Optional[np.ndarray] = None
    dual_start: Optional[np.ndarray] = None
    cone_fixed: Sequence[tuple] = ()  # (cone index, slot, value) applied with set_cone_fixed
    # (row, Q) pairs: row `row` of A also carries x'Qx (Q n x n, diagonal for
    # the SOC lifting); attached to the assembled problem's quadratic-constraint
    # fields, which create_qp_problem has no arguments for
    quadratic_constraints: Sequence[tuple] = ()

    @property
    def n(self) -> int:
        return int(np.asarray(self.c).size)


class Handle:
    """Owns a qp_problem_t* created by ``lib``; frees it on close."""

    def __init__(self, lib, prefix: str, ptr, keep: List[Any]):
        self.lib, self.prefix, self.ptr, self.keep = lib, prefix, ptr, keep

    @property
    def contents(self):
        return self.ptr.contents

    def close(self):
        if self.ptr:
            getattr(self.lib, self.prefix + "qp_problem_free")(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _desc(mat, fmt, keep):
    if mat is None:
        return None
    d = capi.desc_from(mat, fmt)
    keep.append(d)
    return d.ptr


def build(lib, prefix: str, p: QPData) -> Optional[Handle]:
    keep: List[Any] = []
    arr = lambda v: None if v is None else capi.carray(v)
    c = capi.carray(p.c)
    keep.append(c)
    vecs = {}
    for name in ("con_lb", "con_ub", "var_lb", "var_ub", "g"):
        v = arr(getattr(p, name))
        vecs[name] = v
        keep.append(v)
    D = p.D
    d_desc = None
    if D is not None and np.ndim(D) == 1:
        if p.fmt.get("D", "csr") == "csr":
            d_desc = capi.desc_diag(D)  # diag(D) without scipy's diags (many small problems)
            keep.append(d_desc)
        else:
            import scipy.sparse as sp
            D = sp.diags(np.asarray(D, dtype=np.float64)).tocsr()
    const = C.c_double(p.objective_constant)
    cones = capi.cone_specs(p.cones) if p.cones else None
    acones = capi.cone_specs(p.affine_cones) if p.affine_cones else None
    keep += [const, cones, acones]
    f = p.fmt
    ptr = getattr(lib, prefix + "create_qp_problem")(
        capi.as_ptr(c), _desc(p.Q, f.get("Q", "csr"), keep), _desc(p.R, f.get("R", "csr"), keep),
        d_desc.ptr if d_desc is not None else _desc(D, f.get("D", "csr"), keep), _desc(p.A, f.get("A", "csr"), keep),
        capi.as_ptr(vecs["con_lb"]), capi.as_ptr(vecs["con_ub"]),
        capi.as_ptr(vecs["var_lb"]), capi.as_ptr(vecs["var_ub"]), C.byref(const),
        len(p.cones), cones, _desc(p.F, f.get("F", "csr"), keep), capi.as_ptr(vecs["g"]),
        len(p.affine_cones), acones)
    if not ptr:
        return None
    h = Handle(lib, prefix, ptr, [])
    if p.quadratic_constraints:
        attach_quadratic_constraints(ptr, p.quadratic_constraints, p.n)
    for cone, slot, value in p.cone_fixed:
        if getattr(lib, prefix + "set_cone_fixed")(ptr, int(cone), int(slot), float(value)) != 0:
            h.close()
            return None
    if p.primal_start is not None or p.dual_start is not None:
        ps, ds = arr(p.primal_start), arr(p.dual_start)
        getattr(lib, prefix + "set_start_values")(ptr, capi.as_ptr(ps), capi.as_ptr(ds))
    return h


_libc = C.CDLL(None)
_libc.malloc.restype = C.c_void_p
_libc.malloc.argtypes = [C.c_size_t]


def _c_copy(arr, ctype):
    """A malloc'd copy of arr (the problem's free() releases it)."""
    a = np.ascontiguousarray(arr, dtype=np.int32 if ctype is C.c_int else np.float64)
    ptr = _libc.malloc(max(a.nbytes, 1))
    if not ptr:
        raise MemoryError("malloc failed")
    C.memmove(ptr, a.ctypes.data, a.nbytes)
    return C.cast(ptr, C.POINTER(ctype))


def attach_quadratic_constraints(ptr, qcs, n: int) -> None:
    """Sets the quadratic-constraint fields of an assembled qp_problem_t
    (reference include/pdhcg_types.h:123-126): for each (row, Q), the row index
    and Q as an n x n CSR in malloc'd arrays owned by the problem."""
    import scipy.sparse as sp
    K = len(qcs)
    prob = ptr.contents
    rows = _c_copy([int(r) for r, _ in qcs], C.c_int)
    mats = C.cast(_libc.malloc(K * C.sizeof(C.c_void_p)), C.POINTER(C.POINTER(capi.CsrComponent)))
    nnzs = []
    for i, (_, Q) in enumerate(qcs):
        Qc = sp.csr_matrix(Q, shape=(n, n))
        Qc.sort_indices()
        comp = C.cast(_libc.malloc(C.sizeof(capi.CsrComponent)), C.POINTER(capi.CsrComponent))
        comp.contents.row_ptr = _c_copy(Qc.indptr, C.c_int)
        comp.contents.col_ind = _c_copy(Qc.indices, C.c_int)
        comp.contents.val = _c_copy(Qc.data, C.c_double)
        mats[i] = comp
        nnzs.append(Qc.nnz)
    prob.num_quadratic_constraints = K
    prob.quadratic_constraint_row_indices = rows
    prob.quadratic_constraint_matrices = mats
    prob.quadratic_constraint_matrix_num_nonzeros = _c_copy(nnzs, C.c_int)


def default_params(lib, prefix: str = "", **overrides) -> capi.pdhg_parameters_t:
    prm = capi.pdhg_parameters_t()
    getattr(lib, prefix + "set_default_parameters")(C.byref(prm))
    prm.verbose = 0
    for key, val in overrides.items():
        if key in ("eps", "eps_opt", "eps_feas", "iteration_limit", "time_limit"):
            tc = prm.termination_criteria
            if key == "eps":
                tc.eps_optimal_relative = tc.eps_feasible_relative = val
            elif key == "eps_opt":
                tc.eps_optimal_relative = val
            elif key == "eps_feas":
                tc.eps_feasible_relative = val
            elif key == "iteration_limit":
                tc.iteration_limit = int(val)
            else:
                tc.time_sec_limit = float(val)
        else:
            setattr(prm, key, val)
    return prm


def solve(lib, prefix: str, p: QPData, params: Optional[capi.pdhg_parameters_t] = None) -> Optional[capi.Solution]:
    h = build(lib, prefix, p)
    if h is None:
        return None
    try:
        n, m = h.contents.num_variables, h.contents.num_constraints
        res = getattr(lib, prefix + "solve_qp_problem")(h.ptr, C.byref(params) if params is not None else None)
        if not res:
            return None
        m_out = max(m, 1) if m == 0 else m
        sol = capi.result_to_solution(res, n, m if m > 0 else 0)
        getattr(lib, prefix + ("result_free" if prefix else "pdhcg_result_free"))(res)
        return sol
    finally:
        h.close()
