"""CPU tests of the C interface: the library loads, exports every symbol the
headers declare, and the host-side problem assembly (format conversion,
[A; F] stacking, validation, warm starts, fixed cone slots) agrees bit for bit
with the oracle's independent restatement of reference src/pdhcg.c.
No device code runs here.
"""
import ctypes as C
import os
import re

import numpy as np
import pytest
import scipy.sparse as sp

import pdhcg_amd
from pdhcg_amd import QPData, build, capi
from tests.oracle_lib import oracle, oracle_build

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in sorted(f for f in os.listdir(os.path.join(ROOT, "include")) if f.endswith(".h") and f != "pdhcg_types.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b([a-z_][a-z0-9_]*)\s*\(", text):
            if m.group(1) not in ("if", "while", "for", "sizeof", "return"):
                names.add(m.group(1))
    return names


def test_exports_every_declared_symbol():
    L = pdhcg_amd.lib()
    names = declared_symbols()
    assert {"create_qp_problem", "solve_qp_problem", "set_default_parameters", "qp_problem_free",
            "pdhcg_result_free", "set_start_values", "set_cone_fixed", "qcqp_to_socp_qp",
            "solve_qp_problem_distributed", "pdhcg_solver_create", "pdhcg_solver_run", "pdhcg_project_cones",
            "permute_problem", "optimize_partition_cuts", "get_balanced_cuts"} <= names
    for n in names:
        assert hasattr(L, n), n


def test_ctypes_mirror_matches_c_layout(tmp_path):
    """Every member offset and struct size of the ctypes mirror equals what a C
    compiler computes for include/pdhcg_types.h."""
    import shutil
    import subprocess
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    structs = [capi.CsrComponent, capi.cone_blocks_t, capi.cone_spec_t, capi.qp_problem_t,
               capi.restart_parameters_t, capi.termination_criteria_t, capi.inner_solver_parameters_t,
               capi.grid_size_t, capi.pdhg_parameters_t, capi.pdhcg_result_t, capi.matrix_desc_t]
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "pdhcg_types.h"', "int main(void){"]
    for s in structs:
        lines.append(f'printf("{s.__name__} %zu\\n", sizeof({s.__name__}));')
        for f, _ in s._fields_:
            lines.append(f'printf("{s.__name__}.{f} %zu\\n", offsetof({s.__name__}, {f}));')
    lines.append("return 0;}")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run([cc, "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = dict(l.rsplit(" ", 1) for l in subprocess.run([str(exe)], capture_output=True, text=True,
                                                         check=True).stdout.split("\n") if l)
    for s in structs:
        assert int(out[s.__name__]) == C.sizeof(s), s.__name__
        for f, _ in s._fields_:
            assert int(out[f"{s.__name__}.{f}"]) == getattr(s, f).offset, f"{s.__name__}.{f}"


def test_default_parameters_match_oracle():
    a = pdhcg_amd.default_params(pdhcg_amd.lib(), "")
    b = pdhcg_amd.default_params(oracle(), "orc_")
    assert bytes(a) == bytes(b)
    assert a.l_inf_ruiz_iterations == 10 and a.termination_evaluation_frequency == 200
    assert a.termination_criteria.eps_optimal_relative == 1e-4
    assert a.inner_solver_parameters.iteration_limit == 1000
    assert a.restart_params.k_p == 0.99


def _csr(prob, which):
    c = getattr(prob, which).contents
    rows = {"constraint_matrix": prob.num_constraints, "objective_sparse_matrix": prob.num_variables,
            "objective_lowrank_matrix": prob.num_rank_lowrank_obj}[which]
    nnz = getattr(prob, which + "_num_nonzeros")
    rp = np.ctypeslib.as_array(c.row_ptr, shape=(rows + 1,)).copy()
    ci = np.ctypeslib.as_array(c.col_ind, shape=(nnz,)).copy() if nnz else np.zeros(0, np.int32)
    v = np.ctypeslib.as_array(c.val, shape=(nnz,)).copy() if nnz else np.zeros(0)
    return rp, ci, v


def _vec(ptr, n):
    return np.ctypeslib.as_array(ptr, shape=(n,)).copy() if n else np.zeros(0)


def _same_problem(a, b):
    pa, pb = a.contents, b.contents
    assert pa.num_variables == pb.num_variables and pa.num_constraints == pb.num_constraints
    assert pa.num_rank_lowrank_obj == pb.num_rank_lowrank_obj
    for w in ("constraint_matrix", "objective_sparse_matrix", "objective_lowrank_matrix"):
        for x, y in zip(_csr(pa, w), _csr(pb, w)):
            assert x.dtype == y.dtype and np.array_equal(x, y), w
    n, m = pa.num_variables, pa.num_constraints
    for f, k in (("variable_lower_bound", n), ("variable_upper_bound", n), ("objective_vector", n),
                 ("constraint_lower_bound", m), ("constraint_upper_bound", m), ("affine_cone_offset", m)):
        x, y = _vec(getattr(pa, f), k), _vec(getattr(pb, f), k)
        assert np.array_equal(x, y, equal_nan=True), f
    assert pa.objective_constant == pb.objective_constant


@pytest.mark.parametrize("fmt", ["dense", "csr", "csc", "coo"])
def test_assembly_bit_exact_against_oracle(fmt):
    rng = np.random.default_rng(11)
    n, m = 37, 23
    A = sp.random(m, n, density=0.3, random_state=3, format="csr")
    A.data = rng.standard_normal(A.nnz)
    A.data[::7] = 1e-13  # below the dense default tolerance: dropped by dense input only
    Q = sp.random(n, n, density=0.1, random_state=4, format="csr")
    R = rng.standard_normal((4, n))
    p = QPData(c=rng.standard_normal(n), A=A, Q=Q, R=R, D=rng.uniform(0.5, 2, 4), con_lb=-np.ones(m),
               con_ub=np.ones(m), var_lb=None, var_ub=np.full(n, 3.0), objective_constant=1.5,
               fmt={"A": fmt, "Q": fmt, "R": fmt})
    a = build(pdhcg_amd.lib(), "", p)
    b = oracle_build(p)
    _same_problem(a.ptr, b.ptr)
    # the conversion itself: same matrix as scipy's
    rp, ci, v = _csr(a.contents, "constraint_matrix")
    got = sp.csr_matrix((v, ci, rp), shape=(m, n)).toarray()
    want = A.toarray()
    if fmt == "dense":
        want[np.abs(want) <= 1e-12] = 0.0
    assert np.array_equal(got, want)
    for r in range(m):  # columns within a row keep input order; scipy CSR/CSC/COO/dense give ascending
        assert np.all(np.diff(ci[rp[r]:rp[r + 1]]) > 0)


def test_zero_tolerance_filters():
    A = np.array([[1.0, 1e-5, 0.0], [2e-5, 0.0, 3.0]])
    for fmt in ("dense", "csc", "coo", "csr"):
        p = QPData(c=np.zeros(3), A=A, fmt={"A": fmt})
        d = capi.desc_from(A, fmt, zero_tolerance=1e-4)
        L = pdhcg_amd.lib()
        h = L.create_qp_problem(capi.as_ptr(np.zeros(3)), None, None, None, d.ptr, None, None, None, None, None, 0,
                                None, None, None, 0, None)
        O = oracle()
        ho = O.orc_create_qp_problem(capi.as_ptr(np.zeros(3)), None, None, None, d.ptr, None, None, None, None, None,
                                     0, None, None, None, 0, None)
        _same_problem(h, ho)
        nnz = h.contents.constraint_matrix_num_nonzeros
        assert nnz == (2 if fmt != "csr" else 4)  # CSR input is copied verbatim
        L.qp_problem_free(h)
        O.orc_qp_problem_free(ho)


def test_affine_rows_stacked_under_A():
    rng = np.random.default_rng(2)
    n = 9
    A = sp.random(4, n, density=0.5, random_state=5, format="csr")
    F = sp.random(6, n, density=0.5, random_state=6, format="coo")
    g = rng.standard_normal(6)
    p = QPData(c=np.zeros(n), A=A, F=F, g=g, con_lb=-np.ones(4), con_ub=np.ones(4),
               affine_cones=[(capi.CONE_STANDARD_SOC, 0, 1), (capi.CONE_EXPONENTIAL, 3, 1)], fmt={"F": "coo"})
    a = build(pdhcg_amd.lib(), "", p)
    b = oracle_build(p)
    _same_problem(a.ptr, b.ptr)
    pa = a.contents
    assert pa.num_constraints == 10
    rp, ci, v = _csr(pa, "constraint_matrix")
    got = sp.csr_matrix((v, ci, rp), shape=(10, n)).toarray()
    assert np.array_equal(got, np.vstack([A.toarray(), F.toarray()]))
    off = _vec(pa.affine_cone_offset, 10)
    assert np.array_equal(off, np.concatenate([np.zeros(4), g]))
    assert list(np.ctypeslib.as_array(pa.affine_cones.start_idx, shape=(2,))) == [4, 7]
    lb = _vec(pa.constraint_lower_bound, 10)
    assert np.all(np.isneginf(lb[4:]))


def _create(lib, prefix, p):
    return build(lib, prefix, p)


@pytest.mark.parametrize("case", [
    "no_matrices", "bad_q_shape", "csr_bad_col", "overlap", "cone_box", "bad_alpha", "affine_cover", "psd_fixed_none",
    "affine_fixed",
])
def test_validation_rejects_like_oracle(case):
    n = 6
    A = sp.eye(3, n, format="csr")
    kw = dict(c=np.zeros(n), A=A)
    if case == "no_matrices":
        kw = dict(c=np.zeros(n))
    elif case == "bad_q_shape":
        kw["Q"] = sp.eye(n + 1, format="csr")
    elif case == "csr_bad_col":
        bad = sp.csr_matrix(A)
        bad.indices[0] = n + 3
        kw["A"] = bad
        kw["fmt"] = {"A": "csr"}
    elif case == "overlap":
        kw["cones"] = [(capi.CONE_STANDARD_SOC, 0, 1), (capi.CONE_STANDARD_SOC, 2, 1)]
    elif case == "cone_box":
        kw["cones"] = [(capi.CONE_STANDARD_SOC, 0, 1)]
        kw["var_lb"] = np.zeros(n)
    elif case == "bad_alpha":
        kw["cones"] = [(capi.CONE_POWER, 0, 1, 1.5)]
    elif case == "affine_cover":
        kw["F"] = sp.eye(4, n, format="csr")
        kw["affine_cones"] = [(capi.CONE_STANDARD_SOC, 0, 1)]
    elif case == "affine_fixed":
        # reference test/test_affine_cones.c:247-268: fixed slots on an affine cone
        kw["F"] = sp.eye(3, n, format="csr")
        kw["affine_cones"] = [(capi.CONE_STANDARD_SOC, 0, 1, 0.0, [1, 0, 0])]
    elif case == "psd_fixed_none":
        kw["cones"] = [(capi.CONE_PSD, 0, 3)]  # valid: 6 slots, no fixed mask
    p = QPData(**kw)
    if case == "csr_bad_col":
        # build() passes the scipy arrays through unchanged
        pass
    a = _create(pdhcg_amd.lib(), "", p)
    b = _create(oracle(), "orc_", p)
    if case == "psd_fixed_none":
        assert a is not None and b is not None
        _same_problem(a.ptr, b.ptr)
    else:
        assert a is None and b is None


def test_start_values_and_fixed_slots():
    n = 5
    p = QPData(c=np.zeros(n), A=sp.eye(2, n, format="csr"), cones=[(capi.CONE_STANDARD_SOC, 0, 1)])
    for lib, pre in ((pdhcg_amd.lib(), ""), (oracle(), "orc_")):
        h = build(lib, pre, p)
        fix = getattr(lib, pre + "set_cone_fixed")
        assert fix(h.ptr, 0, 2, 2.0) == 0
        assert fix(h.ptr, 0, 3, 1.0) != 0          # slot out of range
        assert fix(h.ptr, 1, 0, 1.0) != 0          # cone out of range
        assert fix(h.ptr, 0, 0, float("nan")) != 0
        pr = h.contents
        assert pr.cones.fixed_mask_size == n
        ps = _vec(pr.primal_start, n)
        assert ps[2] == 2.0
        ss = getattr(lib, pre + "set_start_values")
        x = np.arange(n, dtype=float)
        ss(h.ptr, capi.as_ptr(x), None)  # changes the fixed slot 2 -> rejected
        assert _vec(h.contents.primal_start, n)[2] == 2.0
        x[2] = 2.0
        y = np.array([0.5, -0.5])
        ss(h.ptr, capi.as_ptr(x), capi.as_ptr(y))
        assert np.array_equal(_vec(h.contents.primal_start, n), x)
        assert np.array_equal(_vec(h.contents.dual_start, 2), y)
        h.close()


def test_unsupported_entry_points_return_null():
    L = pdhcg_amd.lib()
    L.qcqp_to_socp_qp.restype = C.c_void_p
    L.qcqp_to_socp_qp.argtypes = [C.c_void_p, C.c_int]
    assert not L.qcqp_to_socp_qp(None, 0)
    L.solve_qp_problem_distributed.restype = C.c_void_p
    L.solve_qp_problem_distributed.argtypes = [C.c_void_p, C.c_void_p]
    assert not L.solve_qp_problem_distributed(None, None)
    assert not L.solve_qp_problem(None, None)


def test_desc_diag_matches_scipy_diags():
    # the numpy-built descriptor of a diagonal D holds the entries scipy's
    # diags(D).tocsr() keeps (nonzero values, NaN kept), in the same order
    import scipy.sparse as sp
    from pdhcg_amd import capi
    rng = np.random.default_rng(5)
    for dv in (np.array([1.0, 0.0, 2.0, np.nan, -3.0]), np.array([3.0]), np.zeros(4), rng.uniform(-1, 1, 57)):
        rp, ci, v = capi.desc_diag(dv).keep
        m = sp.diags(dv).tocsr()
        assert np.array_equal(rp, m.indptr) and np.array_equal(ci, m.indices)
        assert np.array_equal(v, m.data, equal_nan=True)
