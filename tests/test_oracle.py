"""Pins the CPU oracle (the checker of the GPU path) to things other than itself:
the reference's own known answers, scipy's LP solver, an independent numpy
KKT recomputation, closed forms of the scaling passes, numpy's SVD for the
step size, and scipy's CSR->CSC for the transpose.
"""
import ctypes as C

import numpy as np
import pytest
import scipy.optimize as so
import scipy.sparse as sp

from pdhcg_amd import QPData, capi, workloads
from tests.kkt import kkt
from tests.oracle_lib import oracle, oracle_build, oracle_params, oracle_solve


# ---- reference known answers ------------------------------------------------

def test_lp_known_answer():
    # reference test/test_c_api_working.c:72-277
    p = QPData(c=[1.0, 2.0], A=sp.csr_matrix([[1.0, 1.0]]), con_lb=[1.0], con_ub=[1.0], var_lb=[0.0, 0.0],
               var_ub=[1e30, 1e30])
    s = oracle_solve(p)
    assert s.status == "OPTIMAL"
    assert 0.9 < s.x[0] < 1.1 and 0.0 <= s.x[1] < 0.1 and 0.9 < s.y[0] < 1.1 and 0.9 < s.pobj < 1.1


@pytest.mark.parametrize("Q,R,c,obj,x", [
    (None, sp.eye(2), [0.0, 0.0], 0.25, [0.5, 0.5]),                  # test_pdhcg_python.py:36-68
    (sp.diags([2.0, 2.0]), sp.eye(2), [0.0, 0.0], 0.75, [0.5, 0.5]),  # :70-103
    (sp.diags([4.0, 2.0]), None, [1.0, 1.0], 5.0 / 3.0, [1 / 3, 2 / 3]),  # :131-169
])
def test_qp_known_answers(Q, R, c, obj, x):
    p = QPData(c=c, A=np.array([[1.0, 1.0]]), con_lb=[1.0], con_ub=[1.0], var_lb=[0.0, 0.0], Q=Q, R=R,
               fmt={"A": "dense", "Q": "csc", "R": "csc"})
    s = oracle_solve(p, eps=1e-6)
    assert s.status == "OPTIMAL"
    assert abs(s.pobj - obj) < 1e-4
    assert np.allclose(s.x, x, atol=1e-4)


def test_qp_without_constraint_matrix():
    # test_pdhcg_python.py:105-129
    s = oracle_solve(QPData(c=[-2.0, -2.0], Q=sp.diags([2.0, 2.0])), eps=1e-6)
    assert s.status == "OPTIMAL" and abs(s.pobj + 2.0) < 1e-4 and np.allclose(s.x, 1.0, atol=1e-4)


def test_lowrank_middle_identity_equals_no_middle():
    # invariant of reference test/test_lowrank_middle_D.py:60-70 (D = I equals no D)
    p = workloads.random_problem("lowrank", seed=4)
    a = oracle_solve(p, eps=1e-6)
    p.D = np.ones(6)
    b = oracle_solve(p, eps=1e-6)
    assert a.status == b.status == "OPTIMAL"
    assert np.allclose(a.x, b.x, atol=1e-3) and abs(a.pobj - b.pobj) < 1e-5 * (1 + abs(a.pobj))


def test_lowrank_middle_diag_equals_folded_r():
    # invariant of test_lowrank_middle_D.py:73-88: R' diag(d) R == (sqrt(d) R)'(sqrt(d) R)
    p = workloads.random_problem("lowrank", seed=5)
    d = np.array([0.3, 1.7, 0.9, 2.4, 0.5, 1.1])
    folded = QPData(**{**p.__dict__, "R": np.sqrt(d)[:, None] * p.R})
    p.D = d
    a, b = oracle_solve(p, eps=1e-6), oracle_solve(folded, eps=1e-6)
    assert a.status == b.status == "OPTIMAL"
    assert abs(a.pobj - b.pobj) < 1e-5 * (1 + abs(a.pobj))


# ---- independent optimality checks ------------------------------------------

@pytest.mark.parametrize("kind", ["lp", "diag", "sparse", "lowrank", "lowrank_diag", "lowrank_dense_d"])
def test_random_instances_satisfy_kkt(kind):
    p = workloads.random_problem(kind, seed=11)
    s = oracle_solve(p, eps=1e-6, iteration_limit=100000)
    assert s.status == "OPTIMAL"
    k = kkt(p, s.x, s.y)
    assert k["rel_primal"] < 1e-5 and k["rel_dual"] < 1e-5 and k["rel_gap"] < 1e-5, k
    assert abs(k["pobj"] - s.pobj) < 1e-8 * (1 + abs(s.pobj))


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_lp_objective_matches_scipy_highs(seed):
    p = workloads.random_problem("lp", seed=seed, n=40, m=20)
    s = oracle_solve(p, eps=1e-8)
    A = p.A.toarray()
    res = so.linprog(p.c, A_ub=np.vstack([A, -A]), b_ub=np.concatenate([p.con_ub, -p.con_lb]),
                     bounds=list(zip(p.var_lb, p.var_ub)), method="highs")
    assert res.status == 0 and s.status == "OPTIMAL"
    assert abs(s.pobj - res.fun) < 1e-6 * (1 + abs(res.fun))


def test_infeasible_lp_detected():
    # x1 + x2 >= 3 with x in [0, 1]^2: primal infeasible
    p = QPData(c=[1.0, 1.0], A=sp.csr_matrix([[1.0, 1.0]]), con_lb=[3.0], con_ub=[np.inf], var_lb=[0.0, 0.0],
               var_ub=[1.0, 1.0])
    s = oracle_solve(p, iteration_limit=20000)
    assert s.status in ("PRIMAL_INFEASIBLE", "ITERATION_LIMIT")
    assert s.status == "PRIMAL_INFEASIBLE"


def test_iteration_limit_status():
    s = oracle_solve(workloads.random_problem("sparse", seed=5), eps=1e-12, iteration_limit=57)
    assert s.status == "ITERATION_LIMIT" and s.iterations == 57


# ---- closed forms of the scaling and the step size --------------------------

def _scalings(p, **params):
    L = oracle()
    h = oracle_build(p)
    prm = oracle_params(**params)
    s = L.orc_solver_create(h.ptr, C.byref(prm))
    m, n = h.contents.num_constraints, h.contents.num_variables
    cr, vr = np.zeros(m), np.zeros(n)
    out = [C.c_double() for _ in range(4)]
    L.orc_solver_scalings(s, capi.as_ptr(cr), capi.as_ptr(vr), *[C.byref(o) for o in out])
    L.orc_solver_free(s)
    h.close()
    return cr, vr, [o.value for o in out]


def test_ruiz_equilibrates_rows_and_columns():
    rng = np.random.default_rng(3)
    A = sp.random(30, 40, density=0.2, random_state=7, format="csr")
    A.data = rng.standard_normal(A.nnz) * 10.0 ** rng.integers(-3, 4, A.nnz)
    A = A[np.asarray(A.getnnz(axis=1) > 0)][:, np.asarray(A.getnnz(axis=0) > 0)]
    p = QPData(c=rng.standard_normal(A.shape[1]), A=A, con_lb=-np.ones(A.shape[0]), con_ub=np.ones(A.shape[0]))
    cr, vr, _ = _scalings(p, l_inf_ruiz_iterations=60, has_pock_chambolle_alpha=False,
                          bound_objective_rescaling=False)
    S = sp.diags(1 / cr) @ A @ sp.diags(1 / vr)
    assert np.allclose(abs(S).max(axis=1).toarray().ravel(), 1.0, atol=1e-3)
    assert np.allclose(abs(S).max(axis=0).toarray().ravel(), 1.0, atol=1e-3)


def test_diagonal_matrix_scaling_closed_form():
    # one Ruiz pass on a diagonal A maps every entry to +-1 (d_r = d_c = sqrt|a|)
    a = np.array([4.0, -0.25, 9.0])
    p = QPData(c=np.zeros(3), A=sp.diags(a).tocsr(), con_lb=-np.ones(3), con_ub=np.ones(3))
    cr, vr, (cbr, ovr, step, pw) = _scalings(p, l_inf_ruiz_iterations=1, has_pock_chambolle_alpha=False,
                                             bound_objective_rescaling=False)
    assert np.allclose(cr, np.sqrt(np.abs(a)), rtol=0, atol=1e-15)
    assert np.allclose(vr, np.sqrt(np.abs(a)), rtol=0, atol=1e-15)
    assert cbr == 1.0 and ovr == 1.0
    # scaled matrix is diag(+-1): sigma_max = 1, step = 0.998 (power iteration to 1e-4)
    assert abs(step - 0.998) < 1e-3


def test_pock_chambolle_and_bound_objective_closed_form():
    A = np.array([[1.0, 2.0], [0.0, 3.0]])
    cl, cu = np.array([1.0, -2.0]), np.array([1.0, 4.0])
    c = np.array([3.0, 4.0])
    p = QPData(c=c, A=A, con_lb=cl, con_ub=cu, fmt={"A": "dense"})
    cr, vr, (cbr, ovr, _, _) = _scalings(p, l_inf_ruiz_iterations=0, pock_chambolle_alpha=1.0)
    # alpha = 1: row factors sqrt(sum_j |a_ij|), column factors sqrt(sum_i |a_ij|)
    assert np.allclose(cr, np.sqrt([3.0, 3.0]), atol=1e-15)
    assert np.allclose(vr, np.sqrt([1.0, 5.0]), atol=1e-15)
    # bound/objective scaling on the PC-scaled data (equal bounds count once)
    sl, su, sc = cl / cr, cu / cr, c / vr
    b2 = su[0] ** 2 + sl[1] ** 2 + su[1] ** 2
    assert abs(cbr - 1 / (np.sqrt(b2) + 1)) < 1e-15
    assert abs(ovr - 1 / (np.linalg.norm(sc) + 1)) < 1e-15


def test_step_size_from_largest_singular_value():
    p = workloads.random_problem("lp", seed=9, n=50, m=30)
    cr, vr, (cbr, ovr, step, pw) = _scalings(p)
    S = (sp.diags(1 / cr) @ p.A @ sp.diags(1 / vr)).toarray()
    smax = np.linalg.svd(S, compute_uv=False)[0]
    assert abs(step - 0.998 / smax) < 1e-3 * (0.998 / smax)
    assert pw == 1.0


def test_transpose_matches_scipy():
    L = oracle()
    A = sp.random(57, 33, density=0.15, random_state=9, format="csr")
    m, n = A.shape
    rp, ci, v = A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data.copy()
    trp, tci, tv = np.zeros(n + 1, np.int32), np.zeros(A.nnz, np.int32), np.zeros(A.nnz)
    ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int))
    L.orc_csr_transpose(m, n, ip(rp), ip(ci), capi.as_ptr(v), ip(trp), ip(tci), capi.as_ptr(tv))
    T = A.T.tocsr()
    T.sort_indices()
    assert np.array_equal(trp, T.indptr) and np.array_equal(tci, T.indices) and np.array_equal(tv, T.data)


def test_oracle_results_do_not_depend_on_thread_count():
    # the row-parallel products keep each row's summation order
    L = oracle()
    p = workloads.config1(scale=0.02)
    sols = []
    try:
        for threads in (1, 4):
            L.orc_set_threads(threads)
            sols.append(oracle_solve(p, eps=1e-4, iteration_limit=60))
    finally:
        L.orc_set_threads(0)
    a, b = sols
    assert a.iterations == b.iterations
    assert np.array_equal(a.x, b.x) and np.array_equal(a.y, b.y) and a.pobj == b.pobj
