"""Quadratically constrained problems solved through solve_qp_problem on the
GPU: the library lifts the diagonal quadratic constraints into cone blocks
(qcqp_to_socp_qp, reference src/qcqp_transform.c) and returns the result cut
back to the original variables and rows (reference src/solver.cu:44-66,199).

Known answers: the reference's own example (test/test_qcqp_rsoc.c: min
-x + y^2 s.t. x^2 + y^2 <= 1 -> x = 1, y = 0, objective -1) and the closed form
of min c'x s.t. x'x <= r^2 (x = -r c / |c|).  Random diagonal QCQPs with
linear parts: the device solve against the oracle's solve of the same lifted
problem (status, objective within 5 eps) and against scipy's SLSQP on the
original problem.
"""
import ctypes as C

import numpy as np
import pytest
import scipy.optimize as so
import scipy.sparse as sp

import pdhcg_amd
from pdhcg_amd import QPData, capi
from pdhcg_amd.problem import build, solve as lib_solve
from tests.oracle_lib import oracle, oracle_params

pytestmark = pytest.mark.gpu

FORMS = [capi.CONE_ROTATED_SOC, capi.CONE_STANDARD_SOC]


@pytest.mark.parametrize("form", FORMS)
def test_reference_rsoc_example(form):
    p = QPData(c=[-1.0, 0.0], A=sp.csr_matrix((1, 2)), con_lb=[-np.inf], con_ub=[1.0], var_lb=[-np.inf] * 2,
               var_ub=[np.inf] * 2, Q=sp.diags([0.0, 2.0]).tocsr(), quadratic_constraints=[(0, sp.eye(2))])
    s = pdhcg_amd.solve(p, eps=1e-8, iteration_limit=200000, default_cone_type=form)
    assert s is not None and s.status == "OPTIMAL", s and s.status
    assert s.x.size == 2 and s.y.size == 1
    assert abs(s.x[0] - 1.0) < 1e-4 and abs(s.x[1]) < 1e-4 and abs(s.pobj + 1.0) < 1e-4


@pytest.mark.parametrize("form", FORMS)
def test_ball_closed_form(form):
    rng = np.random.default_rng(5)
    n, r = 12, 1.7
    c = rng.standard_normal(n)
    p = QPData(c=c, A=sp.csr_matrix((1, n)), con_lb=[-np.inf], con_ub=[r * r], var_lb=np.full(n, -10.0),
               var_ub=np.full(n, 10.0), quadratic_constraints=[(0, sp.eye(n).tocsr())])
    s = pdhcg_amd.solve(p, eps=1e-8, iteration_limit=200000, default_cone_type=form)
    assert s.status == "OPTIMAL"
    want = -r * c / np.linalg.norm(c)
    assert np.max(np.abs(s.x - want)) < 1e-5
    assert abs(s.pobj - c @ want) < 1e-6 * (1 + abs(c @ want))


def _random_qcqp(seed):
    rng = np.random.default_rng(seed)
    n, m = 10, 5
    A = sp.random(m, n, density=0.4, random_state=rng, format="csr") + sp.csr_matrix(
        (rng.uniform(0.5, 1.0, m), (range(m), range(m))), shape=(m, n))
    x0 = rng.uniform(-0.5, 0.5, n)
    cl, cu = A @ x0 - rng.uniform(0.5, 1.0, m), A @ x0 + rng.uniform(0.5, 1.0, m)
    qcs = []
    for r in (0, 2):
        d = np.zeros(n)
        idx = rng.choice(n, size=5, replace=False)
        d[idx] = rng.uniform(0.5, 2.0, size=5)
        cl[r], cu[r] = -np.inf, (A @ x0)[r] + x0 @ (d * x0) + rng.uniform(0.2, 0.5)
        qcs.append((r, sp.diags(d).tocsr()))
    return QPData(c=rng.standard_normal(n), A=A, con_lb=cl, con_ub=cu, var_lb=np.full(n, -2.0),
                  var_ub=np.full(n, 2.0), Q=sp.diags(rng.uniform(0.1, 0.5, n)).tocsr(), quadratic_constraints=qcs)


def _slsqp(p):
    A = p.A.toarray()
    qd = p.Q.diagonal()
    f = lambda x: p.c @ x + 0.5 * x @ (qd * x)
    cons = []
    qrows = {r: Q.diagonal() for r, Q in p.quadratic_constraints}
    for i in range(A.shape[0]):
        g = (lambda i: (lambda x: A[i] @ x + (x @ (qrows[i] * x) if i in qrows else 0.0)))(i)
        if np.isfinite(p.con_ub[i]):
            cons.append({"type": "ineq", "fun": (lambda g, u: (lambda x: u - g(x)))(g, p.con_ub[i])})
        if np.isfinite(p.con_lb[i]):
            cons.append({"type": "ineq", "fun": (lambda g, l: (lambda x: g(x) - l))(g, p.con_lb[i])})
    r = so.minimize(f, np.zeros(p.n), method="SLSQP", constraints=cons, bounds=list(zip(p.var_lb, p.var_ub)),
                    options={"ftol": 1e-14, "maxiter": 1000})
    assert r.success, r.message
    return r.fun


@pytest.mark.parametrize("form", FORMS)
@pytest.mark.parametrize("seed", [1, 2])
def test_random_qcqp_matches_oracle_and_slsqp(form, seed):
    p = _random_qcqp(seed)
    eps = 1e-7
    g = pdhcg_amd.solve(p, eps=eps, iteration_limit=200000, default_cone_type=form)
    assert g.status == "OPTIMAL" and g.x.size == p.n and g.y.size == p.A.shape[0]
    # the oracle solves the lifted problem the library builds
    L = pdhcg_amd.lib()
    L.qcqp_to_socp_qp.restype = C.POINTER(capi.qp_problem_t)
    L.qcqp_to_socp_qp.argtypes = [C.POINTER(capi.qp_problem_t), C.c_int]
    h = build(L, "", p)
    lifted = L.qcqp_to_socp_qp(h.ptr, form)
    h.close()
    assert lifted
    try:
        O = oracle()
        prm = oracle_params(eps=eps, iteration_limit=200000)
        res = O.orc_solve_qp_problem(lifted, C.byref(prm))
        o = capi.result_to_solution(res, lifted.contents.num_variables, lifted.contents.num_constraints)
        O.orc_result_free(res)
    finally:
        L.qp_problem_free(lifted)
    assert o.status == "OPTIMAL"
    assert abs(g.pobj - o.pobj) <= 5 * eps * (1 + abs(o.pobj)), (g.pobj, o.pobj)
    ref = _slsqp(p)
    assert abs(g.pobj - ref) <= 1e-5 * (1 + abs(ref)), (g.pobj, ref)
    # the returned point satisfies the original quadratic constraints
    for r, Q in p.quadratic_constraints:
        assert (p.A @ g.x)[r] + g.x @ (Q.diagonal() * g.x) <= p.con_ub[r] + 1e-5 * (1 + abs(p.con_ub[r]))
