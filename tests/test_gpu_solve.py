"""GPU parity: the MI355X library (through the pdhcg.h C interface) against
the FP64 CPU oracle and the reference's own known answers.

Bar (FP64 form): same termination status, objective within
5 * eps * (1 + |obj|) of the oracle's, and both returned points pass the
independent numpy KKT check at the solve tolerance (x 10 slack for the
max-norm box term, which the solvers measure in the scaled space).
"""
import ctypes as C

import numpy as np
import pytest
import scipy.sparse as sp

import pdhcg_amd
from pdhcg_amd import QPData, workloads
from tests.kkt import kkt
from tests.oracle_lib import oracle_solve

pytestmark = pytest.mark.gpu


def _check_pair(p, eps, **kw):
    g = pdhcg_amd.solve(p, eps=eps, **kw)
    o = oracle_solve(p, eps=eps, **kw)
    assert g is not None and o is not None
    assert g.status == o.status == "OPTIMAL", (g.status, o.status)
    assert abs(g.pobj - o.pobj) <= 5 * eps * (1 + abs(o.pobj)), (g.pobj, o.pobj)
    for s in (g, o):
        k = kkt(p, s.x, s.y)
        assert k["rel_primal"] < 10 * eps, k
        assert k["rel_dual"] < 10 * eps, k
        assert k["rel_gap"] < 10 * eps, k
    return g, o


def test_lp_known_answer():
    # reference test/test_c_api_working.c:72-277: min x + 2y, x + y = 1, x, y >= 0
    p = QPData(c=[1.0, 2.0], A=sp.csr_matrix([[1.0, 1.0]]), con_lb=[1.0], con_ub=[1.0], var_lb=[0.0, 0.0],
               var_ub=[1e30, 1e30])
    s = pdhcg_amd.solve(p)
    assert s.status == "OPTIMAL"
    assert 0.9 < s.x[0] < 1.1 and 0.0 <= s.x[1] < 0.1
    assert 0.9 < s.y[0] < 1.1
    assert 0.9 < s.pobj < 1.1


@pytest.mark.parametrize("Q,R,obj", [(None, sp.eye(2), 0.25), (sp.diags([2.0, 2.0]), sp.eye(2), 0.75)])
def test_lowrank_known_answers(Q, R, obj):
    # reference test/test_pdhcg_python.py:36-103
    p = QPData(c=[0.0, 0.0], A=np.array([[1.0, 1.0]]), con_lb=[1.0], con_ub=[1.0], var_lb=[0.0, 0.0], Q=Q, R=R,
               fmt={"A": "dense", "R": "csc", "Q": "csc"})
    s = pdhcg_amd.solve(p, eps=1e-6)
    assert s.status == "OPTIMAL"
    assert abs(s.pobj - obj) < 1e-4
    assert abs(s.x[0] - 0.5) < 1e-4 and abs(s.x[1] - 0.5) < 1e-4


def test_qp_without_constraints():
    # reference test/test_pdhcg_python.py:105-129 (dummy free row)
    p = QPData(c=[-2.0, -2.0], Q=sp.diags([2.0, 2.0]))
    s = pdhcg_amd.solve(p, eps=1e-6)
    assert s.status == "OPTIMAL"
    assert abs(s.pobj + 2.0) < 1e-4
    assert np.allclose(s.x, [1.0, 1.0], atol=1e-4)


@pytest.mark.parametrize("kind", ["lp", "diag", "sparse", "lowrank", "lowrank_diag", "lowrank_dense_d"])
@pytest.mark.parametrize("seed", [0, 1])
def test_random_matches_oracle(kind, seed):
    p = workloads.random_problem(kind, seed=seed)
    _check_pair(p, 1e-6)


@pytest.mark.parametrize("n", [193, 200, 256, 257, 700, 1023])
def test_lowrank_column_chunking(n):
    # the dense R x pass splits the columns into 256-wide chunks; sizes around
    # and off that width exercise the partial last chunk
    p = workloads.random_problem("lowrank_diag", seed=n, n=n, m=max(10, n // 4))
    _check_pair(p, 1e-6)


@pytest.mark.parametrize("kind", ["lp", "lowrank_diag", "sparse"])
def test_column_panel_spmv_matches_oracle(kind, monkeypatch):
    # the column-panel SpMV (used when x is larger than an XCD's L2) forced
    # on small problems, including A' with fewer columns (m = 3) than panels
    monkeypatch.setenv("PDHCG_SPMV_PANELS", "1")
    for n, m in ((300, 120), (6, 4), (5, 3)):
        p = workloads.random_problem(kind, seed=n + m, n=n, m=m)
        _check_pair(p, 1e-6)


@pytest.mark.parametrize("shift", ["-1", "1"])
def test_spmv_lane_shifts_match_oracle(shift, monkeypatch):
    # the autotuner's candidate launch shapes (lanes per row halved / doubled
    # for every row class), forced, on problems with short and long rows
    monkeypatch.setenv("PDHCG_SPMV_SHIFT", shift)
    for kind in ("lp", "lowrank_diag"):
        _check_pair(workloads.random_problem(kind, seed=9, n=300, m=120), 1e-6)
    _check_pair(workloads.config0(), 1e-4)


def test_sparse_lowrank_csr_path():
    p = workloads.random_problem("lowrank_sparse_r", seed=3, n=1000, m=150)
    _check_pair(p, 1e-4)


def test_config0_matches_oracle():
    p = workloads.config0()
    g, o = _check_pair(p, 1e-4)
    # same iteration in FP64: the trajectories stay together to termination
    assert abs(g.iterations - o.iterations) <= 200, (g.iterations, o.iterations)


def test_iteration_limit_and_counts():
    p = workloads.random_problem("sparse", seed=5)
    g = pdhcg_amd.solve(p, eps=1e-12, iteration_limit=57)
    o = oracle_solve(p, eps=1e-12, iteration_limit=57)
    assert g.status == o.status == "ITERATION_LIMIT"
    assert g.iterations == o.iterations == 57


def test_stepper_sampled_kernel_timing():
    # pdhcg_solver_timing(s, p) event-times every p-th launch of the
    # instrumented kernels; the A'y and A x_bar SpMVs run once per step
    p = workloads.random_problem("lowrank_diag", seed=4, n=600, m=150)
    st = pdhcg_amd.Stepper(p, eps=1e-12)
    try:
        st.run(5)
        st.synchronize()
        L = st.L
        assert L.pdhcg_solver_timing(st.s, -1) != 0
        assert L.pdhcg_solver_timing(st.s, 3) == 0
        t0, _ = st.counts()
        assert st.run(12) == 0
        st.synchronize()
        t1, _ = st.counts()
        assert t1 - t0 == 12
        for kid in (2, 3):
            ms, cnt = C.c_double(), C.c_int()
            assert L.pdhcg_solver_kernel_time(st.s, kid, C.byref(ms), C.byref(cnt)) == 0
            assert cnt.value == 4 and ms.value > 0.0, (kid, cnt.value, ms.value)
        assert L.pdhcg_solver_timing(st.s, 0) == 0
    finally:
        st.close()
