"""GPU parity of the mixed-precision form against the FP64 CPU oracle.

Mixed precision stores matrix values and iterates in FP32 and accumulates the
SpMVs and low-rank passes in FP32.  The termination test, the infeasibility
rays and the returned point are evaluated in FP64 on FP64 problem data from
the FP32 iterates widened exactly, so a mixed result's residuals are its
error against the FP64 problem.

Bar: the oracle's status, objective within 5 * eps * (1 + |obj|) of the
FP64 oracle's, and the returned point passes the independent numpy KKT check
(tests/kkt.py, FP64) at 10 * eps.  The reported relative primal residual
equals the numpy recomputation to 1e-6 of its value: an FP32 evaluation of
the same point would differ by the FP32 rounding of the products (about
1e-7 of |A||x|, several percent of a 1e-6 residual).
"""
import numpy as np
import pytest

import pdhcg_amd
from pdhcg_amd import PRECISION_FP64, PRECISION_MIXED, workloads
from tests.kkt import kkt
from tests.oracle_lib import oracle_solve

pytestmark = pytest.mark.gpu

KINDS = ["lp", "diag", "sparse", "lowrank", "lowrank_diag", "lowrank_dense_d", "lowrank_sparse_r"]


def _mixed_pair(p, eps, precision=PRECISION_MIXED, obj_tol=5.0, **kw):
    kw.setdefault("iteration_limit", 200000)
    g = pdhcg_amd.solve(p, precision=precision, eps=eps, **kw)
    o = oracle_solve(p, eps=eps, **kw)
    assert g is not None and o is not None
    assert g.status == o.status == "OPTIMAL", (g.status, o.status)
    assert abs(g.pobj - o.pobj) <= obj_tol * eps * (1 + abs(o.pobj)), (g.pobj, o.pobj)
    k = kkt(p, g.x, g.y)
    assert k["rel_primal"] < 10 * eps, k
    assert k["rel_dual"] < 10 * eps, k
    assert k["rel_gap"] < 10 * eps, k
    # the reported primal residual is the FP64 one of the returned point
    assert abs(g.rel_primal - k["rel_primal"]) <= 1e-6 * k["rel_primal"] + 1e-13, (g.rel_primal, k["rel_primal"])
    assert g.rel_primal < eps and g.rel_dual < eps and g.rel_gap < eps
    return g, o


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("seed", [0, 1])
def test_mixed_random_matches_oracle(kind, seed):
    n, m = (1000, 150) if kind == "lowrank_sparse_r" else (60, 30)
    _mixed_pair(workloads.random_problem(kind, seed=seed, n=n, m=m), 1e-6)


@pytest.mark.parametrize("n", [193, 257, 700, 1023])
def test_mixed_lowrank_column_chunking(n):
    # FP32 R x and R'u passes: 256-column tiles with a partial last tile and
    # four FP32 columns per lane in the R'u pass
    _mixed_pair(workloads.random_problem("lowrank_diag", seed=n, n=n, m=max(10, n // 4)), 1e-6)


def test_mixed_config0_matches_oracle():
    _mixed_pair(workloads.config0(), 1e-4)


@pytest.mark.parametrize("kind", ["lp", "lowrank_diag", "sparse"])
def test_mixed_column_panels_match_oracle(kind, monkeypatch):
    monkeypatch.setenv("PDHCG_SPMV_PANELS", "1")
    for n, m in ((300, 120), (6, 4)):
        _mixed_pair(workloads.random_problem(kind, seed=n + m, n=n, m=m), 1e-6)


@pytest.mark.parametrize("shift", ["-1", "1"])
def test_mixed_lane_shifts_match_oracle(shift, monkeypatch):
    monkeypatch.setenv("PDHCG_SPMV_SHIFT", shift)
    for kind in ("lp", "lowrank_diag", "diag"):
        _mixed_pair(workloads.random_problem(kind, seed=11, n=300, m=120), 1e-6)


def test_mixed_infeasible_lp():
    # x1 + x2 >= 3 with x in [0, 1]^2: primal infeasible; the ray test runs in FP64
    import scipy.sparse as sp
    p = pdhcg_amd.QPData(c=[1.0, 1.0], A=sp.csr_matrix([[1.0, 1.0]]), con_lb=[3.0], con_ub=[np.inf],
                         var_lb=[0.0, 0.0], var_ub=[1.0, 1.0])
    g = pdhcg_amd.solve(p, precision=PRECISION_MIXED, iteration_limit=20000)
    o = oracle_solve(p, iteration_limit=20000)
    assert g.status == o.status == "PRIMAL_INFEASIBLE", (g.status, o.status)


def test_mixed_iteration_limit_counts():
    p = workloads.random_problem("sparse", seed=5)
    g = pdhcg_amd.solve(p, precision=PRECISION_MIXED, eps=1e-12, iteration_limit=57)
    assert g.status == "ITERATION_LIMIT" and g.iterations == 57


def test_lowrank_rank_above_tile_limit():
    # rank 300 > the dense tile pass's 256: the CSR R / R' products take it
    rng = np.random.default_rng(3)
    n, m, k = 400, 60, 300
    p = workloads.random_problem("lp", seed=3, n=n, m=m)
    p.R = rng.standard_normal((k, n)) / np.sqrt(n)
    p.D = rng.uniform(0.5, 2.0, k)
    p.fmt = {"R": "dense"}
    for prec in (PRECISION_FP64, PRECISION_MIXED):
        _mixed_pair(p, 1e-6, precision=prec)


@pytest.fixture(scope="module")
def config1_small():
    return workloads.config1(scale=0.01)


@pytest.mark.parametrize("precision", [PRECISION_FP64, PRECISION_MIXED])
def test_config1_reduced_matches_oracle(config1_small, precision):
    # the large low-rank QP's generator at 1/100 of its size (n = 1e4,
    # m = 5e3, nnz = 1e5, R 50 x n), both precisions against the oracle
    _mixed_pair(config1_small, 1e-6, precision=precision)


@pytest.fixture(scope="module")
def config1_full():
    return workloads.config1()


@pytest.mark.parametrize("precision", [PRECISION_FP64, PRECISION_MIXED])
def test_config1_full_size_fp64_kkt(config1_full, precision):
    # n = 1e6, m = 5e5, nnz = 1e7, R 50 x 1e6: solved to 1e-6 and the returned
    # point checked by the numpy FP64 KKT recomputation (no oracle solve at
    # this size; the reduced-size test above carries the oracle comparison)
    p = config1_full
    eps = 1e-6
    g = pdhcg_amd.solve(p, precision=precision, eps=eps)
    assert g is not None and g.status == "OPTIMAL", g and g.status
    k = kkt(p, g.x, g.y)
    assert k["rel_primal"] < 10 * eps and k["rel_dual"] < 10 * eps and k["rel_gap"] < 10 * eps, k
    assert abs(g.rel_primal - k["rel_primal"]) <= 1e-6 * k["rel_primal"] + 1e-13, (g.rel_primal, k["rel_primal"])
    assert abs(g.pobj - k["pobj"]) <= 1e-9 * (1 + abs(k["pobj"])), (g.pobj, k["pobj"])
