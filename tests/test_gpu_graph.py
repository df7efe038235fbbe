"""GPU tests of graph replay of iteration windows (pdhcg_set_graph_window).

A window of plain iterations (no evaluation, restart or fixed-point error)
replayed from a recorded hipGraph runs the same kernels with the same
arithmetic as the launch-by-launch loop: the solution, objective and counts
must be bit-identical with the feature on and off, and the replay must
actually have been used where it applies (LP and diagonal Q without cones).
The SpMV lane shift is pinned: the autotuner chooses it by timing at each
solver creation, and a different shift sums rows in a different order.
"""
import os

import numpy as np
import pytest

import pdhcg_amd
from pdhcg_amd import workloads

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _pinned_spmv(monkeypatch):
    monkeypatch.setenv("PDHCG_SPMV_SHIFT", "0")


def _run(p, graph, precision=pdhcg_amd.PRECISION_FP64, chunk=10**6, **params):
    pdhcg_amd.set_graph_window(graph)
    try:
        st = pdhcg_amd.Stepper(p, precision, **params)
        rc = 0
        while rc == 0:
            rc = st.run(chunk)
        res = st.result()
        stats = st.graph_stats()
        st.close()
    finally:
        pdhcg_amd.set_graph_window(False)
    return res, stats


def _same(a, b):
    assert a.status == b.status
    assert a.iterations == b.iterations
    assert np.array_equal(a.x, b.x) and np.array_equal(a.y, b.y)
    assert a.pobj == b.pobj


@pytest.mark.parametrize("kind,n,seed", [("lp", 400, 1), ("diag", 500, 2), ("lp", 1200, 3), ("diag", 2000, 4)])
def test_graph_windows_bit_identical(kind, n, seed):
    p = workloads.random_problem(kind, seed=seed, n=n, m=n // 2)
    g, (windows, iters) = _run(p, True, eps=1e-6, iteration_limit=20000)
    e, (w0, i0) = _run(p, False, eps=1e-6, iteration_limit=20000)
    assert w0 == 0 and i0 == 0
    assert windows > 0 and iters > 0.5 * g.iterations, (windows, iters, g.iterations)
    _same(g, e)


def test_graph_windows_mixed_precision_and_small_chunks():
    p = workloads.random_problem("diag", seed=7, n=800, m=300)
    g, (windows, _) = _run(p, True, pdhcg_amd.PRECISION_MIXED, chunk=37, eps=1e-6, iteration_limit=20000)
    e, _ = _run(p, False, pdhcg_amd.PRECISION_MIXED, chunk=37, eps=1e-6, iteration_limit=20000)
    assert windows > 0
    _same(g, e)


def test_graph_windows_with_column_panels(monkeypatch):
    monkeypatch.setenv("PDHCG_SPMV_PANELS", "1")
    p = workloads.random_problem("lp", seed=9, n=3000, m=1000)
    g, (windows, _) = _run(p, True, eps=1e-6, iteration_limit=20000)
    e, _ = _run(p, False, eps=1e-6, iteration_limit=20000)
    assert windows > 0
    _same(g, e)


def test_graph_windows_respect_the_iteration_limit():
    p = workloads.random_problem("lp", seed=11, n=600, m=200)
    g, (windows, _) = _run(p, True, eps=1e-12, iteration_limit=433)
    e, _ = _run(p, False, eps=1e-12, iteration_limit=433)
    assert g.status == "ITERATION_LIMIT" and g.iterations == 433
    assert windows > 0
    _same(g, e)


def test_graph_windows_not_used_for_bb_problems():
    p = workloads.random_problem("lowrank_diag", seed=3, n=300, m=100)
    g, (windows, _) = _run(p, True, eps=1e-6, iteration_limit=20000)
    assert windows == 0 and g.status == "OPTIMAL"
