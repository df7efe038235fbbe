"""GPU test of the learned warm start end to end: training data from the
batched solver, the network trained on the device, and warm-started solves of
unseen instances against cold solves of the same instances.

Bar: every warm-started solve reaches the same status as the cold one with
the objective within 5 eps (1 + |obj|) and the independent numpy KKT check
at the tolerance; one warm-started device solve agrees with the oracle's
solve from the same start values; and the warm starts take fewer iterations
(at 1e-4, where the saved initial phase is the largest share).
"""
import numpy as np
import pytest

import pdhcg_amd
from pdhcg_amd import warmstart, workloads
from tests.kkt import kkt
from tests.oracle_lib import oracle_solve

pytestmark = pytest.mark.gpu

EPS = 1e-6


@pytest.fixture(scope="module")
def trained():
    fam = workloads.ParametricFamily()
    Ztr, ptr = fam.sample(1024, 21)
    sols = pdhcg_amd.solve_batch(ptr, eps=1e-4, iteration_limit=100000, time_limit=120.0)
    assert all(s is not None and s.status == "OPTIMAL" for s in sols)
    X = np.stack([s.x for s in sols])
    Y = np.stack([s.y for s in sols])
    ws = warmstart.WarmStart(fam.z_dim, fam.n, fam.m, hidden=1024, seed=0)
    loss = ws.fit(Ztr, X, Y, epochs=200, batch=256, lr=1e-3)
    assert np.isfinite(loss) and loss < 0.5
    Zte, pte = fam.sample(32, 22)
    yield fam, ws, Zte, pte
    ws.close()


def test_warm_start_solves_agree_with_cold(trained):
    fam, ws, Zte, pte = trained
    warm_p = [ws.start(p, z) for p, z in zip(pte, Zte)]
    for p in warm_p:
        assert np.all(p.primal_start >= 0) and np.all(p.primal_start <= 5)
    cold = pdhcg_amd.solve_batch(pte, eps=EPS, iteration_limit=100000, time_limit=120.0)
    warm = pdhcg_amd.solve_batch(warm_p, eps=EPS, iteration_limit=100000, time_limit=120.0)
    for p, c, w in zip(pte, cold, warm):
        assert c.status == w.status == "OPTIMAL"
        assert abs(c.pobj - w.pobj) <= 5 * EPS * (1 + abs(c.pobj)), (c.pobj, w.pobj)
        k = kkt(p, w.x, w.y)
        assert k["rel_primal"] < 10 * EPS and k["rel_dual"] < 10 * EPS and k["rel_gap"] < 10 * EPS, k


def test_warm_start_cuts_iterations(trained):
    fam, ws, Zte, pte = trained
    warm_p = [ws.start(p, z) for p, z in zip(pte, Zte)]
    cold = pdhcg_amd.solve_batch(pte, eps=1e-4, iteration_limit=100000, time_limit=120.0)
    warm = pdhcg_amd.solve_batch(warm_p, eps=1e-4, iteration_limit=100000, time_limit=120.0)
    ic = np.mean([s.iterations for s in cold])
    iw = np.mean([s.iterations for s in warm])
    assert all(s.status == "OPTIMAL" for s in warm)
    assert iw <= 0.85 * ic, (ic, iw)


def test_warm_started_device_solve_matches_oracle(trained):
    fam, ws, Zte, pte = trained
    p = ws.start(pte[0], Zte[0])
    g = pdhcg_amd.solve(p, eps=EPS, iteration_limit=100000)
    o = oracle_solve(p, eps=EPS, iteration_limit=100000)
    assert g.status == o.status == "OPTIMAL"
    assert abs(g.pobj - o.pobj) <= 5 * EPS * (1 + abs(o.pobj))


def test_primal_only_start_solves_agree_with_cold(trained):
    fam, ws, Zte, pte = trained
    warm_p = [ws.start(p, z, dual=False) for p, z in zip(pte, Zte)]
    assert all(p.dual_start is None and p.primal_start is not None for p in warm_p)
    cold = pdhcg_amd.solve_batch(pte, eps=EPS, iteration_limit=100000, time_limit=120.0)
    warm = pdhcg_amd.solve_batch(warm_p, eps=EPS, iteration_limit=100000, time_limit=120.0)
    for p, c, w in zip(pte, cold, warm):
        assert c.status == w.status == "OPTIMAL"
        assert abs(c.pobj - w.pobj) <= 5 * EPS * (1 + abs(c.pobj)), (c.pobj, w.pobj)
        k = kkt(p, w.x, w.y)
        assert k["rel_primal"] < 10 * EPS and k["rel_dual"] < 10 * EPS and k["rel_gap"] < 10 * EPS, k


def test_warm_start_cuts_iterations_at_1e6_and_guard(trained):
    # at the target tolerance: fewer iterations on average for the warm
    # starts (single device solves of the same instances), and the guard
    # keeps exactly the starts whose iteration-0 KKT error beats the cold one
    fam, ws, Zte, pte = trained
    sample = list(range(8))
    cold = [pdhcg_amd.solve(pte[k], eps=EPS, iteration_limit=100000) for k in sample]
    warm = [pdhcg_amd.solve(ws.start(pte[k], Zte[k]), eps=EPS, iteration_limit=100000) for k in sample]
    assert all(c.status == w.status == "OPTIMAL" for c, w in zip(cold, warm))
    assert np.mean([w.iterations for w in warm]) < np.mean([c.iterations for c in cold])
    guarded, kept = ws.guarded([pte[k] for k in sample], Zte[sample], batch=False, eps=EPS)
    e0 = lambda p: (lambda s: max(s.rel_primal, s.rel_dual, s.rel_gap))(
        pdhcg_amd.solve(p, eps=EPS, iteration_limit=0))
    for k, p, keep in zip(sample, guarded, kept):
        w = ws.start(pte[k], Zte[k])
        assert keep == (e0(w) < e0(pte[k]))
        assert (p.primal_start is not None) == keep
