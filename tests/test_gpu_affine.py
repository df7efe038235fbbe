"""GPU parity of affine cone constraints (F x + g in K on rows of [A; F]).

The reference's one-variable known answers (test/test_affine_cones.c:113-215,
same parameters and acceptance bars), and the tracking-error portfolio (an
affine SOC over D^1/2 R x with a diagonal + low-rank objective) against the
oracle, which is itself pinned to scipy's SLSQP on small instances.
"""
import numpy as np
import pytest

import pdhcg_amd
from pdhcg_amd import workloads
from tests.oracle_lib import oracle_solve
from tests.test_oracle_cones import affine_cases, affine_params, check_affine_result

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", range(4))
def test_affine_cone_known_answers(case):
    name, p, norm, preserving, expected = affine_cases()[case]
    g = pdhcg_amd.solve(p, **affine_params(norm, preserving))
    check_affine_result(g, expected)
    o = oracle_solve(p, **affine_params(norm, preserving))
    assert abs(g.pobj - o.pobj) <= 1e-6 * (1 + abs(o.pobj)), (name, g.pobj, o.pobj)


@pytest.mark.parametrize("n,k,factor", [(200, 20, 1.5), (1000, 50, 1.5), (60, 8, 1.0), (400, 120, 1.5)])
def test_portfolio_tracking_matches_oracle(n, k, factor):
    # k = 120: the tracking-error cone has more than 64 vector slots (one workgroup per block)
    p = workloads.portfolio_tracking(n, k, seed=7 + n, te_factor=factor)
    eps = 1e-6
    g = pdhcg_amd.solve(p, eps=eps, iteration_limit=200000)
    o = oracle_solve(p, eps=eps, iteration_limit=200000)
    assert g.status == o.status == "OPTIMAL", (g.status, o.status)
    assert abs(g.pobj - o.pobj) <= 5 * eps * (1 + abs(o.pobj)), (g.pobj, o.pobj)
    te = np.linalg.norm(np.sqrt(p.D) * (p.R @ g.x))
    assert te <= p.g[-1] * (1 + 1e-4)
    assert abs(g.x.sum() - 1.0) <= 1e-5
