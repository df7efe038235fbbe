"""Whole solves with fixed cone slots on the oracle against the reference's
known answers (cases in tests/fixed_cases.py)."""
import numpy as np
import pytest

from pdhcg_amd import capi
from tests import fixed_cases as fc
from tests.oracle_lib import oracle_solve


@pytest.mark.parametrize("name,alpha,obj,fixed,expected", fc.POWER_FIXED, ids=[c[0] for c in fc.POWER_FIXED])
def test_power_fixed_known_answers(name, alpha, obj, fixed, expected):
    s = oracle_solve(fc.power_free(alpha, obj, fixed), **fc.TINY)
    assert s is not None and s.status == "OPTIMAL", name
    assert fc.check_close(s.x, expected, 3e-5), (name, s.x, expected)
    for slot, v in fixed:
        assert s.x[slot] == v


def test_power_infeasible_fully_fixed_is_rejected():
    assert oracle_solve(fc.infeasible_fully_fixed(), **fc.TINY) is None


def test_power_diag_q_fixed_z():
    p, expected = fc.DIAG_Q_FIXED_Z
    s = oracle_solve(p, **fc.TINY)
    assert s.status == "OPTIMAL" and fc.check_close(s.x, expected, 2e-4), s.x


def test_power_sharp_fixed_axis():
    s = oracle_solve(fc.sharp_fixed_axis(), **fc.TINY)
    assert s.status == "OPTIMAL"
    x, y, z = s.x
    bound = x ** 0.97 * y ** 0.03 if x > 0 and y > 0 else 0.0
    assert x == fc.SHARP_START[0]
    assert max(0.0, -x, -y, abs(z) - bound) <= 1e-12
    assert s.rel_primal <= 1e-8


@pytest.mark.parametrize("norm", [capi.NORM_TYPE_L_INF, capi.NORM_TYPE_L2])
@pytest.mark.parametrize("v_dim", [1, 32, 32768])
def test_zero_w_soc_optimal_at_start(norm, v_dim):
    s = oracle_solve(fc.zero_w_soc(v_dim), optimality_norm=norm, **fc.ONE_STEP)
    assert s.status == "OPTIMAL" and s.iterations == 0, (s.status, s.iterations)


@pytest.mark.parametrize("norm", [capi.NORM_TYPE_L_INF, capi.NORM_TYPE_L2])
@pytest.mark.parametrize("kind", sorted(fc.LARGE_STEP))
def test_large_first_step_not_accepted(kind, norm):
    s = oracle_solve(fc.LARGE_STEP[kind](), optimality_norm=norm, **fc.ONE_STEP)
    assert s.status == "ITERATION_LIMIT" and s.iterations == 1, (kind, s.status, s.iterations)


def test_exp_fixed_y_reaches_log2():
    s = oracle_solve(fc.exp_log2(), **fc.EXP_LOG2)
    assert s.status == "OPTIMAL" and s.iterations > 0
    assert abs(s.x[0] - fc.LOG2) <= 1e-7, s.x
