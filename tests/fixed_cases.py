"""Problems with fixed cone slots and their known answers, restated from the
reference's own tests; run against the oracle (test_oracle_fixed.py) and the
device library (test_gpu_fixed.py) through the same pdhcg.h interface.

- test/test_power_cone.c:150-350: fixed power-cone sections with closed-form
  optima, the fully fixed round-off boundary point, the infeasible fully fixed
  point (rejected), the diagonal-Q fixed-z case and the sharp fixed axis.
- test/test_conic_kkt_termination.c:171-575: an SOC with only w = 0 fixed is
  optimal at iteration 0; large first steps from feasible interior warm
  starts of fixed RSOC / SOC / POWER / EXP sections are not accepted after one
  iteration; min -z over (z, 1, 2) in K_exp with y fixed reaches z = log 2.
"""
import math

import numpy as np
import scipy.sparse as sp

from pdhcg_amd import QPData, capi

POW, SOC, RSOC, EXP = capi.CONE_POWER, capi.CONE_STANDARD_SOC, capi.CONE_ROTATED_SOC, capi.CONE_EXPONENTIAL

# solve_tiny (reference test/test_power_cone.c:66-82)
TINY = dict(optimality_norm=capi.NORM_TYPE_L_INF, termination_evaluation_frequency=10, eps=1e-8,
            iteration_limit=1000000, time_limit=30.0)


def _no_rows(n):
    return sp.csr_matrix((0, n))


def power_free(alpha, objective, fixed=()):
    return QPData(c=np.array(objective, float), A=_no_rows(3), cones=[(POW, 0, 1, alpha)],
                  cone_fixed=[(0, s, v) for s, v in fixed])


_lam = 1.0 / (0.3 ** 0.3 * 0.7 ** 0.7)
POWER_FIXED = [
    # name, alpha, objective, fixed (slot, value), expected
    ("fixed z", 0.3, [1.0, 1.0, 0.0], [(2, 1.0)], [0.3 * _lam, 0.7 * _lam, 1.0]),
    ("fixed x", 0.3, [0.0, 0.7, -1.0], [(0, 1.0)], [1.0, 1.0, 1.0]),
    ("fixed y", 0.7, [0.7, 0.0, -1.0], [(1, 1.0)], [1.0, 1.0, 1.0]),
    ("fixed x,z", 0.3, [0.0, 1.0, 0.0], [(0, 1.0), (2, 1.0)], [1.0, 1.0, 1.0]),
    ("fixed y,z", 0.7, [1.0, 0.0, 0.0], [(1, 1.0), (2, 1.0)], [1.0, 1.0, 1.0]),
    ("fixed x,y", 0.3, [0.0, 0.0, -1.0], [(0, 2.0), (1, 3.0)], [2.0, 3.0, 2.0 ** 0.3 * 3.0 ** 0.7]),
    ("fixed zero z", 0.3, [1.0, 1.0, 0.0], [(2, 0.0)], [0.0, 0.0, 0.0]),
    ("fully fixed roundoff boundary", 0.86039292839558623, [0.0, 0.0, 0.0],
     [(0, 4.4414605580442319e83), (1, 2.0775280372919238e-95), (2, 5.6415660092721006e58)],
     [4.4414605580442319e83, 2.0775280372919238e-95, 5.6415660092721006e58]),
]


def power_quadratic(alpha, center, weights, fixed=()):
    """min sum w_i (x_i - center_i)^2 / 2 over K_alpha (reference :27-63)."""
    w = np.array(weights, float)
    return QPData(c=-w * np.array(center, float), Q=sp.diags(w).tocsr(), A=_no_rows(3), cones=[(POW, 0, 1, alpha)],
                  cone_fixed=[(0, s, v) for s, v in fixed])


DIAG_Q_FIXED_Z = (power_quadratic(0.3, [-0.5, 0.2, 1.0], [2.0, 0.7, 1.0], [(2, 1.0)]),
                  [0.3619815616314009, 1.545732644202259, 1.0])

SHARP_START = [0.62720941, -0.01546521, 0.19700471]


def sharp_fixed_axis():
    p = power_quadratic(0.97, [0.62147354, -0.01546521, 0.19700471], [0.131333592, 0.0263303077, 38.3559275],
                        [(0, SHARP_START[0])])
    p.primal_start = np.array(SHARP_START)
    return p


def infeasible_fully_fixed():
    return power_free(0.3, [0.0, 0.0, 0.0], [(0, 1.0), (1, 1.0), (2, 2.0)])


def check_close(x, expected, tol):
    e = np.asarray(expected, float)
    return bool(np.all(np.abs(np.asarray(x) - e) <= tol * (1.0 + np.abs(e))))


# reference test/test_conic_kkt_termination.c:171-575
ONE_STEP = dict(l_inf_ruiz_iterations=0, has_pock_chambolle_alpha=False, bound_objective_rescaling=False,
                termination_evaluation_frequency=1, eps=1e-8, iteration_limit=1)


def zero_w_soc(v_dim):
    c = np.zeros(v_dim + 2)
    c[v_dim] = 1.0
    return QPData(c=c, A=_no_rows(v_dim + 2), cones=[(SOC, 0, v_dim)], cone_fixed=[(0, v_dim, 0.0)],
                  primal_start=np.zeros(v_dim + 2))


def _one_row(col, coef, rhs, objective, cone, fixed, start):
    A = sp.csr_matrix(([coef], ([0], [col])), shape=(1, 3))
    return QPData(c=np.array(objective, float), A=A, con_lb=[rhs], con_ub=[rhs], cones=[cone],
                  cone_fixed=[(0, s, v) for s, v in fixed], primal_start=np.array(start, float), dual_start=[0.0])


LARGE_STEP = {
    "rsoc": lambda: _one_row(1, 1e-9, 1e-9, [-1.0, 0.0, 0.0], (RSOC, 0, 1), [(1, 1.0), (2, 1.0)], [0.5, 1.0, 1.0]),
    "soc": lambda: _one_row(2, 1e-9, 1e-9, [-1.0, 0.0, 0.0], (SOC, 0, 1), [(2, 1.0)], [0.5, 0.0, 1.0]),
    "power": lambda: _one_row(0, 1e-9, 1e-9, [0.0, 0.0, -1.0], (POW, 0, 1, 0.5), [(0, 1.0), (1, 1.0)],
                              [1.0, 1.0, 0.5]),
    "exp": lambda: _one_row(0, 1e-10, 0.0, [0.0, 0.0, 1.0], (EXP, 0, 1), [(1, 1.0)], [0.0, 1.0, 2.0]),
}


def exp_log2():
    """min -z over (z, 1, 2) in K_exp, y fixed; z* = log 2 (reference :506-575)."""
    return _one_row(2, 1e-12, 2e-12, [-1.0, 0.0, 0.0], (EXP, 0, 1), [(1, 1.0)], [0.0, 1.0, 2.0])


EXP_LOG2 = dict(l_inf_ruiz_iterations=0, has_pock_chambolle_alpha=False, bound_objective_rescaling=False, eps=1e-8)
LOG2 = math.log(2.0)
