"""Pins the oracle's fixed-section projections (oracle/orc_section.c).

- Golden vectors: the fixed-mask power-cone cases of reference
  test/test_power_cone_projection.cu:326-403 (weights w enter as the scaling
  d = sqrt(w), so the metric is w, as in that test's harness :116-240).
- Independent optimum: random sections of SOC, RSOC, EXP and POWER blocks
  (fixed values taken from a point of the cone, so the section is not empty)
  against scipy's SLSQP on the same weighted problem, and the variational
  inequality <W(a - p), c - p> <= 0 over feasible points c of the section
  (the reference's check, test/test_fixed_cone_sections.cu:124-330).
- Closed forms: radial shrink for a fixed SOC apex value with uniform
  weights, the RSOC ball with both ends fixed, and the one-free-slot clamps.
- The residual of a block with fixed slots vanishes at a section point with
  rc = 0 and equals the prox residual otherwise.
"""
import ctypes as C

import numpy as np
import pytest
from scipy.optimize import minimize

from pdhcg_amd import capi
from tests.oracle_lib import oracle

dp = C.POINTER(C.c_double)
SOC, RSOC, EXP, POW = capi.CONE_STANDARD_SOC, capi.CONE_ROTATED_SOC, capi.CONE_EXPONENTIAL, capi.CONE_POWER


def _L():
    L = oracle()
    L.orc_section_project.restype = C.c_int
    L.orc_section_project.argtypes = [C.c_int, dp, C.c_int, dp, dp, C.c_char_p, C.c_double]
    L.orc_section_dual_residual.restype = None
    L.orc_section_dual_residual.argtypes = [C.c_int, dp, dp, C.c_int, dp, C.c_char_p, C.c_double, dp]
    return L


def section(t, a, vd, w, fixed, alpha=0.0, metric=None):
    """Projects actual values a (weights w) onto the section; returns actual values."""
    d = np.sqrt(np.asarray(w, float))
    x = np.asarray(a, float) * d
    m = None if metric is None else capi.as_ptr(np.ascontiguousarray(metric, float))
    assert _L().orc_section_project(t, capi.as_ptr(x), vd, capi.as_ptr(d), m, bytes(fixed), alpha) == 0
    return x / d


# ---- reference golden vectors (test_power_cone_projection.cu:326-403) ----
_sharp_y = (0.19700471 / 0.62720941 ** 0.97) ** (1.0 / 0.03)
_ultra_y = 0.5 ** (1.0 / (1.0 - 0.999))


@pytest.mark.parametrize("alpha,inp,w,fx,exp,tol", [
    (0.3, [-0.5, 0.2, 1.0], [2.0, 0.7, 1.0], [0, 0, 1], [0.3619815616314009, 1.545732644202259, 1.0], 3e-5),
    (0.3, [-0.5e200, 0.2e200, 1.0e200], [2.0, 0.7, 1.0], [0, 0, 1],
     [0.3619815616314009e200, 1.545732644202259e200, 1.0e200], 3e-5),
    (0.5, [0.0, 0.0, 1.0], [1e200, 1e200, 1.0], [0, 0, 1], [1.0, 1.0, 1.0], 1e-10),
    (0.5, [0.0, 0.0, 1.0], [1e-200, 1e-200, 1.0], [0, 0, 1], [1.0, 1.0, 1.0], 1e-10),
    (0.3, [1.0, -0.4, 2.0], [1.0, 2.0, 0.5], [1, 0, 0], [1.0, 0.1445424296046515, 0.2582240838862325], 3e-5),
    (0.97, [0.62720941, -0.01546521, 0.19700471], [1, 1, 1], [1, 0, 0], [0.62720941, _sharp_y, 0.19700471], 1e-10),
    (0.999, [1.0, -0.1, 0.5], [1, 1, 1], [1, 0, 0], [1.0, _ultra_y, 0.5], 1e-10),
    (0.99, [0.00093260334688321987, 0.0, 1.0], [1, 1, 1], [1, 0, 0],
     [0.00093260334688321987, 0.0030706556671616642, 0.00094378334963825893], 1e-9),
    (0.7, [-0.3, 1.0, -1.8], [3.0, 1.0, 0.8], [0, 1, 0], [0.1744696096520296, 1.0, -0.2945804034203198], 3e-5),
    (0.3, [1.0, -0.4, 1.0], [1.0, 2.0, 0.5], [1, 0, 1], [1.0, 1.0, 1.0], 1e-12),
    (0.7, [-0.3, 1.0, -1.0], [3.0, 1.0, 0.8], [0, 1, 1], [1.0, 1.0, -1.0], 1e-12),
    (0.3, [2.0, 3.0, 4.0], [2.0, 7.0, 0.25], [1, 1, 0], [2.0, 3.0, 2.656402479886323], 1e-12),
    (0.3, [1.0, 1.0, 0.5], [1, 1, 1], [1, 1, 1], [1.0, 1.0, 0.5], 1e-12),
])
def test_power_section_golden(alpha, inp, w, fx, exp, tol):
    got = section(POW, inp, 1, w, fx, alpha)
    exp = np.array(exp)
    assert np.all(np.abs(got - exp) <= tol * (1.0 + np.abs(exp))), (got, exp)
    d = np.sqrt(np.asarray(w, float))
    for i in range(3):
        if fx[i]:
            assert got[i] == (inp[i] * d[i]) / d[i]


# ---- membership and an independent solver ----
def member(t, p, vd, alpha, tol=1e-9):
    if t == SOC:
        return np.linalg.norm(p[:vd + 1]) <= p[vd + 1] + tol
    if t == RSOC:
        s, u = p[vd], p[vd + 1]
        return s >= -tol and u >= -tol and p[:vd] @ p[:vd] <= 2 * s * u + tol
    if t == EXP:
        x, y, z = p
        if y > 1e-12:
            return y * np.exp(min(x / y, 700)) <= z + tol * (1 + abs(z))
        return abs(y) <= tol and x <= tol and z >= -tol
    x, y, z = p
    return x >= -tol and y >= -tol and abs(z) <= max(x, 0) ** alpha * max(y, 0) ** (1 - alpha) + tol


def cone_point(t, vd, alpha, rng):
    """A random point of the cone, away from its boundary."""
    if t == SOC:
        u = rng.normal(size=vd + 1)
        return np.concatenate([u, [np.linalg.norm(u) + rng.uniform(0.1, 1.0)]])
    if t == RSOC:
        v = rng.normal(size=vd)
        s = rng.uniform(0.2, 1.5)
        return np.concatenate([v, [s, v @ v / (2 * s) + rng.uniform(0.1, 1.0)]])
    if t == EXP:
        y = rng.uniform(0.2, 1.4)
        x = rng.uniform(-1.2, 1.2)
        return np.array([x, y, y * np.exp(x / y) + rng.uniform(0.05, 1.0)])
    x, y = rng.uniform(0.2, 2.0, size=2)
    b = x ** alpha * y ** (1 - alpha)
    return np.array([x, y, rng.uniform(-0.8, 0.8) * b])


def slsqp(t, a, w, fx, vd, alpha, starts=()):
    """Best feasible SLSQP optimum from a and the given starting points (the
    problem is convex: a point SLSQP cannot improve on is the optimum)."""
    free = [i for i in range(len(a)) if not fx[i]]

    def full(z):
        p = a.copy()
        p[free] = z
        return p

    def obj(z):
        return float(np.sum(w[free] * (z - a[free]) ** 2))

    def jac(z):
        return 2 * w[free] * (z - a[free])

    if t == SOC:
        cons = [{"type": "ineq", "fun": lambda z: full(z)[vd + 1] ** 2 - full(z)[:vd + 1] @ full(z)[:vd + 1]},
                {"type": "ineq", "fun": lambda z: full(z)[vd + 1]}]
    elif t == RSOC:
        cons = [{"type": "ineq", "fun": lambda z: 2 * full(z)[vd] * full(z)[vd + 1] - full(z)[:vd] @ full(z)[:vd]},
                {"type": "ineq", "fun": lambda z: full(z)[vd]}, {"type": "ineq", "fun": lambda z: full(z)[vd + 1]}]
    elif t == EXP:
        cons = [{"type": "ineq", "fun": lambda z: full(z)[2] - max(full(z)[1], 1e-12)
                 * np.exp(min(full(z)[0] / max(full(z)[1], 1e-12), 700))},
                {"type": "ineq", "fun": lambda z: full(z)[1]}]
    else:
        def g(z):
            p = full(z)
            return max(p[0], 0) ** alpha * max(p[1], 0) ** (1 - alpha)
        cons = [{"type": "ineq", "fun": lambda z: g(z) - full(z)[2]}, {"type": "ineq", "fun": lambda z: g(z) + full(z)[2]},
                {"type": "ineq", "fun": lambda z: full(z)[0]}, {"type": "ineq", "fun": lambda z: full(z)[1]}]
    best = None
    rng = np.random.default_rng(0)
    z0s = [a[free]] + [np.asarray(x, float)[free] for x in starts]
    z0s += [a[free] + rng.normal(scale=0.5, size=len(free)) for _ in range(4)]
    for z0 in z0s:
        r = minimize(obj, z0, jac=jac, constraints=cons, method="SLSQP", options={"ftol": 1e-15, "maxiter": 1000})
        p = full(r.x)
        if member(t, p, vd, alpha, 1e-7) and (best is None or r.fun < best[0]):
            best = (r.fun, p)
    return best


CASES = []
for _t, _vd in ((SOC, 1), (SOC, 3), (RSOC, 1), (RSOC, 3), (EXP, 1), (POW, 1)):
    for _seed in range(6):
        CASES.append((_t, _vd, _seed))


@pytest.mark.parametrize("t,vd,seed", CASES)
def test_section_matches_independent_optimum(t, vd, seed):
    rng = np.random.default_rng(100 * t + 10 * vd + seed)
    L = vd + 2 if t in (SOC, RSOC) else 3
    alpha = rng.uniform(0.2, 0.8) if t == POW else 0.0
    anchor = cone_point(t, vd, alpha, rng)
    for trial in range(4):
        fx = list((rng.uniform(size=L) < 0.4).astype(np.uint8))
        if not any(fx) or all(fx):
            fx[int(rng.integers(L))] ^= 1
        if all(fx):
            fx[0] = 0
        a = anchor.copy()
        free = [i for i in range(L) if not fx[i]]
        a[free] = anchor[free] + rng.normal(scale=1.5, size=len(free))
        w = rng.uniform(0.3, 3.0, size=L)
        p = section(t, a, vd, w, fx, alpha)
        d = np.sqrt(w)
        for i in range(L):
            if fx[i]:  # untouched in the scaled space
                assert p[i] == (a[i] * d[i]) / d[i]
        assert member(t, p, vd, alpha, 1e-9), (t, fx, a, p)
        f_p = float(np.sum(w[free] * (p[free] - a[free]) ** 2))
        ref = slsqp(t, a, w, fx, vd, alpha, starts=(anchor, p))
        assert ref is not None
        assert f_p <= ref[0] * (1 + 1e-6) + 1e-10, (t, fx, a, p, ref)
        # variational inequality over feasible points of the section
        for k in range(40):
            c = cone_point(t, vd, alpha, rng)
            c[[i for i in range(L) if fx[i]]] = a[[i for i in range(L) if fx[i]]]
            if not member(t, c, vd, alpha, 0.0):
                continue
            assert np.dot(w * (a - p), c - p) <= 1e-8 * (1 + np.abs(a).max() ** 2), (t, fx, a, p, c)


def test_soc_fixed_apex_uniform_weights_is_radial():
    # ||u|| <= z with z and u_0 fixed, uniform weights: the free part shrinks radially
    a = np.array([0.6, 2.0, -1.5, 0.0, 1.0])
    fx = [1, 0, 0, 1, 1]
    p = section(SOC, a, 3, np.ones(5), fx)
    r = np.sqrt(1.0 - 0.36)
    exp = a.copy()
    exp[1:3] = a[1:3] * r / np.linalg.norm(a[1:3])
    assert np.allclose(p, exp, rtol=0, atol=1e-14)


def test_rsoc_both_ends_fixed_is_ball():
    a = np.array([3.0, -4.0, 0.5, 2.0])
    p = section(RSOC, a, 2, np.ones(4), [0, 0, 1, 1])
    exp = a.copy()
    exp[:2] = a[:2] * np.sqrt(2.0) / 5.0
    assert np.allclose(p, exp, rtol=0, atol=1e-14)


def test_one_free_slot_clamps():
    # EXP (x, y) fixed: z >= y exp(x / y); (y, z) fixed: x <= y log(z / y)
    assert section(EXP, [0.5, 2.0, 1.0], 1, np.ones(3), [1, 1, 0])[2] == pytest.approx(2.0 * np.exp(0.25), rel=1e-15)
    assert section(EXP, [3.0, 2.0, 5.0], 1, np.ones(3), [0, 1, 1])[0] == pytest.approx(2.0 * np.log(2.5), rel=1e-15)
    # POWER (x, y) fixed: |z| <= x^a y^(1-a)
    assert section(POW, [2.0, 3.0, -9.0], 1, np.ones(3), [1, 1, 0], 0.4)[2] == pytest.approx(
        -(2.0 ** 0.4) * 3.0 ** 0.6, rel=1e-15)
    # EXP (x, z) fixed with x > 0: y in [y1, y2] around x, both roots of log y + x / y = log z
    p = section(EXP, [1.0, 50.0, 4.0], 1, np.ones(3), [1, 0, 1])
    y = p[1]
    assert y < 50.0 and np.log(y) + 1.0 / y == pytest.approx(np.log(4.0), abs=1e-13)


def test_weighted_metric_argument():
    # metric m scales the weights: the section projection in metric m d^2
    rng = np.random.default_rng(7)
    a = np.array([0.3, -1.2, 0.8, 0.4])
    w = rng.uniform(0.5, 2.0, size=4)
    m = rng.uniform(0.5, 2.0, size=4)
    fx = [0, 0, 1, 0]
    p1 = section(SOC, a, 2, w, fx, metric=m)
    p2 = section(SOC, a, 2, w * m, fx)
    assert np.allclose(p1, p2, rtol=1e-13, atol=1e-15)


@pytest.mark.parametrize("t,vd", [(SOC, 2), (RSOC, 2), (EXP, 1), (POW, 1)])
def test_section_residual(t, vd):
    rng = np.random.default_rng(11 + t)
    L = vd + 2 if t in (SOC, RSOC) else 3
    alpha = 0.35 if t == POW else 0.0
    x = cone_point(t, vd, alpha, rng)
    fx = [0] * L
    fx[L - 1] = 1
    d = rng.uniform(0.5, 2.0, size=L)
    out = np.zeros(L)
    xs = x * d
    _L().orc_section_dual_residual(t, capi.as_ptr(xs), capi.as_ptr(np.zeros(L)), vd, capi.as_ptr(d), bytes(fx), alpha,
                                  capi.as_ptr(out))
    assert np.all(np.abs(out) <= 1e-13 * (1 + np.abs(xs).max()))
    rc = rng.normal(size=L)
    _L().orc_section_dual_residual(t, capi.as_ptr(xs), capi.as_ptr(rc), vd, capi.as_ptr(d), bytes(fx), alpha,
                                  capi.as_ptr(out))
    t_in = np.where(fx, xs, xs - rc)
    proj = section(t, t_in / d, vd, d * d, fx, alpha) * d
    assert np.allclose(out, np.where(fx, 0.0, (xs - proj) * d), rtol=1e-12, atol=1e-14)
