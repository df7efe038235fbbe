"""GPU parity of fixed cone sections.

Operator: pdhcg_project_cones with a fixed-slot mask against the oracle's
section projection (oracle/orc_section.c, pinned by tests/test_oracle_sections.py):
relative agreement 1e-12 of the block's scale, fixed slots bit-identical.
The reference's fixed-mask power golden vectors
(test/test_power_cone_projection.cu:326-403) are also checked on the device.

Solves: the reference's fixed-slot end-to-end cases (tests/fixed_cases.py)
on the device library, and device vs oracle on the same problems (same
status; objectives within 5 eps (1 + |obj|)).
"""
import ctypes as C

import numpy as np
import pytest

import pdhcg_amd
from pdhcg_amd import capi
from tests import fixed_cases as fc
from tests.oracle_lib import oracle_solve
from tests.test_oracle_sections import POW, RSOC, SOC, EXP, cone_point, section as orc_section

pytestmark = pytest.mark.gpu


def _len(t, vd):
    return vd + 2 if t in (SOC, RSOC) else 3


def _section_case(t, rng, sizes, count, end_patterns=False):
    specs, fixed, x, d, blocks, pos = [], [], [], [], [], 0
    for k in range(count):
        vd = int(rng.choice(sizes)) if t in (SOC, RSOC) else 1
        L = _len(t, vd)
        alpha = float(rng.uniform(0.1, 0.9)) if t == POW else 0.0
        anchor = cone_point(t, vd, alpha, rng)
        fx = (rng.uniform(size=L) < rng.choice([0.05, 0.4, 0.9])).astype(np.uint8)
        if end_patterns:  # cycle through the (w/s, z/t) fixing patterns
            fx[vd], fx[vd + 1] = k % 4 // 2, k % 2
        if not fx.any():  # blocks without fixed slots take the plain projection
            fx[int(rng.integers(L))] = 1
        if fx.all():
            fx[0] = 0
        a = anchor.copy()
        free = fx == 0
        a[free] += rng.normal(scale=1.5, size=int(free.sum()))
        dd = rng.uniform(0.3, 3.0, size=L)
        specs.append((t, pos, vd, alpha))
        blocks.append((pos, vd, alpha, fx, a, dd))
        fixed += list(fx) + [0]
        x += list(a * dd) + [rng.normal()]
        d += list(dd) + [1.0]
        pos += L + 1
    return specs, fixed, np.array(x), np.array(d), blocks


def _check_sections(t, specs, fixed, x, d, blocks):
    got = pdhcg_amd.project_cones(specs, x, d, fixed=fixed)
    for pos, vd, alpha, fx, a, dd in blocks:
        L = _len(t, vd)
        exp = orc_section(t, a, vd, dd * dd, fx, alpha) * dd
        g = got[pos:pos + L]
        scale = 1.0 + np.abs(exp).max()
        assert np.all(np.abs(g - exp) <= 1e-12 * scale), (t, vd, fx, a, g, exp)
        assert np.all(g[fx == 1] == x[pos:pos + L][fx == 1])
    # slots outside the blocks are untouched
    gaps = [p + _len(t, vd) for p, vd, *_ in blocks]
    assert np.all(got[gaps] == x[gaps])


@pytest.mark.parametrize("t", [SOC, RSOC, EXP, POW])
def test_section_projection_matches_oracle(t):
    rng = np.random.default_rng(40 + t)
    _check_sections(t, *_section_case(t, rng, [1, 2, 5, 17, 64], 120))


@pytest.mark.parametrize("t", [SOC, RSOC])
def test_long_section_projection_matches_oracle(t):
    # blocks above 64 vector slots: one workgroup per block
    rng = np.random.default_rng(80 + t)
    _check_sections(t, *_section_case(t, rng, [65, 130, 700, 3000], 32, end_patterns=True))


@pytest.mark.parametrize("alpha,inp,w,fx,exp,tol", [
    (0.3, [-0.5, 0.2, 1.0], [2.0, 0.7, 1.0], [0, 0, 1], [0.3619815616314009, 1.545732644202259, 1.0], 3e-5),
    (0.3, [1.0, -0.4, 2.0], [1.0, 2.0, 0.5], [1, 0, 0], [1.0, 0.1445424296046515, 0.2582240838862325], 3e-5),
    (0.99, [0.00093260334688321987, 0.0, 1.0], [1, 1, 1], [1, 0, 0],
     [0.00093260334688321987, 0.0030706556671616642, 0.00094378334963825893], 1e-9),
    (0.7, [-0.3, 1.0, -1.8], [3.0, 1.0, 0.8], [0, 1, 0], [0.1744696096520296, 1.0, -0.2945804034203198], 3e-5),
    (0.3, [2.0, 3.0, 4.0], [2.0, 7.0, 0.25], [1, 1, 0], [2.0, 3.0, 2.656402479886323], 1e-12),
])
def test_power_section_golden_on_device(alpha, inp, w, fx, exp, tol):
    # reference test/test_power_cone_projection.cu:326-403, weights as d = sqrt(w)
    d = np.sqrt(np.array(w))
    got = pdhcg_amd.project_cones([(POW, 0, 1, alpha)], np.array(inp) * d, d, fixed=fx) / d
    exp = np.array(exp)
    assert np.all(np.abs(got - exp) <= tol * (1.0 + np.abs(exp))), (got, exp)


def test_dual_residual_entry_rejects_fixed_blocks():
    L = pdhcg_amd.lib()
    b, _keep = capi.cone_blocks([(SOC, 0, 1)], fixed=[0, 0, 1])
    rc = np.zeros(3)
    out = np.zeros(3)
    assert L.pdhcg_cone_dual_residual(C.byref(b), None, capi.as_ptr(rc), capi.as_ptr(out), 3) == -1


# ---- whole solves (reference known answers, device and oracle) ----
@pytest.mark.parametrize("name,alpha,obj,fixed,expected", fc.POWER_FIXED, ids=[c[0] for c in fc.POWER_FIXED])
def test_power_fixed_known_answers(name, alpha, obj, fixed, expected):
    p = fc.power_free(alpha, obj, fixed)
    g = pdhcg_amd.solve(p, **fc.TINY)
    assert g is not None and g.status == "OPTIMAL", name
    assert fc.check_close(g.x, expected, 3e-5), (name, g.x, expected)
    for slot, v in fixed:
        assert g.x[slot] == v
    o = oracle_solve(p, **fc.TINY)
    assert abs(g.pobj - o.pobj) <= 5e-8 * (1 + abs(o.pobj)), (g.pobj, o.pobj)


def test_power_infeasible_fully_fixed_is_rejected():
    assert pdhcg_amd.solve(fc.infeasible_fully_fixed(), **fc.TINY) is None


def test_power_diag_q_fixed_z():
    p, expected = fc.DIAG_Q_FIXED_Z
    g = pdhcg_amd.solve(p, **fc.TINY)
    assert g.status == "OPTIMAL" and fc.check_close(g.x, expected, 2e-4), g.x


def test_power_sharp_fixed_axis():
    g = pdhcg_amd.solve(fc.sharp_fixed_axis(), **fc.TINY)
    assert g.status == "OPTIMAL"
    x, y, z = g.x
    bound = x ** 0.97 * y ** 0.03 if x > 0 and y > 0 else 0.0
    assert x == fc.SHARP_START[0]
    assert max(0.0, -x, -y, abs(z) - bound) <= 1e-12
    assert g.rel_primal <= 1e-8


@pytest.mark.parametrize("norm", [capi.NORM_TYPE_L_INF, capi.NORM_TYPE_L2])
@pytest.mark.parametrize("v_dim", [1, 32, 64, 32768])
def test_zero_w_soc_optimal_at_start(norm, v_dim):
    g = pdhcg_amd.solve(fc.zero_w_soc(v_dim), optimality_norm=norm, **fc.ONE_STEP)
    assert g.status == "OPTIMAL" and g.iterations == 0, (g.status, g.iterations)


@pytest.mark.parametrize("norm", [capi.NORM_TYPE_L_INF, capi.NORM_TYPE_L2])
@pytest.mark.parametrize("kind", sorted(fc.LARGE_STEP))
def test_large_first_step_not_accepted(kind, norm):
    g = pdhcg_amd.solve(fc.LARGE_STEP[kind](), optimality_norm=norm, **fc.ONE_STEP)
    assert g.status == "ITERATION_LIMIT" and g.iterations == 1, (kind, g.status, g.iterations)


def test_exp_fixed_y_reaches_log2():
    g = pdhcg_amd.solve(fc.exp_log2(), **fc.EXP_LOG2)
    assert g.status == "OPTIMAL" and g.iterations > 0
    assert abs(g.x[0] - fc.LOG2) <= 1e-7, g.x
