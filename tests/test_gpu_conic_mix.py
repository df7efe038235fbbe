"""The conic-mix workload (workloads.conic_mix: SOC, RSOC, exponential, power
and PSD blocks, an affine SOC block F x + g, sparse Q) on the GPU.

- Reduced size (n = 363: 17 exponential, 8 power, 1 PSD block of order 16,
  SOC/RSOC blocks, 90 rows of A, affine SOC rows): device against the oracle
  to 1e-6, both precisions (status, objective within 5 eps).
- Full size (n = 572,102): three FP64 iterations whose returned point lies
  in every cone block (membership checked per block in numpy: SOC norm,
  RSOC product, exponential and power inequalities, PSD eigenvalues).
"""
import numpy as np
import pytest

import pdhcg_amd
from pdhcg_amd import PRECISION_FP64, PRECISION_MIXED, capi, workloads
from tests.oracle_lib import oracle_solve
from tests.test_oracle_cones import _svec_to_mat

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", [PRECISION_FP64, PRECISION_MIXED])
def test_small_conic_mix_matches_oracle(precision):
    p = workloads.conic_mix(scale=0.0005, psd_orders=(16, 16))
    eps = 1e-6
    g = pdhcg_amd.solve(p, precision=precision, eps=eps, iteration_limit=100000)
    o = oracle_solve(p, eps=eps, iteration_limit=100000)
    assert g.status == o.status == "OPTIMAL", (g.status, o.status)
    assert abs(g.pobj - o.pobj) <= 5 * eps * (1 + abs(o.pobj)), (g.pobj, o.pobj)


def _violation(t, v, vd, alpha):
    scale = 1.0 + np.max(np.abs(v))
    if t == capi.CONE_STANDARD_SOC:
        return max(0.0, np.linalg.norm(v[:-1]) - v[-1]) / scale
    if t == capi.CONE_ROTATED_SOC:
        s, u = v[vd], v[vd + 1]
        return max(0.0, v[:vd] @ v[:vd] - 2 * s * u, -s, -u) / scale ** 2
    if t == capi.CONE_EXPONENTIAL:
        x, y, z = v
        if y > 1e-12 * scale:
            return max(0.0, y * np.exp(min(x / y, 700.0)) - z) / scale
        return max(0.0, x, -z, -y) / scale
    if t == capi.CONE_POWER:
        x, y, z = v
        if x < 0 or y < 0:
            return max(-x, -y) / scale
        return max(0.0, abs(z) - x ** alpha * y ** (1 - alpha)) / scale
    return max(0.0, -np.linalg.eigvalsh(_svec_to_mat(v, vd)).min()) / scale


def test_full_size_conic_mix_point_in_cones():
    # three outer iterations: every BB inner step projects all 572,102 slots
    # (the 67 PSD blocks of order 16..128 in one launch); the returned point
    # must lie in every block whatever the iteration count
    p = workloads.conic_mix()
    print(f"conic mix n={p.c.size}: 3 iterations", flush=True)
    s = pdhcg_amd.solve(p, eps=1e-6, time_limit=90.0, iteration_limit=3)
    assert s is not None and s.status in ("OPTIMAL", "TIME_LIMIT", "ITERATION_LIMIT"), s and s.status
    assert np.all(np.isfinite(s.x)) and np.all(np.isfinite(s.y))
    worst = 0.0
    for c in p.cones:
        t, st, vd = c[0], c[1], c[2]
        L = {capi.CONE_STANDARD_SOC: vd + 2, capi.CONE_ROTATED_SOC: vd + 2, capi.CONE_EXPONENTIAL: 3,
             capi.CONE_POWER: 3, capi.CONE_PSD: vd * (vd + 1) // 2}[t]
        worst = max(worst, _violation(t, s.x[st:st + L], vd, c[3] if len(c) > 3 else 0.0))
    print(f"status {s.status}, {s.iterations} iterations, worst cone violation {worst:.2e}", flush=True)
    assert worst <= 1e-9, worst
