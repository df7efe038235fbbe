"""The device scaling passes (scale.hip) against the host passes.

The entry-wise passes of Ruiz, Pock-Chambolle and the bound/objective
scaling, and the transpose of A, run on the device for large problems
(PDHCG_DEVICE_SCALING forces either path).  They must produce the host's
scaled problem bit for bit, so a solve started from either path follows the
same trajectory: the same status, iteration counts and identical iterates.
The SpMV lane shift and column panels are pinned (the autotuner picks them
by timing).  Cases: every objective kind, variable and affine cones (the
cone-preserving merges run on the host between the device passes), PSD
blocks, a Curtis-Reid pass (host) before the device passes, and the large
low-rank workload at reduced scale.
"""
import numpy as np
import pytest

import pdhcg_amd
from pdhcg_amd import PRECISION_FP64, PRECISION_MIXED, workloads
from tests.test_gpu_cones import EXP, POW, PSD, RSOC, SOC, random_conic

pytestmark = pytest.mark.gpu


def _both(p, monkeypatch, precision=PRECISION_FP64, **kw):
    monkeypatch.setenv("PDHCG_SPMV_SHIFT", "0")
    monkeypatch.setenv("PDHCG_SPMV_PANELS", "0")
    out = []
    for flag in ("0", "1"):
        monkeypatch.setenv("PDHCG_DEVICE_SCALING", flag)
        out.append(pdhcg_amd.solve(p, precision=precision, **kw))
    host, dev = out
    assert host is not None and dev is not None
    assert host.status == dev.status, (host.status, dev.status)
    assert (host.iterations, host.inner_iterations) == (dev.iterations, dev.inner_iterations)
    assert np.array_equal(host.x, dev.x) and np.array_equal(host.y, dev.y)
    assert host.pobj == dev.pobj
    return host


@pytest.mark.parametrize("kind", ["lp", "diag", "sparse", "lowrank", "lowrank_diag", "lowrank_dense_d",
                                  "lowrank_sparse_r"])
def test_device_scaling_bit_identical(kind, monkeypatch):
    n, m = (1000, 150) if kind == "lowrank_sparse_r" else (300, 120)
    _both(workloads.random_problem(kind, seed=5, n=n, m=m), monkeypatch, eps=1e-6, iteration_limit=3000)


def test_device_scaling_mixed_precision(monkeypatch):
    _both(workloads.random_problem("lowrank_diag", seed=6, n=400, m=100), monkeypatch, PRECISION_MIXED, eps=1e-6,
          iteration_limit=3000)


@pytest.mark.parametrize("types", [(SOC, RSOC), (EXP, POW), (PSD,)])
def test_device_scaling_with_cones(types, monkeypatch):
    _both(random_conic(31, types), monkeypatch, eps=1e-6, iteration_limit=3000)


def test_device_scaling_affine_cones_and_curtis_reid(monkeypatch):
    p = workloads.conic_mix(scale=0.0005, psd_orders=(16, 16))
    _both(p, monkeypatch, eps=1e-6, iteration_limit=400)
    _both(workloads.random_problem("sparse", seed=8, n=200, m=90), monkeypatch, eps=1e-6, iteration_limit=2000,
          curtis_reid_iterations=5)


def test_device_scaling_large_lowrank(monkeypatch):
    _both(workloads.config1(scale=0.01), monkeypatch, eps=1e-6, iteration_limit=300)
