"""pdhcg_project_cones_noalloc: the allocation-free variant of
pdhcg_project_cones (its buffers live in a fixed device array of the library)
against the allocating entry point, through the C interface.

The two run the same kernels on the same inputs, so the outputs must be
bit-identical.  Two calls in a row (the second smaller, over a scratch
holding the first call's contents), every cone type including PSD blocks,
SOC/RSOC blocks above 64 vector slots and a block with a fixed slot; then a
request larger than the fixed array is refused and the next call still
matches.
"""
import ctypes as C

import numpy as np
import pytest

import pdhcg_amd
from pdhcg_amd import capi

pytestmark = pytest.mark.gpu

SOC, RSOC, EXP, POW, PSD = (capi.CONE_STANDARD_SOC, capi.CONE_ROTATED_SOC, capi.CONE_EXPONENTIAL, capi.CONE_POWER,
                            capi.CONE_PSD)
SLOTS = {SOC: lambda v: v + 2, RSOC: lambda v: v + 2, EXP: lambda v: 3, POW: lambda v: 3, PSD: lambda v: v * (v + 1) // 2}


def _lib():
    L = pdhcg_amd.lib()
    L.pdhcg_project_cones_noalloc.restype = C.c_int
    L.pdhcg_project_cones_noalloc.argtypes = [C.POINTER(capi.cone_blocks_t), capi.c_double_p, capi.c_double_p,
                                              C.c_int]
    return L


def _layout(rng, count, psd_orders, long_sizes):
    blocks = []
    for _ in range(count):
        t = rng.choice([SOC, RSOC, EXP, POW])
        if t in (SOC, RSOC):
            blocks.append((t, int(rng.integers(1, 65))))
        elif t == POW:
            blocks.append((t, 1, float(rng.uniform(0.1, 0.9))))
        else:
            blocks.append((t, 1))
    blocks += [(PSD, k) for k in psd_orders] + [(SOC, v) for v in long_sizes] + [(RSOC, v) for v in long_sizes]
    order = rng.permutation(len(blocks))
    specs, pos = [], 0
    for i in order:
        b = blocks[i]
        specs.append((b[0], pos, b[1]) + tuple(b[2:]))
        pos += SLOTS[b[0]](b[1]) + 1
    return specs, pos


def _noalloc(L, specs, x, d, fixed=None):
    b, _keep = capi.cone_blocks(specs, fixed)
    out = np.ascontiguousarray(x, dtype=np.float64).copy()
    rc = L.pdhcg_project_cones_noalloc(C.byref(b), capi.as_ptr(d), capi.as_ptr(out), len(out))
    return rc, out


def _check(L, rng, count, psd_orders, long_sizes, with_fixed):
    specs, n = _layout(rng, count, psd_orders, long_sizes)
    x = rng.standard_normal(n) * rng.choice([0.1, 1.0, 10.0], size=n)
    d = rng.uniform(0.5, 2.0, n)
    for s in specs:  # one scale per PSD block (the PSD projection is onto d K with uniform d)
        if s[0] == PSD:
            d[s[1]:s[1] + SLOTS[PSD](s[2])] = d[s[1]]
    fixed = None
    if with_fixed:
        fixed = np.zeros(n, dtype=np.int8)
        soc = [s for s in specs if s[0] == SOC and s[2] >= 3]
        fixed[soc[0][1]] = 1  # the first vector slot of one SOC block keeps its value
    print(f"noalloc call: {len(specs)} blocks, n = {n}", flush=True)
    rc, got = _noalloc(L, specs, x, d, fixed)
    assert rc == 0
    want = pdhcg_amd.project_cones(specs, x, d, fixed)
    assert np.array_equal(got, want), np.max(np.abs(got - want))
    return n


def test_noalloc_matches_allocating_entry_two_calls():
    L = _lib()
    rng = np.random.default_rng(2024)
    n1 = _check(L, rng, 3000, (5, 40, 70, 128), (100, 700), True)
    n2 = _check(L, rng, 200, (3, 16), (90,), False)
    assert n2 < n1


def test_noalloc_rejects_oversize_then_works():
    L = _lib()
    n = 3_000_000  # the point and its scaling alone need 48 MB
    specs = [(SOC, 0, 1)]
    rc, out = _noalloc(L, specs, np.ones(n), np.ones(n))
    assert rc == -1
    assert np.array_equal(out, np.ones(n))  # nothing was written back
    _check(L, np.random.default_rng(7), 50, (8,), (), False)
