"""GPU parity of the cone operators and of conic solves.

Operators: pdhcg_project_cones / pdhcg_cone_dual_residual of the MI355X
library against the oracle's per-block routines (which are pinned by the
reference's golden vectors and the Moreau / eigh checks of
test_oracle_cones.py).  Bars: SOC, RSOC, power and PSD within 1e-12
(relative to the block's scale; measured: 3e-14 at worst, mostly 1e-16) —
all of them are bisections or Jacobi sweeps whose stopping tests sit at
1e-13; exponential cone: 1e-12 as well, except where
the reference's Newton iteration from rho = 0 reaches a different root (or
its fallback) because the last bit of the device exp() differs from glibc's;
there both results must be admissible outputs of that routine (at least 99%
of random points agree to 1e-12).

Solves: the reference's conic end-to-end cases (test/test_conic_e2e.c,
test/test_sparse_q_cones.c) and random strictly feasible conic programs,
device vs oracle: same status, objectives within 5 eps (1 + |obj|).
"""
import numpy as np
import pytest
import scipy.sparse as sp

import pdhcg_amd
from pdhcg_amd import QPData, capi
from tests.oracle_lib import oracle_solve
from tests.test_oracle_cones import FREE, FREE_UB, _L, _mat_to_svec, diag_q_cone_problem, in_exp, \
    project as orc_project, psd_test_matrix, sparse_q_soc_problem

pytestmark = pytest.mark.gpu

SOC, RSOC, EXP, POW, PSD = (capi.CONE_STANDARD_SOC, capi.CONE_ROTATED_SOC, capi.CONE_EXPONENTIAL, capi.CONE_POWER,
                            capi.CONE_PSD)


def _slots(t, vd):
    return {SOC: vd + 2, RSOC: vd + 2, EXP: 3, POW: 3, PSD: vd * (vd + 1) // 2}[t]


def _layout(blocks, gap=1):
    """(type, v_dim[, alpha]) list -> cone specs with `gap` free slots between blocks, and n."""
    specs, pos = [], 0
    for b in blocks:
        specs.append((b[0], pos, b[1]) + tuple(b[2:]))
        pos += _slots(b[0], b[1]) + gap
    return specs, pos


def _orc_dual_residual(t, rc, vd, d, alpha):
    out = np.zeros_like(rc)
    assert _L().orc_dual_residual_block(t, capi.as_ptr(np.ascontiguousarray(rc)), vd,
                                        capi.as_ptr(np.ascontiguousarray(d)), alpha, capi.as_ptr(out)) == 0
    return out


def _blocks_for(t, rng, count, sizes=(1, 2, 3, 7, 16, 33, 64)):
    if t in (SOC, RSOC):
        return [(t, int(v)) for v in rng.choice(list(sizes), size=count)]
    if t == POW:
        return [(t, 1, float(a)) for a in rng.uniform(0.05, 0.95, size=count)]
    if t == PSD:
        return [(t, int(v)) for v in rng.choice([1, 2, 3, 5, 8, 12], size=count)]
    return [(t, 1)] * count


def _scaling(specs, n, rng, uniform):
    d = np.ones(n)
    for s in specs:
        L = _slots(s[0], s[2])
        if uniform or s[0] == PSD:
            d[s[1]:s[1] + L] = rng.uniform(0.2, 5.0)
        else:
            d[s[1]:s[1] + L] = rng.uniform(0.2, 5.0, size=L)
    return d


@pytest.mark.parametrize("t", [SOC, RSOC, POW, PSD])
@pytest.mark.parametrize("uniform", [True, False])
def test_projection_matches_oracle(t, uniform):
    rng = np.random.default_rng(100 + 10 * t + uniform)
    specs, n = _layout(_blocks_for(t, rng, 300 if t != PSD else 60))
    d = _scaling(specs, n, rng, uniform)
    x = rng.standard_normal(n) * rng.choice([0.01, 1.0, 100.0], size=n)
    got = pdhcg_amd.project_cones(specs, x, d)
    free = np.ones(n, bool)
    for s in specs:
        L = _slots(s[0], s[2])
        sl = slice(s[1], s[1] + L)
        free[sl] = False
        exp = orc_project(s[0], x[sl], s[2], d[sl], s[3] if len(s) > 3 else 0.0)
        scale = 1.0 + np.max(np.abs(x[sl]))
        assert np.max(np.abs(got[sl] - exp)) <= 1e-12 * scale, (s, got[sl], exp)
    assert np.array_equal(got[free], x[free])


@pytest.mark.parametrize("mode", ["", "1", "2", "3"])
def test_psd_large_orders_match_oracle(mode, monkeypatch):
    # orders 16 .. 128 (the conic-mix workload's range), odd and even, random,
    # clustered and rank-one spectra, several blocks of one order in one
    # launch; mode "" lets the library choose where V' and the packed A live,
    # "1" / "2" / "3" force V' / A / both into the global work buffer
    if mode:
        monkeypatch.setenv("PDHCG_PSD_MODE", mode)
    rng = np.random.default_rng(77 + len(mode))
    blocks = []
    for k in (16, 31, 64, 97, 128):
        for kind in ("random", "clustered", "rank1"):
            blocks.append((k, _mat_to_svec(psd_test_matrix(k, rng, kind))))
        blocks.append((k, _mat_to_svec(psd_test_matrix(k, rng, "random"))))
    specs, pos, xs = [], 0, []
    for k, x in blocks:
        specs.append((PSD, pos, k))
        xs.append(x)
        pos += x.size
    x = np.concatenate(xs)
    got = pdhcg_amd.project_cones(specs, x, np.ones(pos))
    for (k, xb), sp_ in zip(blocks, specs):
        sl = slice(sp_[1], sp_[1] + xb.size)
        exp = orc_project(PSD, xb, k)
        scale = 1.0 + np.max(np.abs(xb))
        assert np.all(np.isfinite(got[sl]))
        assert np.max(np.abs(got[sl] - exp)) <= 1e-12 * scale, (k, np.max(np.abs(got[sl] - exp)))


def _psd_block_set(rng, orders):
    blocks = [(k, _mat_to_svec(psd_test_matrix(k, rng, ("random", "clustered", "rank1")[i % 3])))
              for i, k in enumerate(orders)]
    specs, pos = [], 0
    for k, x in blocks:
        specs.append((PSD, pos, k))
        pos += x.size
    return blocks, specs, pos


def _check_psd(got, blocks, specs, xs):
    for (k, _), sp_ in zip(blocks, specs):
        L = k * (k + 1) // 2
        sl = slice(sp_[1], sp_[1] + L)
        exp = orc_project(PSD, xs[sl].copy(), k)
        scale = 1.0 + np.max(np.abs(xs[sl]))
        assert np.all(np.isfinite(got[sl]))
        assert np.max(np.abs(got[sl] - exp)) <= 1e-12 * scale, (k, np.max(np.abs(got[sl] - exp)))


def test_psd_mixed_orders_one_launch():
    # every order in one launch, largest first: orders 1..3, odd orders
    # (padded), the LDS limits of the layouts (V' and the packed A in LDS up
    # to order 114, one of them above) and order 150
    rng = np.random.default_rng(91)
    blocks, specs, n = _psd_block_set(rng, [1, 2, 3, 16, 17, 64, 99, 113, 114, 115, 116, 128, 137, 138, 139, 140,
                                            141, 150, 7, 128])
    x = np.concatenate([b for _, b in blocks])
    got = pdhcg_amd.project_cones(specs, x, np.ones(n))
    _check_psd(got, blocks, specs, x)


@pytest.mark.parametrize("mode", ["", "1", "2", "3"])
def test_psd_warm_sequence_matches_oracle(mode, monkeypatch):
    # one cone set projects a sequence: a point, two perturbations of it
    # (warm starts close to the answer), an unrelated point (a warm start far
    # from it) and the first point again; each result is that point's
    # projection to 1e-12
    if mode:
        monkeypatch.setenv("PDHCG_PSD_MODE", mode)
    rng = np.random.default_rng(92 + len(mode))
    blocks, specs, n = _psd_block_set(rng, [16, 33, 64, 97, 128])
    x0 = np.concatenate([b for _, b in blocks])
    seq = [x0, x0 + 1e-3 * rng.standard_normal(n), x0 + 1e-2 * rng.standard_normal(n),
           np.concatenate([_mat_to_svec(psd_test_matrix(k, rng, "random")) for k, _ in blocks]), x0]
    got = pdhcg_amd.project_cones_seq(specs, np.stack(seq), np.ones(n))
    for i, xs in enumerate(seq):
        _check_psd(got[i], blocks, specs, xs)
    # the repeated point: the same projection to rounding whatever came before
    assert np.max(np.abs(got[4] - got[0])) <= 1e-12 * (1.0 + np.max(np.abs(x0)))


@pytest.mark.parametrize("t", [SOC, RSOC, POW, PSD])
def test_dual_residual_matches_oracle(t):
    rng = np.random.default_rng(200 + t)
    specs, n = _layout(_blocks_for(t, rng, 200 if t != PSD else 40))
    d = _scaling(specs, n, rng, False)
    rc = rng.standard_normal(n) * 3.0
    got = pdhcg_amd.cone_dual_residual(specs, rc, d)
    for s in specs:
        L = _slots(s[0], s[2])
        sl = slice(s[1], s[1] + L)
        exp = _orc_dual_residual(s[0], rc[sl].copy(), s[2], d[sl].copy(), s[3] if len(s) > 3 else 0.0)
        scale = 1.0 + np.max(np.abs(rc[sl]))
        assert np.max(np.abs(got[sl] - exp)) <= 1e-12 * scale * np.max(d[sl]), (s, got[sl], exp)


def _exp_admissible(r, d, o):
    """True when o is an output the reference's exponential-cone routine can
    give for the point r (scaling d): r itself (inside), 0 (polar), the
    r2 <= 0 special case, the fallback (min(r1, 0), 0, max(r3, 0)), or the
    routine's closing formula (d1 rho u, d2 u, d3 u e^rho) with
    u = (a1 + alpha a3 e^rho) / (rho + alpha e^(2 rho)) > 0 at the rho where
    its Newton iteration stopped (a root of its Newton function, or the end
    of a cycle of clamped +-10 steps) (reference
    src/kernels/pdhcg_exp_cone_kernels.cu:26-131)."""
    r, d, o = (np.asarray(v, float) for v in (r, d, o))
    if np.array_equal(o, r) or not np.any(o):
        return True
    if np.array_equal(o, [min(r[0], 0.0), 0.0, max(r[2], 0.0)]):
        return True
    a1, a2, a3 = r / d
    if a1 <= 0.0 and a2 <= 0.0:
        return np.array_equal(o, [r[0], 0.0, 0.0 if a3 < 0.0 else r[2]])
    u = o[1] / d[1]
    if not u > 0.0:
        return False
    rho = (o[0] / d[0]) / u
    al = (d[2] / d[0]) ** 2
    er = np.exp(rho)
    u_formula = (a1 + al * a3 * er) / (rho + al * er * er)
    return bool(abs(u - u_formula) <= 1e-9 * abs(u) and abs(o[2] - d[2] * u * er) <= 1e-9 * abs(o[2]))


@pytest.mark.parametrize("weighted", [False, True])
def test_exp_projection_matches_oracle(weighted):
    # Device and oracle run the same Newton iteration from rho = 0 without
    # fused multiply-adds; they differ only through the last bits of exp().
    # Every point must agree to 1e-12 (relative to the point) or, where the
    # Newton path reaches a different root or the fallback, both outputs must
    # be admissible results of the routine; at least 99% agree to 1e-12.
    rng = np.random.default_rng(7 + weighted)
    specs, n = _layout([(EXP, 1)] * 4000, gap=0)
    d = rng.uniform(0.2, 5.0, n) if weighted else np.ones(n)
    x = rng.standard_normal(n) * rng.choice([0.5, 2.0, 10.0], size=n)
    got = pdhcg_amd.project_cones(specs, x, d)
    agree = 0
    for s in specs:
        sl = slice(s[1], s[1] + 3)
        exp = orc_project(EXP, x[sl], 1, d[sl])
        if np.max(np.abs(got[sl] - exp)) <= 1e-12 * (1.0 + np.max(np.abs(x[sl]))):
            agree += 1
            continue
        assert _exp_admissible(x[sl], d[sl], got[sl]), (x[sl], d[sl], got[sl], exp)
        assert _exp_admissible(x[sl], d[sl], exp), (x[sl], d[sl], got[sl], exp)
    print(f"exponential cone: {agree}/{len(specs)} points agree with the oracle to 1e-12")
    assert agree >= 0.99 * len(specs), agree


def _device_vs_oracle(p, eps, **kw):
    g = pdhcg_amd.solve(p, eps=eps, **kw)
    o = oracle_solve(p, eps=eps, **kw)
    assert g is not None and o is not None
    assert g.status == o.status == "OPTIMAL", (g.status, o.status)
    assert abs(g.pobj - o.pobj) <= 5 * eps * (1 + abs(o.pobj)), (g.pobj, o.pobj)
    return g, o


def test_soc_e2e():
    p = QPData(c=[0.0, 0.0, 1.0], A=sp.csr_matrix(([1.0, 1.0], ([0, 1], [0, 1])), shape=(2, 3)),
               con_lb=[3.0, 4.0], con_ub=[3.0, 4.0], var_lb=FREE, var_ub=FREE_UB, cones=[(SOC, 0, 1)])
    g, _ = _device_vs_oracle(p, 1e-7)
    assert np.allclose(g.x, [3.0, 4.0, 5.0], rtol=1e-3)


def test_rsoc_e2e():
    p = QPData(c=[0.0, 1.0, 1.0], A=sp.csr_matrix(([1.0], ([0], [0])), shape=(1, 3)), con_lb=[3.0], con_ub=[3.0],
               var_lb=FREE, var_ub=FREE_UB, cones=[(RSOC, 0, 1)])
    g, _ = _device_vs_oracle(p, 1e-7)
    e = 3.0 / np.sqrt(2.0)
    assert abs(g.x[1] - e) < 5e-3 * (1 + e) and abs(g.x[2] - e) < 5e-3 * (1 + e)


def test_exp_e2e():
    p = QPData(c=[-1.0, 0.0, 0.0], A=sp.csr_matrix(([1.0, 1.0], ([0, 1], [1, 2])), shape=(2, 3)),
               con_lb=[1.0, 2.0], con_ub=[1.0, 2.0], var_lb=FREE, var_ub=FREE_UB, cones=[(EXP, 0, 1)])
    g, _ = _device_vs_oracle(p, 1e-6)
    assert abs(g.x[0] - np.log(2.0)) < 1e-3 * (1 + np.log(2.0))


def test_power_e2e():
    p = QPData(c=[0.0, 0.0, -1.0], A=sp.csr_matrix(([1.0, 1.0], ([0, 1], [0, 1])), shape=(2, 3)),
               con_lb=[4.0, 1.0], con_ub=[4.0, 1.0], var_lb=FREE, var_ub=FREE_UB, cones=[(POW, 0, 1, 0.5)])
    g, _ = _device_vs_oracle(p, 1e-7)
    assert abs(g.x[2] - 2.0) < 1e-4


def test_psd_e2e():
    Cm = np.array([[2.0, 1.0], [1.0, 3.0]])
    p = QPData(c=_mat_to_svec(Cm), A=sp.csr_matrix([[1.0, 0.0, 1.0]]), con_lb=[1.0], con_ub=[1.0], var_lb=FREE,
               var_ub=FREE_UB, cones=[(PSD, 0, 2)])
    g, _ = _device_vs_oracle(p, 1e-7)
    assert abs(g.pobj - np.linalg.eigvalsh(Cm)[0]) < 1e-5


def test_sparse_q_with_soc_e2e():
    g, _ = _device_vs_oracle(sparse_q_soc_problem(), 1e-7, iteration_limit=100000)
    assert np.allclose(g.x, [3.0, 4.0, 3.0, 4.0, 5.0], atol=1e-4)
    assert abs(g.pobj + 1.5) < 1e-3


def random_conic(seed, types, q=None, soc_sizes=(1, 2, 3, 7, 16, 33, 64), count=6):
    """Strictly primal and dual feasible conic program: min c'x (+ x'Qx/2),
    A x = A x_int, x in K on the cone slots, 0 <= x <= 10 on the free slots,
    with x_int interior to K and c = A'y0 + s0, s0 interior to K*."""
    rng = np.random.default_rng(seed)
    blocks = []
    for t in types:
        blocks += _blocks_for(t, rng, count, soc_sizes)
    specs, n = _layout(blocks, gap=2)
    xi = rng.uniform(1.0, 2.0, n)
    s0 = rng.uniform(0.5, 1.5, n)
    lb, ub = np.zeros(n), np.full(n, 10.0)
    for s in specs:
        t, st, vd = s[0], s[1], s[2]
        L = _slots(t, vd)
        lb[st:st + L], ub[st:st + L] = -np.inf, np.inf
        if t == SOC:
            v = rng.standard_normal(vd + 1)
            xi[st:st + vd + 1] = v
            xi[st + vd + 1] = np.linalg.norm(v) + 1.0
            w = rng.standard_normal(vd + 1) * 0.3
            s0[st:st + vd + 1] = w
            s0[st + vd + 1] = np.linalg.norm(w) + 1.0
        elif t == RSOC:
            v = rng.standard_normal(vd)
            xi[st:st + vd] = v
            xi[st + vd] = xi[st + vd + 1] = np.sqrt(0.5 * v @ v) + 1.0
            s0[st:st + vd] = 0.1 * v
            s0[st + vd] = s0[st + vd + 1] = 1.0 + 0.1 * np.linalg.norm(v)
        elif t == EXP:
            xi[st:st + 3] = [0.0, 1.0, 2.0]
            s0[st:st + 3] = [-1.0, 1.0, 2.0]
        elif t == POW:
            xi[st:st + 3] = [1.0, 1.0, 0.3]
            s0[st:st + 3] = [1.0, 1.0, -0.2]
        else:
            M = rng.standard_normal((vd, vd))
            xi[st:st + L] = _mat_to_svec(M @ M.T + vd * np.eye(vd))
            s0[st:st + L] = _mat_to_svec(np.eye(vd))
    m = max(2, n // 3)
    A = sp.random(m, n, density=min(1.0, 6.0 / n), random_state=rng, format="csr") + \
        sp.csr_matrix((np.ones(m), (np.arange(m), rng.integers(0, n, m))), shape=(m, n))
    b = A @ xi
    y0 = rng.standard_normal(m)
    c = A.T @ y0 + s0
    return QPData(c=c, A=A.tocsr(), Q=q(n, rng) if q else None, con_lb=b, con_ub=b, var_lb=lb, var_ub=ub,
                  cones=specs)


@pytest.mark.parametrize("types", [(SOC, RSOC), (POW,), (PSD,), (SOC, RSOC, POW, PSD)])
def test_random_conic_lp_matches_oracle(types):
    p = random_conic(31 + len(types), types)
    _device_vs_oracle(p, 1e-6, iteration_limit=200000)


def test_random_conic_qp_matches_oracle():
    # diag(Q) varies inside the cone blocks: the Jacobi-preconditioned inner
    # step followed by the Euclidean cone projection has the wrong fixed point
    # there (reference src/solver_state.cu:323 keeps the preconditioner on with
    # second-order cones), so this case runs with the preconditioner off
    def q(n, rng):
        B = sp.random(n, n, density=3.0 / n, random_state=rng, format="csr")
        return (B.T @ B + sp.eye(n) * 0.1).tocsr()
    p = random_conic(41, (SOC, RSOC), q=q)
    _device_vs_oracle(p, 1e-6, iteration_limit=200000, diag_jacobi_precond=False)


def test_random_conic_qp_preconditioned_matches_oracle():
    # coupling only among the box slots, so diag(Q) is uniform on every cone
    # block and the preconditioned inner step keeps the right fixed point
    specs_holder = {}

    def q(n, rng):
        box = np.ones(n)
        for s in specs_holder["specs"]:
            box[s[1]:s[1] + _slots(s[0], s[2])] = 0.0
        P = sp.diags(box)
        B = sp.random(n, n, density=3.0 / n, random_state=rng, format="csr")
        return (P @ (B.T @ B) @ P + sp.eye(n) * 0.5).tocsr()
    base = random_conic(47, (SOC, RSOC, POW))
    specs_holder["specs"] = base.cones
    p = random_conic(47, (SOC, RSOC, POW), q=q)
    g, _ = _device_vs_oracle(p, 1e-6, iteration_limit=200000)
    assert g.inner_iterations > g.iterations


def test_random_conic_lowrank_qp_matches_oracle():
    p = random_conic(53, (SOC, EXP))
    rng = np.random.default_rng(53)
    p.R = sp.csr_matrix(rng.standard_normal((4, len(p.c))) * 0.3)
    _device_vs_oracle(p, 1e-6, iteration_limit=200000, diag_jacobi_precond=False)


def test_diag_q_with_psd_uses_inner_solve():
    def q(n, rng):
        return sp.diags(rng.uniform(0.5, 2.0, n)).tocsr()
    p = random_conic(43, (PSD,), q=q)
    g, _ = _device_vs_oracle(p, 1e-6, iteration_limit=200000)
    assert g.inner_iterations > g.iterations


# ---- second-order blocks above 64 vector slots (one workgroup per block) ----

LONG = (65, 100, 257, 1000, 4096)


def _long_point(specs, n, rng):
    """Random point whose long blocks are, in turn, outside the cone, inside
    it and in its polar, so that every branch of the projection runs."""
    x = rng.standard_normal(n) * rng.choice([0.01, 1.0, 100.0], size=n)
    for i, sp_ in enumerate(specs):
        t, st, vd = sp_[0], sp_[1], sp_[2]
        nrm = np.linalg.norm(x[st:st + vd + 1])
        if i % 3 == 1:
            x[st + vd + 1] = 2.0 * nrm + 1.0 if t == SOC else nrm * nrm + 1.0
            if t == RSOC:
                x[st + vd] = x[st + vd + 1]
        elif i % 3 == 2:
            x[st + vd + 1] = -(2.0 * nrm + 1.0)
            if t == RSOC:
                x[st + vd] = x[st + vd + 1]
    return x


@pytest.mark.parametrize("t", [SOC, RSOC])
@pytest.mark.parametrize("uniform", [True, False])
def test_long_block_projection_matches_oracle(t, uniform):
    rng = np.random.default_rng(300 + 10 * t + uniform)
    blocks = [(t, v) for v in rng.permutation(list(LONG) * 3 + [3, 64])]
    specs, n = _layout(blocks)
    d = _scaling(specs, n, rng, uniform)
    x = _long_point(specs, n, rng)
    got = pdhcg_amd.project_cones(specs, x, d)
    for s in specs:
        L = _slots(s[0], s[2])
        sl = slice(s[1], s[1] + L)
        exp = orc_project(s[0], x[sl], s[2], d[sl], 0.0)
        scale = 1.0 + np.max(np.abs(x[sl]))
        assert np.max(np.abs(got[sl] - exp)) <= 1e-12 * scale, (s[:3], np.max(np.abs(got[sl] - exp)))


@pytest.mark.parametrize("t", [SOC, RSOC])
@pytest.mark.parametrize("uniform", [True, False])
def test_long_block_dual_residual_matches_oracle(t, uniform):
    rng = np.random.default_rng(400 + 10 * t + uniform)
    blocks = [(t, v) for v in rng.permutation(list(LONG) * 3 + [5])]
    specs, n = _layout(blocks)
    d = _scaling(specs, n, rng, uniform)
    rc = _long_point(specs, n, rng)
    got = pdhcg_amd.cone_dual_residual(specs, rc, d)
    for s in specs:
        L = _slots(s[0], s[2])
        sl = slice(s[1], s[1] + L)
        exp = _orc_dual_residual(s[0], rc[sl].copy(), s[2], d[sl].copy(), 0.0)
        scale = 1.0 + np.max(np.abs(rc[sl]))
        assert np.max(np.abs(got[sl] - exp)) <= 1e-12 * scale * np.max(d[sl]), (s[:3], np.max(np.abs(got[sl] - exp)))


@pytest.mark.parametrize("precision", [pdhcg_amd.PRECISION_FP64, pdhcg_amd.PRECISION_MIXED])
def test_long_blocks_e2e(precision):
    p = random_conic(61, (SOC, RSOC), soc_sizes=(65, 200, 700), count=2)
    eps = 1e-6
    g = pdhcg_amd.solve(p, precision, eps=eps, iteration_limit=200000)
    o = oracle_solve(p, eps=eps, iteration_limit=200000)
    assert g.status == o.status == "OPTIMAL", (g.status, o.status)
    assert abs(g.pobj - o.pobj) <= 5 * eps * (1 + abs(o.pobj)), (g.pobj, o.pobj)


def test_long_blocks_diag_q_e2e():
    def q(n, rng):
        return sp.diags(rng.uniform(0.1, 2.0, n)).tocsr()
    p = random_conic(67, (SOC, RSOC), q=q, soc_sizes=(65, 300), count=2)
    _device_vs_oracle(p, 1e-6, iteration_limit=200000)


def test_fixed_slot_in_long_block_closed_form():
    # A second-order block of v_dim 65 (slots 0..65 the vector, slot 66 the
    # bound) with the bound set to 1 by a row and slot 3 fixed at 0.5: the
    # other 65 vector slots range over the ball of radius sqrt(1 - 0.25), so
    # the optimum is c66 + 0.5 c3 - sqrt(0.75) ||c_rest||.
    n = 67
    rng = np.random.default_rng(9)
    c = rng.standard_normal(n)
    A = sp.csr_matrix(([1.0], ([0], [n - 1])), shape=(1, n))
    p = QPData(c=c, A=A, con_lb=[1.0], con_ub=[1.0], var_lb=[-1e30] * n, var_ub=[1e30] * n,
               cones=[(SOC, 0, 65)], cone_fixed=[(0, 3, 0.5)])
    rest = np.delete(c[:n - 1], 3)
    exact = c[n - 1] + 0.5 * c[3] - np.sqrt(0.75) * np.linalg.norm(rest)
    eps = 1e-7
    g = pdhcg_amd.solve(p, eps=eps, iteration_limit=200000)
    o = oracle_solve(p, eps=eps, iteration_limit=200000)
    assert g.status == o.status == "OPTIMAL", (g.status, o.status)
    assert g.x[3] == 0.5
    assert abs(g.pobj - exact) <= 20 * eps * (1 + abs(exact)), (g.pobj, exact)
    assert abs(g.pobj - o.pobj) <= 20 * eps * (1 + abs(exact)), (g.pobj, o.pobj)


def test_fixed_psd_section_is_rejected():
    # fixed slots in a PSD block are refused, as in the reference
    # (src/pdhcg.c:632-636)
    n = 6
    p = QPData(c=np.zeros(n), A=sp.csr_matrix(np.eye(1, n)), con_lb=[1.0], con_ub=[1.0], var_lb=[-1e30] * n,
               var_ub=[1e30] * n, cones=[(PSD, 0, 3)], cone_fixed=[(0, 1, 0.5)])
    assert pdhcg_amd.solve(p, eps=1e-6) is None


@pytest.mark.parametrize("cone_type", [SOC, RSOC, EXP, POW])
def test_diag_q_cone_prox_matches_oracle(cone_type):
    # diagonal Q with a cone block: the cone step is a projection in the
    # metric diag(1 + tau q)
    p, _ = diag_q_cone_problem(11 + cone_type, cone_type)
    _device_vs_oracle(p, 1e-7, iteration_limit=500000)


def test_random_conic_diag_qp_matches_oracle():
    def q(n, rng):
        return sp.diags(rng.uniform(0.1, 2.0, n)).tocsr()
    p = random_conic(59, (SOC, RSOC, POW), q=q)
    _device_vs_oracle(p, 1e-6, iteration_limit=200000)
