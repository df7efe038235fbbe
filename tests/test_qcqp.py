"""qcqp_to_socp_qp: diagonal quadratic constraints lifted into SOC / RSOC
blocks (reference src/qcqp_transform.c:96-552), host code.

Pinned by construction rather than by a second restatement:
- the lifted form of the reference's own example (test/test_qcqp_rsoc.c:1-8:
  min -x + y^2 s.t. x^2 + y^2 <= 1 -> v = sqrt2 (x, y), (v, s, t) in RSOC
  with s = t = 1), written out by hand;
- for random diagonal constraints with linear parts, a point x maps to a
  lifted point that satisfies every equality row exactly, and its cone
  block is feasible exactly when a'x + x'Qx <= b (both SOC and RSOC forms,
  both senses);
- the reference's rejections (non-diagonal or mixed-sign Q, wrong sense,
  affine cones, no quadratic constraint).
End-to-end solves through solve_qp_problem run on the GPU
(tests/test_gpu_qcqp.py).
"""
import ctypes as C

import numpy as np
import pytest
import scipy.sparse as sp

import pdhcg_amd
from pdhcg_amd import QPData, capi
from pdhcg_amd.problem import build

SQRT2 = np.sqrt(2.0)


def _L():
    L = pdhcg_amd.lib()
    L.qcqp_to_socp_qp.restype = C.POINTER(capi.qp_problem_t)
    L.qcqp_to_socp_qp.argtypes = [C.POINTER(capi.qp_problem_t), C.c_int]
    return L


def _fixed_mask(q, n):
    # cones.is_fixed is declared char*: read the raw pointer, not a C string
    addr = C.c_void_p.from_buffer(q.cones, capi.cone_blocks_t.is_fixed.offset).value
    return np.array((C.c_int8 * n).from_address(addr), dtype=np.int8) if addr else None


def _lift(p, form):
    L = _L()
    h = build(L, "", p)
    assert h is not None
    try:
        out = L.qcqp_to_socp_qp(h.ptr, form)
        if not out:
            return None
        q = out.contents
        n, m = q.num_variables, q.num_constraints
        nnz = q.constraint_matrix_num_nonzeros
        A = q.constraint_matrix.contents
        rp = np.ctypeslib.as_array(A.row_ptr, shape=(m + 1,)).copy()
        M = sp.csr_matrix((np.ctypeslib.as_array(A.val, shape=(max(nnz, 1),))[:nnz].copy(),
                           np.ctypeslib.as_array(A.col_ind, shape=(max(nnz, 1),))[:nnz].copy(), rp), shape=(m, n))
        arr = lambda ptr, k: np.ctypeslib.as_array(ptr, shape=(k,)).copy()
        K = q.cones.num_cones
        res = dict(n=n, m=m, A=M, cl=arr(q.constraint_lower_bound, m), cu=arr(q.constraint_upper_bound, m),
                   lb=arr(q.variable_lower_bound, n), ub=arr(q.variable_upper_bound, n), c=arr(q.objective_vector, n),
                   cones=[(q.cones.type[i], q.cones.start_idx[i], q.cones.v_dim[i]) for i in range(K)],
                   fixed=_fixed_mask(q, n),
                   start=arr(q.primal_start, n) if q.primal_start else None,
                   n_orig=q.num_original_variables, nqc=q.num_quadratic_constraints)
        L.qp_problem_free(out)
        return res
    finally:
        h.close()


def test_reference_rsoc_example_lifted_by_hand():
    # min -x + y^2  s.t.  x^2 + y^2 <= 1  (row 0 of A empty: the aux slots are pinned)
    p = QPData(c=[-1.0, 0.0], A=sp.csr_matrix((1, 2)), con_lb=[-np.inf], con_ub=[1.0], var_lb=[-np.inf] * 2,
               var_ub=[np.inf] * 2, Q=sp.diags([0.0, 2.0]).tocsr(), quadratic_constraints=[(0, sp.eye(2))])
    L = _lift(p, capi.CONE_ROTATED_SOC)
    assert (L["n"], L["m"], L["n_orig"], L["nqc"]) == (6, 4, 2, 0)
    assert L["cones"] == [(capi.CONE_ROTATED_SOC, 2, 2)]
    want = np.zeros((4, 6))
    want[1, 0], want[1, 2] = -SQRT2, 1.0  # v1 = sqrt2 x
    want[2, 1], want[2, 3] = -SQRT2, 1.0  # v2 = sqrt2 y
    assert np.array_equal(L["A"].toarray(), want)
    assert np.array_equal(L["A"].indptr, [0, 0, 2, 4, 4])
    assert np.array_equal(L["cl"], np.zeros(4)) and np.array_equal(L["cu"], np.zeros(4))
    assert np.array_equal(L["fixed"], [0, 0, 0, 0, 1, 1])
    assert np.array_equal(L["start"], [0, 0, 0, 0, 1.0, 1.0])  # (s, t) = (b, 1)
    assert np.array_equal(L["c"], [-1.0, 0, 0, 0, 0, 0])
    assert np.all(np.isinf(L["lb"][2:])) and np.all(np.isinf(L["ub"][2:]))
    # standard form: (w, z) = ((b - 1)/sqrt2, (b + 1)/sqrt2) pinned
    S = _lift(p, capi.CONE_STANDARD_SOC)
    assert S["cones"] == [(capi.CONE_STANDARD_SOC, 2, 2)]
    assert np.allclose(S["start"][4:], [0.0, SQRT2], rtol=0, atol=1e-15)


def _random_qcqp(seed, nsd=False):
    rng = np.random.default_rng(seed)
    n, m = 7, 4
    A = sp.random(m, n, density=0.5, random_state=rng, format="csr") + sp.csr_matrix(
        ([1.0] * m, (range(m), range(m))), shape=(m, n))
    qcs = []
    cl, cu = np.full(m, -np.inf), np.full(m, 5.0)
    for r in (1, 3):
        d = np.zeros(n)
        idx = rng.choice(n, size=4, replace=False)
        d[idx] = rng.uniform(0.2, 2.0, size=4)
        if nsd and r == 3:
            d = -d
            cl[r], cu[r] = -5.0, np.inf
        qcs.append((r, sp.diags(d).tocsr()))
    return QPData(c=rng.standard_normal(n), A=A, con_lb=cl, con_ub=cu, var_lb=np.full(n, -3.0),
                  var_ub=np.full(n, 3.0), quadratic_constraints=qcs), rng


def _lifted_point(p, L, x, form):
    """The lifted point of x: v = sqrt(2|q|) x on the Q support, and the aux
    slots from the coupling rows."""
    A = p.A.toarray()
    z = np.zeros(L["n"])
    z[:x.size] = x
    for (row, Q), (_, start, vd) in zip(p.quadratic_constraints, L["cones"]):
        q = Q.diagonal()
        sup = np.nonzero(q)[0]
        z[start:start + vd] = np.sqrt(2 * np.abs(q[sup])) * x[sup]
        flip = q[sup[0]] < 0
        b = -p.con_lb[row] if flip else p.con_ub[row]
        slack = b - (-1 if flip else 1) * (A[row] @ x)
        if form == capi.CONE_ROTATED_SOC:
            z[start + vd], z[start + vd + 1] = slack, 1.0
        else:
            z[start + vd], z[start + vd + 1] = (slack - 1) / SQRT2, (slack + 1) / SQRT2
    return z


@pytest.mark.parametrize("form", [capi.CONE_ROTATED_SOC, capi.CONE_STANDARD_SOC])
@pytest.mark.parametrize("nsd", [False, True])
def test_lifted_point_feasible_iff_constraint_holds(form, nsd):
    p, rng = _random_qcqp(11 + nsd, nsd)
    L = _lift(p, form)
    m0 = p.A.shape[0]
    assert L["nqc"] == 0 and L["n_orig"] == p.n
    assert L["m"] == m0 + sum(Q.nnz for _, Q in p.quadratic_constraints) + len(p.quadratic_constraints)
    assert np.array_equal(L["A"][:m0, :p.n].toarray() != 0, p.A.toarray() != 0)
    seen = [0, 0]
    for trial in range(200):
        x = rng.uniform(-3, 3, p.n) * rng.choice([0.2, 1.0])
        z = _lifted_point(p, L, x, form)
        Az = L["A"] @ z
        # every row except the unconstrained originals is an equality that the lifted point meets
        qrows = {r for r, _ in p.quadratic_constraints}
        for i in range(L["m"]):
            if i < m0 and i not in qrows:
                assert np.isclose(Az[i], (p.A @ x)[i], rtol=0, atol=1e-12)
            else:
                assert L["cl"][i] == L["cu"][i] and abs(Az[i] - L["cl"][i]) <= 1e-12 * (1 + abs(L["cl"][i]))
        for (row, Q), (t, start, vd) in zip(p.quadratic_constraints, L["cones"]):
            q = Q.diagonal()
            ax = (p.A @ x)[row]
            holds = ax + x @ (q * x) <= p.con_ub[row] if q.sum() > 0 else ax + x @ (q * x) >= p.con_lb[row]
            v, a0, a1 = z[start:start + vd], z[start + vd], z[start + vd + 1]
            if t == capi.CONE_ROTATED_SOC:
                member = v @ v <= 2 * a0 * a1 + 1e-9 and a0 >= -1e-12 and a1 >= 0
            else:
                member = np.hypot(np.linalg.norm(v), a0) <= a1 + 1e-9
            assert member == holds or abs(ax + x @ (q * x) - (p.con_ub[row] if q.sum() > 0 else p.con_lb[row])) < 1e-9
            seen[int(holds)] += 1
    assert min(seen) > 20


def test_rejections():
    base = dict(c=[1.0, 1.0], A=sp.csr_matrix([[1.0, 1.0]]), con_lb=[-np.inf], con_ub=[1.0], var_lb=[-1.0] * 2,
                var_ub=[1.0] * 2)
    L = _L()
    cases = {
        "nondiagonal": QPData(quadratic_constraints=[(0, sp.csr_matrix([[1.0, 0.5], [0.5, 1.0]]))], **base),
        "mixed signs": QPData(quadratic_constraints=[(0, sp.diags([1.0, -1.0]).tocsr())], **base),
        "psd with >=": QPData(**{**base, "con_lb": [0.0], "con_ub": [np.inf]},
                              quadratic_constraints=[(0, sp.eye(2).tocsr())]),
        "nsd with <=": QPData(quadratic_constraints=[(0, -sp.eye(2).tocsr())], **base),
        "affine cones": QPData(quadratic_constraints=[(0, sp.eye(2).tocsr())], F=sp.csr_matrix(np.eye(3, 2)),
                               g=np.zeros(3), affine_cones=[(capi.CONE_STANDARD_SOC, 0, 1)], **base),
    }
    for name, p in cases.items():
        assert _lift(p, capi.CONE_ROTATED_SOC) is None, name
    assert _lift(QPData(**base), capi.CONE_ROTATED_SOC) is None  # nothing to lift
