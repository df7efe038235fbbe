"""Problem permutation and partition cuts (host integer work, no GPU).

Pins: the reference's known answers for the cut optimiser
(test/test_partition_utils.c:17-35) and its permutation invariants
(test/test_cone_permutation.c:19-137); brute-force enumeration of every
legal cut set on small instances (the optimiser must reach the minimum
squared displacement); balanced cuts against a numpy cumulative sum; and a
solve of a permuted problem mapping back to the unpermuted solution.  The
library and the oracle must agree bit for bit (same rand() draws after the
same srand()).
"""
import ctypes as C
import os
import itertools

import numpy as np
import pytest
import scipy.sparse as sp

import pdhcg_amd
from pdhcg_amd import QPData, capi
from pdhcg_amd.problem import build
from tests.oracle_lib import oracle, oracle_build, oracle_solve

libc = C.CDLL(None)
libc.srand.argtypes = [C.c_uint]
ip = C.POINTER(C.c_int)
NO_PERM, FULL, BLOCK = 0, 1, 2


def _libs():
    L, O = pdhcg_amd.lib(), oracle()
    P = C.POINTER(capi.qp_problem_t)
    L.optimize_partition_cuts.restype = C.c_bool
    L.optimize_partition_cuts.argtypes = [C.c_int, C.c_int, ip, ip, C.c_int, ip]
    O.orc_optimize_partition_cuts.restype = C.c_int
    O.orc_optimize_partition_cuts.argtypes = [C.c_int, C.c_int, ip, ip, C.c_int, ip]
    L.get_balanced_cuts.restype = ip
    L.get_balanced_cuts.argtypes = [ip, C.c_int, C.c_int]
    O.orc_balanced_cuts.restype = None
    O.orc_balanced_cuts.argtypes = [ip, C.c_int, C.c_int, ip]
    for f, o in (("generate_cone_aware_permutation", "orc_generate_cone_aware_permutation"),
                 ("generate_affine_cone_aware_row_permutation", "orc_generate_affine_row_permutation")):
        getattr(L, f).restype = None
        getattr(L, f).argtypes = [P, C.c_int, C.c_int, ip]
        getattr(O, o).restype = None
        getattr(O, o).argtypes = [P, C.c_int, C.c_int, ip]
    for f, o in (("validate_cone_permutation", "orc_validate_cone_permutation"),
                 ("validate_affine_cone_row_permutation", "orc_validate_affine_row_permutation")):
        getattr(L, f).restype = C.c_bool
        getattr(L, f).argtypes = [P, ip]
        getattr(O, o).restype = C.c_int
        getattr(O, o).argtypes = [P, ip]
    L.permute_problem_return_new.restype = P
    L.permute_problem_return_new.argtypes = [P, ip, ip]
    O.orc_permute_problem_copy.restype = P
    O.orc_permute_problem_copy.argtypes = [P, ip, ip]
    L.permute_problem.restype = C.c_bool
    L.permute_problem.argtypes = [P, ip, ip]
    L.repermute_solution.restype = None
    L.repermute_solution.argtypes = [C.POINTER(capi.pdhcg_result_t), ip, ip]
    L.free.argtypes = [C.c_void_p]
    return L, O


def _iarr(v):
    return (C.c_int * max(1, len(v)))(*[int(x) for x in v])


def _cuts(fn, total, parts, spans, start):
    fs, fe = _iarr([s[0] for s in spans]), _iarr([s[1] for s in spans])
    cuts = _iarr(start)
    ok = fn(total, parts, fs, fe, len(spans), cuts)
    return bool(ok), list(cuts)[:parts + 1]


def test_partition_known_answers():
    # reference test/test_partition_utils.c:17-35
    L, O = _libs()
    for fn in (L.optimize_partition_cuts, O.orc_optimize_partition_cuts):
        ok, c = _cuts(fn, 12, 3, [(2, 5), (7, 10)], [0, 4, 8, 12])
        assert ok and c[0] == 0 and 0 < c[1] < c[2] < c[3] == 12
        assert not (2 <= c[1] <= 5 or 7 <= c[1] <= 10) and not (2 <= c[2] <= 5 or 7 <= c[2] <= 10)
        ok, c = _cuts(fn, 101, 3, [(2, 9), (11, 99)], [0, 10, 11, 101])
        assert ok and c[1] == 1 and c[2] == 10
        ok, _ = _cuts(fn, 6, 3, [(1, 5)], [0, 2, 4, 6])
        assert not ok


def _legal(b, spans, total):
    return 0 < b < total and not any(f <= b <= l for f, l in spans)


@pytest.mark.parametrize("seed", range(40))
def test_partition_cuts_reach_brute_force_minimum(seed):
    L, O = _libs()
    rng = np.random.default_rng(seed)
    total = int(rng.integers(4, 16))
    parts = int(rng.integers(2, min(5, total) + 1))
    spans = []
    for _ in range(int(rng.integers(0, 4))):
        f = int(rng.integers(1, total))
        spans.append((f, int(min(total - 1, f + rng.integers(0, 4)))))
    start = [0] + sorted(int(x) for x in rng.integers(1, total, parts - 1)) + [total]
    legal = [b for b in range(1, total) if _legal(b, spans, total)]
    best = None
    for combo in itertools.combinations(legal, parts - 1):
        cost = sum((c - t) ** 2 for c, t in zip(combo, start[1:-1]))
        best = cost if best is None else min(best, cost)
    ok_l, cl = _cuts(L.optimize_partition_cuts, total, parts, spans, start)
    ok_o, co = _cuts(O.orc_optimize_partition_cuts, total, parts, spans, start)
    assert ok_l == ok_o == (best is not None)
    assert cl == co
    if best is not None:
        assert all(_legal(b, spans, total) for b in cl[1:-1]) and cl == sorted(set(cl))
        assert sum((c - t) ** 2 for c, t in zip(cl[1:-1], start[1:-1])) == best


@pytest.mark.parametrize("parts", [1, 2, 3, 7])
def test_balanced_cuts(parts):
    L, O = _libs()
    rng = np.random.default_rng(parts)
    w = rng.integers(0, 50, 200)
    p = L.get_balanced_cuts(_iarr(w), len(w), parts)
    got = [p[i] for i in range(parts + 1)]
    L.free(C.cast(p, C.c_void_p))
    oc = _iarr([0] * (parts + 1))
    O.orc_balanced_cuts(_iarr(w), len(w), parts, oc)
    assert got == list(oc)
    cum = np.cumsum(w)
    tgt = cum[-1] / parts
    want = [0] + [int(np.argmax(cum >= q * tgt)) + 1 for q in range(1, parts)] + [len(w)]
    assert got == want


def reference_permutation_problem():
    # reference test/test_cone_permutation.c:19-79
    n = 21
    cones = [(capi.CONE_STANDARD_SOC, 1, 1, 0.0, [0, 0, 1]), (capi.CONE_POWER, 5, 1, 0.3, [1, 0, 0]),
             (capi.CONE_EXPONENTIAL, 9, 1, 0.0, [0, 1, 0]), (capi.CONE_ROTATED_SOC, 12, 2, 0.0, [0, 0, 1, 1]),
             (capi.CONE_PSD, 17, 2)]
    affine = [(capi.CONE_STANDARD_SOC, 0, 1), (capi.CONE_EXPONENTIAL, 3, 1), (capi.CONE_PSD, 6, 2)]
    return QPData(c=np.arange(n, dtype=float), A=sp.csr_matrix((1, n)), con_lb=[0.0], con_ub=[0.0],
                  F=sp.csr_matrix((9, n)), g=[10.0, 11.0, 12.0, 20.0, 21.0, 22.0, 30.0, 31.0, 32.0], cones=cones,
                  affine_cones=affine)


def _arr(ptr, k, dtype=float):
    return np.array([ptr[i] for i in range(k)], dtype=dtype)


def _fixed_mask(blocks, n):
    # is_fixed is mirrored as c_char_p (bytes up to the first zero); read the raw pointer instead
    addr = C.c_void_p.from_buffer(blocks, capi.cone_blocks_t.is_fixed.offset).value
    return np.ctypeslib.as_array(C.cast(addr, C.POINTER(C.c_int8)), shape=(n,)).astype(int)


def _blocks(b):
    return [(b.type[k], b.start_idx[k], b.v_dim[k]) for k in range(b.num_cones)]


def _block_len(t, vd):
    return {capi.CONE_STANDARD_SOC: vd + 2, capi.CONE_ROTATED_SOC: vd + 2, capi.CONE_EXPONENTIAL: 3,
            capi.CONE_POWER: 3, capi.CONE_PSD: vd * (vd + 1) // 2}[t]


def test_reference_permutation_invariants():
    L, O = _libs()
    p = reference_permutation_problem()
    h, ho = build(L, "", p), oracle_build(p)
    n, m = 21, 10
    rows_id = _iarr(range(m))
    perms = {}
    for name, fn, ofn, method, bs, seed, size in (
            ("cols_full", L.generate_cone_aware_permutation, O.orc_generate_cone_aware_permutation, FULL, 1, 7, n),
            ("cols_block", L.generate_cone_aware_permutation, O.orc_generate_cone_aware_permutation, BLOCK, 2, 11, n),
            ("rows_full", L.generate_affine_cone_aware_row_permutation, O.orc_generate_affine_row_permutation, FULL,
             1, 13, m)):
        a, b = _iarr([0] * size), _iarr([0] * size)
        libc.srand(seed)
        fn(h.ptr, method, bs, a)
        libc.srand(seed)
        ofn(ho.ptr, method, bs, b)
        assert list(a) == list(b), name
        assert sorted(a) == list(range(size))
        perms[name] = a
    for name in ("cols_full", "cols_block"):
        assert L.validate_cone_permutation(h.ptr, perms[name])
        assert O.orc_validate_cone_permutation(ho.ptr, perms[name])
    assert L.validate_affine_cone_row_permutation(h.ptr, perms["rows_full"])
    # columns: every cone's objective slots and fixed flags move with its block
    q = L.permute_problem_return_new(h.ptr, rows_id, perms["cols_full"])
    assert q
    src, dst = h.contents, q.contents
    c_new = _arr(dst.objective_vector, n)
    fix_old, fix_new = _fixed_mask(src.cones, n), _fixed_mask(dst.cones, n)
    for k in range(5):
        t, s0, vd = _blocks(src.cones)[k]
        t1, s1, _ = _blocks(dst.cones)[k]
        assert t1 == t
        if t == capi.CONE_POWER:
            assert dst.cones.power_alpha[k] == src.cones.power_alpha[k]
        for j in range(_block_len(t, vd)):
            assert c_new[s1 + j] == float(s0 + j)
            assert fix_new[s1 + j] == fix_old[s0 + j]
    L.qp_problem_free(q)
    # rows: affine offsets move with their blocks
    cols_id = _iarr(range(n))
    q = L.permute_problem_return_new(h.ptr, perms["rows_full"], cols_id)
    assert q
    g_old, g_new = _arr(src.affine_cone_offset, m), _arr(q.contents.affine_cone_offset, m)
    for k in range(3):
        t, s0, vd = _blocks(src.affine_cones)[k]
        s1 = q.contents.affine_cones.start_idx[k]
        for j in range(_block_len(t, vd)):
            assert g_new[s1 + j] == g_old[s0 + j]
    L.qp_problem_free(q)
    bad_rows = _iarr([0, 2, 1] + list(range(3, m)))   # swaps two rows inside the first affine block
    assert not L.validate_affine_cone_row_permutation(h.ptr, bad_rows)
    assert not O.orc_validate_affine_row_permutation(ho.ptr, bad_rows)
    assert not L.permute_problem(h.ptr, bad_rows, cols_id)
    dup = _iarr([0, 0] + list(range(2, n)))
    assert not L.validate_cone_permutation(h.ptr, dup)
    assert not O.orc_validate_cone_permutation(ho.ptr, dup)
    h.close()
    ho.close()


def _dump(pp, which_csr):
    pr = pp.contents
    out = {}
    n, m, k = pr.num_variables, pr.num_constraints, pr.num_rank_lowrank_obj
    for name, rows, nnz in (("constraint_matrix", m, pr.constraint_matrix_num_nonzeros),
                            ("objective_sparse_matrix", n, pr.objective_sparse_matrix_num_nonzeros),
                            ("objective_lowrank_matrix", k, pr.objective_lowrank_matrix_num_nonzeros)):
        c = getattr(pr, name)
        if c and nnz > 0:
            out[name] = (_arr(c.contents.row_ptr, rows + 1, int), _arr(c.contents.col_ind, nnz, int),
                         _arr(c.contents.val, nnz))
    for name, size in (("objective_vector", n), ("variable_lower_bound", n), ("variable_upper_bound", n),
                       ("constraint_lower_bound", m), ("constraint_upper_bound", m), ("affine_cone_offset", m)):
        out[name] = _arr(getattr(pr, name), size)
    out["cones"] = _blocks(pr.cones)
    return out


def _random_permutable(seed):
    rng = np.random.default_rng(seed)
    n, m = 40, 12
    cones = [(capi.CONE_STANDARD_SOC, 3, 2), (capi.CONE_EXPONENTIAL, 10, 1), (capi.CONE_PSD, 20, 3)]
    lb, ub = np.zeros(n), np.full(n, 5.0)
    for t, s, vd in cones:
        L = _block_len(t, vd)
        lb[s:s + L], ub[s:s + L] = -np.inf, np.inf
    B = sp.random(6, n, density=0.3, random_state=seed, format="csr")
    return QPData(c=rng.standard_normal(n), A=sp.random(m, n, density=0.3, random_state=seed + 1, format="csr"),
                  con_lb=-np.ones(m), con_ub=np.ones(m), var_lb=lb, var_ub=ub, Q=(B.T @ B).tocsr(),
                  R=rng.standard_normal((3, n)), cones=cones)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_permute_problem_bit_exact(seed):
    L, O = _libs()
    p = _random_permutable(seed)
    h, ho = build(L, "", p), oracle_build(p)
    n, m = 40, 12
    cp, rp = _iarr([0] * n), _iarr([0] * m)
    libc.srand(seed)
    L.generate_cone_aware_permutation(h.ptr, BLOCK, 3, cp)
    libc.srand(seed + 100)
    L.generate_affine_cone_aware_row_permutation(h.ptr, FULL, 1, rp)
    a = L.permute_problem_return_new(h.ptr, rp, cp)
    b = O.orc_permute_problem_copy(ho.ptr, rp, cp)
    da, db = _dump(a, True), _dump(b, True)
    assert da.keys() == db.keys()
    for key in da:
        if isinstance(da[key], tuple):
            for x, y in zip(da[key], db[key]):
                assert np.array_equal(x, y), key
        elif isinstance(da[key], list):
            assert da[key] == db[key], key
        else:
            assert np.array_equal(da[key], db[key], equal_nan=True), key
    # the permuted matrices are the original ones with rows and columns reordered
    A0 = p.A.toarray()
    rpa, ci, v = da["constraint_matrix"]
    Ap = sp.csr_matrix((v, ci, rpa), shape=(m, n)).toarray()
    assert np.array_equal(Ap, A0[np.array(list(rp))][:, np.array(list(cp))])
    L.qp_problem_free(a)
    O.orc_qp_problem_free(b)
    h.close()
    ho.close()


def test_permuted_solve_maps_back():
    # the oracle solves the original and the permuted LP; repermute_solution
    # maps the permuted solution back onto the original one
    L, O = _libs()
    rng = np.random.default_rng(5)
    n, m = 30, 12
    A = sp.random(m, n, density=0.4, random_state=5, format="csr")
    x0 = rng.uniform(0, 1, n)
    p = QPData(c=rng.standard_normal(n), A=A, con_lb=A @ x0 - 1, con_ub=A @ x0 + 1, var_lb=np.zeros(n),
               var_ub=np.ones(n))
    cp = np.array(rng.permutation(n))
    rp = np.array(rng.permutation(m))
    pp = QPData(c=p.c[cp], A=A[rp][:, cp].tocsr(), con_lb=p.con_lb[rp], con_ub=p.con_ub[rp], var_lb=p.var_lb[cp],
                var_ub=p.var_ub[cp])
    s0 = oracle_solve(p, eps=1e-9)
    s1 = oracle_solve(pp, eps=1e-9)
    assert s0.status == s1.status == "OPTIMAL"
    assert abs(s0.pobj - s1.pobj) <= 1e-7 * (1 + abs(s0.pobj))
    res = capi.pdhcg_result_t()
    xb, yb, rb = np.array(s1.x), np.array(s1.y), np.array(s1.reduced_cost)
    res.num_variables, res.num_constraints = n, m
    res.primal_solution, res.dual_solution, res.reduced_cost = capi.as_ptr(xb), capi.as_ptr(yb), capi.as_ptr(rb)
    L.repermute_solution(C.byref(res), _iarr(rp), _iarr(cp))
    assert np.allclose(xb, s0.x, atol=1e-5)


def test_reference_partition_client():
    # the reference's own test/test_partition_utils.c (known answers of
    # optimize_partition_cuts), compiled unchanged by oracle/Makefile
    # (ref-clients) and linked to this library; host code only
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref",
                       "test_partition_utils")
    if not os.path.exists(exe):
        pytest.skip(f"{exe} not built (needs the reference tree at build time)")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
