"""Problem-file readers (read_mps_file, read_cbf_file; reference
src/mps_parser.c, src/cbf_parser.c) and their Python front end
(pdhcg_amd.pdhcg.read_problem_file, Model.read_file), host code.

Pinned by hand-derived expectations:
- an MPS/QPS file exercising every section (N/L/G/E rows, an integer
  MARKER, the objective row's RHS as the negated constant, RANGES on G, L and
  E rows of both signs, the seven bound types, QUADOBJ filled symmetrically,
  QCMATRIX rows, OBJSENSE MAX) against the problem worked out from the MPS
  rules written in the test;
- a random LP written as MPS (values printed exactly) and read back, plain
  and gzip-compressed: identical arrays;
- the reference's CBF fixture (tests/data/cbf_q3_smoke.cbf, kept in
  tests/golden/) and an affine-cone CBF model, with the component orders of
  CBF Q (t, v) -> (v, w, z), QR (s, t, v) -> (v, s, t), EXP reversed;
- rejections (integer and PSD sections, unknown cones, truncated data).
"""
import ctypes as C
import gzip
import os

import numpy as np
import pytest
import scipy.sparse as sp

import pdhcg_amd
from pdhcg_amd import capi
from pdhcg_amd.pdhcg import ConeType, Model, read_problem_file

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _read(path, kind):
    L = pdhcg_amd.lib()
    fn = L.read_cbf_file if kind == "cbf" else L.read_mps_file
    fn.restype = C.POINTER(capi.qp_problem_t)
    fn.argtypes = [C.c_char_p]
    ptr = fn(str(path).encode())
    if not ptr:
        return None
    q = ptr.contents
    n, m = q.num_variables, q.num_constraints
    arr = lambda p_, k: np.ctypeslib.as_array(p_, shape=(k,)).copy() if (p_ and k > 0) else np.zeros(k)

    def csr(comp, rows, cols, nnz):
        if not comp:
            return None
        c = comp.contents
        rp = np.ctypeslib.as_array(c.row_ptr, shape=(rows + 1,)).copy()
        if nnz == 0:
            return sp.csr_matrix((rows, cols)), rp
        return sp.csr_matrix((arr(c.val, nnz), arr(c.col_ind, nnz), rp), shape=(rows, cols)), rp

    out = dict(n=n, m=m, c=arr(q.objective_vector, n), c0=q.objective_constant, lb=arr(q.variable_lower_bound, n),
               ub=arr(q.variable_upper_bound, n), cl=arr(q.constraint_lower_bound, m),
               cu=arr(q.constraint_upper_bound, m), off=arr(q.affine_cone_offset, m) if q.affine_cone_offset else None,
               A=csr(q.constraint_matrix, m, n, q.constraint_matrix_num_nonzeros),
               Q=csr(q.objective_sparse_matrix, n, n, q.objective_sparse_matrix_num_nonzeros),
               cones=[(q.cones.type[i], q.cones.start_idx[i], q.cones.v_dim[i]) for i in range(q.cones.num_cones)],
               acones=[(q.affine_cones.type[i], q.affine_cones.start_idx[i], q.affine_cones.v_dim[i])
                       for i in range(q.affine_cones.num_cones)],
               qcs=[(q.quadratic_constraint_row_indices[k],
                     csr(q.quadratic_constraint_matrices[k], n, n, q.quadratic_constraint_matrix_num_nonzeros[k])[0])
                    for k in range(q.num_quadratic_constraints)])
    addr = C.c_void_p.from_buffer(q.cones, capi.cone_blocks_t.is_fixed.offset).value
    out["fixed"] = np.array((C.c_int8 * n).from_address(addr)) if addr else None
    out["start"] = arr(q.primal_start, n) if q.primal_start else None
    L.qp_problem_free(ptr)
    return out


MPS = """NAME          TESTQP
* a comment line
OBJSENSE
    MAX
ROWS
 N  obj
 L  lim1
 G  lim2
 E  eq1
 E  eq2
 N  free2
COLUMNS
    MARKER                 'MARKER'                 'INTORG'
    x1        obj       1.0        lim1      1.0
    x1        lim2      2.0
    MARKER                 'MARKER'                 'INTEND'
    x2        obj       -2.0       eq1       3.0
    x2        free2     9.0
    x3        lim1      -1.0       eq2       1.0
    x4        obj       0.5
    x5        eq1       1.0
    x6        obj       1.0
    x7        lim2      1.0
RHS
    RHS       obj       7.5        lim1      4.0
    RHS       lim2      1.0        eq1       2.0
    RHS       eq2       3.0
RANGES
    RNG       lim1      2.5        lim2      -1.5
    RNG       eq1       0.5        eq2       -2.0
BOUNDS
 UP BND       x1        4.0
 LO BND       x2        -1.0
 FX BND       x3        2.0
 FR BND       x4
 MI BND       x5
 PL BND       x6
 BV BND       x7
QUADOBJ
    x1        x1        2.0
    x1        x2        0.5
    x2        x2        1.0
QCMATRIX   lim1
    x1        x1        1.0
    x3        x3        2.0
ENDATA
"""


@pytest.mark.parametrize("maximize", [False, True])
def test_mps_every_section(tmp_path, maximize):
    # A line holding only "MAX" reads as a (unknown) section header, so the
    # reference's reader keeps minimizing for it (src/mps_parser.c:457-492);
    # "MAX" followed by a second token inside OBJSENSE selects maximization.
    f = tmp_path / "t.mps"
    f.write_text(MPS.replace("    MAX\n", "    MAX  SENSE\n") if maximize else MPS)
    P = _read(f, "mps")
    inf = np.inf
    sgn = -1.0 if maximize else 1.0
    assert (P["n"], P["m"]) == (7, 4)  # the second N row is dropped
    # maximization negates c, Q, the constant (RHS 7.5 on the objective row is -7.5) and the QC matrices
    assert np.array_equal(P["c"], sgn * np.array([1.0, -2.0, 0.0, 0.5, 0.0, 1.0, 0.0]))
    assert P["c0"] == sgn * -7.5
    # rows lim1 (L), lim2 (G), eq1 (E), eq2 (E)
    A = P["A"][0].toarray()
    want = np.zeros((4, 7))
    want[0, 0], want[1, 0], want[2, 1], want[0, 2], want[3, 2], want[2, 4], want[1, 6] = 1, 2, 3, -1, 1, 1, 1
    assert np.array_equal(A, want)
    # RHS then RANGES: L [4 - 2.5, 4]; G [1, 1 + 1.5]; E(+0.5) [2, 2.5]; E(-2) [1, 3]
    assert np.array_equal(P["cl"], [1.5, 1.0, 2.0, 1.0]) and np.array_equal(P["cu"], [4.0, 2.5, 2.5, 3.0])
    assert np.array_equal(P["lb"], [0.0, -1.0, 2.0, -inf, -inf, 0.0, 0.0])
    assert np.array_equal(P["ub"], [4.0, inf, 2.0, inf, inf, inf, 1.0])
    Q = P["Q"][0].toarray()
    wantQ = np.zeros((7, 7))
    wantQ[0, 0], wantQ[0, 1], wantQ[1, 0], wantQ[1, 1] = 2.0, 0.5, 0.5, 1.0
    assert np.array_equal(Q, sgn * wantQ)
    assert len(P["qcs"]) == 1 and P["qcs"][0][0] == 0
    assert np.array_equal(P["qcs"][0][1].toarray(), sgn * np.diag([1.0, 0, 2.0, 0, 0, 0, 0]))
    assert np.array_equal(P["off"], np.zeros(4)) and P["cones"] == [] and P["acones"] == []


def _write_mps(path, c, A, cl, cu, lb, ub):
    lines = ["NAME RT", "ROWS", " N obj"]
    m, n = A.shape
    for i in range(m):
        lines.append(f" {'E' if cl[i] == cu[i] else 'G'} r{i}")
    lines.append("COLUMNS")
    Ac = A.tocsc()
    for j in range(n):
        lines.append(f"    x{j} obj {float(c[j])!r}")
        for k in range(Ac.indptr[j], Ac.indptr[j + 1]):
            lines.append(f"    x{j} r{Ac.indices[k]} {float(Ac.data[k])!r}")
    lines.append("RHS")
    lines += [f"    RHS r{i} {float(cl[i])!r}" for i in range(m)]
    lines.append("RANGES")
    lines += [f"    RNG r{i} {float(cu[i] - cl[i])!r}" for i in range(m) if cl[i] != cu[i]]
    lines.append("BOUNDS")
    for j in range(n):
        lines.append(f" LO BND x{j} {float(lb[j])!r}")
        lines.append(f" UP BND x{j} {float(ub[j])!r}")
    lines.append("ENDATA")
    path.write_text("\n".join(lines) + "\n")


def test_mps_round_trip_plain_and_gzip(tmp_path):
    rng = np.random.default_rng(3)
    n, m = 40, 25
    A = sp.random(m, n, density=0.2, random_state=rng, format="csr")
    A.data = rng.standard_normal(A.nnz)
    c, lb = rng.standard_normal(n), -rng.uniform(0, 2, n)
    ub = lb + rng.uniform(0.5, 3.0, n)
    cl = rng.standard_normal(m)
    cu = np.where(rng.random(m) < 0.3, cl, cl + rng.uniform(0.1, 1.0, m))
    f = tmp_path / "rt.mps"
    _write_mps(f, c, A, cl, cu, lb, ub)
    gz = tmp_path / "rt.mps.gz"
    gz.write_bytes(gzip.compress(f.read_bytes()))
    for path in (f, gz):
        P = _read(path, "mps")
        assert (P["n"], P["m"]) == (n, m)
        assert np.array_equal(P["c"], c) and np.array_equal(P["lb"], lb) and np.array_equal(P["ub"], ub)
        assert np.array_equal(P["cl"], cl)
        assert np.allclose(P["cu"], cu, rtol=0, atol=1e-15)  # cl + (cu - cl)
        assert (P["A"][0] != A).nnz == 0


def test_reference_cbf_fixture():
    # tests/data/cbf_q3_smoke.cbf of the reference: min t s.t. v1 = 3, v2 = 4, (t, v1, v2) in Q
    path = os.path.join(GOLDEN, "cbf_q3_smoke.cbf")
    P = _read(path, "cbf")
    assert (P["n"], P["m"]) == (3, 2)
    assert P["cones"] == [(capi.CONE_STANDARD_SOC, 0, 1)]
    assert np.array_equal(P["c"], [0.0, 0.0, 1.0])  # CBF t (slot 0) -> z (last slot)
    assert np.array_equal(P["A"][0].toarray(), [[1.0, 0.0, 0.0], [0.0, 1.0, 0.0]])
    assert np.array_equal(P["cl"], [3.0, 4.0]) and np.array_equal(P["cu"], [3.0, 4.0])
    assert np.all(np.isinf(P["lb"])) and np.all(np.isinf(P["ub"]))
    raw = read_problem_file(path)
    assert raw["cones"]["types"][0] == ConeType.SOC and raw["cones"]["starts"][0] == 0
    m = Model.read_file(path)
    assert m.num_vars == 3 and m.num_constrs == 2 and len(m._variable_cones) == 1


CBF_MIX = """VER
3
OBJSENSE
MAX
VAR
7 3
L+ 2
QR 3
EXP 2
"""


def test_cbf_cone_orders_and_affine_rows(tmp_path):
    # variables: L+ (2), QR (s, t, v) -> internal (v, s, t), and a bad EXP of dimension 2 first
    bad = tmp_path / "bad.cbf"
    bad.write_text(CBF_MIX)
    assert _read(bad, "cbf") is None  # EXP needs dimension 3
    model = """VER
3
OBJSENSE
MAX
VAR
8 3
L+ 2
QR 3
EXP 3
CON
5 3
Q 3
L- 1
EXP 1
OBJACOORD
3
0 1.0
2 2.0
5 3.0
OBJBCOORD
4.0
ACOORD
4
0 3 1.0
1 4 1.0
3 0 2.0
3 0 1.0
BCOORD
2
2 5.0
3 -1.0
"""
    f = tmp_path / "mix.cbf"
    f.write_text(model.replace("EXP 1\n", "L= 1\n").replace("CON\n5 3", "CON\n5 3"))
    P = _read(f, "cbf")
    assert P["n"] == 8 and P["m"] == 5
    # columns: L+ block (0, 1), then QR at 2 (v at 2, s at 3, t at 4), EXP at 5 reversed
    assert P["cones"] == [(capi.CONE_ROTATED_SOC, 2, 1), (capi.CONE_EXPONENTIAL, 5, 1)]
    # MAX negates c but not the constant; CBF var 2 (QR's s) -> column 3; var 5 (EXP x0) -> column 7
    want_c = np.zeros(8)
    want_c[0], want_c[3], want_c[7] = -1.0, -2.0, -3.0
    assert np.array_equal(P["c"], want_c) and P["c0"] == 4.0
    assert np.array_equal(P["lb"][:2], [0.0, 0.0]) and np.all(np.isinf(P["lb"][2:]))
    # rows: Q block (t, v1, v2) -> rows (v1, v2, t) = (1, 2, 0)... CBF row r -> internal row
    A = P["A"][0].toarray()
    assert A[2, 4] == 1.0   # CBF row 0 (Q's t) -> row 2, CBF var 3 (QR's t) -> column 4
    assert A[0, 2] == 1.0   # CBF row 1 -> row 0, CBF var 4 (QR's v) -> column 2
    assert A[3, 0] == 3.0   # CBF row 3 (L-), duplicates summed
    assert P["acones"] == [(capi.CONE_STANDARD_SOC, 0, 1)]
    assert np.array_equal(P["off"], [0.0, 5.0, 0.0, 0.0, 0.0])  # BCOORD of the Q block's v2 (CBF row 2 -> 1)
    assert P["cu"][3] == 1.0 and P["cl"][3] == -np.inf         # L-: A x + b <= 0 -> A x <= 1
    assert P["cl"][4] == 0.0 and P["cu"][4] == 0.0             # L=


@pytest.mark.parametrize("section", ["INT\n1\n0\n", "PSDVAR\n1\n2\n", "FOO\n"])
def test_cbf_rejections(tmp_path, section):
    f = tmp_path / "r.cbf"
    f.write_text("VER\n3\nVAR\n1 1\nF 1\n" + section)
    assert _read(f, "cbf") is None
    g = tmp_path / "t.cbf"
    g.write_text("VER\n3\nVAR\n1 1\nF 1\nOBJACOORD\n2\n0 1.0\n")
    assert _read(g, "cbf") is None  # truncated
