"""Seeded synthetic problem families used by the benchmark and the tests.

config0 : small random convex QP (sparse Q = B'B + diag), n=2000, m=1000.
config1 : large sparse QP with a dense low-rank term, power-law (Zipf 1.5)
          row lengths in A, diagonal Q, R (k x n) dense N(0, 1/n), diagonal D.
Both follow the instance descriptions of the benchmark configurations, with
the documented deviation that x0 ~ U[0, 5] for config1 (a start inside the
[0, 5] box, so that the row bounds built around A x0 are feasible).
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

from . import capi
from .problem import QPData


def _unique_columns_per_row(rng, lengths: np.ndarray, n: int) -> np.ndarray:
    """Column indices, uniform without replacement within each row, rows in order."""
    nnz = int(lengths.sum())
    rows = np.repeat(np.arange(lengths.size, dtype=np.int64), lengths)
    cols = rng.integers(0, n, nnz, dtype=np.int64)
    for _ in range(64):
        key = rows * n + cols
        order = np.argsort(key, kind="stable")
        ks = key[order]
        dup = np.zeros(nnz, dtype=bool)
        dup[order[1:][ks[1:] == ks[:-1]]] = True
        if not dup.any():
            break
        cols[dup] = rng.integers(0, n, int(dup.sum()), dtype=np.int64)
    # sort columns inside each row
    key = rows * n + cols
    order = np.argsort(key, kind="stable")
    return cols[order].astype(np.int32)


def random_sparse(rng, m: int, n: int, nnz: int) -> sp.csr_matrix:
    """Exactly nnz positions uniform without replacement, N(0,1) values."""
    flat = rng.choice(m * n, size=nnz, replace=False)
    r, c = np.divmod(flat, n)
    return sp.csr_matrix((rng.standard_normal(nnz), (r, c)), shape=(m, n))


def config0(seed: int = 1000, n: int = 2000, m: int = 1000) -> QPData:
    """Plain FP64 small random convex QP (tolerance 1e-4 in the benchmark)."""
    rng = np.random.default_rng(seed)
    A = random_sparse(rng, m, n, int(round(0.005 * m * n)))
    kb = max(1, n // 10)
    B = random_sparse(rng, kb, n, int(round(0.01 * kb * n)))
    Q = (B.T @ B + sp.diags(rng.uniform(0.1, 1.0, n))).tocsr()
    Q.sort_indices()
    c = rng.standard_normal(n)
    x0 = rng.uniform(-1.0, 1.0, n)
    ax = A @ x0
    return QPData(c=c, A=A, Q=Q, con_lb=ax - rng.uniform(0, 1, m), con_ub=ax + rng.uniform(0, 1, m),
                  var_lb=np.full(n, -10.0), var_ub=np.full(n, 10.0))


def zipf_lengths(rng, m: int, nnz: int, a: float = 1.5, max_len: int = 5000) -> np.ndarray:
    """Row lengths ~ Zipf(a) truncated to [1, max_len], rescaled to sum exactly nnz."""
    raw = rng.zipf(a, size=m)
    while True:
        bad = raw > max_len
        if not bad.any():
            break
        raw[bad] = rng.zipf(a, size=int(bad.sum()))
    lens = np.maximum(1, np.round(raw * (nnz / raw.sum()))).astype(np.int64)
    lens = np.minimum(lens, max_len)
    diff = nnz - int(lens.sum())
    order = np.argsort(-lens, kind="stable")
    i = 0
    while diff != 0:
        r = order[i % m]
        step = 1 if diff > 0 else -1
        if 1 <= lens[r] + step <= max_len:
            lens[r] += step
            diff -= step
        i += 1
    return lens


def config1(scale: float = 1.0, seed: int = 1001, rank: int = 50) -> QPData:
    """Large sparse QP + dense low-rank term; scale shrinks n, m and nnz together."""
    rng = np.random.default_rng(seed)
    n = max(10, int(round(1_000_000 * scale)))
    m = max(5, int(round(500_000 * scale)))
    nnz = max(m, int(round(10_000_000 * scale)))
    lens = zipf_lengths(rng, m, nnz, 1.5, min(5000, n))
    cols = _unique_columns_per_row(rng, lens, n)
    rp = np.zeros(m + 1, dtype=np.int64)
    np.cumsum(lens, out=rp[1:])
    vals = rng.standard_normal(int(rp[-1]))
    A = sp.csr_matrix((vals, cols, rp.astype(np.int32)), shape=(m, n))
    qd = rng.uniform(0.0, 1.0, n)
    Q = sp.diags(qd).tocsr()
    R = rng.standard_normal((rank, n)) * np.sqrt(1.0 / n)
    D = rng.uniform(0.5, 2.0, rank)
    c = rng.standard_normal(n)
    x0 = rng.uniform(0.0, 5.0, n)
    ax = A @ x0
    return QPData(c=c, A=A, Q=Q, R=R, D=D, con_lb=ax - rng.uniform(0, 1, m), con_ub=ax + rng.uniform(0, 1, m),
                  var_lb=np.zeros(n), var_ub=np.full(n, 5.0), fmt={"R": "dense"})


def random_problem(kind: str, seed: int = 0, n: int = 60, m: int = 30) -> QPData:
    """Small random instances of each quadratic-objective type."""
    rng = np.random.default_rng(seed)
    A = random_sparse(rng, m, n, max(1, int(0.2 * m * n)))
    x0 = rng.uniform(-1, 1, n)
    ax = A @ x0
    base = dict(c=rng.standard_normal(n), A=A, con_lb=ax - rng.uniform(0, 1, m), con_ub=ax + rng.uniform(0, 1, m),
                var_lb=np.full(n, -10.0), var_ub=np.full(n, 10.0))
    if kind == "lp":
        return QPData(**base)
    if kind == "diag":
        return QPData(Q=sp.diags(rng.uniform(0.1, 2.0, n)).tocsr(), **base)
    if kind == "sparse":
        B = random_sparse(rng, max(2, n // 5), n, max(1, int(0.3 * (n // 5) * n)))
        return QPData(Q=(B.T @ B + sp.diags(rng.uniform(0.1, 1.0, n))).tocsr(), **base)
    if kind == "lowrank":
        return QPData(R=rng.standard_normal((6, n)) * 0.5, fmt={"R": "dense"}, **base)
    if kind == "lowrank_diag":
        return QPData(Q=sp.diags(rng.uniform(0.0, 1.0, n)).tocsr(), R=rng.standard_normal((5, n)) / np.sqrt(n),
                      D=rng.uniform(0.5, 2.0, 5), fmt={"R": "dense"}, **base)
    if kind == "lowrank_dense_d":
        M = rng.standard_normal((4, 4))
        return QPData(R=rng.standard_normal((4, n)) * 0.5, D=M.T @ M + 0.2 * np.eye(4), fmt={"R": "dense"}, **base)
    if kind == "lowrank_sparse_r":
        # R with 5% density: stored as CSR (R and R') by the device solver
        R = random_sparse(rng, 8, n, max(1, int(0.05 * 8 * n)))
        B = random_sparse(rng, max(2, n // 20), n, max(1, int(0.05 * (n // 20) * n)))
        return QPData(Q=(B.T @ B + sp.diags(rng.uniform(0.1, 1.0, n))).tocsr(), R=R, **base)
    raise ValueError(kind)


def portfolio_tracking(n: int, k: int, seed=1003, te_factor: float = 1.5) -> QPData:
    """Long-only portfolio with a tracking-error limit as an affine SOC.

    min 1/2 x'(diag(q) + R' D R) x + c'x  s.t.  1'x = 1,  0 <= x <= 0.1,
    || D^1/2 R x || <= tau  (rows F = [D^1/2 R; 0; 0], g = (0, ..., 0, 0, tau)).
    R is k x n with N(0, 1/k) entries, D = diag U[.01, .1], q ~ U[1e-4, 1e-2],
    c = -(0.05 + 0.02 xi), tau = te_factor x the tracking error of equal weights.
    """
    rng = np.random.default_rng(seed)
    R = rng.standard_normal((k, n)) * np.sqrt(1.0 / k)
    D = rng.uniform(0.01, 0.1, k)
    q = rng.uniform(1e-4, 1e-2, n)
    c = -(0.05 + 0.02 * rng.standard_normal(n))
    eq = np.full(n, 1.0 / n)
    tau = te_factor * float(np.linalg.norm(np.sqrt(D) * (R @ eq)))
    F = np.vstack([np.sqrt(D)[:, None] * R, np.zeros((2, n))])
    g = np.zeros(k + 2)
    g[-1] = tau
    return QPData(c=c, A=sp.csr_matrix(np.ones((1, n))), con_lb=[1.0], con_ub=[1.0], var_lb=np.zeros(n),
                  var_ub=np.full(n, 0.1), Q=sp.diags(q).tocsr(), R=R, D=D, F=sp.csr_matrix(F), g=g,
                  affine_cones=[(1, 0, k)], fmt={"R": "dense"})


def portfolio_batch(count: int, seed: int = 1003, n_range=(50, 2000), k_range=(10, 50),
                    te_factor: float = 1.5) -> list:
    """The batched portfolio workload: `count` independent instances of
    portfolio_tracking with n ~ U{n_range} assets and k ~ U{k_range} factors;
    instance i draws its data from the seed pair (seed, i)."""
    rng = np.random.default_rng(seed)
    ns = rng.integers(n_range[0], n_range[1] + 1, size=count)
    ks = rng.integers(k_range[0], k_range[1] + 1, size=count)
    return [portfolio_tracking(int(n), int(k), seed=(seed, i), te_factor=te_factor)
            for i, (n, k) in enumerate(zip(ns, ks))]


class ParametricFamily:
    """Parametric instances of the large sparse QP + low-rank family for the
    learned warm start: one shared A (Zipf 1.5 row lengths), diagonal Q, R
    (rank x n, N(0, 1/n)) and D; a parameter vector z ~ N(0, I) of z_dim
    entries moves the objective and the row bounds:

        c(z)  = c_bar + W_c z + N(0, 0.01^2)
        x0(z) = clip(2.5 + W_x z, 0, 5)        (clipped before the bounds)
        l, u  = A x0(z) -/+ s + N(0, 0.01^2),  s ~ U[0.1, 1] fixed

    with W_c entries N(0, c_spread^2 / z_dim) and W_x entries
    N(0, x_spread^2 / z_dim) (each coordinate of W z then has standard
    deviation c_spread or x_spread; x_spread is kept small so that the row
    bounds move less than their width s).  x0(z) lies in the box [0, 5]
    and A x0(z) within [l, u] unless a noise draw exceeds s (at least ten
    standard deviations), so every instance is feasible; `clipped` counts the
    entries of x0 that the clip moved.
    """

    def __init__(self, n: int = 2000, m: int = 1000, nnz: int = 20000, rank: int = 50, z_dim: int = 16,
                 seed: int = 1004, c_spread: float = 0.1, x_spread: float = 0.1):
        rng = np.random.default_rng(seed)
        self.n, self.m, self.z_dim = n, m, z_dim
        lens = zipf_lengths(rng, m, nnz, 1.5, min(5000, n))
        cols = _unique_columns_per_row(rng, lens, n)
        rp = np.zeros(m + 1, dtype=np.int64)
        np.cumsum(lens, out=rp[1:])
        self.A = sp.csr_matrix((rng.standard_normal(int(rp[-1])), cols, rp.astype(np.int32)), shape=(m, n))
        self.Q = sp.diags(rng.uniform(0.0, 1.0, n)).tocsr()
        self.R = rng.standard_normal((rank, n)) * np.sqrt(1.0 / n)
        self.D = rng.uniform(0.5, 2.0, rank)
        self.c_bar = rng.standard_normal(n)
        self.W_c = rng.standard_normal((n, z_dim)) * (c_spread / np.sqrt(z_dim))
        self.W_x = rng.standard_normal((n, z_dim)) * (x_spread / np.sqrt(z_dim))
        self.s = rng.uniform(0.1, 1.0, m)
        self.clipped = 0

    def instance(self, z, noise_seed) -> QPData:
        rng = np.random.default_rng(noise_seed)
        z = np.asarray(z, dtype=np.float64)
        c = self.c_bar + self.W_c @ z + 0.01 * rng.standard_normal(self.n)
        raw = 2.5 + self.W_x @ z
        x0 = np.clip(raw, 0.0, 5.0)
        self.clipped += int(np.count_nonzero(x0 != raw))
        ax = self.A @ x0
        lb = ax - self.s + 0.01 * rng.standard_normal(self.m)
        ub = ax + self.s + 0.01 * rng.standard_normal(self.m)
        return QPData(c=c, A=self.A, Q=self.Q, R=self.R, D=self.D, con_lb=lb, con_ub=ub, var_lb=np.zeros(self.n),
                      var_ub=np.full(self.n, 5.0), fmt={"R": "dense"})

    def sample(self, count: int, seed: int):
        """(Z, problems): count parameter vectors and their instances;
        instance i draws z and its noise from the seed pair (seed, i)."""
        Z = np.empty((count, self.z_dim), dtype=np.float32)
        problems = []
        for i in range(count):
            rng = np.random.default_rng((seed, i))
            Z[i] = rng.standard_normal(self.z_dim)
            problems.append(self.instance(Z[i], (seed, i, 1)))
        return Z, problems


def conic_mix(scale: float = 1.0, seed: int = 1002, psd_orders=(16, 128), with_affine: bool = True) -> QPData:
    """Conic mix: SOC, RSOC, exponential, power and PSD blocks plus an affine
    SOC block F x + g, with a sparse positive definite Q.

    The stated parts of this family (30% of n in SOC blocks of U{3..64} slots,
    10% in RSOC blocks of U{3..32} slots, 100k exponential cones, 50k power
    cones with alpha ~ U(0.1, 0.9), 200 PSD blocks of order U{16..128}) add up
    to about 1.8M variables, not 600k.  Resolution: the exponential, power
    and PSD counts are scaled by one factor (1/3) so that n is about 600k at
    scale 1: 33,333 exponential cones, 16,667 power cones and 67 PSD blocks
    (about 212k svec slots); SOC and RSOC blocks then fill 30% and 10% of n.
    `scale` multiplies every count.  Block order is shuffled.

    A: m = n / 4 rows with 2M nnz (scale 1) at uniform positions, N(0, 1);
    row bounds A x_int -/+ U[0, 1] around a point x_int strictly inside every
    cone.  F: p = n / 6 rows, 2M nnz N(0, 1), grouped into affine SOC blocks
    of U{3..64} rows; g = s0 - F x_int with s0 = (0, ..., 0, 2) per block, so
    x_int is strictly feasible for the affine cones too.  Q = B'B + diag(U[0.1,
    1]) with B of n / 30 rows and 10 entries per row (nnz(Q) about 2M at
    scale 1); the diagonal part makes the objective strongly convex, so the
    problem is bounded for any c.  c ~ N(0, 1).  Cone slots are free.
    """
    rng = np.random.default_rng(seed)
    n_exp = max(1, int(round(33_333 * scale)))
    n_pow = max(1, int(round(16_667 * scale)))
    n_psd = max(1, int(round(67 * scale)))
    orders = rng.integers(psd_orders[0], psd_orders[1] + 1, size=n_psd)
    fixed_slots = 3 * n_exp + 3 * n_pow + int(sum(k * (k + 1) // 2 for k in orders))
    n_target = fixed_slots / 0.6
    blocks = [("exp", 1, 0.0)] * n_exp + [("pow", 1, float(a)) for a in rng.uniform(0.1, 0.9, size=n_pow)]
    blocks += [("psd", int(k), 0.0) for k in orders]
    soc_slots = 0
    while soc_slots < 0.3 * n_target:
        size = int(rng.integers(3, 65))
        blocks.append(("soc", size - 2, 0.0))
        soc_slots += size
    rsoc_slots = 0
    while rsoc_slots < 0.1 * n_target:
        size = int(rng.integers(3, 33))
        blocks.append(("rsoc", size - 2, 0.0))
        rsoc_slots += size
    order = rng.permutation(len(blocks))
    kinds = {"soc": capi.CONE_STANDARD_SOC, "rsoc": capi.CONE_ROTATED_SOC, "exp": capi.CONE_EXPONENTIAL,
             "pow": capi.CONE_POWER, "psd": capi.CONE_PSD}
    slots = {"soc": lambda v: v + 2, "rsoc": lambda v: v + 2, "exp": lambda v: 3, "pow": lambda v: 3,
             "psd": lambda v: v * (v + 1) // 2}
    n = int(sum(slots[blocks[i][0]](blocks[i][1]) for i in order))
    xi = np.empty(n)
    cones = []
    pos = 0
    for i in order:
        kind, vd, alpha = blocks[i]
        L = slots[kind](vd)
        if kind == "soc":
            v = rng.standard_normal(vd + 1)
            xi[pos:pos + vd + 1] = v
            xi[pos + vd + 1] = np.linalg.norm(v) + 1.0
        elif kind == "rsoc":
            v = rng.standard_normal(vd)
            xi[pos:pos + vd] = v
            xi[pos + vd] = xi[pos + vd + 1] = np.sqrt(0.5 * v @ v) + 1.0
        elif kind == "exp":
            xi[pos:pos + 3] = [0.0, 1.0, 2.0]
        elif kind == "pow":
            xi[pos:pos + 3] = [1.0, 1.0, 0.3]
        else:
            M = rng.standard_normal((vd, vd)) / np.sqrt(vd)
            S = M @ M.T + np.eye(vd)
            r, c = np.tril_indices(vd)
            order_cm = np.lexsort((r, c))  # svec: lower triangle, column-major
            vals = S[r[order_cm], c[order_cm]] * np.where(r[order_cm] == c[order_cm], 1.0, np.sqrt(2.0))
            xi[pos:pos + L] = vals
        cones.append((kinds[kind], pos, vd, alpha) if kind == "pow" else (kinds[kind], pos, vd))
        pos += L
    nnz_a = max(n, int(round(2_000_000 * scale)))
    m = max(2, n // 4)
    A = random_sparse(rng, m, n, min(nnz_a, m * n))
    ax = A @ xi
    kb = max(2, n // 30)
    B = sp.csr_matrix((rng.standard_normal(10 * kb), (np.repeat(np.arange(kb), 10), rng.integers(0, n, 10 * kb))),
                      shape=(kb, n))
    Q = (B.T @ B + sp.diags(rng.uniform(0.1, 1.0, n))).tocsr()
    Q.sort_indices()
    data = dict(c=rng.standard_normal(n), A=A, Q=Q, con_lb=ax - rng.uniform(0, 1, m),
                con_ub=ax + rng.uniform(0, 1, m), var_lb=np.full(n, -np.inf), var_ub=np.full(n, np.inf),
                cones=cones)
    if with_affine:
        p_target = max(3, n // 6)
        acones, rows = [], 0
        while rows < p_target:
            size = int(rng.integers(3, 65))
            acones.append((capi.CONE_STANDARD_SOC, rows, size - 2))
            rows += size
        F = random_sparse(rng, rows, n, min(max(rows, int(round(2_000_000 * scale))), rows * n))
        s0 = np.zeros(rows)
        for (_, st, vd) in acones:
            s0[st + vd + 1] = 2.0
        data.update(F=F, g=s0 - F @ xi, affine_cones=acones)
    return QPData(**data)
