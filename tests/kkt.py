"""Independent KKT recomputation in numpy for returned (x, y).

Given the original (unscaled) problem data and a primal/dual pair, computes
relative primal infeasibility, relative dual infeasibility (with the optimal
box multiplier) and the relative duality gap using the definitions of the
reference (relative to 1 + ||b||_inf and 1 + max(||c||, ||Qx||, ||A'y||)).
It shares no code with either solver.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp


def _Qx(p, x):
    q = np.zeros_like(x)
    if p.Q is not None:
        q = q + sp.csr_matrix(p.Q) @ x
    if p.R is not None:
        R = p.R.toarray() if sp.issparse(p.R) else np.asarray(p.R)
        u = R @ x
        if p.D is not None:
            D = np.diag(p.D) if np.ndim(p.D) == 1 else (p.D.toarray() if sp.issparse(p.D) else np.asarray(p.D))
            u = D @ u
        q = q + R.T @ u
    return q


def kkt(p, x, y):
    n = x.size
    A = sp.csr_matrix(p.A) if p.A is not None else sp.csr_matrix((0, n))
    c = np.asarray(p.c, dtype=float)
    lb = np.full(n, -np.inf) if p.var_lb is None else np.asarray(p.var_lb, float).copy()
    ub = np.full(n, np.inf) if p.var_ub is None else np.asarray(p.var_ub, float).copy()
    lb[lb <= -1e30] = -np.inf
    ub[ub >= 1e30] = np.inf
    m = A.shape[0]
    cl = np.full(m, -np.inf) if p.con_lb is None else np.asarray(p.con_lb, float)
    cu = np.full(m, np.inf) if p.con_ub is None else np.asarray(p.con_ub, float)
    ax = A @ x
    pres = ax - np.clip(ax, cl, cu)
    box = x - np.clip(x, lb, ub)
    bnorm = 0.0
    for lo, hi in zip(cl, cu):
        if np.isfinite(lo) and lo != hi:
            bnorm = max(bnorm, abs(lo))
        if np.isfinite(hi):
            bnorm = max(bnorm, abs(hi))
    rel_primal = max(np.max(np.abs(pres), initial=0.0), np.max(np.abs(box), initial=0.0)) / (1.0 + bnorm)
    qx = _Qx(p, x)
    aty = A.T @ y[:m] if m > 0 else np.zeros(n)
    grad = c + qx - aty
    # box multiplier: the gradient on variables at (within act_tol of) the bound
    # it pushes against, zero on interior variables
    act_tol = 1e-5 * (1.0 + np.abs(np.where(np.isfinite(lb), lb, 0.0)) + np.abs(np.where(np.isfinite(ub), ub, 0.0)))
    at_lb = np.isfinite(lb) & (x - lb <= act_tol)
    at_ub = np.isfinite(ub) & (ub - x <= act_tol)
    lam = np.where((grad > 0) & at_lb, grad, 0.0) + np.where((grad < 0) & at_ub, grad, 0.0)
    dres = grad - lam
    den = 1.0 + max(np.max(np.abs(c)), np.max(np.abs(qx), initial=0.0), np.max(np.abs(aty), initial=0.0))
    rel_dual = np.max(np.abs(dres), initial=0.0) / den
    pobj = c @ x + 0.5 * x @ qx + p.objective_constant
    ymx = y[:m]
    dual_rows = np.sum(np.where(ymx > 0, np.where(np.isfinite(cl), cl, 0.0) * ymx, 0.0)) + \
        np.sum(np.where(ymx < 0, np.where(np.isfinite(cu), cu, 0.0) * ymx, 0.0))
    dual_box = np.sum(np.where(lam > 0, np.where(np.isfinite(lb), lb, 0.0) * lam, 0.0)) + \
        np.sum(np.where(lam < 0, np.where(np.isfinite(ub), ub, 0.0) * lam, 0.0))
    dobj = dual_rows + dual_box - 0.5 * x @ qx + p.objective_constant
    rel_gap = abs(pobj - dobj) / (1.0 + abs(pobj) + abs(dobj))
    return dict(rel_primal=rel_primal, rel_dual=rel_dual, rel_gap=rel_gap, pobj=pobj, dobj=dobj)
