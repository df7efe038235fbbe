"""Learned warm start: host side of the MLP in libpdhcg_amd.so.

The network (pdhcg_mlp_* in include/pdhcg_mlp.h) maps the parameter vector z
of an instance of a parametric problem family to a starting point; the
starting point enters the solve through set_start_values (reference
src/pdhcg.c:569-618), exactly where a user-supplied start would.  The
reference has no learned warm start.

Targets are standardised per coordinate (mean and standard deviation over the
training set) before training and mapped back after prediction; the primal
part is clipped to the variable bounds of the instance.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import replace
from typing import Optional

import numpy as np

_declared = set()


def _lib():
    from . import lib
    L = lib()
    if id(L) not in _declared:
        fp, dp = C.POINTER(C.c_float), C.POINTER(C.c_double)
        L.pdhcg_mlp_create.restype = C.c_void_p
        L.pdhcg_mlp_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_uint]
        L.pdhcg_mlp_free.restype = None
        L.pdhcg_mlp_free.argtypes = [C.c_void_p]
        L.pdhcg_mlp_num_params.restype = C.c_int
        L.pdhcg_mlp_num_params.argtypes = [C.c_void_p]
        for name in ("pdhcg_mlp_get_params", "pdhcg_mlp_set_params"):
            getattr(L, name).restype = C.c_int
            getattr(L, name).argtypes = [C.c_void_p, fp]
        L.pdhcg_mlp_loss_grad.restype = C.c_int
        L.pdhcg_mlp_loss_grad.argtypes = [C.c_void_p, fp, fp, C.c_int, dp, fp]
        L.pdhcg_mlp_train.restype = C.c_int
        L.pdhcg_mlp_train.argtypes = [C.c_void_p, fp, fp, C.c_int, C.c_int, C.c_int, C.c_float, C.c_uint, dp]
        L.pdhcg_mlp_predict.restype = C.c_int
        L.pdhcg_mlp_predict.argtypes = [C.c_void_p, fp, C.c_int, fp]
        L.pdhcg_mfma_gemm.restype = C.c_int
        L.pdhcg_mfma_gemm.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, fp, fp, fp]
        _declared.add(id(L))
    return L


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


def _fp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def mfma_gemm(A, B, ta: bool = False, tb: bool = False) -> np.ndarray:
    """op(A) @ op(B) through the device BF16 matrix-core product (FP32 out)."""
    L = _lib()
    A, B = _f32(A), _f32(B)
    M, K = (A.shape[1], A.shape[0]) if ta else A.shape
    K2, N = (B.shape[1], B.shape[0]) if tb else B.shape
    if K != K2:
        raise ValueError("inner dimensions differ")
    out = np.empty((M, N), dtype=np.float32)
    if L.pdhcg_mfma_gemm(int(ta), int(tb), M, N, K, _fp(A), _fp(B), _fp(out)) != 0:
        raise RuntimeError("pdhcg_mfma_gemm failed")
    return out


class MLP:
    """in -> hidden -> hidden -> out network with GELU, trained on the device."""

    def __init__(self, in_dim: int, hidden: int, out_dim: int, seed: int = 0):
        self.L = _lib()
        self.dims = (in_dim, hidden, out_dim)
        self.h = self.L.pdhcg_mlp_create(in_dim, hidden, out_dim, seed)
        if not self.h:
            raise RuntimeError("pdhcg_mlp_create failed")

    @property
    def num_params(self) -> int:
        return int(self.L.pdhcg_mlp_num_params(self.h))

    def get_params(self) -> np.ndarray:
        out = np.empty(self.num_params, dtype=np.float32)
        if self.L.pdhcg_mlp_get_params(self.h, _fp(out)) != 0:
            raise RuntimeError("pdhcg_mlp_get_params failed")
        return out

    def set_params(self, params) -> None:
        p = _f32(params)
        if p.size != self.num_params:
            raise ValueError("parameter vector has the wrong length")
        if self.L.pdhcg_mlp_set_params(self.h, _fp(p)) != 0:
            raise RuntimeError("pdhcg_mlp_set_params failed")

    def loss_grad(self, X, T):
        X, T = _f32(X), _f32(T)
        g = np.empty(self.num_params, dtype=np.float32)
        loss = C.c_double()
        if self.L.pdhcg_mlp_loss_grad(self.h, _fp(X), _fp(T), X.shape[0], C.byref(loss), _fp(g)) != 0:
            raise RuntimeError("pdhcg_mlp_loss_grad failed")
        return loss.value, g

    def train(self, X, T, epochs: int, batch: int = 256, lr: float = 1e-3, seed: int = 0) -> float:
        X, T = _f32(X), _f32(T)
        loss = C.c_double()
        if self.L.pdhcg_mlp_train(self.h, _fp(X), _fp(T), X.shape[0], epochs, batch, lr, seed, C.byref(loss)) != 0:
            raise RuntimeError("pdhcg_mlp_train failed")
        return loss.value

    def predict(self, X) -> np.ndarray:
        X = _f32(X)
        Y = np.empty((X.shape[0], self.dims[2]), dtype=np.float32)
        if self.L.pdhcg_mlp_predict(self.h, _fp(X), X.shape[0], _fp(Y)) != 0:
            raise RuntimeError("pdhcg_mlp_predict failed")
        return Y

    def close(self):
        if self.h:
            self.L.pdhcg_mlp_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class WarmStart:
    """Maps family parameters z to (x, y) starting points."""

    def __init__(self, z_dim: int, n: int, m: int, hidden: int = 1024, seed: int = 0):
        self.n, self.m = n, m
        self.net = MLP(z_dim, hidden, n + m, seed)
        self.mean: Optional[np.ndarray] = None
        self.scale: Optional[np.ndarray] = None

    def fit(self, Z, X, Y, epochs: int = 100, batch: int = 256, lr: float = 1e-3, seed: int = 0) -> float:
        """Trains on parameters Z (count x z_dim) and solutions X (count x n),
        Y (count x m); returns the last epoch's mean loss (standardised)."""
        T = np.hstack([np.asarray(X, dtype=np.float64), np.asarray(Y, dtype=np.float64)])
        self.mean = T.mean(axis=0)
        sd = T.std(axis=0)
        self.scale = np.where(sd > 1e-12, sd, 1.0)
        return self.net.train(Z, (T - self.mean) / self.scale, epochs, batch, lr, seed)

    def predict(self, Z):
        """Predicted (X, Y) for parameters Z, in the problem's units."""
        if self.mean is None:
            raise RuntimeError("fit the model first")
        T = self.net.predict(Z).astype(np.float64) * self.scale + self.mean
        return T[:, : self.n], T[:, self.n:]

    def start(self, problem, z, dual: bool = True):
        """A copy of ``problem`` (QPData) carrying the predicted start values.

        With ``dual=False`` only the primal start is set and the dual starts
        at zero.  The dual prediction is the less accurate one; it lowers the
        mean iteration count but, on a few instances, raises theirs above the
        cold solve's, and a batch's wall time follows its slowest instance."""
        x, y = self.predict(np.asarray(z, dtype=np.float32).reshape(1, -1))
        lb = -np.inf if problem.var_lb is None else problem.var_lb
        ub = np.inf if problem.var_ub is None else problem.var_ub
        x = np.clip(x[0], lb, ub)
        return replace(problem, primal_start=x, dual_start=y[0] if dual else None)

    def guarded(self, problems, Z, dual: bool = True, batch: bool = True, **params):
        """Predicted starts with the guard: instance i keeps its predicted
        start only where the relative KKT error at that point (the largest of
        the relative primal, dual and gap residuals the termination test
        uses) is below the cold start's; otherwise it starts cold.  Both
        errors come from the device solver at iteration 0 (iteration_limit
        0), batched when ``batch``.  Returns (problems, kept) with kept[i]
        True where the predicted start was kept."""
        from . import solve, solve_batch
        Z = np.asarray(Z, dtype=np.float32).reshape(len(problems), -1)
        warm = [self.start(p, z, dual) for p, z in zip(problems, Z)]
        kw = dict(params, iteration_limit=0)
        if batch:
            r = solve_batch(list(problems) + warm, **kw)
        else:
            r = [solve(p, **kw) for p in list(problems) + warm]
        err = lambda s_: np.inf if s_ is None else max(s_.rel_primal, s_.rel_dual, s_.rel_gap)
        count = len(problems)
        kept = np.array([err(r[count + i]) < err(r[i]) for i in range(count)], dtype=bool)
        return [warm[i] if kept[i] else problems[i] for i in range(count)], kept

    def close(self):
        self.net.close()
