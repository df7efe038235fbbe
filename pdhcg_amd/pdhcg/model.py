"""Model: the reference's object interface to a QP (python/pdhcg/model.py),
solved by libpdhcg_amd.so through solve_once.

    minimize    1/2 x'(Q + R' D R) x + c'x + c0
    subject to  l_c <= A x <= u_c,   F x + g in K,   l_v <= x <= u_v,
                x in K_var on the variable cone blocks

Same constructor, setters, parameter names (and aliases such as TimeLimit or
OptimalityTol), model sense and result attributes (X, Pi, ObjVal, Status,
...), so code written for the reference's Model runs unchanged.
"""
from __future__ import annotations

import warnings
from typing import Any, Optional, Union

import numpy as np
import scipy.sparse as sp

from . import PDHCG
from .cones import ConeSpec
from .core import get_default_params, read_problem_file, solve_once

ArrayLike = Any


def _dense(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


def _csr(a) -> sp.csr_matrix:
    m = sp.csr_matrix(a, dtype=np.float64)
    m.sort_indices()
    return m


def _matrix(name: str, value, cols: int):
    """None, or value as float64 CSR / dense with `cols` columns."""
    if value is None:
        return None
    if not isinstance(value, (np.ndarray, sp.spmatrix)) and not sp.issparse(value):
        raise TypeError(f"{name} must be a numpy.ndarray or scipy.sparse matrix")
    if len(value.shape) != 2:
        raise ValueError(f"{name} must be 2D, got shape {value.shape}")
    if value.shape[1] != cols:
        raise ValueError(f"{name} shape {value.shape} does not match number of variables ({cols})")
    return _csr(value) if sp.issparse(value) else _dense(value)


class _ParamsView:
    """model.Params["Name"] / model.Params.Name access with the aliases."""

    def __init__(self, model: "Model"):
        object.__setattr__(self, "_m", model)

    def __getattr__(self, name: str):
        return self._m.getParam(name)

    def __setattr__(self, name: str, value):
        self._m.setParam(name, value)

    def __getitem__(self, name: str):
        return self._m.getParam(name)

    def __setitem__(self, name: str, value):
        self._m.setParam(name, value)

    def keys(self):
        return self._m._params.keys()


class Model:
    def __init__(self, objective_vector: ArrayLike,
                 constraint_matrix: Optional[Union[np.ndarray, sp.spmatrix]] = None,
                 constraint_lower_bound: Optional[ArrayLike] = None,
                 constraint_upper_bound: Optional[ArrayLike] = None,
                 objective_matrix: Optional[Union[np.ndarray, sp.spmatrix]] = None,
                 objective_matrix_low_rank: Optional[Union[np.ndarray, sp.spmatrix]] = None,
                 objective_matrix_low_rank_middle: Optional[ArrayLike] = None,
                 variable_lower_bound: Optional[ArrayLike] = None,
                 variable_upper_bound: Optional[ArrayLike] = None,
                 objective_constant: float = 0.0,
                 affine_cone_matrix: Optional[Union[np.ndarray, sp.spmatrix]] = None,
                 affine_cone_offset: Optional[ArrayLike] = None,
                 affine_cones: Optional[ConeSpec] = None,
                 variable_cones: Optional[ConeSpec] = None):
        self.num_vars = 0
        self.num_constrs = 0
        self.num_affine_constrs = 0
        # the variable count comes from the first matrix given, else from c
        for mat, name in ((constraint_matrix, "constraint_matrix"), (objective_matrix, "objective_matrix"),
                          (objective_matrix_low_rank, "objective_matrix_low_rank"),
                          (affine_cone_matrix, "affine_cone_matrix")):
            if mat is None:
                continue
            if not hasattr(mat, "shape") or len(mat.shape) != 2:
                raise ValueError(f"{name} must be a 2D numpy.ndarray or scipy.sparse matrix.")
            if self.num_vars == 0:
                self.num_vars = int(mat.shape[1])
            elif mat.shape[1] != self.num_vars:
                raise ValueError(f"{name} dimensions mismatch variables ({self.num_vars})")
        if constraint_matrix is not None:
            self.num_constrs = int(constraint_matrix.shape[0])
        if self.num_vars == 0 and objective_vector is not None:
            shape = np.asarray(objective_vector).shape
            if len(shape) != 1:
                raise ValueError(f"objective_vector must be 1D, got shape {shape}")
            self.num_vars = int(shape[0])
        self.ModelSense = PDHCG.MINIMIZE
        self._params: dict = dict(get_default_params())
        self.Params = _ParamsView(self)
        self.constr_lb = self.constr_ub = None
        self.R = self.D = None
        self.setObjectiveVector(objective_vector)
        self.setObjectiveConstant(objective_constant)
        self.setObjectiveMatrix(objective_matrix)
        self.setObjectiveMatrixLowRank(objective_matrix_low_rank)
        self.setObjectiveMatrixLowRankMiddle(objective_matrix_low_rank_middle)
        self.setConstraintMatrix(constraint_matrix)
        self.setConstraintLowerBound(constraint_lower_bound)
        self.setConstraintUpperBound(constraint_upper_bound)
        self.setVariableLowerBound(variable_lower_bound)
        self.setVariableUpperBound(variable_upper_bound)
        self._variable_cones: Optional[ConeSpec] = None
        self.affine_F = self.affine_g = None
        self._affine_cones: Optional[ConeSpec] = None
        self.setVariableCones(variable_cones)
        self.setAffineConeConstraints(affine_cone_matrix, affine_cone_offset, affine_cones)
        self._primal_start: Optional[np.ndarray] = None
        self._dual_start: Optional[np.ndarray] = None
        self._clear_solution_cache()

    @classmethod
    def read_file(cls, path: str) -> "Model":
        """Reads .mps / .qps / .cbf (optionally .gz); cones in the file are kept."""
        raw = read_problem_file(path)
        to_csr = lambda d: None if d is None else sp.csr_matrix((d["data"], d["indices"], d["indptr"]),
                                                                 shape=d["shape"])
        A = to_csr(raw.get("A"))
        if A is not None and A.shape[0] == 0:
            A = None
        m = cls(objective_vector=raw["c"], constraint_matrix=A,
                constraint_lower_bound=raw["constr_lb"] if A is not None else None,
                constraint_upper_bound=raw["constr_ub"] if A is not None else None,
                objective_matrix=to_csr(raw.get("Q")), variable_lower_bound=raw["var_lb"],
                variable_upper_bound=raw["var_ub"], objective_constant=raw.get("obj_const", 0.0),
                affine_cone_matrix=to_csr(raw.get("affine_F")), affine_cone_offset=raw.get("affine_g"),
                affine_cones=ConeSpec.from_columnar(raw["affine_cones"]) if raw.get("affine_cones") else None,
                variable_cones=ConeSpec.from_columnar(raw["cones"]) if raw.get("cones") else None)
        if raw.get("primal_start") is not None:
            m._primal_start = np.asarray(raw["primal_start"], dtype=np.float64)
        return m

    # ---- data
    def setObjectiveVector(self, c: ArrayLike) -> None:
        self.c = _dense(c)
        if self.c.ndim != 1:
            raise ValueError(f"setObjectiveVector: c must be 1D, got shape {self.c.shape}")
        if self.c.size != self.num_vars:
            raise ValueError(f"setObjectiveVector: length {self.c.size} != self.num_vars ({self.num_vars})")
        self._clear_solution_cache()

    def setObjectiveConstant(self, c0: float) -> None:
        self.c0 = float(c0)
        self._clear_solution_cache()

    def setObjectiveMatrix(self, Q_like: ArrayLike) -> None:
        self.Q = _matrix("setObjectiveMatrix: Q", Q_like, self.num_vars)
        self._clear_solution_cache()

    def setObjectiveMatrixLowRank(self, R_like: ArrayLike) -> None:
        self.R = _matrix("setObjectiveMatrixLowRank: R", R_like, self.num_vars)
        self._clear_solution_cache()

    def setObjectiveMatrixLowRankMiddle(self, D_like) -> None:
        """D in Q + R'DR: None (identity), a length-rank vector (diagonal), or a
        rank x rank dense or sparse matrix."""
        if D_like is None:
            self.D = None
            self._clear_solution_cache()
            return
        if getattr(self, "R", None) is None:
            raise ValueError("setObjectiveMatrixLowRankMiddle: D is only meaningful when R is set; "
                             "call setObjectiveMatrixLowRank first.")
        rank = int(self.R.shape[0])
        if sp.issparse(D_like):
            if D_like.shape != (rank, rank):
                raise ValueError(f"setObjectiveMatrixLowRankMiddle: sparse D shape {D_like.shape} must be "
                                 f"({rank}, {rank})")
            self.D = _csr(D_like)
        else:
            d = _dense(D_like)
            if (d.ndim == 1 and d.size != rank) or (d.ndim == 2 and d.shape != (rank, rank)) or d.ndim > 2:
                raise ValueError(f"setObjectiveMatrixLowRankMiddle: D must be length {rank} or {rank} x {rank}")
            self.D = d
        self._clear_solution_cache()

    def setConstraintMatrix(self, A_like: ArrayLike) -> None:
        if A_like is None:
            self.A = None
            self.num_constrs = 0
            self.constr_lb = self.constr_ub = None
            self._clear_solution_cache()
            return
        self.A = _matrix("setConstraintMatrix: A", A_like, self.num_vars)
        self.num_constrs = int(self.A.shape[0])
        for name in ("constr_lb", "constr_ub"):
            v = getattr(self, name, None)
            if v is not None and np.asarray(v).size != self.num_constrs:
                raise ValueError(f"setConstraintMatrix: {name} length {np.asarray(v).size} != rows "
                                 f"{self.num_constrs}")
        self._clear_solution_cache()

    def _bound(self, v, count, name):
        if v is None:
            return None
        a = _dense(v).ravel()
        if a.size != count:
            raise ValueError(f"{name}: length {a.size} != {count}")
        return a

    def setConstraintLowerBound(self, constr_lb: Optional[ArrayLike]) -> None:
        self.constr_lb = self._bound(constr_lb, self.num_constrs, "setConstraintLowerBound")
        self._clear_solution_cache()

    def setConstraintUpperBound(self, constr_ub: Optional[ArrayLike]) -> None:
        self.constr_ub = self._bound(constr_ub, self.num_constrs, "setConstraintUpperBound")
        self._clear_solution_cache()

    def setVariableLowerBound(self, lb: Optional[ArrayLike]) -> None:
        self.lb = self._bound(lb, self.num_vars, "setVariableLowerBound")
        self._clear_solution_cache()

    def setVariableUpperBound(self, ub: Optional[ArrayLike]) -> None:
        self.ub = self._bound(ub, self.num_vars, "setVariableUpperBound")
        self._clear_solution_cache()

    def setAffineConeConstraints(self, F_like, g: Optional[ArrayLike], cones: Optional[ConeSpec]) -> None:
        """F x + g in K over F's rows; the cones must cover every row of F."""
        if F_like is None:
            if g is not None or cones is not None:
                raise ValueError("affine_cone_matrix is required when affine offset or cones are provided")
            self.affine_F = self.affine_g = None
            self._affine_cones = None
            self.num_affine_constrs = 0
            self._clear_solution_cache()
            return
        F = _matrix("setAffineConeConstraints: F", F_like, self.num_vars)
        if not isinstance(cones, ConeSpec):
            raise TypeError("setAffineConeConstraints: cones must be a ConeSpec")
        if len(cones) == 0:
            raise ValueError("setAffineConeConstraints: cones must cover every row of F")
        p = int(F.shape[0])
        cones.validate_ambient(p, allow_fixed=False, require_cover=True)
        self.affine_F, self.affine_g = F, self._bound(g, p, "setAffineConeConstraints: affine offset")
        self.num_affine_constrs = p
        self._affine_cones = cones
        self._clear_solution_cache()

    def setVariableCones(self, cones: Optional[ConeSpec]) -> None:
        if cones is not None and not isinstance(cones, ConeSpec):
            raise TypeError("setVariableCones: cones must be a ConeSpec")
        if cones is not None and len(cones) > 0:
            cones.validate_ambient(self.num_vars, allow_fixed=True)
            self._variable_cones = cones
        else:
            self._variable_cones = None
        self._clear_solution_cache()

    def setWarmStart(self, primal: Optional[ArrayLike] = None, dual: Optional[ArrayLike] = None) -> None:
        """Primal (n) and dual (rows of A, then rows of F) starting points; a
        wrong length is ignored with a warning."""
        self._primal_start = None
        if primal is not None:
            a = _dense(primal).ravel()
            if a.size == self.num_vars:
                self._primal_start = a
            else:
                warnings.warn(f"Warm start primal size mismatch (expected {self.num_vars}, got {a.size}).",
                              RuntimeWarning)
        self._dual_start = None
        if dual is not None:
            a = _dense(dual).ravel()
            want = self.num_constrs + self.num_affine_constrs
            if a.size == want:
                self._dual_start = a
            else:
                warnings.warn(f"Warm start dual size mismatch (expected {want}, got {a.size}).", RuntimeWarning)

    def clearWarmStart(self) -> None:
        self.setWarmStart(None, None)

    # ---- parameters
    def setParam(self, name: str, value: Any) -> None:
        self._params[PDHCG._PARAM_ALIAS.get(name, name)] = value

    def getParam(self, name: str) -> Any:
        return self._params.get(PDHCG._PARAM_ALIAS.get(name, name))

    def setParams(self, /, **kwargs) -> None:
        for k, v in kwargs.items():
            self.setParam(k, v)

    # ---- solve
    def optimize(self):
        self._clear_solution_cache()
        if self.ModelSense not in (PDHCG.MINIMIZE, PDHCG.MAXIMIZE):
            raise ValueError("model_sense must be PDHCG.MINIMIZE or PDHCG.MAXIMIZE")
        sign = 1 if self.ModelSense == PDHCG.MINIMIZE else -1
        info = solve_once(self.Q, self.R, self.A, sign * self.c, sign * self.c0, self.lb, self.ub, self.constr_lb,
                          self.constr_ub, zero_tolerance=0.0, params=self._params, primal_start=self._primal_start,
                          dual_start=self._dual_start, D=self.D, cones=self._variable_cones, affine_F=self.affine_F,
                          affine_g=self.affine_g, affine_cones=self._affine_cones)
        self._x = np.asarray(info["X"])
        self._y = np.asarray(info["Pi"])
        self._objval = sign * info["PrimalObj"]
        self._dualobj = sign * info["DualObj"]
        self._gap = info["ObjectiveGap"]
        self._rel_gap = info["RelativeObjectiveGap"]
        self._status = str(info["Status"])
        self._status_code = int(info["StatusCode"])
        self._iter = int(info["Iterations"])
        self._runtime = info["RuntimeSec"]
        self._rescale_time = info["RescalingTimeSec"]
        self._rel_p_res = info["RelativePrimalResidual"]
        self._rel_d_res = info["RelativeDualResidual"]
        self._max_p_ray = info["MaxPrimalRayInfeas"]
        self._max_d_ray = info["MaxDualRayInfeas"]
        self._p_ray_lin_obj = sign * info["PrimalRayLinObj"]
        self._d_ray_obj = sign * info["DualRayObj"]

    def _clear_solution_cache(self) -> None:
        self._x = self._y = None
        self._objval = self._dualobj = None
        self._gap = self._rel_gap = None
        self._status = self._status_code = None
        self._iter = None
        self._runtime = self._rescale_time = None
        self._rel_p_res = self._rel_d_res = None
        self._max_p_ray = self._max_d_ray = None
        self._p_ray_lin_obj = self._d_ray_obj = None

    # ---- results
    X = property(lambda self: self._x)
    Pi = property(lambda self: self._y)
    ObjVal = property(lambda self: self._objval)
    DualObj = property(lambda self: self._dualobj)
    Gap = property(lambda self: self._gap)
    RelGap = property(lambda self: self._rel_gap)
    Status = property(lambda self: self._status)
    StatusCode = property(lambda self: self._status_code)
    IterCount = property(lambda self: self._iter)
    Runtime = property(lambda self: self._runtime)
    RescalingTime = property(lambda self: self._rescale_time)
    RelPrimalResidual = property(lambda self: self._rel_p_res)
    RelDualResidual = property(lambda self: self._rel_d_res)
    MaxPrimalRayInfeas = property(lambda self: self._max_p_ray)
    MaxDualRayInfeas = property(lambda self: self._max_d_ray)
    PrimalRayLinObj = property(lambda self: self._p_ray_lin_obj)
    DualRayObj = property(lambda self: self._d_ray_obj)
    PrimalInfeas = property(lambda self: self._rel_p_res)
    DualInfeas = property(lambda self: self._rel_d_res)
