"""The reference's Python interface (python/pdhcg: Model, ConeSpec,
ConeType, PDHCG constants, and the solve_once / get_default_params /
read_problem_file functions of its compiled module) over libpdhcg_amd.so.

    from pdhcg_amd import pdhcg
    m = pdhcg.Model(objective_vector=c, constraint_matrix=A, ...)
    m.setParams(OptimalityTol=1e-6)
    m.optimize()
    m.X, m.ObjVal, m.Status
"""
from . import PDHCG
from .cones import ConeSpec, ConeType
from .core import get_default_params, read_problem_file, solve_once
from .model import Model

__all__ = ["ConeSpec", "ConeType", "Model", "PDHCG", "get_default_params", "read_problem_file", "solve_once"]
