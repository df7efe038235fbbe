"""The Python interface mirror (pdhcg_amd.pdhcg: ConeSpec, ConeType, Model,
PDHCG, get_default_params, solve_once, read_problem_file) on the CPU: the
metadata and argument checks of the reference's python/pdhcg and its
tests/test_cones.py, tests/test_model_cones.py scenarios that need no
device.  The solves run on the GPU (tests/test_gpu_pdhcg_python.py).
"""
import numpy as np
import pytest
import scipy.sparse as sp

from pdhcg_amd import pdhcg
from pdhcg_amd.pdhcg import ConeSpec, ConeType, Model, PDHCG


def test_cone_spec_broadcasts_columnar_metadata():
    cones = ConeSpec(ConeType.EXP, 3 * np.arange(4, dtype=np.int32))
    assert len(cones) == 4 and cones.types.dtype == np.int32 and cones.starts.flags.c_contiguous
    np.testing.assert_array_equal(cones.types, np.full(4, ConeType.EXP, dtype=np.int32))
    np.testing.assert_array_equal(cones.v_dims, np.ones(4, dtype=np.int32))
    np.testing.assert_array_equal(cones.power_alphas, np.zeros(4))
    cones.validate_ambient(12, allow_fixed=True)


def test_cone_spec_heterogeneous_and_fixed_mask():
    cones = ConeSpec(np.array([ConeType.SOC, ConeType.EXP], dtype=np.int32), np.array([0, 3], dtype=np.int32),
                     fixed_mask=np.array([0, 0, 0, 0, 1, 0], dtype=np.uint8))
    np.testing.assert_array_equal(cones.types, [ConeType.SOC, ConeType.EXP])
    np.testing.assert_array_equal(cones.starts, [0, 3])
    np.testing.assert_array_equal(cones.fixed_mask, [0, 0, 0, 0, 1, 0])
    assert ConeSpec(["soc", "rsoc", "power"], [0, 3, 6], power_alphas=[0.0, 0.0, 0.3]).types.tolist() == [1, 0, 3]


def test_cone_spec_validation():
    with pytest.raises(ValueError, match="alphas"):
        ConeSpec(ConeType.POWER, np.array([0], dtype=np.int32), power_alphas=np.nan)
    with pytest.raises(ValueError, match="ambient"):
        ConeSpec(ConeType.SOC, np.array([1], dtype=np.int32), v_dims=2).validate_ambient(4, allow_fixed=True)
    with pytest.raises(ValueError, match="v_dim == 1"):
        ConeSpec(ConeType.EXP, [0], v_dims=2)
    with pytest.raises(ValueError):
        ConeSpec("cube", [0])
    ConeSpec("psd", np.array([0], dtype=np.int32), v_dims=3).validate_ambient(6, allow_fixed=True, require_cover=True)
    fixed = ConeSpec(ConeType.PSD, [0], v_dims=2, fixed_mask=np.array([0, 1, 0], dtype=np.uint8))
    with pytest.raises(ValueError, match="PSD"):
        fixed.validate_ambient(3, allow_fixed=True)
    with pytest.raises(ValueError, match="cover"):
        ConeSpec(ConeType.SOC, [0], v_dims=1).validate_ambient(4, allow_fixed=False, require_cover=True)


def test_default_params_and_aliases():
    d = pdhcg.get_default_params()
    assert d["eps_optimal_relative"] == 1e-4 and d["eps_feasible_relative"] == 1e-4
    assert d["termination_evaluation_frequency"] == 200 and d["l_inf_ruiz_iterations"] == 10
    assert d["optimality_norm"] == "linf" and d["inner_iter_limit"] == 1000
    assert set(PDHCG._PARAM_ALIAS.values()) <= set(d)
    m = Model(objective_vector=np.zeros(2), objective_matrix=sp.eye(2))
    m.setParams(OptimalityTol=1e-7, TimeLimit=5.0)
    m.Params["IterationLimit"] = 1234
    m.Params.RuizIters = 3
    assert m.getParam("eps_optimal_relative") == 1e-7 and m.Params["TimeLimit"] == 5.0
    assert m.getParam("iteration_limit") == 1234 and m.getParam("l_inf_ruiz_iterations") == 3
    prm = pdhcg.core.params_struct(m._params)
    assert prm.termination_criteria.eps_optimal_relative == 1e-7 and prm.termination_criteria.iteration_limit == 1234
    assert prm.l_inf_ruiz_iterations == 3
    with pytest.raises(ValueError):
        pdhcg.core.params_struct({"optimality_norm": 2})


def test_model_argument_checks():
    with pytest.raises(TypeError, match="ConeSpec"):
        Model(objective_vector=np.zeros(3), variable_cones=[{"type": "soc", "start_idx": 0, "v_dim": 1}])
    with pytest.raises(ValueError, match="mismatch"):
        Model(objective_vector=np.zeros(3), constraint_matrix=np.ones((1, 3)), objective_matrix=np.eye(2))
    with pytest.raises(ValueError):
        Model(objective_vector=np.zeros(2), constraint_matrix=np.ones((1, 2)), constraint_lower_bound=[1.0, 2.0])
    m = Model(objective_vector=np.zeros(2), objective_matrix_low_rank=np.ones((1, 2)))
    with pytest.raises(ValueError):
        m.setObjectiveMatrixLowRankMiddle(np.ones(2))
    with pytest.warns(RuntimeWarning):
        m.setWarmStart(primal=np.zeros(5))
    assert m.X is None and m.Status is None and m.ModelSense == PDHCG.MINIMIZE
