"""The Python interface mirror (pdhcg_amd.pdhcg) solving on the GPU: the
scenarios and expected answers of the reference's test/test_pdhcg_python.py
(low-rank, mixed sparse + low-rank, unconstrained and simple QPs) and
tests/test_model_cones.py (variable SOC and PSD blocks, affine SOC and PSD
rows), plus model sense and warm start round trips.
"""
import numpy as np
import pytest
import scipy.sparse as sp

from pdhcg_amd.pdhcg import ConeSpec, ConeType, Model, PDHCG, solve_once

pytestmark = pytest.mark.gpu

PARAMS = {"Presolve": False, "LogLevel": 0, "OptimalityTol": 1e-6, "FeasibilityTol": 1e-6, "TimeLimit": 10.0}


def _solve(m):
    m.setParams(**PARAMS)
    m.optimize()
    assert m.Status == "OPTIMAL", (m.Status, m.ObjVal, m.X)
    return m


@pytest.mark.parametrize("Q,obj", [(None, 0.25), (sp.diags([2.0, 2.0], format="csc"), 0.75)])
def test_lowrank_qp(Q, obj):
    m = _solve(Model(objective_vector=np.zeros(2), constraint_matrix=np.array([[1.0, 1.0]]),
                     constraint_lower_bound=[1.0], constraint_upper_bound=[1.0], objective_matrix=Q,
                     objective_matrix_low_rank=sp.eye(2, format="csc"), variable_lower_bound=[0.0, 0.0]))
    assert abs(m.ObjVal - obj) < 1e-4 and np.allclose(m.X, [0.5, 0.5], atol=1e-4)


def test_qp_without_constraint_matrix():
    m = _solve(Model(objective_vector=np.array([-2.0, -2.0]), objective_matrix=sp.diags([2.0, 2.0], format="csc")))
    assert abs(m.ObjVal + 2.0) < 1e-4 and np.allclose(m.X, [1.0, 1.0], atol=1e-4)


def test_simple_qp_and_maximize():
    kw = dict(objective_matrix=sp.diags([4.0, 2.0], format="csc"), constraint_matrix=np.array([[1.0, 1.0]]),
              constraint_lower_bound=[1.0], constraint_upper_bound=[1.0], variable_lower_bound=[0.0, 0.0])
    m = _solve(Model(objective_vector=np.array([1.0, 1.0]), **kw))
    assert abs(m.ObjVal - 5.0 / 3.0) < 1e-4 and np.allclose(m.X, [1 / 3, 2 / 3], atol=1e-4)
    # MAXIMIZE negates the linear objective and the constant only, as the
    # reference's Model.optimize does (python/pdhcg/model.py:711-716): the
    # solver minimises -c'x + x'Qx/2, and ObjVal is the negated optimum
    n = Model(objective_vector=np.array([-1.0, -1.0]), **kw)
    n.ModelSense = PDHCG.MAXIMIZE
    _solve(n)
    assert abs(n.ObjVal + 5.0 / 3.0) < 1e-4 and np.allclose(n.X, [1 / 3, 2 / 3], atol=1e-4)
    # an LP maximum: max x1 + 2 x2 on the simplex edge x1 + x2 = 1, x >= 0
    lp = Model(objective_vector=np.array([1.0, 2.0]), constraint_matrix=np.array([[1.0, 1.0]]),
               constraint_lower_bound=[1.0], constraint_upper_bound=[1.0], variable_lower_bound=[0.0, 0.0])
    lp.ModelSense = PDHCG.MAXIMIZE
    _solve(lp)
    assert abs(lp.ObjVal - 2.0) < 1e-4 and np.allclose(lp.X, [0.0, 1.0], atol=1e-4)
    # a warm start at the solution: optimal again, in few iterations
    w = Model(objective_vector=np.array([1.0, 1.0]), **kw)
    w.setWarmStart(primal=m.X, dual=m.Pi)
    _solve(w)
    assert w.IterCount <= m.IterCount


def test_model_variable_soc():
    m = _solve(Model(objective_vector=np.array([0.0, 0.0, 1.0]),
                     constraint_matrix=sp.csr_matrix([[1.0, 0.0, 0.0], [0.0, 1.0, 0.0]]),
                     constraint_lower_bound=np.array([3.0, 4.0]), constraint_upper_bound=np.array([3.0, 4.0]),
                     variable_cones=ConeSpec(ConeType.SOC, np.array([0], dtype=np.int32), v_dims=1)))
    np.testing.assert_allclose(m.X, [3.0, 4.0, 5.0], atol=2e-4)


@pytest.mark.parametrize("sparse", [False, True])
def test_model_affine_soc(sparse):
    F = np.array([[0.0], [0.0], [1.0]])
    m = _solve(Model(objective_vector=np.array([1.0]), affine_cone_matrix=sp.csr_matrix(F) if sparse else F,
                     affine_cone_offset=np.array([3.0, 4.0, 0.0]),
                     affine_cones=ConeSpec(ConeType.SOC, np.array([0], dtype=np.int32), v_dims=1)))
    np.testing.assert_allclose(m.X, [5.0], atol=2e-4)


def test_model_variable_psd():
    s2 = np.sqrt(2.0)
    m = _solve(Model(objective_vector=np.array([0.0, 0.0, 1.0]),
                     constraint_matrix=sp.csr_matrix([[1.0, 0.0, 0.0], [0.0, 1.0, 0.0]]),
                     constraint_lower_bound=np.array([1.0, 2.0 * s2]), constraint_upper_bound=np.array([1.0, 2.0 * s2]),
                     variable_cones=ConeSpec(ConeType.PSD, np.array([0], dtype=np.int32), v_dims=2)))
    np.testing.assert_allclose(m.X, [1.0, 2.0 * s2, 4.0], atol=8e-4)


def test_model_affine_psd():
    s2 = np.sqrt(2.0)
    m = _solve(Model(objective_vector=np.array([1.0]), affine_cone_matrix=sp.csr_matrix([[1.0], [0.0], [0.0]]),
                     affine_cone_offset=np.array([0.0, s2, 1.0]),
                     affine_cones=ConeSpec(ConeType.PSD, np.array([0], dtype=np.int32), v_dims=2)))
    np.testing.assert_allclose(m.X, [1.0], atol=8e-4)


def test_solve_once_dictionary_and_legacy_cones():
    info = solve_once(sp.diags([2.0, 2.0]).tocsr(), None, None, np.array([-2.0, -2.0]),
                      params={"eps_optimal_relative": 1e-6, "eps_feasible_relative": 1e-6, "verbose": 0})
    assert info["Status"] == "OPTIMAL" and info["StatusCode"] == 0
    assert abs(info["PrimalObj"] + 2.0) < 1e-4 and info["X"].shape == (2,)
    for key in ("Pi", "DualObj", "ObjectiveGap", "RelativeObjectiveGap", "Iterations", "RescalingTimeSec",
                "RuntimeSec", "RelativePrimalResidual", "RelativeDualResidual", "MaxPrimalRayInfeas",
                "MaxDualRayInfeas", "PrimalRayLinObj", "DualRayObj"):
        assert key in info
    with pytest.raises(ValueError, match="ConeSpec"):
        solve_once(None, None, sp.csr_matrix((0, 3)), np.zeros(3), cones=[{"type": "soc", "start_idx": 0, "v_dim": 1}])


def test_model_read_file_and_cli(tmp_path):
    # the reference's CBF fixture through Model.read_file, and the same file
    # through the command-line solver (summary and solution files)
    import os
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    path = os.path.join(root, "tests", "golden", "cbf_q3_smoke.cbf")
    m = Model.read_file(path)
    _solve(m)
    np.testing.assert_allclose(m.X, [3.0, 4.0, 5.0], atol=2e-4)
    cli = os.path.join(root, "pdhcg_amd", "pdhcg-cli")
    r = subprocess.run([cli, "--eps_opt", "1e-6", "--eps_feas", "1e-6", "-v", "0", path, str(tmp_path)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    summary = (tmp_path / "cbf_q3_smoke_summary.txt").read_text()
    assert "Termination Reason: OPTIMAL" in summary and "Columns: 3" in summary and "Rows: 2" in summary
    x = np.loadtxt(tmp_path / "cbf_q3_smoke_primal_solution.txt")
    y = np.loadtxt(tmp_path / "cbf_q3_smoke_dual_solution.txt")
    np.testing.assert_allclose(x, [3.0, 4.0, 5.0], atol=2e-4)
    assert y.shape == (2,)
