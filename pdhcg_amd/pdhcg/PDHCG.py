"""Constants of the Python interface (reference python/pdhcg/PDHCG.py):
model sense, status codes, and the user-facing parameter names mapped to the
keys of get_default_params()."""

MINIMIZE = 1
MAXIMIZE = -1

OPTIMAL = 0
PRIMAL_INFEASIBLE = 1
DUAL_INFEASIBLE = 2
TIME_LIMIT = 3
ITERATION_LIMIT = 4
UNSPECIFIED = -1

# the same aliases as the reference's (an interface table)
_PARAM_ALIAS = {
    "TimeLimit": "time_sec_limit",
    "IterationLimit": "iteration_limit",
    "LogLevel": "verbose",
    "Verbosity": "verbose",
    "OutputFlag": "verbose",
    "LogToConsole": "verbose",
    "TermCheckFreq": "termination_evaluation_frequency",
    "OptimalityTol": "eps_optimal_relative",
    "FeasibilityTol": "eps_feasible_relative",
    "CurtisReidIters": "curtis_reid_iterations",
    "RuizIters": "l_inf_ruiz_iterations",
    "UsePCAlpha": "has_pock_chambolle_alpha",
    "PCAlpha": "pock_chambolle_alpha",
    "BoundObjRescaling": "bound_objective_rescaling",
    "UseConePreservingScaling": "use_cone_preserving_scaling",
    "RestartArtificialThresh": "artificial_restart_threshold",
    "RestartSufficientReduction": "sufficient_reduction_for_restart",
    "RestartNecessaryReduction": "necessary_reduction_for_restart",
    "RestartKp": "k_p",
    "ReflectionCoeff": "reflection_coefficient",
    "OptimalityNorm": "optimality_norm",
    "SVMaxIter": "sv_max_iter",
    "SVTol": "sv_tol",
    "InnerIterLimit": "inner_iter_limit",
    "InnerInitTol": "inner_init_tol",
    "InnerMinTol": "inner_min_tol",
    "DiagJacobiPrecond": "diag_jacobi_precond",
    "Presolve": "presolve",
}
