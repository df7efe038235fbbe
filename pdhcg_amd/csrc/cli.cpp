// Command-line solver over libpdhcg_amd.so (reference src/cli.c:79-395):
//   pdhcg [OPTIONS] <problem_file> <output_dir>
// reads .mps / .qps / .cbf (optionally .gz), solves on the GPU and writes
// <name>_summary.txt, <name>_primal_solution.txt and <name>_dual_solution.txt
// into output_dir.  Same options and file formats as the reference CLI, plus
// --precision fp64|mixed.
#include <getopt.h>
#include <libgen.h>

#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "../../include/cbf_parser.h"
#include "../../include/mps_parser.h"
#include "../../include/pdhcg.h"
#include "../../include/pdhcg_amd.h"

namespace {

const char *reason_name(termination_reason_t r) {
    switch (r) {
    case TERMINATION_REASON_OPTIMAL: return "OPTIMAL";
    case TERMINATION_REASON_PRIMAL_INFEASIBLE: return "PRIMAL_INFEASIBLE";
    case TERMINATION_REASON_DUAL_INFEASIBLE: return "DUAL_INFEASIBLE";
    case TERMINATION_REASON_INFEASIBLE_OR_UNBOUNDED: return "INFEASIBLE_OR_UNBOUNDED";
    case TERMINATION_REASON_TIME_LIMIT: return "TIME_LIMIT";
    case TERMINATION_REASON_ITERATION_LIMIT: return "ITERATION_LIMIT";
    case TERMINATION_REASON_USER_INTERRUPT: return "USER_INTERRUPT";
    case TERMINATION_REASON_FEAS_POLISH_SUCCESS: return "FEAS_POLISH_SUCCESS";
    default: return "UNSPECIFIED";
    }
}

// basename up to its first '.'
std::string instance_name(const char *path) {
    std::string copy(path);
    std::string base = basename(&copy[0]);
    const size_t dot = base.find('.');
    return dot == std::string::npos ? base : base.substr(0, dot);
}

bool is_cbf(const std::string &f) {
    std::string s = f;
    if (s.size() > 3 && s.compare(s.size() - 3, 3, ".gz") == 0) s.resize(s.size() - 3);
    return s.size() > 4 && s.compare(s.size() - 4, 4, ".cbf") == 0;
}

void write_vector(const double *v, int n, const std::string &path) {
    if (!v) return;
    FILE *f = std::fopen(path.c_str(), "w");
    if (!f) {
        std::perror("Error opening solution file");
        return;
    }
    for (int i = 0; i < n; ++i) std::fprintf(f, "%.10g\n", v[i]);
    std::fclose(f);
}

void write_summary(const pdhcg_result_t *r, const std::string &path) {
    FILE *f = std::fopen(path.c_str(), "w");
    if (!f) {
        std::perror("Error opening summary file");
        return;
    }
    std::fprintf(f, "Termination Reason: %s\n", reason_name(r->termination_reason));
    std::fprintf(f, "Runtime (sec): %e\n", r->cumulative_time_sec);
    std::fprintf(f, "Iterations Count: %d\n", r->total_count);
    if (r->total_inner_count > 0) std::fprintf(f, "Inner Iterations Count: %d\n", r->total_inner_count);
    std::fprintf(f, "Primal Objective Value: %e\n", r->primal_objective_value);
    std::fprintf(f, "Dual Objective Value: %e\n", r->dual_objective_value);
    std::fprintf(f, "Absolute Primal Residual: %e\n", r->absolute_primal_residual);
    std::fprintf(f, "Absolute Dual Residual: %e\n", r->absolute_dual_residual);
    std::fprintf(f, "Relative Primal Residual: %e\n", r->relative_primal_residual);
    std::fprintf(f, "Relative Dual Residual: %e\n", r->relative_dual_residual);
    std::fprintf(f, "Absolute Objective Gap: %e\n", r->objective_gap);
    std::fprintf(f, "Relative Objective Gap: %e\n", r->relative_objective_gap);
    std::fprintf(f, "Rows: %d\n", r->num_constraints);
    std::fprintf(f, "Columns: %d\n", r->num_variables);
    std::fprintf(f, "Nonzeros: %d\n", r->num_nonzeros);
    std::fclose(f);
}

void usage(const char *prog) {
    std::fprintf(stderr,
                 "Usage: %s [OPTIONS] <problem_file> <output_dir>\n\n"
                 "Arguments:\n"
                 "  <problem_file>           Input problem: .mps, .qps, .cbf, and gzip-compressed variants.\n"
                 "  <output_dir>             Directory for <basename>_summary.txt, <basename>_primal_solution.txt\n"
                 "                           and <basename>_dual_solution.txt.\n\n"
                 "Options:\n"
                 "  -h, --help               Display this help message.\n"
                 "  -v, --verbose <int>      Verbosity level: 0=silent, 1=summary, 2=detailed (default: 1).\n"
                 "      --time_limit <sec>   Time limit in seconds (default: 3600.0).\n"
                 "      --iter_limit <int>   Iteration limit (default: %d).\n"
                 "      --eps_opt <float>    Relative optimality tolerance (default: 1e-4).\n"
                 "      --eps_feas <float>   Relative feasibility tolerance (default: 1e-4).\n"
                 "      --eps_infeas_detect  Infeasibility detection tolerance (default: 1e-12).\n"
                 "      --curtis_reid_iter   Iterations for Curtis-Reid scaling (default: 0, disabled).\n"
                 "      --l_inf_ruiz_iter    Iterations for L-inf Ruiz rescaling (default: 10).\n"
                 "      --no_pock_chambolle  Disable Pock-Chambolle rescaling (default: enabled).\n"
                 "      --pock_chambolle_alpha Value for Pock-Chambolle alpha (default: 1.0).\n"
                 "      --no_bound_obj_rescaling Disable bound objective rescaling.\n"
                 "      --no_cone_preserving_scaling Disable cone-preserving scaling.\n"
                 "      --eval_freq <int>    Termination evaluation frequency (default: 200).\n"
                 "      --artificial_restart_threshold Artificial restart threshold (default: 0.36).\n"
                 "      --sufficient_reduction_for_restart Sufficient reduction factor (default: 0.2).\n"
                 "      --necessary_reduction_for_restart Necessary reduction factor (default: 0.8).\n"
                 "      --sv_max_iter <int>  Max iterations for singular value estimation (default: 5000).\n"
                 "      --sv_tol <float>     Tolerance for singular value estimation (default: 1e-4).\n"
                 "      --opt_norm <type>    Norm for optimality criteria: l2 or linf (default: linf).\n"
                 "      --inner_iter_limit   Max iterations for the inner solver (default: 1000).\n"
                 "      --inner_init_tol     Initial tolerance for the inner solver (default: 1e-3).\n"
                 "      --inner_min_tol      Minimum tolerance for the inner solver (default: 1e-9).\n"
                 "      --no_diag_precond    Disable the Jacobi preconditioner of the inner subproblem.\n"
                 "      --soc_form <form>    QCQP cone formulation: 'rotated' or 'standard' (default: rotated).\n"
                 "      --precision <p>      fp64 (default) or mixed: FP32 iterations, FP64 termination.\n",
                 prog, INT_MAX);
}

}  // namespace

int main(int argc, char *argv[]) {
    pdhg_parameters_t p;
    set_default_parameters(&p);
    int precision = PDHCG_PRECISION_FP64;
    static const struct option longs[] = {{"help", no_argument, 0, 'h'},
                                          {"verbose", required_argument, 0, 'v'},
                                          {"time_limit", required_argument, 0, 1001},
                                          {"iter_limit", required_argument, 0, 1002},
                                          {"eps_opt", required_argument, 0, 1003},
                                          {"eps_feas", required_argument, 0, 1004},
                                          {"eps_infeas_detect", required_argument, 0, 1005},
                                          {"eps_feas_polish", required_argument, 0, 1006},
                                          {"feasibility_polishing", no_argument, 0, 'f'},
                                          {"l_inf_ruiz_iter", required_argument, 0, 1007},
                                          {"pock_chambolle_alpha", required_argument, 0, 1008},
                                          {"no_pock_chambolle", no_argument, 0, 1009},
                                          {"no_bound_obj_rescaling", no_argument, 0, 1010},
                                          {"sv_max_iter", required_argument, 0, 1011},
                                          {"sv_tol", required_argument, 0, 1012},
                                          {"eval_freq", required_argument, 0, 1013},
                                          {"opt_norm", required_argument, 0, 1014},
                                          {"inner_iter_limit", required_argument, 0, 1015},
                                          {"inner_init_tol", required_argument, 0, 1016},
                                          {"inner_min_tol", required_argument, 0, 1017},
                                          {"presolve", required_argument, 0, 1018},
                                          {"no_diag_precond", no_argument, 0, 1019},
                                          {"soc_form", required_argument, 0, 1020},
                                          {"no_cone_preserving_scaling", no_argument, 0, 1021},
                                          {"artificial_restart_threshold", required_argument, 0, 1022},
                                          {"sufficient_reduction_for_restart", required_argument, 0, 1023},
                                          {"necessary_reduction_for_restart", required_argument, 0, 1024},
                                          {"curtis_reid_iter", required_argument, 0, 1025},
                                          {"precision", required_argument, 0, 1026},
                                          {0, 0, 0, 0}};
    int opt;
    while ((opt = getopt_long(argc, argv, "hv:fp", longs, nullptr)) != -1) {
        switch (opt) {
        case 'h': usage(argv[0]); return 0;
        case 'v': p.verbose = std::atoi(optarg); break;
        case 'f': p.feasibility_polishing = true; break;
        case 1001: p.termination_criteria.time_sec_limit = std::atof(optarg); break;
        case 1002: p.termination_criteria.iteration_limit = std::atoi(optarg); break;
        case 1003: p.termination_criteria.eps_optimal_relative = std::atof(optarg); break;
        case 1004: p.termination_criteria.eps_feasible_relative = std::atof(optarg); break;
        case 1005: p.termination_criteria.eps_infeasible = std::atof(optarg); break;
        case 1006: p.termination_criteria.eps_feas_polish_relative = std::atof(optarg); break;
        case 1007: p.l_inf_ruiz_iterations = std::atoi(optarg); break;
        case 1008: p.pock_chambolle_alpha = std::atof(optarg); break;
        case 1009: p.has_pock_chambolle_alpha = false; break;
        case 1010: p.bound_objective_rescaling = false; break;
        case 1011: p.sv_max_iter = std::atoi(optarg); break;
        case 1012: p.sv_tol = std::atof(optarg); break;
        case 1013: p.termination_evaluation_frequency = std::atoi(optarg); break;
        case 1014:
            if (std::strcmp(optarg, "l2") == 0) p.optimality_norm = NORM_TYPE_L2;
            else if (std::strcmp(optarg, "linf") == 0) p.optimality_norm = NORM_TYPE_L_INF;
            else {
                std::fprintf(stderr, "Error: opt_norm must be 'l2' or 'linf'\n");
                return 1;
            }
            break;
        case 1015: p.inner_solver_parameters.iteration_limit = std::atoi(optarg); break;
        case 1016: p.inner_solver_parameters.initial_tolerance = std::atof(optarg); break;
        case 1017: p.inner_solver_parameters.min_tolerance = std::atof(optarg); break;
        case 1018: p.presolve = std::atoi(optarg) != 0; break;
        case 1019: p.diag_jacobi_precond = false; break;
        case 1020:
            if (std::strcmp(optarg, "rotated") == 0) p.default_cone_type = CONE_ROTATED_SOC;
            else if (std::strcmp(optarg, "standard") == 0) p.default_cone_type = CONE_STANDARD_SOC;
            else {
                std::fprintf(stderr, "Error: soc_form must be 'rotated' or 'standard'\n");
                return 1;
            }
            break;
        case 1021: p.use_cone_preserving_scaling = false; break;
        case 1022: p.restart_params.artificial_restart_threshold = std::atof(optarg); break;
        case 1023: p.restart_params.sufficient_reduction_for_restart = std::atof(optarg); break;
        case 1024: p.restart_params.necessary_reduction_for_restart = std::atof(optarg); break;
        case 1025: p.curtis_reid_iterations = std::atoi(optarg); break;
        case 1026:
            if (std::strcmp(optarg, "mixed") == 0) precision = PDHCG_PRECISION_MIXED;
            else if (std::strcmp(optarg, "fp64") != 0) {
                std::fprintf(stderr, "Error: precision must be 'fp64' or 'mixed'\n");
                return 1;
            }
            break;
        default: return 1;
        }
    }
    if (argc - optind != 2) {
        std::fprintf(stderr, "Error: You must specify an input file and an output directory.\n\n");
        usage(argv[0]);
        return 1;
    }
    const std::string file = argv[optind], outdir = argv[optind + 1];
    const std::string name = instance_name(file.c_str());
    qp_problem_t *prob = is_cbf(file) ? read_cbf_file(file.c_str()) : read_mps_file(file.c_str());
    if (!prob) {
        std::fprintf(stderr, "Failed to read or parse the file.\n");
        return 1;
    }
    pdhcg_set_precision(precision);
    pdhcg_result_t *r = solve_qp_problem(prob, &p);
    if (!r) {
        std::fprintf(stderr, "Solver failed.\n");
    } else {
        const std::string stem = outdir + "/" + name;
        write_summary(r, stem + "_summary.txt");
        write_vector(r->primal_solution, r->num_variables, stem + "_primal_solution.txt");
        write_vector(r->dual_solution, r->num_constraints, stem + "_dual_solution.txt");
        pdhcg_result_free(r);
    }
    qp_problem_free(prob);
    return 0;
}
