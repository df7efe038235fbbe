/*
 * ORACLE (test infrastructure, see oracle.h): internal working-problem and
 * solver-state types of the FP64 CPU restatement.
 */
#pragma once

#include "oracle.h"

#include <stdbool.h>

typedef struct {
    int rows, cols, nnz;
    int *rp, *ci;
    double *v;
} orc_csr;

/* Scaled working problem (the reference's rescale_info_t::scaled_problem and
 * processed_qp_problem_t, reference internal/internal_types.h:268-310). */
typedef struct {
    int n, m, k;
    orc_csr A, Q, R;
    double *lb, *ub, *c, *cl, *cu, *g;
    double obj_const;
    quad_obj_type_t qtype;
    int dkind; /* 0 none, 1 diag, 2 dense (reference lowrank_middle_kind_t) */
    double *ddiag, *ddense;
    double *qdiag; /* DIAG_Q: diag(Q) + 1e-12 */
    double *con_rescale, *var_rescale;
    double cbr, ovr;
    double rescale_time;
} orc_scaled;

int orc_cone_block_length(const cone_blocks_t *b, int k);
void orc_csr_free(orc_csr *a);
orc_scaled *orc_rescale(const pdhg_parameters_t *params, const qp_problem_t *work);
void orc_scaled_free(orc_scaled *s);
