/*
 * ORACLE (test infrastructure, see oracle.h): per-block cone projections and
 * dual-cone residuals on the scaled space (d = the block's variable scaling).
 */
#pragma once

void orc_soc_project(double *x, int v_dim, const double *d, double *warm);
void orc_soc_dual_residual(const double *rc, int v_dim, const double *d, double *warm, double *out);
void orc_rsoc_project(double *x, int v_dim, const double *d, double *warm);
void orc_rsoc_dual_residual(const double *rc, int v_dim, const double *d, double *warm, double *out);
void orc_exp_project_point(double r1, double r2, double r3, double d1, double d2, double d3, double *xo, double *yo,
                           double *zo);
void orc_exp_project(double *x, const double *d);
void orc_exp_dual_residual(const double *rc, const double *d, double *out);
void orc_power_project_point(double rx, double ry, double rz, double wx, double wy, double wz, double alpha,
                             double *xo, double *yo, double *zo);
void orc_power_project(double *x, const double *d, double alpha);
void orc_power_dual_residual(const double *rc, const double *d, double alpha, double *out);
void orc_power_violation(const double *x, const double *d, double alpha, double hscale, double *absv, double *relv);
void orc_psd_project(double *x, int order);
void orc_psd_dual_residual(const double *rc, int order, const double *d, double *out);
