/*
 * ORACLE (test infrastructure, see oracle.h): cone projections and dual-cone
 * residuals on the scaled space, FP64, one block at a time.
 *
 * Second-order cones (reference src/kernels/pdhcg_soc_cone_kernels.cu:267-469,
 * :471-726): closed form when the block scaling is uniform, bisection on the
 * multiplier (60 steps, relative width 1e-13, warm start) otherwise.
 * Rotated cones (reference src/kernels/pdhcg_rsoc_cone_kernels.cu:466-945):
 * the same after the 45-degree rotation of (s, t).
 * Exponential cone (reference src/kernels/pdhcg_exp_cone_kernels.cu:26-131,
 * :526-614): in-cone / polar / special-case tests, then Newton on rho with
 * steps clamped to +-10 (100 steps, tolerance 1e-13).
 * Power cone (reference src/kernels/pdhcg_power_cone_kernels.cu:25-132,
 * :435-669): homogeneity normalisation, 80-step bisection in log rho.
 * PSD cone (reference src/kernels/pdhcg_psd_cone_kernels.cu:65-136): svec ->
 * dense, eigen-decomposition, sum of max(lambda, 0) u u', dense -> svec.  The
 * eigen-decomposition here is a cyclic Jacobi (Golub & Van Loan, sec. 8.5)
 * run to full FP64 convergence; the reference calls cuSOLVER syevj/syevd.
 * Fixed cone sections: orc_section.c.
 */
#include "orc_cones.h"
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define INV_SQRT2 0.7071067811865475
#define SQRT2 1.41421356237309504880

/* ---------------- second-order cone ---------------- */

/* f(lam) of the weighted SOC multiplier equation */
static double soc_f(const double *v, int k, const double *dh, double w, double dw, double z, double lam) {
    double sum = 0.0;
    for (int m = 0; m < k; ++m) {
        double dh2 = dh[m] * dh[m];
        double t = v[m] * dh[m] / (dh2 + 2.0 * lam);
        sum += t * t;
    }
    double tw = w * dw / (dw * dw + 2.0 * lam);
    sum += tw * tw;
    double zt = z / (1.0 - 2.0 * lam);
    return sum - zt * zt;
}

/* Bisection for the weighted SOC multiplier; dh[m] = d_m / d_z, dw = d_w / d_z.
 * Returns lam (and updates *warm). */
static double soc_bisect(const double *v, int k, const double *dh, double w, double dw, double z, double *warm) {
    double lo, hi;
    int z_pos = z > 0.0;
    if (z_pos) {
        lo = 0.0;
        hi = 0.5 - 1e-14;
    } else {
        lo = 0.5 + 1e-14;
        hi = 1.0;
        for (int d = 0; d < 60; ++d) {
            if (soc_f(v, k, dh, w, dw, z, hi) > 0.0) break;
            lo = hi;
            hi *= 2.0;
        }
    }
    double wl = *warm;
    if (wl > lo && wl < hi) {
        double f = soc_f(v, k, dh, w, dw, z, wl);
        if (fabs(f) < 1e-12) return wl;
        if (z_pos) { if (f > 0.0) lo = wl; else hi = wl; }
        else { if (f > 0.0) hi = wl; else lo = wl; }
    }
    for (int it = 0; it < 60; ++it) {
        double lam = 0.5 * (lo + hi);
        double f = soc_f(v, k, dh, w, dw, z, lam);
        if (z_pos) { if (f > 0.0) lo = lam; else hi = lam; }
        else { if (f > 0.0) hi = lam; else lo = lam; }
        if ((hi - lo) / (1.0 + hi + lo) < 1e-13) break;
    }
    double lam = 0.5 * (lo + hi);
    *warm = lam;
    return lam;
}

/* Projection of (v, w, z) onto D K_soc with d the block's scaling. */
void orc_soc_project(double *x, int k, const double *d, double *warm) {
    double *v = x, w = x[k], z = x[k + 1];
    double dz = d[k + 1], dw = d[k] / dz;
    int uniform = dw == 1.0;
    for (int m = 0; m < k && uniform; ++m) uniform = d[m] == dz;
    if (uniform) {
        double ss = w * w;
        for (int m = 0; m < k; ++m) ss += v[m] * v[m];
        double r = sqrt(ss);
        if (r <= z) return;
        if (r <= -z) {
            memset(x, 0, (size_t)(k + 2) * sizeof(double));
            return;
        }
        double sc = (z + r) / (2.0 * r);
        for (int m = 0; m < k; ++m) v[m] *= sc;
        x[k] = sc * w;
        x[k + 1] = sc * r;
        return;
    }
    double *dh = (double *)malloc((size_t)(k > 0 ? k : 1) * sizeof(double));
    double ri = (w / dw) * (w / dw), rp = (w * dw) * (w * dw);
    for (int m = 0; m < k; ++m) {
        dh[m] = d[m] / dz;
        ri += (v[m] / dh[m]) * (v[m] / dh[m]);
        rp += (v[m] * dh[m]) * (v[m] * dh[m]);
    }
    if (sqrt(ri) <= z) { free(dh); return; }
    if (sqrt(rp) <= -z) {
        memset(x, 0, (size_t)(k + 2) * sizeof(double));
        free(dh);
        return;
    }
    double lam = soc_bisect(v, k, dh, w, dw, z, warm);
    x[k + 1] = z / (1.0 - 2.0 * lam);
    x[k] = w * dw * dw / (dw * dw + 2.0 * lam);
    for (int m = 0; m < k; ++m) v[m] = v[m] * dh[m] * dh[m] / (dh[m] * dh[m] + 2.0 * lam);
    free(dh);
}

/* Dual residual of a SOC block: (rc - P_{K*}(rc)) d with rc = c_eff - A'y
 * (reference :471-726).  Writes the k+2 residual entries to out. */
void orc_soc_dual_residual(const double *rc, int k, const double *d, double *warm, double *out) {
    double rw = rc[k], rz = rc[k + 1];
    double dz = d[k + 1], ew = dz / d[k];
    int uniform = ew == 1.0;
    for (int m = 0; m < k && uniform; ++m) uniform = d[m] == dz;
    if (uniform) {
        double ss = rw * rw;
        for (int m = 0; m < k; ++m) ss += rc[m] * rc[m];
        double r = sqrt(ss), vf, pw, pz;
        if (r <= rz) { vf = 0.0; pw = rw; pz = rz; }
        else if (r <= -rz) { vf = 1.0; pw = 0.0; pz = 0.0; }
        else {
            double sc = (rz + r) / (2.0 * r);
            vf = 1.0 - sc;
            pw = sc * rw;
            pz = sc * r;
        }
        for (int m = 0; m < k; ++m) out[m] = rc[m] * vf * d[m];
        out[k] = (rw - pw) * d[k];
        out[k + 1] = (rz - pz) * d[k + 1];
        return;
    }
    double *e = (double *)malloc((size_t)(k > 0 ? k : 1) * sizeof(double));
    double ri = (rw / ew) * (rw / ew), rp = (rw * ew) * (rw * ew);
    for (int m = 0; m < k; ++m) {
        e[m] = dz / d[m];
        ri += (rc[m] / e[m]) * (rc[m] / e[m]);
        rp += (rc[m] * e[m]) * (rc[m] * e[m]);
    }
    if (sqrt(ri) <= rz) {
        for (int m = 0; m < k + 2; ++m) out[m] = 0.0;
        free(e);
        return;
    }
    if (sqrt(rp) <= -rz) {
        for (int m = 0; m < k + 2; ++m) out[m] = rc[m] * d[m];
        free(e);
        return;
    }
    double lam = soc_bisect(rc, k, e, rw, ew, rz, warm);
    double pz = rz / (1.0 - 2.0 * lam), pw = rw * ew * ew / (ew * ew + 2.0 * lam);
    for (int m = 0; m < k; ++m) {
        double pm = rc[m] * e[m] * e[m] / (e[m] * e[m] + 2.0 * lam);
        out[m] = (rc[m] - pm) * d[m];
    }
    out[k] = (rw - pw) * d[k];
    out[k + 1] = (rz - pz) * d[k + 1];
    free(e);
}

/* ---------------- rotated second-order cone ---------------- */

void orc_rsoc_project(double *x, int k, const double *d, double *warm) {
    double *v = x, s = x[k], t = x[k + 1];
    double w = (s - t) * INV_SQRT2, z = (s + t) * INV_SQRT2;
    double dst = sqrt(d[k] * d[k + 1]);
    int uniform = 1;
    for (int m = 0; m < k && uniform; ++m) uniform = d[m] == dst;
    if (uniform) {
        double ss = w * w;
        for (int m = 0; m < k; ++m) ss += v[m] * v[m];
        double r = sqrt(ss);
        if (r <= z) return;
        if (r <= -z) {
            memset(x, 0, (size_t)(k + 2) * sizeof(double));
            return;
        }
        double sc = (z + r) / (2.0 * r);
        for (int m = 0; m < k; ++m) v[m] *= sc;
        double wn = sc * w, zn = sc * r;
        x[k] = (zn + wn) * INV_SQRT2;
        x[k + 1] = (zn - wn) * INV_SQRT2;
        return;
    }
    double *dh = (double *)malloc((size_t)(k > 0 ? k : 1) * sizeof(double));
    double ri = w * w, rp = w * w;
    for (int m = 0; m < k; ++m) {
        dh[m] = d[m] / dst;
        ri += (v[m] / dh[m]) * (v[m] / dh[m]);
        rp += (v[m] * dh[m]) * (v[m] * dh[m]);
    }
    if (sqrt(ri) <= z) { free(dh); return; }
    if (sqrt(rp) <= -z) {
        memset(x, 0, (size_t)(k + 2) * sizeof(double));
        free(dh);
        return;
    }
    double lam = soc_bisect(v, k, dh, w, 1.0, z, warm);
    double wn = w / (1.0 + 2.0 * lam), zn = z / (1.0 - 2.0 * lam);
    for (int m = 0; m < k; ++m) v[m] = v[m] * dh[m] * dh[m] / (dh[m] * dh[m] + 2.0 * lam);
    x[k] = (zn + wn) * INV_SQRT2;
    x[k + 1] = (zn - wn) * INV_SQRT2;
    free(dh);
}

void orc_rsoc_dual_residual(const double *rc, int k, const double *d, double *warm, double *out) {
    double rs = rc[k], rt = rc[k + 1];
    double rw = (rs - rt) * INV_SQRT2, rz = (rs + rt) * INV_SQRT2;
    double dst = sqrt(d[k] * d[k + 1]);
    int uniform = 1;
    for (int m = 0; m < k && uniform; ++m) uniform = d[m] == dst;
    if (uniform) {
        double ss = rw * rw;
        for (int m = 0; m < k; ++m) ss += rc[m] * rc[m];
        double r = sqrt(ss), vf, ps, pt;
        if (r <= rz) { vf = 0.0; ps = rs; pt = rt; }
        else if (r <= -rz) { vf = 1.0; ps = 0.0; pt = 0.0; }
        else {
            double sc = (rz + r) / (2.0 * r);
            vf = 1.0 - sc;
            double wn = sc * rw, zn = sc * r;
            ps = (zn + wn) * INV_SQRT2;
            pt = (zn - wn) * INV_SQRT2;
        }
        for (int m = 0; m < k; ++m) out[m] = rc[m] * vf * d[m];
        out[k] = (rs - ps) * d[k];
        out[k + 1] = (rt - pt) * d[k + 1];
        return;
    }
    double *e = (double *)malloc((size_t)(k > 0 ? k : 1) * sizeof(double));
    double ri = rw * rw, rp = rw * rw;
    for (int m = 0; m < k; ++m) {
        e[m] = dst / d[m];
        ri += (rc[m] / e[m]) * (rc[m] / e[m]);
        rp += (rc[m] * e[m]) * (rc[m] * e[m]);
    }
    if (sqrt(ri) <= rz) {
        for (int m = 0; m < k + 2; ++m) out[m] = 0.0;
        free(e);
        return;
    }
    if (sqrt(rp) <= -rz) {
        for (int m = 0; m < k + 2; ++m) out[m] = rc[m] * d[m];
        free(e);
        return;
    }
    double lam = soc_bisect(rc, k, e, rw, 1.0, rz, warm);
    double wn = rw / (1.0 + 2.0 * lam), zn = rz / (1.0 - 2.0 * lam);
    for (int m = 0; m < k; ++m) {
        double pm = rc[m] * e[m] * e[m] / (e[m] * e[m] + 2.0 * lam);
        out[m] = (rc[m] - pm) * d[m];
    }
    out[k] = (rs - (zn + wn) * INV_SQRT2) * d[k];
    out[k + 1] = (rt - (zn - wn) * INV_SQRT2) * d[k + 1];
    free(e);
}

/* ---------------- exponential cone ---------------- */

void orc_exp_project_point(double r1, double r2, double r3, double d1, double d2, double d3, double *xo, double *yo,
                           double *zo) {
    const double E = 2.718281828459045;
    double a1 = r1 / d1, a2 = r2 / d2, a3 = r3 / d3;
    if (a2 > 0.0) {
        double ratio = a1 / a2;
        if (ratio < 700.0 && a2 * exp(ratio) <= a3) { *xo = r1; *yo = r2; *zo = r3; return; }
    } else if (a2 == 0.0 && a1 <= 0.0 && a3 >= 0.0) {
        *xo = r1; *yo = r2; *zo = r3;
        return;
    }
    if (r1 > 0.0) {
        double ratio = (d2 * r2) / (d1 * r1);
        if (ratio < 700.0 && d1 * r1 * exp(ratio) + E * d3 * r3 <= 0.0) { *xo = *yo = *zo = 0.0; return; }
    } else if (r1 == 0.0 && r2 <= 0.0 && r3 <= 0.0) {
        *xo = *yo = *zo = 0.0;
        return;
    }
    if (a1 <= 0.0 && a2 <= 0.0) {
        *xo = r1;
        *yo = 0.0;
        *zo = a3 < 0.0 ? 0.0 : r3;
        return;
    }
    double al = (d3 / d1) * (d3 / d1), be = (d3 / d2) * (d3 / d2);
    double rho = 0.0;
    int diverged = 0;
    for (int it = 0; it < 100; ++it) {
        if (rho >= 299.0 || rho <= -299.0) { diverged = 1; break; }
        double er = exp(rho), e2 = er * er, om = 1.0 - rho;
        double at = al - be * rho * om;
        double bt = be * a1 * om - al * a2;
        double f = a1 - a2 * rho + a3 * er * at + e2 * bt;
        double da = -be * (1.0 - 2.0 * rho), db = -be * a1;
        double df = -a2 + a3 * er * (at + da) + e2 * (2.0 * bt + db);
        if (fabs(df) < 1e-300) break;
        double st = f / df;
        if (st > 10.0) st = 10.0;
        if (st < -10.0) st = -10.0;
        rho -= st;
        if (fabs(st) < 1e-13 * (1.0 + fabs(rho))) break;
    }
    double er = exp(rho);
    double den = rho + al * er * er;
    double u2 = (a1 + al * a3 * er) / den;
    if (diverged || !isfinite(u2) || u2 <= 0.0) {
        *xo = r1 < 0.0 ? r1 : 0.0;
        *yo = 0.0;
        *zo = r3 > 0.0 ? r3 : 0.0;
        return;
    }
    *xo = d1 * (rho * u2);
    *yo = d2 * u2;
    *zo = d3 * (u2 * er);
}

void orc_exp_project(double *x, const double *d) {
    double a, b, c;
    orc_exp_project_point(x[0], x[1], x[2], d[0], d[1], d[2], &a, &b, &c);
    x[0] = a; x[1] = b; x[2] = c;
}

/* dist(rc, K_exp*) via Moreau with inverse scaling (reference :562-614). */
void orc_exp_dual_residual(const double *rc, const double *d, double *out) {
    double a, b, c;
    orc_exp_project_point(-rc[0], -rc[1], -rc[2], 1.0 / d[0], 1.0 / d[1], 1.0 / d[2], &a, &b, &c);
    out[0] = -a * d[0];
    out[1] = -b * d[1];
    out[2] = -c * d[2];
}

/* ---------------- power cone ---------------- */

static double pos_root(double r, double q) {
    if (!(q > 0.0)) return fmax(r, 0.0);
    double disc = hypot(r, 2.0 * sqrt(q));
    if (r >= 0.0) return 0.5 * (r + disc);
    return (2.0 * q) / (disc - r);
}

static void pow_normalized(double rx, double ry, double rz, double wx, double wy, double wz, double al, double *xo,
                           double *yo, double *zo) {
    double az = fabs(rz), sg = rz >= 0.0 ? 1.0 : -1.0, om = 1.0 - al;
    if (az == 0.0) { *xo = fmax(rx, 0.0); *yo = fmax(ry, 0.0); *zo = 0.0; return; }
    if (rx > 0.0 && ry > 0.0 && al * log(rx) + om * log(ry) >= log(az)) { *xo = rx; *yo = ry; *zo = rz; return; }
    if (rx <= 0.0 && ry <= 0.0) {
        double u = rx < 0.0 ? (-wx * rx) / al : 0.0;
        double v = ry < 0.0 ? (-wy * ry) / om : 0.0;
        if (u > 0.0 && v > 0.0 && al * log(u) + om * log(v) >= log(wz) + log(az)) { *xo = *yo = *zo = 0.0; return; }
    }
    double cx = 4.0 * al * (wz / wx), cy = 4.0 * om * (wz / wy);
    double lo = log(DBL_MIN), hi = log(az);
    if (!(hi > lo)) { *xo = fmax(rx, 0.0); *yo = fmax(ry, 0.0); *zo = 0.0; return; }
    for (int it = 0; it < 80; ++it) {
        double lr = lo + 0.5 * (hi - lo), rho = exp(lr);
        double x = pos_root(rx, 0.25 * cx * rho * (az - rho));
        double y = pos_root(ry, 0.25 * cy * rho * (az - rho));
        if (x > 0.0 && y > 0.0 && al * log(x) + om * log(y) > lr) lo = lr;
        else hi = lr;
    }
    double rho = exp(lo + 0.5 * (hi - lo));
    *xo = pos_root(rx, 0.25 * cx * rho * (az - rho));
    *yo = pos_root(ry, 0.25 * cy * rho * (az - rho));
    *zo = sg * fmin(rho, exp(al * log(*xo) + om * log(*yo)));
}

void orc_power_project_point(double rx, double ry, double rz, double wx, double wy, double wz, double al, double *xo,
                             double *yo, double *zo) {
    double sc = fmax(fabs(rx), fmax(fabs(ry), fabs(rz)));
    if (!(sc > 0.0) || !isfinite(sc)) {
        pow_normalized(rx, ry, rz, wx, wy, wz, al, xo, yo, zo);
        return;
    }
    double a, b, c;
    pow_normalized(rx / sc, ry / sc, rz / sc, wx, wy, wz, al, &a, &b, &c);
    *xo = a * sc;
    *yo = b * sc;
    *zo = c * sc;
}

void orc_power_project(double *x, const double *d, double al) {
    double a, b, c;
    orc_power_project_point(x[0] / d[0], x[1] / d[1], x[2] / d[2], d[0] * d[0], d[1] * d[1], d[2] * d[2], al, &a, &b,
                            &c);
    x[0] = a * d[0];
    x[1] = b * d[1];
    x[2] = c * d[2];
}

void orc_power_dual_residual(const double *rc, const double *d, double al, double *out) {
    double a, b, c;
    orc_power_project_point(-rc[0] * d[0], -rc[1] * d[1], -rc[2] * d[2], 1.0, 1.0, 1.0, al, &a, &b, &c);
    out[0] = -a;
    out[1] = -b;
    out[2] = -c;
}

/* Primal violation of a power block, unscaled (reference :621-669). */
void orc_power_violation(const double *x3, const double *d, double al, double hscale, double *absv, double *relv) {
    double x = x3[0] / d[0], y = x3[1] / d[1], z = x3[2] / d[2];
    if (!isfinite(x) || !isfinite(y) || !isfinite(z)) { *absv = *relv = INFINITY; return; }
    double viol = fmax(-x, -y), az = fabs(z);
    if (az > 0.0) {
        double bound = 0.0;
        if (x > 0.0 && y > 0.0) {
            double lb = al * log(x) + (1.0 - al) * log(y), lz = log(az);
            if (lb + 64.0 * DBL_EPSILON * (1.0 + fabs(lb) + fabs(lz)) >= lz) {
                viol = fmax(viol, 0.0);
                *absv = viol;
                *relv = viol / (hscale + fmax(fabs(x), fmax(fabs(y), az)));
                return;
            }
            bound = exp(lb);
        }
        viol = fmax(viol, az - bound);
    }
    viol = fmax(viol, 0.0);
    *absv = viol;
    *relv = viol / (hscale + fmax(fabs(x), fmax(fabs(y), az)));
}

/* ---------------- PSD cone ---------------- */

/* Cyclic Jacobi on a dense symmetric order-k matrix (row-major, overwritten):
 * eigenvalues to w, eigenvectors to the columns of V. */
static void jacobi_eig(double *A, int k, double *w, double *V) {
    for (int i = 0; i < k * k; ++i) V[i] = 0.0;
    for (int i = 0; i < k; ++i) V[i * k + i] = 1.0;
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0, tot = 0.0;
        for (int i = 0; i < k; ++i)
            for (int j = 0; j < k; ++j) {
                tot += A[i * k + j] * A[i * k + j];
                if (i != j) off += A[i * k + j] * A[i * k + j];
            }
        if (off <= 1e-30 * tot || off == 0.0) break;
        for (int p = 0; p < k - 1; ++p)
            for (int q = p + 1; q < k; ++q) {
                double apq = A[p * k + q];
                if (apq == 0.0) continue;
                double app = A[p * k + p], aqq = A[q * k + q];
                double theta = (aqq - app) / (2.0 * apq);
                double t = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                for (int r = 0; r < k; ++r) {
                    double arp = A[r * k + p], arq = A[r * k + q];
                    A[r * k + p] = c * arp - s * arq;
                    A[r * k + q] = s * arp + c * arq;
                }
                for (int r = 0; r < k; ++r) {
                    double apr = A[p * k + r], aqr = A[q * k + r];
                    A[p * k + r] = c * apr - s * aqr;
                    A[q * k + r] = s * apr + c * aqr;
                }
                for (int r = 0; r < k; ++r) {
                    double vrp = V[r * k + p], vrq = V[r * k + q];
                    V[r * k + p] = c * vrp - s * vrq;
                    V[r * k + q] = s * vrp + c * vrq;
                }
            }
    }
    for (int i = 0; i < k; ++i) w[i] = A[i * k + i];
}

/* svec layout: lower triangle, column-major, off-diagonals times sqrt(2). */
void orc_psd_project(double *x, int order) {
    int k = order;
    if (k == 1) {
        x[0] = fmax(x[0], 0.0);
        return;
    }
    double *A = (double *)malloc((size_t)k * k * sizeof(double));
    double *V = (double *)malloc((size_t)k * k * sizeof(double));
    double *w = (double *)malloc((size_t)k * sizeof(double));
    int slot = 0;
    for (int c = 0; c < k; ++c)
        for (int r = c; r < k; ++r, ++slot) {
            double v = x[slot];
            if (r != c) v *= INV_SQRT2;
            A[r * k + c] = v;
            A[c * k + r] = v;
        }
    jacobi_eig(A, k, w, V);
    slot = 0;
    for (int c = 0; c < k; ++c)
        for (int r = c; r < k; ++r, ++slot) {
            double v = 0.0;
            for (int e = 0; e < k; ++e) v += fmax(w[e], 0.0) * V[r * k + e] * V[c * k + e];
            if (r != c) v *= SQRT2;
            x[slot] = v;
        }
    free(A);
    free(V);
    free(w);
}

/* (rc - P_K(rc)) d on a PSD block (self-dual; reference :148-199). */
void orc_psd_dual_residual(const double *rc, int order, const double *d, double *out) {
    int L = order * (order + 1) / 2;
    double *p = (double *)malloc((size_t)L * sizeof(double));
    memcpy(p, rc, (size_t)L * sizeof(double));
    orc_psd_project(p, order);
    for (int i = 0; i < L; ++i) out[i] = (rc[i] - p[i]) * d[i];
    free(p);
}

/* ---------------- generic entry points ---------------- */

int orc_project_block(int type, double *x, int v_dim, const double *d, double alpha) {
    double warm = 0.0;
    switch (type) {
    case CONE_STANDARD_SOC: orc_soc_project(x, v_dim, d, &warm); return 0;
    case CONE_ROTATED_SOC: orc_rsoc_project(x, v_dim, d, &warm); return 0;
    case CONE_EXPONENTIAL: orc_exp_project(x, d); return 0;
    case CONE_POWER: orc_power_project(x, d, alpha); return 0;
    case CONE_PSD: orc_psd_project(x, v_dim); return 0;
    default: return -1;
    }
}

int orc_dual_residual_block(int type, const double *rc, int v_dim, const double *d, double alpha, double *out) {
    double warm = 0.0;
    switch (type) {
    case CONE_STANDARD_SOC: orc_soc_dual_residual(rc, v_dim, d, &warm, out); return 0;
    case CONE_ROTATED_SOC: orc_rsoc_dual_residual(rc, v_dim, d, &warm, out); return 0;
    case CONE_EXPONENTIAL: orc_exp_dual_residual(rc, d, out); return 0;
    case CONE_POWER: orc_power_dual_residual(rc, d, alpha, out); return 0;
    case CONE_PSD: orc_psd_dual_residual(rc, v_dim, d, out); return 0;
    default: return -1;
    }
}
