/*
 * ORACLE (test infrastructure, see oracle.h): projections onto the fixed
 * sections of the variable cones, and the residual used for blocks that have
 * fixed slots.
 *
 * A block with fixed slots (set_cone_fixed) keeps those slots at their value;
 * the free slots are moved to the nearest point, in the weighted norm, such
 * that the whole block lies in the cone.  With actual values a_i = v_i / d_i
 * and weights w_i = metric_i d_i^2 (metric = 1 + tau q on the diagonal-Q cone
 * step, 1 otherwise; reference internal/cone_projection_utils.h:22-43) the
 * problem is
 *      min  sum_{i free} w_i (p_i - a_i)^2   s.t.  p in K,  p_i = a_i (fixed).
 * The objective is strictly convex on the free slots and the section is
 * convex, so the minimiser is unique; the reference computes it by case
 * analysis and bisection (src/kernels/pdhcg_soc_cone_kernels.cu:23-260,
 * pdhcg_rsoc_cone_kernels.cu:78-465, pdhcg_exp_cone_kernels.cu:285-525,
 * pdhcg_power_cone_kernels.cu:134-433) and this file computes the same point
 * by the following formulations:
 *
 *  - every root is found by bracketed bisection on a monotone function, with
 *    geometric midpoints while a bracket spans more than a factor 4 on one
 *    side of zero, so a root near zero is resolved to full relative
 *    precision;
 *  - SOC {||u|| <= z}, u = (v, w): z fixed -> weighted ball projection of
 *    the free part of u (multiplier l: u_i = w_i a_i / (w_i + l));
 *    z free -> u_i(n) = w_i a_i / (w_i + n), N(n)^2 = Phi + sum u_i(n)^2
 *    (Phi = squared norm of the fixed part), and the root of the decreasing
 *    G(n) = N(n) (1 - n / w_z) - a_z gives z = N(n) (both branches of the
 *    reference's bisection and its a_z = 0 case in one function; the apex
 *    when Phi = 0 and -w_z a_z >= ||w a||);
 *  - RSOC {||v||^2 <= 2 s t, s, t >= 0}: s and t fixed -> ball of radius^2
 *    2 s t - Phi; one end E fixed -> root of the decreasing
 *    H(l) = Phi + sum u_i(l)^2 - 2 E (a_f + l E / w_f) for the free end f;
 *    neither fixed -> the weighted SOC section after s' = sqrt(w_s) s,
 *    t' = sqrt(w_t) t, v' = (w_s w_t)^(1/4) v and the rotation
 *    (w, z) = ((s' - t'), (s' + t')) / sqrt(2);
 *  - EXP and POWER: two fixed slots leave an interval for the free one
 *    (clamp); one fixed slot leaves a 2-D set {q >= g(p)} with g convex, whose
 *    projection is found through the multiplier m of q >= g(p): p(m) is the
 *    root of the increasing w_p (p - a_p) + m g'(p) and m the root of the
 *    decreasing g(p(m)) - a_q - m / w_q (the derivative of the concave dual);
 *    fixed values at zero leave boxes.
 *
 * The residual of a block with fixed slots (reference
 * src/kernels/pdhcg_soc_cone_kernels.cu:489-507 and the matching code of the
 * other cone files) is the prox residual (x - P(x - rc)) d on the free slots,
 * with P the section projection in the metric d^2, and 0 on the fixed slots.
 */
#include "orc_cones.h"
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define SQRT_HALF 0.70710678118654752440

typedef double (*orc_fn)(double, const void *);

/* Root of an increasing f on [lo, hi] with f(lo) <= 0 <= f(hi) (f is not
 * evaluated at the ends). */
static double root_increasing(orc_fn f, const void *ctx, double lo, double hi) {
    for (int it = 0; it < 5000; ++it) {
        double mid;
        if (lo >= 0.0 && hi > 4.0 * lo) mid = lo > 0.0 ? sqrt(lo) * sqrt(hi) : hi * 0.0625;
        else if (hi <= 0.0 && lo < 4.0 * hi) mid = hi < 0.0 ? -(sqrt(-lo) * sqrt(-hi)) : lo * 0.0625;
        else mid = lo + 0.5 * (hi - lo);
        if (!(mid > lo && mid < hi)) break;
        if (f(mid, ctx) > 0.0) hi = mid;
        else lo = mid;
    }
    return lo + 0.5 * (hi - lo);
}

/* ---------------- weighted ball and SOC sections ---------------- */

typedef struct {
    const double *a, *w;
    const char *fx;
    int n;
    double phi, r2, wz, az;
} ball_ctx;

static double shrunk_norm2(const ball_ctx *c, double l) {
    double s = 0.0;
    for (int i = 0; i < c->n; ++i)
        if (!c->fx[i]) {
            double u = c->w[i] * c->a[i] / (c->w[i] + l);
            s += u * u;
        }
    return s;
}

/* decreasing in l: sum u_i(l)^2 - r2; negated for root_increasing */
static double ball_neg(double l, const void *p) {
    const ball_ctx *c = (const ball_ctx *)p;
    return c->r2 - shrunk_norm2(c, l);
}

/* Free slots of a[0..n) (weights w) projected onto sum_free u^2 <= r2. */
static void ball_section(const double *a, const double *w, const char *fx, int n, double r2, double *out) {
    double f2 = 0.0, p2 = 0.0;
    for (int i = 0; i < n; ++i)
        if (!fx[i]) {
            f2 += a[i] * a[i];
            p2 += (w[i] * a[i]) * (w[i] * a[i]);
            out[i] = a[i];
        }
    if (f2 <= r2) return;
    if (!(r2 > 0.0)) {
        for (int i = 0; i < n; ++i)
            if (!fx[i]) out[i] = 0.0;
        return;
    }
    ball_ctx c = {a, w, fx, n, 0.0, r2, 0.0, 0.0};
    /* at l = ||w a|| / r every u_i <= w_i a_i / l, so sum u^2 <= r2 */
    double hi = sqrt(p2) / sqrt(r2) * (1.0 + 1e-12) + DBL_MIN;
    double l = root_increasing(ball_neg, &c, 0.0, hi);
    for (int i = 0; i < n; ++i)
        if (!fx[i]) out[i] = w[i] * a[i] / (w[i] + l);
}

/* G(n) = N(n) (1 - n / w_z) - a_z, decreasing; negated */
static double soc_neg(double l, const void *p) {
    const ball_ctx *c = (const ball_ctx *)p;
    double N = sqrt(c->phi + shrunk_norm2(c, l));
    return c->az - N * (1.0 - l / c->wz);
}

/* SOC section over u = a[0..n) (free/fixed per fx) and a free z = az with
 * weight wz; writes the free u slots to out and returns z. */
static double soc_free_z(const double *a, const double *w, const char *fx, int n, double az, double wz, double *out) {
    double phi = 0.0, f2 = 0.0, p2 = 0.0;
    int nfree = 0;
    for (int i = 0; i < n; ++i) {
        if (fx[i]) phi += a[i] * a[i];
        else {
            f2 += a[i] * a[i];
            p2 += (w[i] * a[i]) * (w[i] * a[i]);
            out[i] = a[i];
            ++nfree;
        }
    }
    if (az >= 0.0 && phi + f2 <= az * az) return az;
    if (nfree == 0) return fmax(az, sqrt(phi));
    if (phi == 0.0 && -wz * az >= sqrt(p2)) {
        for (int i = 0; i < n; ++i)
            if (!fx[i]) out[i] = 0.0;
        return 0.0;
    }
    ball_ctx c = {a, w, fx, n, phi, 0.0, wz, az};
    double l;
    if (az == 0.0) l = wz;
    else if (az > 0.0) l = root_increasing(soc_neg, &c, 0.0, wz);
    else {
        double lo = wz, hi = 2.0 * wz;
        for (int it = 0; it < 2100 && soc_neg(hi, &c) <= 0.0; ++it) {
            lo = hi;
            hi *= 2.0;
        }
        l = root_increasing(soc_neg, &c, lo, hi);
    }
    double s = phi;
    for (int i = 0; i < n; ++i)
        if (!fx[i]) {
            out[i] = w[i] * a[i] / (w[i] + l);
            s += out[i] * out[i];
        }
    return sqrt(s);
}

/* SOC block of vd + 2 slots: u = slots 0..vd, z = slot vd + 1. */
static void soc_section(const double *a, const double *w, const char *fx, int vd, double *out) {
    int n = vd + 1;
    for (int i = 0; i <= n; ++i) out[i] = a[i];
    if (fx[n]) {
        double phi = 0.0;
        for (int i = 0; i < n; ++i)
            if (fx[i]) phi += a[i] * a[i];
        ball_section(a, w, fx, n, fmax(0.0, a[n] * a[n] - phi), out);
    } else {
        out[n] = soc_free_z(a, w, fx, n, a[n], w[n], out);
    }
}

/* H(l) = Phi + sum u_i(l)^2 - 2 E (a_f + l E / w_f), decreasing; negated */
typedef struct {
    ball_ctx b;
    double E, af, wf;
} rsoc_ctx;

static double rsoc_end_neg(double l, const void *p) {
    const rsoc_ctx *c = (const rsoc_ctx *)p;
    return 2.0 * c->E * (c->af + l * c->E / c->wf) - (c->b.phi + shrunk_norm2(&c->b, l));
}

/* RSOC block: v = slots 0..vd-1, s = vd, t = vd + 1. */
static void rsoc_section(const double *a, const double *w, const char *fx, int vd, double *out) {
    const int si = vd, ti = vd + 1;
    for (int i = 0; i < vd + 2; ++i) out[i] = a[i];
    double phi = 0.0, f2 = 0.0;
    int nfree = 0;
    for (int i = 0; i < vd; ++i) {
        if (fx[i]) phi += a[i] * a[i];
        else {
            f2 += a[i] * a[i];
            ++nfree;
        }
    }
    if (fx[si] && fx[ti]) {
        ball_section(a, w, fx, vd, fmax(0.0, 2.0 * a[si] * a[ti] - phi), out);
        return;
    }
    if (fx[si] || fx[ti]) {
        const int e = fx[si] ? si : ti, f = fx[si] ? ti : si;
        const double E = a[e], af = a[f], wf = w[f];
        if (!(E > 0.0)) {
            for (int i = 0; i < vd; ++i)
                if (!fx[i]) out[i] = 0.0;
            out[f] = fmax(af, 0.0);
            return;
        }
        if (af >= 0.0 && phi + f2 <= 2.0 * E * af) return;
        if (nfree == 0) {
            out[f] = fmax(af, phi / (2.0 * E));
            return;
        }
        rsoc_ctx c = {{a, w, fx, vd, phi, 0.0, 0.0, 0.0}, E, af, wf};
        /* H(l) <= Phi + f2 - 2 E a_f - 2 l E^2 / w_f */
        double hi = wf * (phi + f2 - 2.0 * E * af) / (2.0 * E * E) * (1.0 + 1e-12) + DBL_MIN;
        double l = root_increasing(rsoc_end_neg, &c, 0.0, hi);
        for (int i = 0; i < vd; ++i)
            if (!fx[i]) out[i] = w[i] * a[i] / (w[i] + l);
        out[f] = af + l * E / wf;
        return;
    }
    if (a[si] >= 0.0 && a[ti] >= 0.0 && phi + f2 <= 2.0 * a[si] * a[ti]) return;
    /* rotate into a weighted SOC section */
    const double ss = sqrt(w[si]), st = sqrt(w[ti]), r = ss * st, sr = sqrt(r);
    double *b = (double *)malloc((size_t)(vd + 2) * sizeof(double));
    double *bw = (double *)malloc((size_t)(vd + 2) * sizeof(double));
    double *bo = (double *)malloc((size_t)(vd + 2) * sizeof(double));
    char *bf = (char *)malloc((size_t)(vd + 2));
    for (int i = 0; i < vd; ++i) {
        b[i] = sr * a[i];
        bw[i] = w[i] / r;
        bf[i] = fx[i];
    }
    const double sp = ss * a[si], tp = st * a[ti];
    b[vd] = (sp - tp) * SQRT_HALF;
    b[vd + 1] = (sp + tp) * SQRT_HALF;
    bw[vd] = bw[vd + 1] = 1.0;
    bf[vd] = bf[vd + 1] = 0;
    for (int i = 0; i < vd + 2; ++i) bo[i] = b[i];
    double z = soc_free_z(b, bw, bf, vd + 1, b[vd + 1], 1.0, bo);
    double wr = bo[vd];
    for (int i = 0; i < vd; ++i)
        if (!fx[i]) out[i] = bo[i] / sr;
    out[si] = fmax(0.0, (z + wr) * SQRT_HALF) / ss;
    out[ti] = fmax(0.0, (z - wr) * SQRT_HALF) / st;
    free(b);
    free(bw);
    free(bo);
    free(bf);
}

/* ---------------- 2-D epigraph sections (EXP, POWER) ---------------- */

enum { CURVE_EXP_Y, CURVE_EXP_X, CURVE_EXP_Z, CURVE_POW_Z, CURVE_POW_AXIS };

typedef struct {
    int kind;
    double c, e; /* curve parameters */
} curve_t;

/* g and g' of each curve (p in the domain) */
static double curve_g(const curve_t *cv, double p) {
    switch (cv->kind) {
    case CURVE_EXP_Y: /* z >= Y exp(x / Y), p = x */
        return cv->c * exp(p / cv->c);
    case CURVE_EXP_X: /* z >= y exp(X / y), p = y >= 0 */
        if (p <= 0.0) return cv->c <= 0.0 ? 0.0 : INFINITY;
        if (cv->c / p < -745.0) return 0.0;
        return p * exp(cv->c / p);
    case CURVE_EXP_Z: /* -x >= y log(y / Z), p = y >= 0 */
        return p > 0.0 ? p * log(p / cv->c) : 0.0;
    case CURVE_POW_Z: /* y >= |Z|^(1/(1-a)) x^(-a/(1-a)), p = x > 0; c = |Z|^(1/(1-a)), e = a/(1-a) */
        return p > 0.0 ? cv->c * pow(p, -cv->e) : INFINITY;
    default: /* q >= (|p| / c)^e, e > 1 */
        return pow(fabs(p) / cv->c, cv->e);
    }
}

static double curve_dg(const curve_t *cv, double p) {
    switch (cv->kind) {
    case CURVE_EXP_Y:
        return exp(p / cv->c);
    case CURVE_EXP_X:
        if (p <= 0.0) return cv->c < 0.0 ? 0.0 : (cv->c == 0.0 ? 1.0 : -INFINITY);
        if (cv->c / p < -700.0) return 0.0;
        return exp(cv->c / p) * (1.0 - cv->c / p);
    case CURVE_EXP_Z:
        return p > 0.0 ? log(p / cv->c) + 1.0 : -INFINITY;
    case CURVE_POW_Z:
        return p > 0.0 ? -cv->e * cv->c * pow(p, -cv->e - 1.0) : -INFINITY;
    default: {
        double t = fabs(p) / cv->c;
        double d = cv->e * pow(t, cv->e - 1.0) / cv->c;
        return p < 0.0 ? -d : d;
    }
    }
}

typedef struct {
    const curve_t *cv;
    double lo;
    int closed, bounded;
    double ap, aq, wp, wq, mu;
} epi_ctx;

static double epi_stationary(double p, const void *q) {
    const epi_ctx *c = (const epi_ctx *)q;
    return c->wp * (p - c->ap) + c->mu * curve_dg(c->cv, p);
}

/* p(mu) = argmin_{p in dom} w_p (p - a_p)^2 / 2 + mu g(p) */
static double epi_p_of_mu(epi_ctx *c, double mu) {
    c->mu = mu;
    if (mu == 0.0) return c->bounded ? fmax(c->ap, c->lo) : c->ap;
    if (c->bounded && c->closed && epi_stationary(c->lo, c) >= 0.0) return c->lo;
    double L, R;
    if (c->bounded) L = c->lo;
    else {
        L = fmin(c->ap, 0.0) - 1.0;
        for (int it = 0; it < 2100 && epi_stationary(L, c) > 0.0; ++it) L -= fmax(1.0, fabs(L));
    }
    R = (c->bounded ? fmax(c->ap, c->lo) : c->ap) + 1.0;
    for (int it = 0; it < 2100 && epi_stationary(R, c) < 0.0; ++it) R += fmax(1.0, fabs(R));
    return root_increasing(epi_stationary, c, L, R);
}

/* F(mu) = g(p(mu)) - a_q - mu / w_q, decreasing; negated */
static double epi_dual_neg(double mu, const void *q) {
    epi_ctx *c = (epi_ctx *)q;
    double p = epi_p_of_mu(c, mu);
    return c->aq + mu / c->wq - curve_g(c->cv, p);
}

/* Projection of (ap, aq) with weights (wp, wq) onto {p in dom, q >= g(p)};
 * dom = R (bounded = 0) or (lo, inf) / [lo, inf) (closed). */
static void epi_project(const curve_t *cv, int bounded, double lo, int closed, double ap, double aq, double wp,
                        double wq, double *po, double *qo) {
    epi_ctx c = {cv, lo, closed, bounded, ap, aq, wp, wq, 0.0};
    /* mu = 0: p is a_p moved into the domain; feasible if q = a_q clears g */
    double p0 = epi_p_of_mu(&c, 0.0);
    if (aq >= curve_g(cv, p0)) {
        *po = p0;
        *qo = aq;
        return;
    }
    double hi = wq * (1.0 + fabs(aq));
    for (int it = 0; it < 2100 && epi_dual_neg(hi, &c) <= 0.0; ++it) hi *= 2.0;
    double mu = root_increasing(epi_dual_neg, &c, 0.0, hi);
    double p = epi_p_of_mu(&c, mu);
    *po = p;
    *qo = curve_g(cv, p);
}

static int exp_member(double x, double y, double z) {
    if (y > 0.0 && z > 0.0) {
        double lhs = log(y) + x / y, rhs = log(z);
        return lhs <= rhs + 64.0 * DBL_EPSILON * (1.0 + fabs(lhs) + fabs(rhs));
    }
    return y == 0.0 && x <= 0.0 && z >= 0.0;
}

typedef struct {
    double X, Z;
} exp_xz_ctx;

/* log y + X / y - log Z: decreasing below y = X, increasing above */
static double exp_xz_phi(double y, const void *p) {
    const exp_xz_ctx *c = (const exp_xz_ctx *)p;
    return log(y) + c->X / y - log(c->Z);
}
static double exp_xz_phi_neg(double y, const void *p) { return -exp_xz_phi(y, p); }

/* EXP block (x, y, z), closure of {y > 0, y exp(x / y) <= z}. */
static void exp_section(const double *a, const double *w, const char *fx, double *out) {
    double x = a[0], y = a[1], z = a[2];
    out[0] = x;
    out[1] = y;
    out[2] = z;
    if (exp_member(x, y, z) || (fx[0] && fx[1] && fx[2])) return;
    if (fx[0] && fx[1]) {
        out[2] = fmax(z, y > 0.0 ? y * exp(x / y) : 0.0);
    } else if (fx[1] && fx[2]) {
        out[0] = fmin(x, y > 0.0 ? y * (log(z) - log(y)) : 0.0);
    } else if (fx[0] && fx[2]) {
        if (!(z > 0.0)) {
            out[1] = 0.0;
        } else {
            exp_xz_ctx c = {x, z};
            double lo = 0.0, hi;
            double top = fmax(fmax(x, 0.0), 1.0) + z;
            for (int it = 0; it < 2100 && exp_xz_phi(top, &c) <= 0.0; ++it) top *= 2.0;
            if (x > 0.0) {
                lo = root_increasing(exp_xz_phi_neg, &c, 0.0, x);
                hi = root_increasing(exp_xz_phi, &c, x, top);
            } else {
                hi = root_increasing(exp_xz_phi, &c, 0.0, top);
            }
            out[1] = fmin(fmax(y, lo), hi);
        }
    } else if (fx[1]) {
        if (!(y > 0.0)) {
            out[0] = fmin(x, 0.0);
            out[2] = fmax(z, 0.0);
        } else {
            curve_t cv = {CURVE_EXP_Y, y, 0.0};
            epi_project(&cv, 0, 0.0, 0, x, z, w[0], w[2], &out[0], &out[2]);
        }
    } else if (fx[0]) {
        curve_t cv = {CURVE_EXP_X, x, 0.0};
        epi_project(&cv, 1, 0.0, x <= 0.0, y, z, w[1], w[2], &out[1], &out[2]);
    } else if (fx[2]) {
        if (!(z > 0.0)) {
            out[0] = fmin(x, 0.0);
            out[1] = 0.0;
        } else {
            curve_t cv = {CURVE_EXP_Z, z, 0.0};
            double q;
            epi_project(&cv, 1, 0.0, 1, y, -x, w[1], w[0], &out[1], &q);
            out[0] = -q;
        }
    }
}

static int pow_member(double x, double y, double z, double al) {
    if (x < 0.0 || y < 0.0) return 0;
    if (z == 0.0) return 1;
    if (x == 0.0 || y == 0.0) return 0;
    double lhs = log(fabs(z)), rhs = al * log(x) + (1.0 - al) * log(y);
    return lhs <= rhs + 64.0 * DBL_EPSILON * (1.0 + fabs(lhs) + fabs(rhs));
}

/* POWER block (x, y, z), {x, y >= 0, |z| <= x^al y^(1-al)}. */
static void pow_section(const double *a, const double *w, const char *fx, double al, double *out) {
    double x = a[0], y = a[1], z = a[2];
    out[0] = x;
    out[1] = y;
    out[2] = z;
    if (pow_member(x, y, z, al) || (fx[0] && fx[1] && fx[2])) return;
    if (fx[0] && fx[1]) {
        double B = pow(x, al) * pow(y, 1.0 - al);
        out[2] = fmin(fmax(z, -B), B);
    } else if (fx[1] && fx[2]) {
        out[0] = fmax(x, y > 0.0 ? pow(fabs(z) / pow(y, 1.0 - al), 1.0 / al) : 0.0);
    } else if (fx[0] && fx[2]) {
        out[1] = fmax(y, x > 0.0 ? pow(fabs(z) / pow(x, al), 1.0 / (1.0 - al)) : 0.0);
    } else if (fx[2]) {
        if (z == 0.0) {
            out[0] = fmax(x, 0.0);
            out[1] = fmax(y, 0.0);
        } else {
            curve_t cv = {CURVE_POW_Z, pow(fabs(z), 1.0 / (1.0 - al)), al / (1.0 - al)};
            epi_project(&cv, 1, 0.0, 0, x, y, w[0], w[1], &out[0], &out[1]);
        }
    } else if (fx[0]) {
        if (!(x > 0.0)) {
            out[2] = 0.0;
            out[1] = fmax(y, 0.0);
        } else {
            curve_t cv = {CURVE_POW_AXIS, pow(x, al), 1.0 / (1.0 - al)};
            epi_project(&cv, 0, 0.0, 0, z, y, w[2], w[1], &out[2], &out[1]);
        }
    } else if (fx[1]) {
        if (!(y > 0.0)) {
            out[2] = 0.0;
            out[0] = fmax(x, 0.0);
        } else {
            curve_t cv = {CURVE_POW_AXIS, pow(y, 1.0 - al), 1.0 / al};
            epi_project(&cv, 0, 0.0, 0, z, x, w[2], w[0], &out[2], &out[0]);
        }
    }
}

/* ---------------- entry points ---------------- */

int orc_section_has_fixed(const char *fixed, int len) {
    if (!fixed) return 0;
    for (int i = 0; i < len; ++i)
        if (fixed[i]) return 1;
    return 0;
}

int orc_section_project(int type, double *v, int vd, const double *d, const double *metric, const char *fixed,
                        double alpha) {
    int L;
    switch (type) {
    case CONE_STANDARD_SOC:
    case CONE_ROTATED_SOC: L = vd + 2; break;
    case CONE_EXPONENTIAL:
    case CONE_POWER: L = 3; break;
    default: return -1;
    }
    double *a = (double *)malloc((size_t)L * sizeof(double));
    double *w = (double *)malloc((size_t)L * sizeof(double));
    double *o = (double *)malloc((size_t)L * sizeof(double));
    /* scale-free problem: values by their largest magnitude, weights by the
       largest free weight (the cones are invariant under positive scaling) */
    double s = 0.0, wmax = 0.0;
    for (int i = 0; i < L; ++i) {
        a[i] = v[i] / d[i];
        double m = metric ? metric[i] : 1.0;
        w[i] = fmax(m * d[i] * d[i], DBL_MIN);
        s = fmax(s, fabs(a[i]));
        if (!fixed[i]) wmax = fmax(wmax, w[i]);
    }
    if (s > 0.0 && isfinite(s)) {
        for (int i = 0; i < L; ++i) {
            a[i] /= s;
            if (wmax > 0.0) w[i] /= wmax;
        }
        switch (type) {
        case CONE_STANDARD_SOC: soc_section(a, w, fixed, vd, o); break;
        case CONE_ROTATED_SOC: rsoc_section(a, w, fixed, vd, o); break;
        case CONE_EXPONENTIAL: exp_section(a, w, fixed, o); break;
        case CONE_POWER: pow_section(a, w, fixed, alpha, o); break;
        }
        for (int i = 0; i < L; ++i)
            if (!fixed[i]) v[i] = o[i] * s * d[i];
    }
    free(a);
    free(w);
    free(o);
    return 0;
}

void orc_section_dual_residual(int type, const double *x, const double *rc, int vd, const double *d, const char *fixed,
                               double alpha, double *out) {
    int L = (type == CONE_STANDARD_SOC || type == CONE_ROTATED_SOC) ? vd + 2 : 3;
    double *t = (double *)malloc((size_t)L * sizeof(double));
    for (int i = 0; i < L; ++i) t[i] = fixed[i] ? x[i] : x[i] - rc[i];
    orc_section_project(type, t, vd, d, NULL, fixed, alpha);
    for (int i = 0; i < L; ++i) out[i] = fixed[i] ? 0.0 : (x[i] - t[i]) * d[i];
    free(t);
}
