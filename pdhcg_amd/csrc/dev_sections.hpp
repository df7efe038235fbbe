// Fixed cone sections on the device: blocks with slots pinned by
// set_cone_fixed keep those slots, and the free slots are moved to the
// nearest point (in the weighted norm) such that the block lies in the cone
// (reference src/kernels/pdhcg_soc_cone_kernels.cu:23-260,
// pdhcg_rsoc_cone_kernels.cu:78-465, pdhcg_exp_cone_kernels.cu:285-525,
// pdhcg_power_cone_kernels.cu:134-433).
//
// With actual values a_i = v_i / d_i and weights w_i = metric_i d_i^2 the
// problem  min sum_free w_i (p_i - a_i)^2  over the section has a unique
// minimiser, computed here by one monotone root per case:
//  - SOC, apex value fixed: weighted ball projection of the free vector part;
//    apex free: with u_i(n) = w_i a_i / (w_i + n) and N(n)^2 = Phi + sum u^2
//    (Phi the squared fixed part), the root of the decreasing
//    N(n) (1 - n / w_z) - a_z;
//  - RSOC: both ends fixed -> ball; one end E fixed -> root of the
//    decreasing Phi + sum u^2 - 2 E (a_f + l E / w_f); neither -> the SOC
//    section after scaling the ends by sqrt(w) and rotating them by 45 deg;
//  - EXP / POWER: one free slot -> clamp to its interval; two free slots ->
//    {q >= g(p)} with g convex: the multiplier m of that constraint is the
//    root of the decreasing g(p(m)) - a_q - m / w_q, p(m) the root of the
//    increasing w_p (p - a_p) + m g'(p); zero fixed values leave boxes.
// Roots come from bracketed bisection that takes geometric midpoints while
// a bracket spans more than a factor 4 on one side of zero.  One thread per
// block, FP64, values normalised by the block's largest magnitude and
// weights by the largest free weight.  Blocks without fixed slots never get
// here.
#pragma once

#include <cfloat>

#include "dev_common.hpp"

namespace pdhcg {

constexpr int kMaxSectionSlots = 66;  // SOC/RSOC blocks of up to 64 vector slots
constexpr double kSqrtHalf = 0.70710678118654752440;

template <class F> __device__ double sec_root_increasing(F f, double lo, double hi) {
    for (int it = 0; it < 5000; ++it) {
        double mid;
        if (lo >= 0.0 && hi > 4.0 * lo) mid = lo > 0.0 ? sqrt(lo) * sqrt(hi) : hi * 0.0625;
        else if (hi <= 0.0 && lo < 4.0 * hi) mid = hi < 0.0 ? -(sqrt(-lo) * sqrt(-hi)) : lo * 0.0625;
        else mid = lo + 0.5 * (hi - lo);
        if (!(mid > lo && mid < hi)) break;
        if (f(mid) > 0.0) hi = mid;
        else lo = mid;
    }
    return lo + 0.5 * (hi - lo);
}

__device__ inline double sec_shrunk_norm2(const double *a, const double *w, const char *fx, int n, double l) {
    double s = 0.0;
    for (int i = 0; i < n; ++i)
        if (!fx[i]) {
            const double u = w[i] * a[i] / (w[i] + l);
            s += u * u;
        }
    return s;
}

__device__ inline void sec_ball(const double *a, const double *w, const char *fx, int n, double r2, double *out) {
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
    const double hi = sqrt(p2) / sqrt(r2) * (1.0 + 1e-12) + DBL_MIN;
    const double l = sec_root_increasing([&](double t) { return r2 - sec_shrunk_norm2(a, w, fx, n, t); }, 0.0, hi);
    for (int i = 0; i < n; ++i)
        if (!fx[i]) out[i] = w[i] * a[i] / (w[i] + l);
}

// Section of {||u|| <= z} with u = a[0..n) and z free (value az, weight wz);
// free u slots go to out, z is returned.
__device__ inline double sec_soc_free_z(const double *a, const double *w, const char *fx, int n, double az, double wz,
                                        double *out) {
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
    auto neg = [&](double l) { return az - sqrt(phi + sec_shrunk_norm2(a, w, fx, n, l)) * (1.0 - l / wz); };
    double l;
    if (az == 0.0) l = wz;
    else if (az > 0.0) l = sec_root_increasing(neg, 0.0, wz);
    else {
        double lo = wz, hi = 2.0 * wz;
        for (int it = 0; it < 2100 && neg(hi) <= 0.0; ++it) {
            lo = hi;
            hi *= 2.0;
        }
        l = sec_root_increasing(neg, lo, hi);
    }
    double s = phi;
    for (int i = 0; i < n; ++i)
        if (!fx[i]) {
            out[i] = w[i] * a[i] / (w[i] + l);
            s += out[i] * out[i];
        }
    return sqrt(s);
}

__device__ inline void sec_soc(const double *a, const double *w, const char *fx, int vd, double *out) {
    const int n = vd + 1;
    for (int i = 0; i <= n; ++i) out[i] = a[i];
    if (fx[n]) {
        double phi = 0.0;
        for (int i = 0; i < n; ++i)
            if (fx[i]) phi += a[i] * a[i];
        sec_ball(a, w, fx, n, fmax(0.0, a[n] * a[n] - phi), out);
    } else {
        out[n] = sec_soc_free_z(a, w, fx, n, a[n], w[n], out);
    }
}

// a and w are overwritten when neither end is fixed (rotated in place).
__device__ inline void sec_rsoc(double *a, double *w, const char *fx, int vd, double *out) {
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
        sec_ball(a, w, fx, vd, fmax(0.0, 2.0 * a[si] * a[ti] - phi), out);
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
        const double hi = wf * (phi + f2 - 2.0 * E * af) / (2.0 * E * E) * (1.0 + 1e-12) + DBL_MIN;
        const double l = sec_root_increasing(
            [&](double t) { return 2.0 * E * (af + t * E / wf) - (phi + sec_shrunk_norm2(a, w, fx, vd, t)); }, 0.0, hi);
        for (int i = 0; i < vd; ++i)
            if (!fx[i]) out[i] = w[i] * a[i] / (w[i] + l);
        out[f] = af + l * E / wf;
        return;
    }
    if (a[si] >= 0.0 && a[ti] >= 0.0 && phi + f2 <= 2.0 * a[si] * a[ti]) return;
    const double ss = sqrt(w[si]), st = sqrt(w[ti]), r = ss * st, sr = sqrt(r);
    // neither end is fixed here, so fx also marks the rotated (w, z) pair free
    const double sp = ss * a[si], tp = st * a[ti];
    for (int i = 0; i < vd; ++i) {
        a[i] = sr * a[i];
        w[i] = w[i] / r;
    }
    a[vd] = (sp - tp) * kSqrtHalf;
    a[vd + 1] = (sp + tp) * kSqrtHalf;
    w[vd] = w[vd + 1] = 1.0;
    for (int i = 0; i < vd + 2; ++i) out[i] = a[i];
    const double z = sec_soc_free_z(a, w, fx, vd + 1, a[vd + 1], 1.0, out);
    const double wr = out[vd];
    for (int i = 0; i < vd; ++i)
        if (!fx[i]) out[i] = out[i] / sr;
    out[si] = fmax(0.0, (z + wr) * kSqrtHalf) / ss;
    out[ti] = fmax(0.0, (z - wr) * kSqrtHalf) / st;
}

// ---- 2-D epigraph sections {q >= g(p)} of the 3-D cones ----
enum SecCurve { kCurveExpY, kCurveExpX, kCurveExpZ, kCurvePowZ, kCurvePowAxis };

struct SecCurveFn {
    int kind;
    double c, e;
    __device__ double g(double p) const {
        switch (kind) {
        case kCurveExpY: return c * exp(p / c);
        case kCurveExpX:
            if (p <= 0.0) return c <= 0.0 ? 0.0 : INFINITY;
            if (c / p < -745.0) return 0.0;
            return p * exp(c / p);
        case kCurveExpZ: return p > 0.0 ? p * log(p / c) : 0.0;
        case kCurvePowZ: return p > 0.0 ? c * pow(p, -e) : INFINITY;
        default: return pow(fabs(p) / c, e);
        }
    }
    __device__ double dg(double p) const {
        switch (kind) {
        case kCurveExpY: return exp(p / c);
        case kCurveExpX:
            if (p <= 0.0) return c < 0.0 ? 0.0 : (c == 0.0 ? 1.0 : -INFINITY);
            if (c / p < -700.0) return 0.0;
            return exp(c / p) * (1.0 - c / p);
        case kCurveExpZ: return p > 0.0 ? log(p / c) + 1.0 : -INFINITY;
        case kCurvePowZ: return p > 0.0 ? -e * c * pow(p, -e - 1.0) : -INFINITY;
        default: {
            const double d = e * pow(fabs(p) / c, e - 1.0) / c;
            return p < 0.0 ? -d : d;
        }
        }
    }
};

struct SecEpi {
    SecCurveFn cv;
    int bounded, closed;
    double lo, ap, aq, wp, wq;

    __device__ double p_of_mu(double mu) const {
        if (mu == 0.0) return bounded ? fmax(ap, lo) : ap;
        auto st = [&](double p) { return wp * (p - ap) + mu * cv.dg(p); };
        if (bounded && closed && st(lo) >= 0.0) return lo;
        double L, R;
        if (bounded) L = lo;
        else {
            L = fmin(ap, 0.0) - 1.0;
            for (int it = 0; it < 2100 && st(L) > 0.0; ++it) L -= fmax(1.0, fabs(L));
        }
        R = (bounded ? fmax(ap, lo) : ap) + 1.0;
        for (int it = 0; it < 2100 && st(R) < 0.0; ++it) R += fmax(1.0, fabs(R));
        return sec_root_increasing(st, L, R);
    }

    __device__ void project(double *po, double *qo) const {
        const double p0 = p_of_mu(0.0);
        if (aq >= cv.g(p0)) {
            *po = p0;
            *qo = aq;
            return;
        }
        auto neg = [&](double mu) { return aq + mu / wq - cv.g(p_of_mu(mu)); };
        double hi = wq * (1.0 + fabs(aq));
        for (int it = 0; it < 2100 && neg(hi) <= 0.0; ++it) hi *= 2.0;
        const double mu = sec_root_increasing(neg, 0.0, hi);
        const double p = p_of_mu(mu);
        *po = p;
        *qo = cv.g(p);
    }
};

__device__ inline void sec_epi(SecCurveFn cv, int bounded, double lo, int closed, double ap, double aq, double wp,
                               double wq, double *po, double *qo) {
    SecEpi e{cv, bounded, closed, lo, ap, aq, wp, wq};
    e.project(po, qo);
}

__device__ inline bool sec_exp_member(double x, double y, double z) {
    if (y > 0.0 && z > 0.0) {
        const double lhs = log(y) + x / y, rhs = log(z);
        return lhs <= rhs + 64.0 * DBL_EPSILON * (1.0 + fabs(lhs) + fabs(rhs));
    }
    return y == 0.0 && x <= 0.0 && z >= 0.0;
}

__device__ inline void sec_exp(const double *a, const double *w, const char *fx, double *out) {
    const double x = a[0], y = a[1], z = a[2];
    out[0] = x;
    out[1] = y;
    out[2] = z;
    if (sec_exp_member(x, y, z) || (fx[0] && fx[1] && fx[2])) return;
    if (fx[0] && fx[1]) {
        out[2] = fmax(z, y > 0.0 ? y * exp(x / y) : 0.0);
    } else if (fx[1] && fx[2]) {
        out[0] = fmin(x, y > 0.0 ? y * (log(z) - log(y)) : 0.0);
    } else if (fx[0] && fx[2]) {
        if (!(z > 0.0)) {
            out[1] = 0.0;
        } else {
            auto phi = [&](double t) { return log(t) + x / t - log(z); };
            double lo = 0.0, hi;
            double top = fmax(fmax(x, 0.0), 1.0) + z;
            for (int it = 0; it < 2100 && phi(top) <= 0.0; ++it) top *= 2.0;
            if (x > 0.0) {
                lo = sec_root_increasing([&](double t) { return -phi(t); }, 0.0, x);
                hi = sec_root_increasing(phi, x, top);
            } else {
                hi = sec_root_increasing(phi, 0.0, top);
            }
            out[1] = fmin(fmax(y, lo), hi);
        }
    } else if (fx[1]) {
        if (!(y > 0.0)) {
            out[0] = fmin(x, 0.0);
            out[2] = fmax(z, 0.0);
        } else {
            sec_epi(SecCurveFn{kCurveExpY, y, 0.0}, 0, 0.0, 0, x, z, w[0], w[2], &out[0], &out[2]);
        }
    } else if (fx[0]) {
        sec_epi(SecCurveFn{kCurveExpX, x, 0.0}, 1, 0.0, x <= 0.0, y, z, w[1], w[2], &out[1], &out[2]);
    } else if (fx[2]) {
        if (!(z > 0.0)) {
            out[0] = fmin(x, 0.0);
            out[1] = 0.0;
        } else {
            double q;
            sec_epi(SecCurveFn{kCurveExpZ, z, 0.0}, 1, 0.0, 1, y, -x, w[1], w[0], &out[1], &q);
            out[0] = -q;
        }
    }
}

__device__ inline bool sec_pow_member(double x, double y, double z, double al) {
    if (x < 0.0 || y < 0.0) return false;
    if (z == 0.0) return true;
    if (x == 0.0 || y == 0.0) return false;
    const double lhs = log(fabs(z)), rhs = al * log(x) + (1.0 - al) * log(y);
    return lhs <= rhs + 64.0 * DBL_EPSILON * (1.0 + fabs(lhs) + fabs(rhs));
}

__device__ inline void sec_pow(const double *a, const double *w, const char *fx, double al, double *out) {
    const double x = a[0], y = a[1], z = a[2];
    out[0] = x;
    out[1] = y;
    out[2] = z;
    if (sec_pow_member(x, y, z, al) || (fx[0] && fx[1] && fx[2])) return;
    if (fx[0] && fx[1]) {
        const double B = pow(x, al) * pow(y, 1.0 - al);
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
            sec_epi(SecCurveFn{kCurvePowZ, pow(fabs(z), 1.0 / (1.0 - al)), al / (1.0 - al)}, 1, 0.0, 0, x, y, w[0],
                    w[1], &out[0], &out[1]);
        }
    } else if (fx[0]) {
        if (!(x > 0.0)) {
            out[2] = 0.0;
            out[1] = fmax(y, 0.0);
        } else {
            sec_epi(SecCurveFn{kCurvePowAxis, pow(x, al), 1.0 / (1.0 - al)}, 0, 0.0, 0, z, y, w[2], w[1], &out[2],
                    &out[1]);
        }
    } else if (fx[1]) {
        if (!(y > 0.0)) {
            out[2] = 0.0;
            out[0] = fmax(x, 0.0);
        } else {
            sec_epi(SecCurveFn{kCurvePowAxis, pow(y, 1.0 - al), 1.0 / al}, 0, 0.0, 0, z, x, w[2], w[0], &out[2],
                    &out[0]);
        }
    }
}

// Projects the block's slots p[0..len) (scaled space, d the scaling) onto
// the section in the metric (1 + tau q) d^2 (q != null) or d^2; fx = the
// block's fixed flags.
template <class X>
__device__ inline void sec_project_block(double *p, int type, int vd, const double *d, const X *q, double tau,
                                         const char *fx, double alpha) {
    const int len = (type == CONE_STANDARD_SOC || type == CONE_ROTATED_SOC) ? vd + 2 : 3;
    double a[kMaxSectionSlots], w[kMaxSectionSlots], o[kMaxSectionSlots];
    double s = 0.0, wmax = 0.0;
    for (int i = 0; i < len; ++i) {
        a[i] = p[i] / d[i];
        double m = 1.0;
        if (q) {
            m = 1.0 + tau * (double)q[i];
            if (!(m > 0.0)) m = 1.0;
        }
        w[i] = fmax(m * d[i] * d[i], DBL_MIN);
        s = fmax(s, fabs(a[i]));
        if (!fx[i]) wmax = fmax(wmax, w[i]);
    }
    if (!(s > 0.0) || !isfinite(s)) return;
    for (int i = 0; i < len; ++i) {
        a[i] /= s;
        if (wmax > 0.0) w[i] /= wmax;
    }
    switch (type) {
    case CONE_STANDARD_SOC: sec_soc(a, w, fx, vd, o); break;
    case CONE_ROTATED_SOC: sec_rsoc(a, w, fx, vd, o); break;
    case CONE_EXPONENTIAL: sec_exp(a, w, fx, o); break;
    default: sec_pow(a, w, fx, alpha, o); break;
    }
    for (int i = 0; i < len; ++i)
        if (!fx[i]) p[i] = o[i] * s * d[i];
}

// The blocks with fixed slots: start, v_dim, power exponent and cone type.
struct SectionList {
    int count = 0;
    int *start = nullptr, *vdim = nullptr, *type = nullptr;
    double *alpha = nullptr;
};

__device__ inline int sec_block_len(int type, int vd) {
    return (type == CONE_STANDARD_SOC || type == CONE_ROTATED_SOC) ? vd + 2 : 3;
}

// v <- section projection of every listed block; q != null gives the
// diagonal-Q metric 1 + tau q (non-positive metric entries read as 1).
template <class X>
__global__ void __launch_bounds__(kBlock) section_project_kernel(X *v, const double *d, const X *q, double tau,
                                                                  const char *fixed, SectionList S, const int *gate) {
    if (gate && *gate == 0) return;
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= S.count) return;
    const int st = S.start[b], t = S.type[b], vd = S.vdim[b], len = sec_block_len(t, vd);
    double p[kMaxSectionSlots];
    for (int i = 0; i < len; ++i) p[i] = (double)v[st + i];
    sec_project_block<X>(p, t, vd, d + st, q ? q + st : nullptr, tau, fixed + st, S.alpha[b]);
    for (int i = 0; i < len; ++i)
        if (!fixed[st + i]) v[st + i] = (X)p[i];
}

// Residual of the listed blocks (reference
// src/kernels/pdhcg_soc_cone_kernels.cu:489-507 and its siblings): with
// rc = c + qx - aty, dres = (x - P(x - rc)) d on free slots, 0 on fixed ones,
// P the section projection in the metric d^2.
template <class X>
__global__ void __launch_bounds__(kBlock) section_residual_kernel(double *dres, const X *x, const X *c, const X *qx,
                                                                   const X *aty, const double *d, const char *fixed,
                                                                   SectionList S) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= S.count) return;
    const int st = S.start[b], t = S.type[b], vd = S.vdim[b], len = sec_block_len(t, vd);
    double p[kMaxSectionSlots];
    for (int i = 0; i < len; ++i) {
        const double xi = (double)x[st + i];
        const double rc = (double)c[st + i] + (qx ? (double)qx[st + i] : 0.0) - (double)aty[st + i];
        p[i] = fixed[st + i] ? xi : xi - rc;
    }
    sec_project_block<X>(p, t, vd, d + st, nullptr, 0.0, fixed + st, S.alpha[b]);
    for (int i = 0; i < len; ++i) dres[st + i] = fixed[st + i] ? 0.0 : ((double)x[st + i] - p[i]) * d[st + i];
}

// ---- long SOC / RSOC sections: one workgroup per block ----
//
// Blocks above 64 vector slots with fixed slots run sec_soc / sec_rsoc with
// the whole workgroup: the per-slot arrays live in a global scratch area,
// every sum is a fixed-order workgroup reduction broadcast to all threads
// (so each root finder takes the same steps in every thread), and barriers
// separate a single thread's writes from the reads that follow them.

// Workgroup sum and maximum; every thread receives the result.
struct SecWg {
    double lds[16];
    double bc;
    __device__ double sum(double v) {
        double t[1] = {v};
        block_reduce<1, 0>(t, lds);
        if (threadIdx.x == 0) bc = t[0];
        __syncthreads();
        const double r = bc;
        __syncthreads();
        return r;
    }
    __device__ double max(double v) {
        double t[1] = {v};
        block_reduce<0, 1>(t, lds);
        if (threadIdx.x == 0) bc = t[0];
        __syncthreads();
        const double r = bc;
        __syncthreads();
        return r;
    }
};

__device__ inline double secw_shrunk_norm2(const double *a, const double *w, const char *fx, int n, double l,
                                           SecWg &g) {
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x)
        if (!fx[i]) {
            const double u = w[i] * a[i] / (w[i] + l);
            s += u * u;
        }
    return g.sum(s);
}

__device__ inline void secw_ball(const double *a, const double *w, const char *fx, int n, double r2, double *out,
                                 SecWg &g) {
    double f2 = 0.0, p2 = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x)
        if (!fx[i]) {
            f2 += a[i] * a[i];
            p2 += (w[i] * a[i]) * (w[i] * a[i]);
            out[i] = a[i];
        }
    f2 = g.sum(f2);
    p2 = g.sum(p2);
    if (f2 <= r2) return;
    if (!(r2 > 0.0)) {
        for (int i = threadIdx.x; i < n; i += blockDim.x)
            if (!fx[i]) out[i] = 0.0;
        __syncthreads();
        return;
    }
    const double hi = sqrt(p2) / sqrt(r2) * (1.0 + 1e-12) + DBL_MIN;
    const double l =
        sec_root_increasing([&](double t) { return r2 - secw_shrunk_norm2(a, w, fx, n, t, g); }, 0.0, hi);
    for (int i = threadIdx.x; i < n; i += blockDim.x)
        if (!fx[i]) out[i] = w[i] * a[i] / (w[i] + l);
    __syncthreads();
}

// sec_soc_free_z by the workgroup; out's free slots are written by the thread
// that owns them (i = threadIdx.x mod blockDim.x), the returned z is uniform.
__device__ inline double secw_soc_free_z(const double *a, const double *w, const char *fx, int n, double az,
                                         double wz, double *out, SecWg &g) {
    double phi = 0.0, f2 = 0.0, p2 = 0.0, nf = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        if (fx[i]) phi += a[i] * a[i];
        else {
            f2 += a[i] * a[i];
            p2 += (w[i] * a[i]) * (w[i] * a[i]);
            out[i] = a[i];
            nf += 1.0;
        }
    }
    phi = g.sum(phi);
    f2 = g.sum(f2);
    p2 = g.sum(p2);
    nf = g.sum(nf);
    if (az >= 0.0 && phi + f2 <= az * az) return az;
    if (nf == 0.0) return fmax(az, sqrt(phi));
    if (phi == 0.0 && -wz * az >= sqrt(p2)) {
        for (int i = threadIdx.x; i < n; i += blockDim.x)
            if (!fx[i]) out[i] = 0.0;
        __syncthreads();
        return 0.0;
    }
    auto neg = [&](double l) { return az - sqrt(phi + secw_shrunk_norm2(a, w, fx, n, l, g)) * (1.0 - l / wz); };
    double l;
    if (az == 0.0) l = wz;
    else if (az > 0.0) l = sec_root_increasing(neg, 0.0, wz);
    else {
        double lo = wz, hi = 2.0 * wz;
        for (int it = 0; it < 2100 && neg(hi) <= 0.0; ++it) {
            lo = hi;
            hi *= 2.0;
        }
        l = sec_root_increasing(neg, lo, hi);
    }
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x)
        if (!fx[i]) {
            out[i] = w[i] * a[i] / (w[i] + l);
            s += out[i] * out[i];
        }
    return sqrt(phi + g.sum(s));
}

__device__ inline void secw_soc(const double *a, const double *w, const char *fx, int vd, double *out, SecWg &g) {
    const int n = vd + 1;
    for (int i = threadIdx.x; i <= n; i += blockDim.x) out[i] = a[i];
    if (fx[n]) {
        double phi = 0.0;
        for (int i = threadIdx.x; i < n; i += blockDim.x)
            if (fx[i]) phi += a[i] * a[i];
        phi = g.sum(phi);
        secw_ball(a, w, fx, n, fmax(0.0, a[n] * a[n] - phi), out, g);
    } else {
        const double z = secw_soc_free_z(a, w, fx, n, a[n], w[n], out, g);
        __syncthreads();
        if (threadIdx.x == 0) out[n] = z;
    }
    __syncthreads();
}

// a and w are overwritten when neither end is fixed (rotated in place).
__device__ inline void secw_rsoc(double *a, double *w, const char *fx, int vd, double *out, SecWg &g) {
    const int si = vd, ti = vd + 1;
    for (int i = threadIdx.x; i < vd + 2; i += blockDim.x) out[i] = a[i];
    double phi = 0.0, f2 = 0.0, nf = 0.0;
    for (int i = threadIdx.x; i < vd; i += blockDim.x) {
        if (fx[i]) phi += a[i] * a[i];
        else {
            f2 += a[i] * a[i];
            nf += 1.0;
        }
    }
    phi = g.sum(phi);
    f2 = g.sum(f2);
    nf = g.sum(nf);
    if (fx[si] && fx[ti]) {
        secw_ball(a, w, fx, vd, fmax(0.0, 2.0 * a[si] * a[ti] - phi), out, g);
        return;
    }
    if (fx[si] || fx[ti]) {
        const int e = fx[si] ? si : ti, f = fx[si] ? ti : si;
        const double E = a[e], af = a[f], wf = w[f];
        if (!(E > 0.0)) {
            for (int i = threadIdx.x; i < vd; i += blockDim.x)
                if (!fx[i]) out[i] = 0.0;
            __syncthreads();
            if (threadIdx.x == 0) out[f] = fmax(af, 0.0);
            __syncthreads();
            return;
        }
        if (af >= 0.0 && phi + f2 <= 2.0 * E * af) return;
        if (nf == 0.0) {
            if (threadIdx.x == 0) out[f] = fmax(af, phi / (2.0 * E));
            __syncthreads();
            return;
        }
        const double hi = wf * (phi + f2 - 2.0 * E * af) / (2.0 * E * E) * (1.0 + 1e-12) + DBL_MIN;
        const double l = sec_root_increasing(
            [&](double t) { return 2.0 * E * (af + t * E / wf) - (phi + secw_shrunk_norm2(a, w, fx, vd, t, g)); },
            0.0, hi);
        for (int i = threadIdx.x; i < vd; i += blockDim.x)
            if (!fx[i]) out[i] = w[i] * a[i] / (w[i] + l);
        __syncthreads();
        if (threadIdx.x == 0) out[f] = af + l * E / wf;
        __syncthreads();
        return;
    }
    if (a[si] >= 0.0 && a[ti] >= 0.0 && phi + f2 <= 2.0 * a[si] * a[ti]) return;
    const double ss = sqrt(w[si]), st = sqrt(w[ti]), r = ss * st, sr = sqrt(r);
    const double sp = ss * a[si], tp = st * a[ti];
    __syncthreads();  // every thread has read a[si], a[ti], w[si], w[ti]
    for (int i = threadIdx.x; i < vd; i += blockDim.x) {
        a[i] = sr * a[i];
        w[i] = w[i] / r;
    }
    if (threadIdx.x == 0) {
        a[vd] = (sp - tp) * kSqrtHalf;
        a[vd + 1] = (sp + tp) * kSqrtHalf;
        w[vd] = w[vd + 1] = 1.0;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < vd + 2; i += blockDim.x) out[i] = a[i];
    const double z = secw_soc_free_z(a, w, fx, vd + 1, a[vd + 1], 1.0, out, g);
    __syncthreads();
    const double wr = out[vd];
    for (int i = threadIdx.x; i < vd; i += blockDim.x)
        if (!fx[i]) out[i] = out[i] / sr;
    __syncthreads();
    if (threadIdx.x == 0) {
        out[si] = fmax(0.0, (z + wr) * kSqrtHalf) / ss;
        out[ti] = fmax(0.0, (z - wr) * kSqrtHalf) / st;
    }
    __syncthreads();
}

// sec_project_block for a long SOC/RSOC block by the workgroup: p (global,
// vd + 2 values) is projected in place on its free slots; a, w, o are
// scratch areas of vd + 2 values each.
template <class X>
__device__ void secw_project_block(double *p, int type, int vd, const double *d, const X *q, double tau,
                                   const char *fx, double *a, double *w, double *o, SecWg &g) {
    const int len = vd + 2;
    double smax = 0.0, wmax = 0.0;
    for (int i = threadIdx.x; i < len; i += blockDim.x) {
        a[i] = p[i] / d[i];
        double m = 1.0;
        if (q) {
            m = 1.0 + tau * (double)q[i];
            if (!(m > 0.0)) m = 1.0;
        }
        w[i] = fmax(m * d[i] * d[i], DBL_MIN);
        smax = fmax(smax, fabs(a[i]));
        if (!fx[i]) wmax = fmax(wmax, w[i]);
    }
    smax = g.max(smax);
    wmax = g.max(wmax);
    if (!(smax > 0.0) || !isfinite(smax)) return;
    for (int i = threadIdx.x; i < len; i += blockDim.x) {
        a[i] /= smax;
        if (wmax > 0.0) w[i] /= wmax;
    }
    __syncthreads();
    if (type == CONE_STANDARD_SOC) secw_soc(a, w, fx, vd, o, g);
    else secw_rsoc(a, w, fx, vd, o, g);
    __syncthreads();
    for (int i = threadIdx.x; i < len; i += blockDim.x)
        if (!fx[i]) p[i] = o[i] * smax * d[i];
    __syncthreads();
}

// section_project_kernel for long blocks: one workgroup per block, scratch
// holds 4 (vd + 2) values per block from off[b].
template <class X>
__global__ void __launch_bounds__(kBlock) section_long_project_kernel(X *v, const double *d, const X *q, double tau,
                                                                       const char *fixed, SectionList S,
                                                                       double *scratch, const long long *off,
                                                                       const int *gate) {
    if (gate && *gate == 0) return;
    __shared__ SecWg g;
    const int b = blockIdx.x;
    const int st = S.start[b], t = S.type[b], vd = S.vdim[b], len = vd + 2;
    double *p = scratch + off[b], *a = p + len, *w = a + len, *o = w + len;
    for (int i = threadIdx.x; i < len; i += blockDim.x) p[i] = (double)v[st + i];
    __syncthreads();
    secw_project_block<X>(p, t, vd, d + st, q ? q + st : nullptr, tau, fixed + st, a, w, o, g);
    for (int i = threadIdx.x; i < len; i += blockDim.x)
        if (!fixed[st + i]) v[st + i] = (X)p[i];
}

// section_residual_kernel for long blocks (layout as section_long_project_kernel).
template <class X>
__global__ void __launch_bounds__(kBlock) section_long_residual_kernel(double *dres, const X *x, const X *c,
                                                                        const X *qx, const X *aty, const double *d,
                                                                        const char *fixed, SectionList S,
                                                                        double *scratch, const long long *off) {
    __shared__ SecWg g;
    const int b = blockIdx.x;
    const int st = S.start[b], t = S.type[b], vd = S.vdim[b], len = vd + 2;
    double *p = scratch + off[b], *a = p + len, *w = a + len, *o = w + len;
    for (int i = threadIdx.x; i < len; i += blockDim.x) {
        const double xi = (double)x[st + i];
        const double rc = (double)c[st + i] + (qx ? (double)qx[st + i] : 0.0) - (double)aty[st + i];
        p[i] = fixed[st + i] ? xi : xi - rc;
    }
    __syncthreads();
    secw_project_block<X>(p, t, vd, d + st, (const X *)nullptr, 0.0, fixed + st, a, w, o, g);
    for (int i = threadIdx.x; i < len; i += blockDim.x)
        dres[st + i] = fixed[st + i] ? 0.0 : ((double)x[st + i] - p[i]) * d[st + i];
}

}  // namespace pdhcg
