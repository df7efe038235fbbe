// Cone projections and dual-cone residuals on the device (gfx950), applied
// in the scaled space: the projection is onto D K where d is the block's
// variable scaling.  Algorithms follow the reference kernels
//   SOC  reference src/kernels/pdhcg_soc_cone_kernels.cu:267-469 (projection),
//        :471-726 (dual residual): closed form for uniform block scaling,
//        bisection on the multiplier otherwise (60 steps, width 1e-13, warm start)
//   RSOC reference src/kernels/pdhcg_rsoc_cone_kernels.cu:466-945 (rotated form)
//   EXP  reference src/kernels/pdhcg_exp_cone_kernels.cu:26-131,:526-614
//        (Newton on rho from 0, steps clamped to +-10, and its fallback)
//   POW  reference src/kernels/pdhcg_power_cone_kernels.cu:25-132,:435-669
//        (80-step bisection in log rho after homogeneity normalisation)
//   PSD  reference src/kernels/pdhcg_psd_cone_kernels.cu:65-136 (svec <->
//        dense, eigen-decomposition, positive part); the eigen-decomposition
//        here is a cyclic two-sided Jacobi run by one workgroup per matrix.
// All cone arithmetic is FP64, whatever the storage type X of the vectors.
#pragma once

#include <cfloat>

#include "dev_common.hpp"

namespace pdhcg {

constexpr double kInvSqrt2 = 0.7071067811865475;
constexpr double kSqrt2 = 1.41421356237309504880;

// One list of cone blocks of a single type on the device.
struct ConeList {
    int count = 0;
    int *start = nullptr;   // first slot of each block
    int *vdim = nullptr;    // SOC/RSOC vector dimension, PSD order
    double *alpha = nullptr;// power exponent
    double *warm = nullptr; // projection warm starts (one per block)
    double *rwarm = nullptr;// dual-residual warm starts
};

// ---------------- second-order cones ----------------

template <class Get>
__device__ double soc_multiplier_f(Get dh, const double *pv, int k, double w, double dw, double z, double lam) {
    double sum = 0.0;
    for (int m = 0; m < k; ++m) {
        const double h = dh(m);
        const double t = pv[m] * h / (h * h + 2.0 * lam);
        sum += t * t;
    }
    const double tw = w * dw / (dw * dw + 2.0 * lam);
    sum += tw * tw;
    const double zt = z / (1.0 - 2.0 * lam);
    return sum - zt * zt;
}

// Bisection on the multiplier of the weighted SOC projection.
template <class Get>
__device__ double soc_multiplier(Get dh, const double *pv, int k, double w, double dw, double z, double *warm) {
    double lo, hi;
    const bool zpos = z > 0.0;
    if (zpos) {
        lo = 0.0;
        hi = 0.5 - 1e-14;
    } else {
        lo = 0.5 + 1e-14;
        hi = 1.0;
        for (int it = 0; it < 60; ++it) {
            if (soc_multiplier_f(dh, pv, k, w, dw, z, hi) > 0.0) break;
            lo = hi;
            hi *= 2.0;
        }
    }
    const double wl = *warm;
    if (wl > lo && wl < hi) {
        const double f = soc_multiplier_f(dh, pv, k, w, dw, z, wl);
        if (fabs(f) < 1e-12) return wl;
        if (zpos == (f > 0.0)) lo = wl;
        else hi = wl;
    }
    for (int it = 0; it < 60; ++it) {
        const double lam = 0.5 * (lo + hi);
        const double f = soc_multiplier_f(dh, pv, k, w, dw, z, lam);
        if (zpos == (f > 0.0)) lo = lam;
        else hi = lam;
        if ((hi - lo) / (1.0 + hi + lo) < 1e-13) break;
    }
    const double lam = 0.5 * (lo + hi);
    *warm = lam;
    return lam;
}

// Projects slots (v[0..k), a, b) held in FP64 registers/scratch `pv` (length
// k + 2) in place.  rotated: (a, b) = (s, t), else (w, z).  d: block scaling.
__device__ inline void soc_project_block(double *pv, int k, const double *d, bool rotated, double *warm) {
    const double a = pv[k], b = pv[k + 1];
    const double w = rotated ? (a - b) * kInvSqrt2 : a;
    const double z = rotated ? (a + b) * kInvSqrt2 : b;
    const double dref = rotated ? sqrt(d[k] * d[k + 1]) : d[k + 1];
    const double dw = rotated ? 1.0 : d[k] / dref;
    bool uniform = rotated || dw == 1.0;
    for (int m = 0; m < k && uniform; ++m) uniform = d[m] == dref;
    double wn, zn;
    if (uniform) {
        double ss = w * w;
        for (int m = 0; m < k; ++m) ss += pv[m] * pv[m];
        const double r = sqrt(ss);
        if (r <= z) return;
        if (r <= -z) {
            for (int m = 0; m < k + 2; ++m) pv[m] = 0.0;
            return;
        }
        const double sc = (z + r) / (2.0 * r);
        for (int m = 0; m < k; ++m) pv[m] *= sc;
        wn = sc * w;
        zn = sc * r;
    } else {
        auto dh = [&](int m) { return d[m] / dref; };
        double ri = (w / dw) * (w / dw), rp = (w * dw) * (w * dw);
        for (int m = 0; m < k; ++m) {
            const double h = dh(m);
            ri += (pv[m] / h) * (pv[m] / h);
            rp += (pv[m] * h) * (pv[m] * h);
        }
        if (sqrt(ri) <= z) return;
        if (sqrt(rp) <= -z) {
            for (int m = 0; m < k + 2; ++m) pv[m] = 0.0;
            return;
        }
        const double lam = soc_multiplier(dh, pv, k, w, dw, z, warm);
        for (int m = 0; m < k; ++m) {
            const double h = dh(m);
            pv[m] = pv[m] * h * h / (h * h + 2.0 * lam);
        }
        wn = rotated ? w / (1.0 + 2.0 * lam) : w * dw * dw / (dw * dw + 2.0 * lam);
        zn = z / (1.0 - 2.0 * lam);
    }
    if (rotated) {
        pv[k] = (zn + wn) * kInvSqrt2;
        pv[k + 1] = (zn - wn) * kInvSqrt2;
    } else {
        pv[k] = wn;
        pv[k + 1] = zn;
    }
}

// Dual residual (rc - P_{K*}(rc)) d of a SOC/RSOC block; rc in `pr`
// (length k + 2), result written to `out` (FP64).
__device__ inline void soc_dual_residual_block(const double *pr, int k, const double *d, bool rotated, double *warm,
                                        double *out) {
    const double a = pr[k], b = pr[k + 1];
    const double rw = rotated ? (a - b) * kInvSqrt2 : a;
    const double rz = rotated ? (a + b) * kInvSqrt2 : b;
    const double dref = rotated ? sqrt(d[k] * d[k + 1]) : d[k + 1];
    const double ew = rotated ? 1.0 : dref / d[k];
    bool uniform = rotated || ew == 1.0;
    for (int m = 0; m < k && uniform; ++m) uniform = d[m] == dref;
    double pa, pb;  // projected (a, b) slots
    if (uniform) {
        double ss = rw * rw;
        for (int m = 0; m < k; ++m) ss += pr[m] * pr[m];
        const double r = sqrt(ss);
        double vf, pw, pz;
        if (r <= rz) {
            vf = 0.0;
            pw = rw;
            pz = rz;
        } else if (r <= -rz) {
            vf = 1.0;
            pw = 0.0;
            pz = 0.0;
        } else {
            const double sc = (rz + r) / (2.0 * r);
            vf = 1.0 - sc;
            pw = sc * rw;
            pz = sc * r;
        }
        for (int m = 0; m < k; ++m) out[m] = pr[m] * vf * d[m];
        if (rotated) {
            if (r <= rz) {
                pa = a;
                pb = b;
            } else {
                pa = (pz + pw) * kInvSqrt2;
                pb = (pz - pw) * kInvSqrt2;
            }
        } else {
            pa = pw;
            pb = pz;
        }
        out[k] = (a - pa) * d[k];
        out[k + 1] = (b - pb) * d[k + 1];
        return;
    }
    auto e = [&](int m) { return dref / d[m]; };
    double ri = (rw / ew) * (rw / ew), rp = (rw * ew) * (rw * ew);
    for (int m = 0; m < k; ++m) {
        const double h = e(m);
        ri += (pr[m] / h) * (pr[m] / h);
        rp += (pr[m] * h) * (pr[m] * h);
    }
    if (sqrt(ri) <= rz) {
        for (int m = 0; m < k + 2; ++m) out[m] = 0.0;
        return;
    }
    if (sqrt(rp) <= -rz) {
        for (int m = 0; m < k + 2; ++m) out[m] = pr[m] * d[m];
        return;
    }
    const double lam = soc_multiplier(e, pr, k, rw, ew, rz, warm);
    for (int m = 0; m < k; ++m) {
        const double h = e(m);
        out[m] = (pr[m] - pr[m] * h * h / (h * h + 2.0 * lam)) * d[m];
    }
    const double zn = rz / (1.0 - 2.0 * lam);
    const double wn = rotated ? rw / (1.0 + 2.0 * lam) : rw * ew * ew / (ew * ew + 2.0 * lam);
    if (rotated) {
        pa = (zn + wn) * kInvSqrt2;
        pb = (zn - wn) * kInvSqrt2;
    } else {
        pa = wn;
        pb = zn;
    }
    out[k] = (a - pa) * d[k];
    out[k + 1] = (b - pb) * d[k + 1];
}

// ---------------- exponential cone ----------------

// Multiply-adds are not contracted here: the Newton iteration from rho = 0
// can settle on different roots when the last bit of a step differs, and the
// host restatement of this routine is compiled without contraction too.
__device__ inline void exp_project_point(double r1, double r2, double r3, double d1, double d2, double d3, double *o) {
#pragma clang fp contract(off)
    const double E = 2.718281828459045;
    const double a1 = r1 / d1, a2 = r2 / d2, a3 = r3 / d3;
    if (a2 > 0.0) {
        const double ratio = a1 / a2;
        if (ratio < 700.0 && a2 * exp(ratio) <= a3) {
            o[0] = r1; o[1] = r2; o[2] = r3;
            return;
        }
    } else if (a2 == 0.0 && a1 <= 0.0 && a3 >= 0.0) {
        o[0] = r1; o[1] = r2; o[2] = r3;
        return;
    }
    if (r1 > 0.0) {
        const double ratio = (d2 * r2) / (d1 * r1);
        if (ratio < 700.0 && d1 * r1 * exp(ratio) + E * d3 * r3 <= 0.0) {
            o[0] = o[1] = o[2] = 0.0;
            return;
        }
    } else if (r1 == 0.0 && r2 <= 0.0 && r3 <= 0.0) {
        o[0] = o[1] = o[2] = 0.0;
        return;
    }
    if (a1 <= 0.0 && a2 <= 0.0) {
        o[0] = r1;
        o[1] = 0.0;
        o[2] = a3 < 0.0 ? 0.0 : r3;
        return;
    }
    const double al = (d3 / d1) * (d3 / d1), be = (d3 / d2) * (d3 / d2);
    double rho = 0.0;
    bool diverged = false;
    for (int it = 0; it < 100; ++it) {
        if (rho >= 299.0 || rho <= -299.0) {
            diverged = true;
            break;
        }
        const double er = exp(rho), e2 = er * er, om = 1.0 - rho;
        const double at = al - be * rho * om;
        const double bt = be * a1 * om - al * a2;
        const double f = a1 - a2 * rho + a3 * er * at + e2 * bt;
        const double df = -a2 + a3 * er * (at + (-be * (1.0 - 2.0 * rho))) + e2 * (2.0 * bt + (-be * a1));
        if (fabs(df) < 1e-300) break;
        double st = f / df;
        st = fmin(10.0, fmax(-10.0, st));
        rho -= st;
        if (fabs(st) < 1e-13 * (1.0 + fabs(rho))) break;
    }
    const double er = exp(rho);
    const double u2 = (a1 + al * a3 * er) / (rho + al * er * er);
    if (diverged || !isfinite(u2) || u2 <= 0.0) {
        o[0] = r1 < 0.0 ? r1 : 0.0;
        o[1] = 0.0;
        o[2] = r3 > 0.0 ? r3 : 0.0;
        return;
    }
    o[0] = d1 * (rho * u2);
    o[1] = d2 * u2;
    o[2] = d3 * (u2 * er);
}

// ---------------- power cone ----------------

__device__ __forceinline__ double positive_root(double r, double q) {
    if (!(q > 0.0)) return fmax(r, 0.0);
    const double disc = hypot(r, 2.0 * sqrt(q));
    return r >= 0.0 ? 0.5 * (r + disc) : (2.0 * q) / (disc - r);
}

__device__ inline void power_project_normalized(double rx, double ry, double rz, double wx, double wy, double wz, double al,
                                         double *o) {
    const double az = fabs(rz), sg = rz >= 0.0 ? 1.0 : -1.0, om = 1.0 - al;
    if (az == 0.0) {
        o[0] = fmax(rx, 0.0); o[1] = fmax(ry, 0.0); o[2] = 0.0;
        return;
    }
    if (rx > 0.0 && ry > 0.0 && al * log(rx) + om * log(ry) >= log(az)) {
        o[0] = rx; o[1] = ry; o[2] = rz;
        return;
    }
    if (rx <= 0.0 && ry <= 0.0) {
        const double u = rx < 0.0 ? (-wx * rx) / al : 0.0;
        const double v = ry < 0.0 ? (-wy * ry) / om : 0.0;
        if (u > 0.0 && v > 0.0 && al * log(u) + om * log(v) >= log(wz) + log(az)) {
            o[0] = o[1] = o[2] = 0.0;
            return;
        }
    }
    const double cx = 4.0 * al * (wz / wx), cy = 4.0 * om * (wz / wy);
    double lo = log(DBL_MIN), hi = log(az);
    if (!(hi > lo)) {
        o[0] = fmax(rx, 0.0); o[1] = fmax(ry, 0.0); o[2] = 0.0;
        return;
    }
    for (int it = 0; it < 80; ++it) {
        const double lr = lo + 0.5 * (hi - lo), rho = exp(lr);
        const double x = positive_root(rx, 0.25 * cx * rho * (az - rho));
        const double y = positive_root(ry, 0.25 * cy * rho * (az - rho));
        if (x > 0.0 && y > 0.0 && al * log(x) + om * log(y) > lr) lo = lr;
        else hi = lr;
    }
    const double rho = exp(lo + 0.5 * (hi - lo));
    o[0] = positive_root(rx, 0.25 * cx * rho * (az - rho));
    o[1] = positive_root(ry, 0.25 * cy * rho * (az - rho));
    o[2] = sg * fmin(rho, exp(al * log(o[0]) + om * log(o[1])));
}

__device__ inline void power_project_point(double rx, double ry, double rz, double wx, double wy, double wz, double al,
                                    double *o) {
    const double sc = fmax(fabs(rx), fmax(fabs(ry), fabs(rz)));
    if (!(sc > 0.0) || !isfinite(sc)) {
        power_project_normalized(rx, ry, rz, wx, wy, wz, al, o);
        return;
    }
    power_project_normalized(rx / sc, ry / sc, rz / sc, wx, wy, wz, al, o);
    o[0] *= sc; o[1] *= sc; o[2] *= sc;
}

// ---------------- per-block kernels (one thread per block) ----------------

constexpr int kMaxThreadCone = 64;  // longest SOC/RSOC block handled in registers/stack

// type: 0 rotated, 1 standard SOC, 2 exponential, 3 power (cone_type_t values)
template <class X, int Type>
__global__ void __launch_bounds__(kBlock) cone_project_kernel(X *v, const double *d, ConeList L, int use_rwarm,
                                                               const int *gate) {
    if (gate && *gate == 0) return;
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= L.count) return;
    const int s = L.start[b];
    double *warm = use_rwarm ? L.rwarm + b : L.warm + b;
    if (Type == CONE_STANDARD_SOC || Type == CONE_ROTATED_SOC) {
        const int k = L.vdim[b];
        double pv[kMaxThreadCone + 2];
        for (int m = 0; m < k + 2; ++m) pv[m] = (double)v[s + m];
        soc_project_block(pv, k, d + s, Type == CONE_ROTATED_SOC, warm);
        for (int m = 0; m < k + 2; ++m) v[s + m] = (X)pv[m];
    } else if (Type == CONE_EXPONENTIAL) {
        double o[3];
        exp_project_point((double)v[s], (double)v[s + 1], (double)v[s + 2], d[s], d[s + 1], d[s + 2], o);
        v[s] = (X)o[0]; v[s + 1] = (X)o[1]; v[s + 2] = (X)o[2];
    } else {
        double o[3];
        const double d0 = d[s], d1 = d[s + 1], d2 = d[s + 2];
        power_project_point((double)v[s] / d0, (double)v[s + 1] / d1, (double)v[s + 2] / d2, d0 * d0, d1 * d1, d2 * d2,
                            L.alpha[b], o);
        v[s] = (X)(o[0] * d0); v[s + 1] = (X)(o[1] * d1); v[s + 2] = (X)(o[2] * d2);
    }
}

// Cone step of the diagonal-Q proximal update: the projection onto D K in
// the metric diag(w), w = 1 + tau q, computed as W^-1/2 P_{W^1/2 D K}(W^1/2 v)
// (reference src/cone_dispatch.cu:59-85 with the *_diag_q kernels, e.g.
// pdhcg_exp_cone_kernels.cu:624-700).  Not used for PSD blocks.
template <class X, int Type>
__global__ void __launch_bounds__(kBlock) cone_project_diag_q_kernel(X *v, const double *d, const X *q, double tau,
                                                                      ConeList L) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= L.count) return;
    const int s = L.start[b];
    const bool soc = Type == CONE_STANDARD_SOC || Type == CONE_ROTATED_SOC;
    const int len = soc ? L.vdim[b] + 2 : 3;
    auto root_w = [&](int m) {
        const double w = 1.0 + tau * (double)q[s + m];
        return sqrt(w > 0.0 ? w : 1.0);
    };
    double pv[kMaxThreadCone + 2], e[kMaxThreadCone + 2];
    for (int m = 0; m < len; ++m) {
        const double sw = root_w(m);
        pv[m] = (double)v[s + m] * sw;
        e[m] = sw * d[s + m];
    }
    if (soc) {
        soc_project_block(pv, len - 2, e, Type == CONE_ROTATED_SOC, L.warm + b);
    } else if (Type == CONE_EXPONENTIAL) {
        double o[3];
        exp_project_point(pv[0], pv[1], pv[2], e[0], e[1], e[2], o);
        pv[0] = o[0]; pv[1] = o[1]; pv[2] = o[2];
    } else {
        double o[3];
        power_project_point(pv[0] / e[0], pv[1] / e[1], pv[2] / e[2], e[0] * e[0], e[1] * e[1], e[2] * e[2],
                            L.alpha[b], o);
        pv[0] = o[0] * e[0]; pv[1] = o[1] * e[1]; pv[2] = o[2] * e[2];
    }
    for (int m = 0; m < len; ++m) v[s + m] = (X)(pv[m] / root_w(m));
}

// Dual-cone residual of rc = c_eff - A'y on one block, written into dres
// (reference src/cone_dispatch.cu:88-115).  c_eff = c + qx when qx != null.
template <class X, int Type>
__global__ void __launch_bounds__(kBlock) cone_dual_residual_kernel(double *dres, const X *c, const X *qx,
                                                                     const X *aty, const double *d, ConeList L) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= L.count) return;
    const int s = L.start[b];
    auto rc = [&](int i) { return (double)c[i] + (qx ? (double)qx[i] : 0.0) - (double)aty[i]; };
    if (Type == CONE_STANDARD_SOC || Type == CONE_ROTATED_SOC) {
        const int k = L.vdim[b];
        double pr[kMaxThreadCone + 2], out[kMaxThreadCone + 2];
        for (int m = 0; m < k + 2; ++m) pr[m] = rc(s + m);
        soc_dual_residual_block(pr, k, d + s, Type == CONE_ROTATED_SOC, L.rwarm + b, out);
        for (int m = 0; m < k + 2; ++m) dres[s + m] = out[m];
    } else if (Type == CONE_EXPONENTIAL) {
        double o[3];
        exp_project_point(-rc(s), -rc(s + 1), -rc(s + 2), 1.0 / d[s], 1.0 / d[s + 1], 1.0 / d[s + 2], o);
        dres[s] = -o[0] * d[s];
        dres[s + 1] = -o[1] * d[s + 1];
        dres[s + 2] = -o[2] * d[s + 2];
    } else {
        double o[3];
        power_project_point(-rc(s) * d[s], -rc(s + 1) * d[s + 1], -rc(s + 2) * d[s + 2], 1.0, 1.0, 1.0, L.alpha[b], o);
        dres[s] = -o[0];
        dres[s + 1] = -o[1];
        dres[s + 2] = -o[2];
    }
}

// ---------------- long SOC / RSOC blocks (one workgroup per block) ----------------
//
// Blocks with more than kMaxThreadCone vector slots (the reference's warp,
// block and grid methods, src/kernels/pdhcg_soc_cone_kernels.cu:727-810,
// 910-1148, 1725-2443) run the same algorithms as soc_project_block and
// soc_dual_residual_block with one workgroup per block: every sum over the
// block, including each evaluation of the multiplier function inside the
// bisection, is a fixed-order workgroup reduction broadcast to all threads,
// so all threads take the same branches.  The slots stay in global memory.

// Sums NV per-thread values over the workgroup; every thread receives the totals.
template <int NV> struct WgSum {
    double lds[NV * 16];
    double bc[NV];
    __device__ void operator()(double (&v)[NV]) {
        block_reduce<NV, 0>(v, lds);
        if (threadIdx.x == 0)
#pragma unroll
            for (int i = 0; i < NV; ++i) bc[i] = v[i];
        __syncthreads();
#pragma unroll
        for (int i = 0; i < NV; ++i) v[i] = bc[i];
        __syncthreads();
    }
};

// soc_multiplier with f summed over the workgroup; p(m), h(m) give slot m's
// value and relative scaling, sum(x) the workgroup total of x (every thread
// receives it).  Called by every thread of the workgroup.
template <class Pf, class Hf, class Sum>
__device__ double soc_multiplier_wg(Pf p, Hf h, int k, double w, double dw, double z, double *warm, Sum sum) {
    auto f = [&](double lam) {
        double acc = 0.0;
        for (int m = threadIdx.x; m < k; m += blockDim.x) {
            const double hm = h(m);
            const double t = p(m) * hm / (hm * hm + 2.0 * lam);
            acc += t * t;
        }
        acc = sum(acc);
        const double tw = w * dw / (dw * dw + 2.0 * lam);
        const double zt = z / (1.0 - 2.0 * lam);
        return acc + tw * tw - zt * zt;
    };
    double lo, hi;
    const bool zpos = z > 0.0;
    if (zpos) {
        lo = 0.0;
        hi = 0.5 - 1e-14;
    } else {
        lo = 0.5 + 1e-14;
        hi = 1.0;
        for (int it = 0; it < 60; ++it) {
            if (f(hi) > 0.0) break;
            lo = hi;
            hi *= 2.0;
        }
    }
    const double wl = *warm;
    if (wl > lo && wl < hi) {
        const double fv = f(wl);
        if (fabs(fv) < 1e-12) return wl;
        if (zpos == (fv > 0.0)) lo = wl;
        else hi = wl;
    }
    for (int it = 0; it < 60; ++it) {
        const double lam = 0.5 * (lo + hi);
        const double fv = f(lam);
        if (zpos == (fv > 0.0)) lo = lam;
        else hi = lam;
        if ((hi - lo) / (1.0 + hi + lo) < 1e-13) break;
    }
    const double lam = 0.5 * (lo + hi);
    __syncthreads();  // every thread has read *warm
    if (threadIdx.x == 0) *warm = lam;
    return lam;
}

// Projection of one long SOC/RSOC block onto E K (E = diag(e)) by the whole
// workgroup, the algorithm of soc_project_block: P(m) reads slot m (m < k + 2),
// E(m) its scaling, Set(m, v) writes it; sum(x) is a workgroup sum.  With
// write_all, a point already in the cone is written back as well.  Every
// slot is read before any is written (each reduction ends with a barrier).
template <bool Rotated, class Pf, class Ef, class Setf, class Sum>
__device__ void soc_project_coop(Pf P, Ef E, Setf Set, int k, double *warm, Sum sum, bool write_all) {
    const double a = P(k), bb = P(k + 1);
    const double w = Rotated ? (a - bb) * kInvSqrt2 : a;
    const double z = Rotated ? (a + bb) * kInvSqrt2 : bb;
    const double dref = Rotated ? sqrt(E(k) * E(k + 1)) : E(k + 1);
    const double dw = Rotated ? 1.0 : E(k) / dref;
    bool uniform = Rotated || dw == 1.0;
    {
        double nu = 0.0;
        for (int m = threadIdx.x; m < k; m += blockDim.x)
            if (E(m) != dref) nu = 1.0;
        uniform = uniform && sum(nu) == 0.0;
    }
    // final transform: mode 0 keep, 1 zero, 2 uniform scale sc, 3 weighted with lam
    int mode;
    double sc = 0.0, lam = 0.0, wn = 0.0, zn = 0.0;
    if (uniform) {
        double ss = 0.0;
        for (int m = threadIdx.x; m < k; m += blockDim.x) {
            const double pm = P(m);
            ss += pm * pm;
        }
        const double r = sqrt(w * w + sum(ss));
        if (r <= z) mode = 0;
        else if (r <= -z) mode = 1;
        else {
            mode = 2;
            sc = (z + r) / (2.0 * r);
            wn = sc * w;
            zn = sc * r;
        }
    } else {
        auto h = [&](int m) { return E(m) / dref; };
        double r0 = 0.0, r1 = 0.0;  // sum (p/h)^2, sum (p h)^2 over the vector slots
        for (int m = threadIdx.x; m < k; m += blockDim.x) {
            const double hm = h(m), pm = P(m);
            r0 += (pm / hm) * (pm / hm);
            r1 += (pm * hm) * (pm * hm);
        }
        r0 = sum(r0);
        r1 = sum(r1);
        const double ri = (w / dw) * (w / dw) + r0, rp = (w * dw) * (w * dw) + r1;
        if (sqrt(ri) <= z) mode = 0;
        else if (sqrt(rp) <= -z) mode = 1;
        else {
            mode = 3;
            lam = soc_multiplier_wg(P, h, k, w, dw, z, warm, sum);
            wn = Rotated ? w / (1.0 + 2.0 * lam) : w * dw * dw / (dw * dw + 2.0 * lam);
            zn = z / (1.0 - 2.0 * lam);
        }
    }
    if (mode == 0 && !write_all) return;
    for (int m = threadIdx.x; m < k + 2; m += blockDim.x) {
        double pm = P(m);
        if (mode == 1) pm = 0.0;
        else if (mode >= 2 && m < k) {
            if (mode == 2) pm *= sc;
            else {
                const double hm = E(m) / dref;
                pm = pm * hm * hm / (hm * hm + 2.0 * lam);
            }
        } else if (mode >= 2) {
            if (Rotated) pm = m == k ? (zn + wn) * kInvSqrt2 : (zn - wn) * kInvSqrt2;
            else pm = m == k ? wn : zn;
        }
        Set(m, pm);
    }
}

// Projection of long blocks onto D K (q == null) or, for the diagonal-Q prox,
// onto D K in the metric diag(w), w = 1 + tau q, as W^-1/2 P_{W^1/2 D K}(W^1/2 v)
// (cone_project_kernel / cone_project_diag_q_kernel).  Grid: L.count workgroups.
template <class X, bool Rotated>
__global__ void __launch_bounds__(kBlock) soc_long_project_kernel(X *v, const double *d, const X *q, double tau,
                                                                   ConeList L, int use_rwarm, const int *gate) {
    if (gate && *gate == 0) return;
    __shared__ WgSum<1> red;
    const int b = blockIdx.x;
    const int s = L.start[b], k = L.vdim[b];
    double *warm = use_rwarm ? L.rwarm + b : L.warm + b;
    auto sw = [&](int m) {
        if (!q) return 1.0;
        const double w = 1.0 + tau * (double)q[s + m];
        return sqrt(w > 0.0 ? w : 1.0);
    };
    auto P = [&](int m) { return (double)v[s + m] * sw(m); };
    auto E = [&](int m) { return d[s + m] * sw(m); };
    auto Set = [&](int m, double val) { v[s + m] = (X)(val / sw(m)); };
    auto sum = [&](double x) {
        double t[1] = {x};
        red(t);
        return t[0];
    };
    soc_project_coop<Rotated>(P, E, Set, k, warm, sum, q != nullptr);
}

// Dual residual of long blocks (cone_dual_residual_kernel): (rc - P_{K*}(rc)) d
// with rc = c + qx - aty.  Grid: L.count workgroups.
template <class X, bool Rotated>
__global__ void __launch_bounds__(kBlock) soc_long_dual_residual_kernel(double *dres, const X *c, const X *qx,
                                                                         const X *aty, const double *d, ConeList L) {
    __shared__ WgSum<2> red2;
    __shared__ WgSum<1> red1;
    const int b = blockIdx.x;
    const int s = L.start[b], k = L.vdim[b];
    auto sum1 = [&](double x) {
        double t[1] = {x};
        red1(t);
        return t[0];
    };
    auto rc = [&](int m) {
        const int i = s + m;
        return (double)c[i] + (qx ? (double)qx[i] : 0.0) - (double)aty[i];
    };
    auto D = [&](int m) { return d[s + m]; };
    const double a = rc(k), bb = rc(k + 1);
    const double rw = Rotated ? (a - bb) * kInvSqrt2 : a;
    const double rz = Rotated ? (a + bb) * kInvSqrt2 : bb;
    const double dref = Rotated ? sqrt(D(k) * D(k + 1)) : D(k + 1);
    const double ew = Rotated ? 1.0 : dref / D(k);
    bool uniform = Rotated || ew == 1.0;
    {
        double nu[2] = {0.0, 0.0};
        for (int m = threadIdx.x; m < k; m += blockDim.x)
            if (D(m) != dref) nu[0] = 1.0;
        red2(nu);
        uniform = uniform && nu[0] == 0.0;
    }
    double pa, pb;
    if (uniform) {
        double ss[2] = {0.0, 0.0};
        for (int m = threadIdx.x; m < k; m += blockDim.x) {
            const double pm = rc(m);
            ss[0] += pm * pm;
        }
        red2(ss);
        const double r = sqrt(rw * rw + ss[0]);
        double vf, pw, pz;
        if (r <= rz) {
            vf = 0.0;
            pw = rw;
            pz = rz;
        } else if (r <= -rz) {
            vf = 1.0;
            pw = 0.0;
            pz = 0.0;
        } else {
            const double sc = (rz + r) / (2.0 * r);
            vf = 1.0 - sc;
            pw = sc * rw;
            pz = sc * r;
        }
        for (int m = threadIdx.x; m < k; m += blockDim.x) dres[s + m] = rc(m) * vf * D(m);
        if (Rotated) {
            if (r <= rz) {
                pa = a;
                pb = bb;
            } else {
                pa = (pz + pw) * kInvSqrt2;
                pb = (pz - pw) * kInvSqrt2;
            }
        } else {
            pa = pw;
            pb = pz;
        }
    } else {
        auto h = [&](int m) { return dref / D(m); };
        double rr[2] = {0.0, 0.0};
        for (int m = threadIdx.x; m < k; m += blockDim.x) {
            const double hm = h(m), pm = rc(m);
            rr[0] += (pm / hm) * (pm / hm);
            rr[1] += (pm * hm) * (pm * hm);
        }
        red2(rr);
        const double ri = (rw / ew) * (rw / ew) + rr[0], rp = (rw * ew) * (rw * ew) + rr[1];
        if (sqrt(ri) <= rz) {
            for (int m = threadIdx.x; m < k + 2; m += blockDim.x) dres[s + m] = 0.0;
            return;
        }
        if (sqrt(rp) <= -rz) {
            for (int m = threadIdx.x; m < k + 2; m += blockDim.x) dres[s + m] = rc(m) * D(m);
            return;
        }
        const double lam = soc_multiplier_wg(rc, h, k, rw, ew, rz, L.rwarm + b, sum1);
        for (int m = threadIdx.x; m < k; m += blockDim.x) {
            const double hm = h(m), pm = rc(m);
            dres[s + m] = (pm - pm * hm * hm / (hm * hm + 2.0 * lam)) * D(m);
        }
        const double zn = rz / (1.0 - 2.0 * lam);
        const double wn = Rotated ? rw / (1.0 + 2.0 * lam) : rw * ew * ew / (ew * ew + 2.0 * lam);
        if (Rotated) {
            pa = (zn + wn) * kInvSqrt2;
            pb = (zn - wn) * kInvSqrt2;
        } else {
            pa = wn;
            pb = zn;
        }
    }
    if (threadIdx.x == 0) {
        dres[s + k] = (a - pa) * D(k);
        dres[s + k + 1] = (bb - pb) * D(k + 1);
    }
}

// Power-cone primal violation (reference pdhcg_power_cone_kernels.cu:621-669),
// reduced over blocks: out = [sum abs^2, sum rel^2, max abs, max rel].
template <class X>
__global__ void __launch_bounds__(kBlock) power_violation_kernel(const X *xp, const double *d, ConeList L,
                                                                  double hscale, GridRed gr) {
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    for (int b = blockIdx.x * blockDim.x + threadIdx.x; b < L.count; b += gridDim.x * blockDim.x) {
        const int s = L.start[b];
        const double x = (double)xp[s] / d[s], y = (double)xp[s + 1] / d[s + 1], z = (double)xp[s + 2] / d[s + 2];
        double av, rv;
        if (!isfinite(x) || !isfinite(y) || !isfinite(z)) {
            av = rv = INFINITY;
        } else {
            double viol = fmax(-x, -y);
            const double az = fabs(z);
            bool done = false;
            if (az > 0.0) {
                double bound = 0.0;
                if (x > 0.0 && y > 0.0) {
                    const double al = L.alpha[b];
                    const double lb = al * log(x) + (1.0 - al) * log(y), lz = log(az);
                    if (lb + 64.0 * DBL_EPSILON * (1.0 + fabs(lb) + fabs(lz)) >= lz) {
                        viol = fmax(viol, 0.0);
                        done = true;
                    } else {
                        bound = exp(lb);
                    }
                }
                if (!done) viol = fmax(viol, az - bound);
            }
            viol = fmax(viol, 0.0);
            av = viol;
            rv = viol / (hscale + fmax(fabs(x), fmax(fabs(y), az)));
        }
        v[0] += av * av;
        v[1] += rv * rv;
        v[2] = fmax(v[2], av);
        v[3] = fmax(v[3], rv);
    }
    grid_partials<2, 2>(gr, v);
}

// ---------------- PSD cone: one workgroup per matrix ----------------
//
// Parallel two-sided Jacobi (Golub & Van Loan, sec. 8.5) on the dense
// symmetric matrix of an svec block, every PSD block of the problem in one
// launch (one workgroup per block, largest orders first).  The order k is
// padded to an even kp (a zero row and column: a zero eigenvalue that
// contributes nothing).
//
// Warm start: each block keeps the eigenvector rows V' of its previous
// projection (identity before the first) in its slice of the work buffer.
// The sweeps start from B = V' A V'^T with V' as the initial rotation, so a
// matrix close to the previous one is nearly diagonal already and needs fewer
// sweeps.  The result is the projection of A whatever V' holds; V' only
// changes the rounding (the nearest PSD matrix is unique).  A projection that
// does not converge resets V' to the identity.
//
// A sweep is kp - 1 rounds of the circle (round-robin tournament) ordering;
// the kp / 2 pairs of a round are disjoint, so their rotations are applied at
// once.  A round is two phases separated by barriers:
//   1. each pair (p, q) gets its rotation (c, s) from its 2x2 diagonal block;
//   2. every 2x2 block A[{p,q},{r,s}] (pairs P <= R) becomes J_P' B J_R,
//      computed by one thread; rows p, q of V' become J_P' V'[{p,q}, :].
// The matrix is held as its packed lower triangle (entry (i, j), i >= j, at
// i (i + 1) / 2 + j), so every entry belongs to exactly one block and is
// read and written once per round; a diagonal block's off-diagonal entry is
// set to 0.  Where the two live (LDS or the work buffer) is the block's mode:
// 0 both in LDS, 1 A in LDS and V' in the work buffer, 2 V' in LDS and A in
// the work buffer, 3 both in the work buffer.  The dense input and the
// transform's temporary stay in the work buffer.  Sweeps run until the off-diagonal mass
// is at most 1e-30 of the total.  A matrix not converged after kPsdMaxSweeps
// sweeps (only non-finite input does that) is written as NaN, as the
// reference writes NaN when its eigensolver fails (reference
// src/kernels/pdhcg_psd_cone_kernels.cu:263-486).

constexpr int kPsdMaxSweeps = 60;
// Workgroup sizes of the projection (one instance each) and, per size, the
// 2x2 block updates and V' row-pair updates a thread loads before it stores
// any (phase 2); the larger workgroup hides more latency per round and fits
// 128 VGPRs with the smaller chunks.
constexpr int kPsdThreads = 512, kPsdThreadsLarge = 1024;
// V' columns per wave when V' is held in registers during the sweeps (the
// 1024-thread form up to order 128)
constexpr int kPsdRegCols = 8, kPsdRegMinOrder = 96;
template <int NT> struct PsdChunks {
    static constexpr int a = NT >= 1024 ? 2 : 4, v = NT >= 1024 ? 4 : 8;
};
__host__ __device__ inline bool psd_v_in_lds(int mode) { return mode == 0 || mode == 2; }
__host__ __device__ inline bool psd_a_in_lds(int mode) { return mode == 0 || mode == 1; }

// One PSD block of a launch.
struct PsdBlock {
    int start;       // first svec slot in v
    int order;       // k
    int mode;        // where A and V' live during the sweeps (see above)
    int pad;
    long long work;  // offset (doubles) of the block's slice of the work buffer
};

// Workgroup LDS (dynamic) of a projection of order k in a mode: V' and the
// packed A (by mode), the round's rotations and pairs, and the clamped
// eigenvalues.
__host__ __device__ inline size_t psd_lds_bytes(int k, int mode) {
    const size_t kp = (size_t)(k + (k & 1)), np = kp / 2, kt = kp * (kp + 1) / 2;
    const size_t mats = (psd_v_in_lds(mode) ? kp * kp : 0) + (psd_a_in_lds(mode) ? kt : 0);
    return (mats + 2 * np + kp) * sizeof(double) + 2 * np * sizeof(int);
}

// Work doubles per block: the kept V' of the two projection sites (the
// iterate and the dual residual), the dense input, the transform's
// temporary and the packed A when it is not in LDS.
__host__ __device__ inline size_t psd_work_doubles(int k, int mode) {
    const size_t kp = (size_t)(k + (k & 1)), kt = kp * (kp + 1) / 2;
    return 4 * kp * kp + (psd_a_in_lds(mode) ? 0 : kt);
}

// Row r and column c (r >= c) of svec slot `slot` of an order-k block
// (lower triangle, column-major).
__device__ __forceinline__ void psd_slot_rc(int slot, int k, int &r, int &c) {
    c = 0;
    int rem = slot;
    while (rem >= k - c) {
        rem -= k - c;
        ++c;
    }
    r = c + rem;
}

// Packed lower-triangle index of (i, j), either order.
__device__ __forceinline__ int psd_tri(int i, int j) {
    return i >= j ? i * (i + 1) / 2 + j : j * (j + 1) / 2 + i;
}

// Row of packed index e (the largest i with i (i + 1) / 2 <= e).
__device__ __forceinline__ int psd_tri_row(int e) {
    int i = (int)((sqrt(8.0 * e + 1.0) - 1.0) * 0.5);
    if ((i + 1) * (i + 2) / 2 <= e) ++i;
    if (i * (i + 1) / 2 > e) --i;
    return i;
}

// Sets the kept V' of every block of a launch to the identity.
static __global__ void psd_identity_kernel(const PsdBlock *blocks, double *work) {
    const PsdBlock b = blocks[blockIdx.x];
    const size_t kp = (size_t)(b.order + (b.order & 1));
    double *V = work + b.work;
    for (size_t i = threadIdx.x; i < 2 * kp * kp; i += blockDim.x) {
        const size_t e = i % (kp * kp);
        V[i] = (e / kp == e % kp) ? 1.0 : 0.0;
    }
}

// 4 x 4 register tile of C = X Y over kp x kp row-major operands, rows
// i0.., columns j0..; entries outside kp read as zero.  Lanes of a wave share
// the tile row (X reads broadcast) and take consecutive column tiles (Y reads
// contiguous); four steps of l are loaded before they are multiplied, so
// the loads of a Y in the work buffer overlap.
__device__ __forceinline__ void psd_tile_xy(const double *X, const double *Y, int kp, int i0, int j0,
                                            double acc[4][4]) {
    for (int u = 0; u < 4; ++u)
        for (int w = 0; w < 4; ++w) acc[u][w] = 0.0;
    const double *xr[4];
    bool xi[4], yj[4];
    for (int u = 0; u < 4; ++u) {
        xi[u] = i0 + u < kp;
        yj[u] = j0 + u < kp;
        xr[u] = X + (size_t)(xi[u] ? i0 + u : i0) * kp;
    }
    constexpr int LU = 4;
    int l = 0;
    for (; l + LU <= kp; l += LU) {
        double a[LU][4], b[LU][4];
#pragma unroll
        for (int d = 0; d < LU; ++d) {
            const double *yl = Y + (size_t)(l + d) * kp + j0;
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                a[d][u] = xi[u] ? xr[u][l + d] : 0.0;
                b[d][u] = yj[u] ? yl[u] : 0.0;
            }
        }
#pragma unroll
        for (int d = 0; d < LU; ++d)
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int w = 0; w < 4; ++w) acc[u][w] = fma(a[d][u], b[d][w], acc[u][w]);
    }
    for (; l < kp; ++l) {
        const double *yl = Y + (size_t)l * kp + j0;
        double a[4], b[4];
        for (int u = 0; u < 4; ++u) {
            a[u] = xi[u] ? xr[u][l] : 0.0;
            b[u] = yj[u] ? yl[u] : 0.0;
        }
        for (int u = 0; u < 4; ++u)
            for (int w = 0; w < 4; ++w) acc[u][w] = fma(a[u], b[w], acc[u][w]);
    }
}

template <class X, int NT>
__global__ void __launch_bounds__(NT) psd_project_kernel(X *v, const PsdBlock *blocks, double *work,
                                                                   int site, const int *gate) {
    if (gate && *gate == 0) return;
    extern __shared__ double psd_lds[];
    constexpr int kPsdChunk = PsdChunks<NT>::a, kPsdChunkV = PsdChunks<NT>::v;
    __shared__ double red[2 * (NT / kWave)];
    __shared__ int stop;
    const PsdBlock blk = blocks[blockIdx.x];
    const int start = blk.start, k = blk.order, mode = blk.mode;
    const int tid = threadIdx.x, nt = blockDim.x;
    if (k == 1) {
        if (tid == 0) v[start] = (X)fmax((double)v[start], 0.0);
        return;
    }
    const int kp = k + (k & 1), np = kp / 2, m = kp - 1;
    const size_t kk = (size_t)kp * kp;
    const int kt = kp * (kp + 1) / 2;
    double *wb = work + blk.work;
    double *Vkeep = wb + (size_t)site * kk;  // V' kept between calls at this site
    double *F = wb + 2 * kk;                  // dense A
    double *S = wb + 3 * kk;                  // temporary of the warm transform
    const bool vlds = psd_v_in_lds(mode), alds = psd_a_in_lds(mode);
    double *Vt = vlds ? psd_lds : Vkeep;
    double *A = alds ? psd_lds + (vlds ? kk : 0) : wb + 4 * kk;
    double *cs = psd_lds + (vlds ? kk : 0) + (alds ? kt : 0);
    double *lam = cs + 2 * np;
    int *pq = reinterpret_cast<int *>(lam + kp);
    const int L = k * (k + 1) / 2;

    // svec -> dense F (off-diagonal entries / sqrt 2) with zero padding; V' in
    for (size_t i = tid; i < kk; i += nt) {
        F[i] = 0.0;
        if (vlds) Vt[i] = Vkeep[i];
    }
    __syncthreads();
    for (int slot = tid; slot < L; slot += nt) {
        int r, c;
        psd_slot_rc(slot, k, r, c);
        double val = (double)v[start + slot];
        if (r != c) val *= kInvSqrt2;
        F[(size_t)r * kp + c] = val;
        F[(size_t)c * kp + r] = val;
    }
    __syncthreads();

    // warm transform B = V' F V'^T: S = (V' F)^T = F V'^T, then the lower
    // triangle of B = V' S into the packed A
    const int nb = (kp + 3) / 4;
    for (int t = tid; t < nb * nb; t += nt) {
        const int ti = t / nb, tj = t - ti * nb;
        double acc[4][4];
        psd_tile_xy(Vt, F, kp, 4 * ti, 4 * tj, acc);
        for (int u = 0; u < 4; ++u)
            for (int w = 0; w < 4; ++w)
                if (4 * ti + u < kp && 4 * tj + w < kp) S[(size_t)(4 * tj + w) * kp + 4 * ti + u] = acc[u][w];
    }
    __syncthreads();
    for (int t = tid; t < nb * nb; t += nt) {
        const int ti = t / nb, tj = t - ti * nb;
        if (tj > ti) continue;
        double acc[4][4];
        psd_tile_xy(Vt, S, kp, 4 * ti, 4 * tj, acc);
        for (int u = 0; u < 4; ++u)
            for (int w = 0; w < 4; ++w) {
                const int i = 4 * ti + u, j = 4 * tj + w;
                if (i < kp && j <= i) A[i * (i + 1) / 2 + j] = acc[u][w];
            }
    }
    __syncthreads();

    // phase-2 work split, fixed for the whole projection: threads per V'
    // row pair (tpp) and per block row pair (tpr), and the strides that
    // cover more pairs than threads
    const int tpp = nt >= np ? nt / np : 1, pstride = nt / tpp;
    const int nrp = (np + 1) / 2, tpr = nt >= nrp ? nt / nrp : 1, rstride = nt / tpr;
    // V' in registers (the 1024-thread form at orders 97 to 128): lane L of
    // every wave holds rows lu, ld of V' (the pair at position L of the
    // round robin) for its wave's kPsdRegCols columns.  After each round
    // the rows move one position along the circle (the up rows one lane
    // right, the down rows one lane left, position 0's up row fixed), the
    // same pairing as the circle ordering, so V' needs no memory round trip
    // in a round.  Elsewhere V' stays in Vt and pairs come from the formula
    // (below order 97 the LDS update is the cheaper: the shifts cost the
    // same whatever the order).
    const bool vreg = NT >= kPsdThreadsLarge && np <= kWave && kp > kPsdRegMinOrder &&
                      kp <= (NT / kWave) * kPsdRegCols;
    const int lane = tid % kWave, wcol = (tid / kWave) * kPsdRegCols;
    int lu = 2 * lane, ld = 2 * lane + 1;
    double vu[kPsdRegCols], vd[kPsdRegCols];
#pragma unroll
    for (int j = 0; j < kPsdRegCols; ++j) {
        const int c = wcol + j;
        const bool ok = vreg && lane < np && c < kp;
        vu[j] = ok ? Vt[(size_t)lu * kp + c] : 0.0;
        vd[j] = ok ? Vt[(size_t)ld * kp + c] : 0.0;
    }
    bool converged = false;
    for (int sweep = 0; sweep <= kPsdMaxSweeps; ++sweep) {
        double off = 0.0, tot = 0.0;
        for (int e = tid; e < kt; e += nt) {
            const double a = A[e];
            const int i = psd_tri_row(e);
            if (e - i * (i + 1) / 2 == i) tot += a * a;
            else off += 2.0 * (a * a);
        }
        off = wave_sum(off);
        tot = wave_sum(tot);
        if (tid % kWave == 0) {
            red[2 * (tid / kWave)] = off;
            red[2 * (tid / kWave) + 1] = tot;
        }
        __syncthreads();
        if (tid == 0) {
            double o = 0.0, s = 0.0;
            for (int w = 0; w < nt / kWave; ++w) {
                o += red[2 * w];
                s += red[2 * w + 1];
            }
            s += o;
            stop = (o <= 1e-30 * s || o == 0.0) ? 1 : 0;
        }
        __syncthreads();
        if (stop) {
            converged = true;
            break;
        }
        if (sweep == kPsdMaxSweeps) break;
        for (int t = 0; t < m; ++t) {
            // phase 1: the round's pairs and rotations
            for (int i = tid; i < np; i += nt) {
                int a, b;
                if (vreg) {
                    a = lu;  // i == lane: wave 0 holds position i
                    b = ld;
                } else if (i == 0) {
                    a = t;
                    b = m;
                } else {
                    a = (t + i) % m;
                    b = (t - i + m) % m;
                }
                const int p = a < b ? a : b, q = a < b ? b : a;
                const double apq = A[q * (q + 1) / 2 + p];
                double c = 1.0, s = 0.0;
                if (apq != 0.0) {
                    const double theta = (A[q * (q + 1) / 2 + q] - A[p * (p + 1) / 2 + p]) / (2.0 * apq);
                    const double tt = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                    c = 1.0 / sqrt(tt * tt + 1.0);
                    s = tt * c;
                }
                cs[2 * i] = c;
                cs[2 * i + 1] = s;
                pq[2 * i] = p;
                pq[2 * i + 1] = q;
            }
            __syncthreads();
            // phase 2: the 2x2 blocks, then the rows of V' (in Vt or in
            // registers).
            // The 2x2 blocks: a block row pair holds block rows
            // rp and np - 1 - rp (np + 1 blocks; the middle row alone when np
            // is odd); its tpr threads take every tpr-th block, loading
            // kPsdChunk blocks before storing any so their round trips overlap.
            for (int rp = tid / tpr; rp < nrp; rp += rstride) {
                const int len = ((np & 1) && rp == np / 2) ? rp + 1 : np + 1;
                for (int j0 = tid % tpr; j0 < len; j0 += kPsdChunk * tpr) {
                    int off4[kPsdChunk][4], lo[kPsdChunk], hi[kPsdChunk];
                    double e[kPsdChunk][4];
                    bool live[kPsdChunk];
#pragma unroll
                    for (int j = 0; j < kPsdChunk; ++j) {
                        const int jj = j0 + j * tpr;
                        live[j] = jj < len;
                        // block (hi, lo), hi >= lo: rows from pair lo, columns from pair hi
                        hi[j] = !live[j] ? 0 : (jj <= rp ? rp : np - 1 - rp);
                        lo[j] = !live[j] ? 0 : (jj <= rp ? jj : jj - rp - 1);
                        const int p = pq[2 * lo[j]], q = pq[2 * lo[j] + 1], r = pq[2 * hi[j]], s2 = pq[2 * hi[j] + 1];
                        off4[j][0] = psd_tri(p, r);
                        off4[j][1] = psd_tri(p, s2);
                        off4[j][2] = psd_tri(q, r);
                        off4[j][3] = psd_tri(q, s2);
                        for (int u = 0; u < 4; ++u) e[j][u] = live[j] ? A[off4[j][u]] : 0.0;
                    }
#pragma unroll
                    for (int j = 0; j < kPsdChunk; ++j) {
                        if (!live[j]) continue;
                        const double c1 = cs[2 * lo[j]], s1 = cs[2 * lo[j] + 1];
                        const double c2 = cs[2 * hi[j]], s2 = cs[2 * hi[j] + 1];
                        const double b00 = e[j][0], b01 = e[j][1], b10 = e[j][2], b11 = e[j][3];
                        const double m00 = c1 * b00 - s1 * b10, m01 = c1 * b01 - s1 * b11;
                        const double m10 = s1 * b00 + c1 * b10, m11 = s1 * b01 + c1 * b11;
                        if (lo[j] == hi[j]) {
                            // (p, q) = (r, s): entries (p,p), (q,p) (= off4 1 and 2), (q,q)
                            A[off4[j][0]] = c1 * m00 - s1 * m01;
                            A[off4[j][3]] = s1 * m10 + c1 * m11;
                            A[off4[j][1]] = 0.0;
                            continue;
                        }
                        A[off4[j][0]] = c2 * m00 - s2 * m01;
                        A[off4[j][1]] = s2 * m00 + c2 * m01;
                        A[off4[j][2]] = c2 * m10 - s2 * m11;
                        A[off4[j][3]] = s2 * m10 + c2 * m11;
                    }
                }
            }
            // rows p, q of V' in Vt: the tpp threads of a pair take every
            // tpp-th column, kPsdChunkV columns loaded before any is stored
            for (int P = vreg ? np : tid / tpp; P < np; P += pstride) {
                const double c = cs[2 * P], s = cs[2 * P + 1];
                if (s == 0.0) continue;
                double *vp = Vt + (size_t)pq[2 * P] * kp, *vq = Vt + (size_t)pq[2 * P + 1] * kp;
                for (int i0 = tid % tpp; i0 < kp; i0 += kPsdChunkV * tpp) {
                    double va[kPsdChunkV], vb[kPsdChunkV];
#pragma unroll
                    for (int j = 0; j < kPsdChunkV; ++j) {
                        const int i = i0 + j * tpp;
                        va[j] = i < kp ? vp[i] : 0.0;
                        vb[j] = i < kp ? vq[i] : 0.0;
                    }
#pragma unroll
                    for (int j = 0; j < kPsdChunkV; ++j) {
                        const int i = i0 + j * tpp;
                        if (i >= kp) continue;
                        vp[i] = c * va[j] - s * vb[j];
                        vq[i] = s * va[j] + c * vb[j];
                    }
                }
            }
            if (vreg) {
                // this round's rotation of the lane's two rows (p = min(lu, ld))
                const double c = lane < np ? cs[2 * lane] : 1.0, s = lane < np ? cs[2 * lane + 1] : 0.0;
                if (s != 0.0) {
                    const bool up_is_p = lu < ld;
#pragma unroll
                    for (int j = 0; j < kPsdRegCols; ++j) {
                        const double a = up_is_p ? vu[j] : vd[j], b = up_is_p ? vd[j] : vu[j];
                        const double na = c * a - s * b, nb2 = s * a + c * b;
                        vu[j] = up_is_p ? na : nb2;
                        vd[j] = up_is_p ? nb2 : na;
                    }
                }
                // next round's positions (a single pair stays in place);
                // columns past kp are skipped by whole waves
                const int from_prev = lane > 0 ? lane - 1 : 0, from_next = lane + 1 < kWave ? lane + 1 : lane;
#pragma unroll
                for (int j = 0; j < kPsdRegCols; ++j) {
                    if (wcol + j >= kp) break;
                    const double up_prev = __shfl(vu[j], from_prev, kWave);
                    const double dn_next = __shfl(vd[j], from_next, kWave);
                    const double dn0 = __shfl(vd[j], 0, kWave);
                    const double nu = lane == 0 ? vu[j] : (lane == 1 ? dn0 : up_prev);
                    vd[j] = lane == np - 1 ? (np > 1 ? vu[j] : vd[j]) : dn_next;
                    vu[j] = nu;
                }
                const int lu_prev = __shfl(lu, from_prev, kWave), ld_next = __shfl(ld, from_next, kWave);
                const int ld0 = __shfl(ld, 0, kWave);
                const int nlu = lane == 0 ? lu : (lane == 1 ? ld0 : lu_prev);
                ld = lane == np - 1 ? (np > 1 ? lu : ld) : ld_next;
                lu = nlu;
            }
            __syncthreads();
        }
    }
    if (vreg) {
        // V' back to Vt, rows at the lanes' final positions
#pragma unroll
        for (int j = 0; j < kPsdRegCols; ++j) {
            const int c = wcol + j;
            if (lane < np && c < kp) {
                Vt[(size_t)lu * kp + c] = vu[j];
                Vt[(size_t)ld * kp + c] = vd[j];
            }
        }
        __syncthreads();
    }
    // keep V' for the next call (identity after a failure)
    if (vlds || !converged)
        for (size_t i = tid; i < kk; i += nt) Vkeep[i] = converged ? Vt[i] : (i / kp == i % kp ? 1.0 : 0.0);
    for (int e = tid; e < kp; e += nt) lam[e] = fmax(A[e * (e + 1) / 2 + e], 0.0);
    __syncthreads();
    // svec of V' ^T max(diag(A), 0) V' on the lower 4 x 4 tiles of k x k
    const int nk = (k + 3) / 4;
    for (int t = tid; t < nk * nk; t += nt) {
        const int ti = t / nk, tj = t - ti * nk;
        if (tj > ti) continue;
        const int r0 = 4 * ti, c0 = 4 * tj;
        double acc[4][4];
        for (int u = 0; u < 4; ++u)
            for (int w = 0; w < 4; ++w) acc[u][w] = 0.0;
        if (converged) {
            for (int e = 0; e < kp; ++e) {
                const double le = lam[e];
                if (le == 0.0) continue;
                const double *ve = Vt + (size_t)e * kp;
                double a[4], b[4];
                for (int u = 0; u < 4; ++u) {
                    a[u] = r0 + u < k ? le * ve[r0 + u] : 0.0;
                    b[u] = c0 + u < k ? ve[c0 + u] : 0.0;
                }
                for (int u = 0; u < 4; ++u)
                    for (int w = 0; w < 4; ++w) acc[u][w] = fma(a[u], b[w], acc[u][w]);
            }
        }
        for (int u = 0; u < 4; ++u)
            for (int w = 0; w < 4; ++w) {
                const int r = r0 + u, c = c0 + w;
                if (r >= k || c > r) continue;
                double val = converged ? acc[u][w] : NAN;
                if (r != c) val *= kSqrt2;
                v[start + c * k - c * (c - 1) / 2 + (r - c)] = (X)val;
            }
    }
}

// Elementwise helpers on cone slots, driven by a slot mask (kMaskCone on
// SOC/RSOC/EXP/POWER slots, kMaskPsd on PSD slots, 0 elsewhere).
constexpr char kMaskCone = 1;
constexpr char kMaskPsd = 2;

template <class X>
__global__ void __launch_bounds__(kBlock) cone_reflect_kernel(X *xr, const X *xp, const X *x, const char *mask, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        if (mask[i]) xr[i] = (X)(2.0 * (double)xp[i] - (double)x[i]);
}

// ---------------- affine cones on constraint rows ----------------
// Rows of [A; F] whose value F x + g must lie in a cone.  The cones are
// applied in the scaled space as diag(d) K with d = 1 / row scaling
// (reference src/solver_state.cu:621-638).

// Dual step, first half, fused into the A x_bar SpMV: v = A x_bar + g is
// stored, and p = clamp(v - y / sigma, l, u) is written for the projection
// (reference pdhcg_kernels.cu:252-264).
template <class X> struct EpiDualPrepare {
    X *v_out, *p_out;
    const X *y, *g, *lo, *hi;
    double sigma;
    template <class Acc> __device__ void operator()(int r, Acc s) const {
        const double v = (double)s + (double)g[r];
        v_out[r] = (X)v;
        p_out[r] = (X)fmax((double)lo[r], fmin(v - (double)y[r] / sigma, (double)hi[r]));
    }
};

// Dual step, second half after the cone projection of p:
// y+ = y - sigma v + sigma p, reflection (reference pdhcg_kernels.cu:266-281).
template <class X>
__global__ void __launch_bounds__(kBlock) affine_dual_finish_kernel(const X *y, X *yp, X *yr, const X *v, const X *p,
                                                                     double sigma, int store, int m) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gridDim.x * blockDim.x) {
        const double cur = (double)y[i];
        const double nx = cur - sigma * (double)v[i] + sigma * (double)p[i];
        if (store) yp[i] = (X)nx;
        yr[i] = (X)(2.0 * nx - cur);
    }
}

// KKT rows with affine cones, first half: pres = clamp(A x + g) (projected
// onto the cones next), the dual-membership point starts at zero, and the
// dual-objective row terms are reduced: out = [sum row terms]
// (reference pdhcg_kernels.cu:660-676).
template <class X>
__global__ void __launch_bounds__(kBlock) affine_rows_prepare_kernel(const X *ax, const X *g, const X *cl, const X *cu,
                                                                      const X *yp, const X *cl_fin, const X *cu_fin,
                                                                      X *pres, X *amem, int m, GridRed gr) {
    double v[1] = {0.0};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gridDim.x * blockDim.x) {
        const double val = (double)ax[i] + (double)g[i];
        pres[i] = (X)fmax((double)cl[i], fmin(val, (double)cu[i]));
        amem[i] = (X)0.0;
        const double y = (double)yp[i];
        v[0] += fmax(y, 0.0) * (double)cl_fin[i] + fmin(y, 0.0) * (double)cu_fin[i] - (double)g[i] * y;
    }
    grid_partials<1, 0>(gr, v);
}

// Complementarity |<y, F x + g>| / cbr per affine block, and the dual
// membership point -y on its rows; one wave per block, fixed-order sum
// (reference pdhcg_affine_cone_kernels.cu:40-75).
template <class X>
__global__ void __launch_bounds__(kBlock) affine_complementarity_kernel(const X *ax, const X *g, const X *yp, X *amem,
                                                                         const int *start, const int *len, int count,
                                                                         double cbr, double *comp) {
    const int b = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
    const int lane = threadIdx.x % kWave;
    if (b >= count) return;
    const int s = start[b], L = len[b];
    double dot = 0.0;
    for (int j = lane; j < L; j += kWave) {
        const double y = (double)yp[s + j];
        amem[s + j] = (X)(-y);
        dot += y * ((double)ax[s + j] + (double)g[s + j]);
    }
    dot = wave_sum(dot);
    if (lane == 0) comp[b] = fabs(dot) / cbr;
}

// KKT rows with affine cones, second half: pres = (v - P(clamp v)) r and the
// dual membership scaled by d; out = [sum pres^2, sum amem^2, max |pres|,
// max |amem|] (reference pdhcg_affine_cone_kernels.cu:23-38).
template <class X>
__global__ void __launch_bounds__(kBlock) affine_rows_finish_kernel(const X *ax, const X *g, const X *pres,
                                                                     const X *amem, const double *con_rescale,
                                                                     const double *d, int m, GridRed gr) {
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gridDim.x * blockDim.x) {
        const double pr = ((double)ax[i] + (double)g[i] - (double)pres[i]) * con_rescale[i];
        const double am = (double)amem[i] * d[i];
        v[0] += pr * pr;
        v[1] += am * am;
        v[2] = fmax(v[2], abs_or_inf(pr));
        v[3] = fmax(v[3], abs_or_inf(am));
    }
    grid_partials<2, 2>(gr, v);
}

// out = [sum a^2, max |a|] of a double vector.
static __global__ void __launch_bounds__(kBlock) sq_max_kernel(const double *a, int n, GridRed gr) {
    double v[2] = {0.0, 0.0};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        v[0] += a[i] * a[i];
        v[1] = fmax(v[1], abs_or_inf(a[i]));
    }
    grid_partials<1, 1>(gr, v);
}

// PSD dual residual, first half: rc = c_eff - A'y on PSD slots, kept in dres
// and copied into p for the in-place projection.
template <class X>
__global__ void __launch_bounds__(kBlock) psd_rc_kernel(X *p, double *dres, const X *c, const X *qx, const X *aty,
                                                         const char *mask, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        if (mask[i] == kMaskPsd) {
            const double rc = (double)c[i] + (qx ? (double)qx[i] : 0.0) - (double)aty[i];
            dres[i] = rc;
            p[i] = (X)rc;
        }
}

// Second half: dres = (rc - P_psd(rc)) d (the PSD cone is self-dual).
template <class X>
__global__ void __launch_bounds__(kBlock) psd_dres_kernel(double *dres, const X *p, const double *d, const char *mask,
                                                           int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        if (mask[i] == kMaskPsd) dres[i] = (dres[i] - (double)p[i]) * d[i];
}

}  // namespace pdhcg
