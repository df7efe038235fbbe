// PDHG step kernels for gfx950: fused proximal/reflection/Halpern epilogues,
// the device-resident projected Barzilai-Borwein inner solve, the low-rank
// R'DR product, and the FP64 reductions of restart, fixed-point error, KKT
// residuals and infeasibility rays.  Formulas restate reference
// src/kernels/pdhcg_kernels.cu and src/pdhg_core_op.cu (cited per kernel).
#pragma once

#include "dev_spmv.hpp"

namespace pdhcg {

// Scalars that live on the device between launches (FP64 throughout).
struct DevScalars {
    double red[32];             // generic reduction results (several launches write disjoint ranges)
    double bb_nrm;              // sqrt(s's) or sqrt(s'Ms) of the current BB step
    double bb_stms;             // s'Ms (Jacobi-preconditioned BB)
    double bb_tmp;              // s'(g_new - g_old)
    double bb_alpha;
    double tol_scale;           // sqrt(mean(M)) of the Jacobi metric
    double pi_value, pi_old, pi_inv;  // power-iteration scalars
    int bb_active;
    int bb_iter;
    int inner_total;
    int pi_active;
    int pi_iter;
};

template <class T> __device__ __forceinline__ T clampv(T v, T lo, T hi) { return fmax(lo, fmin(v, hi)); }

// ---------------------------------------------------------------------------
// SpMV epilogues for the PDHG step
// ---------------------------------------------------------------------------

// Dual proximal step on row r of A x_bar (reference pdhcg_kernels.cu:197-244),
// optionally followed by the Halpern update of that row (:283-306).
template <class X> struct EpiDualStep {
    X *y, *yp, *yr;
    const X *g, *cl, *cu, *y0;
    double sigma, w, rho;
    int store_pdhg, halpern;
    template <class Acc> __device__ void operator()(int r, Acc s) const {
        const double cur = (double)y[r];
        const double v = (double)s + (double)g[r];
        const double pj = fmax((double)cl[r], fmin(v - cur / sigma, (double)cu[r]));
        const double nx = cur - sigma * v + sigma * pj;
        if (store_pdhg) yp[r] = (X)nx;
        const double ref = 2.0 * nx - cur;
        if (halpern) {
            const double refl = rho * ref + (1.0 - rho) * cur;
            y[r] = (X)(w * refl + (1.0 - w) * (double)y0[r]);
        } else {
            yr[r] = (X)ref;
        }
    }
};

// LP / diagonal-Q primal proximal step on variable r of A'y
// (reference pdhcg_kernels.cu:111-195), optional Halpern update.
template <class X> struct EpiPrimalStep {
    X *x, *xp, *xr, *dslack;
    const X *c, *lb, *ub, *x0, *qdiag;  // qdiag null for LP
    double tau, w, rho;
    int store_pdhg, halpern;
    template <class Acc> __device__ void operator()(int r, Acc s) const {
        const double xi = (double)x[r];
        double t = xi - tau * ((double)c[r] - (double)s);
        if (qdiag) t = t / (1.0 + tau * (double)qdiag[r]);
        const double pj = fmax((double)lb[r], fmin(t, (double)ub[r]));
        const double ref = 2.0 * pj - xi;
        xr[r] = (X)ref;
        if (store_pdhg) {
            xp[r] = (X)pj;
            if (!qdiag) dslack[r] = (X)((pj - t) / tau);
        }
        if (halpern) {
            const double refl = rho * ref + (1.0 - rho) * xi;
            x[r] = (X)(w * refl + (1.0 - w) * (double)x0[r]);
        }
    }
};

// ---------------------------------------------------------------------------
// Elementwise kernels
// ---------------------------------------------------------------------------

// z <- w (rho z_bar + (1-rho) z) + (1-w) z0 on [x; y] (reference pdhcg_kernels.cu:283-306).
template <class X>
__global__ void __launch_bounds__(kBlock) halpern_kernel(const X *x0, X *x, const X *xr, const X *y0, X *y,
                                                          const X *yr, int n, int m, double w, double rho) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < (long long)n + m;
         i += (long long)gridDim.x * blockDim.x) {
        if (i < n) {
            const double refl = rho * (double)xr[i] + (1.0 - rho) * (double)x[i];
            x[i] = (X)(w * refl + (1.0 - w) * (double)x0[i]);
        } else {
            const long long j = i - n;
            const double refl = rho * (double)yr[j] + (1.0 - rho) * (double)y[j];
            y[j] = (X)(w * refl + (1.0 - w) * (double)y0[j]);
        }
    }
}

template <class X> __global__ void __launch_bounds__(kBlock) clamp_kernel(X *x, const X *lb, const X *ub, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        x[i] = (X)fmax((double)lb[i], fmin((double)x[i], (double)ub[i]));
}

template <class X> __global__ void __launch_bounds__(kBlock) mul_kernel(const X *a, const X *b, X *out, int n, const int *gate) {
    if (gate && *gate == 0) return;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        out[i] = (X)((double)a[i] * (double)b[i]);
}

// a <- a * s where s is a device scalar (power iteration normalisation).
template <class X> __global__ void __launch_bounds__(kBlock) scale_dev_kernel(X *a, const double *s, int n, const int *gate) {
    if (gate && *gate == 0) return;
    const double f = *s;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) a[i] = (X)((double)a[i] * f);
}

template <class X> __global__ void __launch_bounds__(kBlock) scale_kernel(X *a, double f, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) a[i] = (X)((double)a[i] * f);
}

template <class X> __global__ void __launch_bounds__(kBlock) copy_gated_kernel(X *dst, const X *src, int n, const int *gate) {
    if (gate && *gate == 0) return;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) dst[i] = src[i];
}

// ---------------------------------------------------------------------------
// Reductions (FP64)
// ---------------------------------------------------------------------------

// out = [sum a_i b_i]
template <class X>
__global__ void __launch_bounds__(kBlock) dot_kernel(const X *a, const X *b, int n, GridRed g, const int *gate) {
    if (gate && *gate == 0) return;
    double v[1] = {0.0};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        v[0] += (double)a[i] * (double)b[i];
    grid_reduce<1, 0>(g, v);
}

// dx = a - b (n), dy = c - d (m); out = [|dx|^2, |dy|^2]
// (reference compute_delta_solution_kernel, pdhcg_kernels.cu:329-348, plus the two norms).
template <class X>
__global__ void __launch_bounds__(kBlock) delta_norms_kernel(const X *a, const X *b, X *dx, const X *c, const X *d,
                                                              X *dy, int n, int m, GridRed g) {
    double v[2] = {0.0, 0.0};
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < (long long)n + m;
         i += (long long)gridDim.x * blockDim.x) {
        if (i < n) {
            const double t = (double)a[i] - (double)b[i];
            dx[i] = (X)t;
            v[0] += (double)(X)t * (double)(X)t;
        } else {
            const long long j = i - n;
            const double t = (double)c[j] - (double)d[j];
            dy[j] = (X)t;
            v[1] += (double)(X)t * (double)(X)t;
        }
    }
    grid_reduce<2, 0>(g, v);
}

// ---------------------------------------------------------------------------
// Projected Barzilai-Borwein inner solve (reference pdhg_core_op.cu:495-669,
// kernels pdhcg_kernels.cu:392-405,497-636), device resident: the stopping
// test runs in the finisher of the reducing kernels and clears bb_active,
// which turns every later launch of the same solve into a no-op.
// ---------------------------------------------------------------------------

// M = diag(H) + 1/tau (non-positive -> 1), M^-1, tol_scale = sqrt(sum|M| / n).
template <class X>
__global__ void __launch_bounds__(kBlock) bb_refresh_kernel(const double *diag_h, double inv_tau, X *m_diag, X *m_inv,
                                                             int n, GridRed g, DevScalars *ds) {
    double v[1] = {0.0};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double mm = diag_h[i] + inv_tau;
        if (mm <= 0.0) mm = 1.0;
        m_diag[i] = (X)mm;
        m_inv[i] = (X)(1.0 / mm);
        v[0] += fabs(mm);
    }
    if (grid_reduce<1, 0>(g, v) && threadIdx.x == 0) ds->tol_scale = n > 0 ? sqrt(v[0] / (double)n) : 1.0;
}

// First projected gradient step from x_k (reference :525-559 with kernels :511-534,591-615).
template <class X>
__global__ void __launch_bounds__(kBlock) bb_init_kernel(const X *aty, X *grad, X *dirn, const X *x, X *xp,
                                                          const X *c, const X *qx, const X *lb, const X *ub,
                                                          const X *m_inv, const X *m_diag, double alpha0, int n,
                                                          GridRed g, DevScalars *ds, double tol, int limit,
                                                          int defer) {
    const bool pre = m_inv != nullptr;
    const double a0 = pre ? alpha0 * (ds->tol_scale * ds->tol_scale) : alpha0;
    double v[1] = {0.0};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const double gr = (double)qx[i] + (double)c[i] - (double)aty[i];
        const double xi = (double)x[i];
        double nx = pre ? xi - a0 * (double)m_inv[i] * gr : xi - a0 * gr;
        nx = fmax((double)lb[i], fmin(nx, (double)ub[i]));
        xp[i] = (X)nx;
        grad[i] = (X)gr;
        const double s = (double)(X)((double)(X)nx - xi);
        dirn[i] = (X)s;
        v[0] += pre ? s * ((double)m_diag[i] * s) : s * s;
    }
    if (defer) return;  // cones: the step is re-measured after the projection
    if (grid_reduce<1, 0>(g, v) && threadIdx.x == 0) {
        ds->bb_iter = 1;
        if (pre) ds->bb_stms = v[0];
        ds->bb_nrm = sqrt(v[0]);
        ds->bb_active = (1 < limit && !(ds->bb_nrm <= tol)) ? 1 : 0;
    }
}

// g_new = Qx + c - A'y + (x_p - x)/tau; tmp = s'(g_new - g); alpha
// (reference :599-650 with kernels :536-555,392-400).
template <class X>
__global__ void __launch_bounds__(kBlock) bb_gradient_kernel(const X *xp, const X *x, const X *c, const X *aty,
                                                              const X *qx, X *grad, const X *dirn, double inv_tau,
                                                              int n, int precond, GridRed g, DevScalars *ds) {
    if (ds->bb_active == 0) return;
    double v[1] = {0.0};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const double last = (double)grad[i];
        const double cur = (double)qx[i] + (double)c[i] - (double)aty[i] + inv_tau * ((double)xp[i] - (double)x[i]);
        const double dg = (double)(X)(cur - last);
        grad[i] = (X)cur;
        v[0] += (double)dirn[i] * dg;
    }
    if (grid_reduce<1, 0>(g, v) && threadIdx.x == 0) {
        ds->bb_tmp = v[0];
        ds->bb_alpha = precond ? ds->bb_stms / v[0] : (ds->bb_nrm * ds->bb_nrm) / v[0];
    }
}

// x_p <- P(x_p - alpha [M^-1] g), s = step, next stopping test
// (reference kernels :557-575,617-636 and loop head :575-597).
template <class X>
__global__ void __launch_bounds__(kBlock) bb_direction_kernel(X *xp, const X *grad, X *dirn, const X *lb, const X *ub,
                                                               const X *m_inv, const X *m_diag, int n, GridRed g,
                                                               DevScalars *ds, double tol, int limit, int defer) {
    if (ds->bb_active == 0) return;
    const bool pre = m_inv != nullptr;
    const double a = ds->bb_alpha;
    double v[1] = {0.0};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const double cur = (double)xp[i];
        double nx = pre ? cur - a * (double)m_inv[i] * (double)grad[i] : cur - a * (double)grad[i];
        nx = fmax((double)lb[i], fmin(nx, (double)ub[i]));
        const double s = (double)(X)((double)(X)nx - cur);
        dirn[i] = (X)s;
        xp[i] = (X)nx;
        v[0] += pre ? s * ((double)m_diag[i] * s) : s * s;
    }
    if (defer) return;
    if (grid_reduce<1, 0>(g, v) && threadIdx.x == 0) {
        const int it = ds->bb_iter + 1;
        ds->bb_iter = it;
        if (pre) ds->bb_stms = v[0];
        ds->bb_nrm = sqrt(v[0]);
        ds->bb_active = (it < limit && !(ds->bb_nrm <= tol)) ? 1 : 0;
    }
}

// With cones: after the cone projection of x_p the step is s = x_p - base
// (base = x_k for the first step, the pre-step snapshot afterwards) and the
// stopping test uses that s (reference :561-566,652-657 and :575-597).
template <class X>
__global__ void __launch_bounds__(kBlock) bb_cone_step_kernel(const X *xp, const X *base, X *dirn, const X *m_diag,
                                                               int n, int first, GridRed g, DevScalars *ds, double tol,
                                                               int limit) {
    if (!first && ds->bb_active == 0) return;
    const bool pre = m_diag != nullptr;
    double v[1] = {0.0};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const double s = (double)(X)((double)xp[i] - (double)base[i]);
        dirn[i] = (X)s;
        v[0] += pre ? s * ((double)m_diag[i] * s) : s * s;
    }
    if (grid_reduce<1, 0>(g, v) && threadIdx.x == 0) {
        const int it = first ? 1 : ds->bb_iter + 1;
        ds->bb_iter = it;
        if (pre) ds->bb_stms = v[0];
        ds->bb_nrm = sqrt(v[0]);
        ds->bb_active = (it < limit && !(ds->bb_nrm <= tol)) ? 1 : 0;
    }
}

// Projected-gradient point for the conic stationarity residual
// (reference pdhcg_kernels.cu:75-91): p = P_box(x - t (c_eff - A'y)).
template <class X>
__global__ void __launch_bounds__(kBlock) projected_gradient_point_kernel(X *pgp, const X *xp, const X *c, const X *qx,
                                                                           const X *aty, const X *lb, const X *ub,
                                                                           double t, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const double g = (double)c[i] + (qx ? (double)qx[i] : 0.0) - (double)aty[i];
        pgp[i] = (X)fmax((double)lb[i], fmin((double)xp[i] - t * g, (double)ub[i]));
    }
}

// dres = the larger-magnitude of dres and (x - p)/t D_c (reference
// pdhcg_kernels.cu:93-109), then the dual-residual norms of the whole vector:
// out = [sum dres^2, max |dres|].
template <class X>
__global__ void __launch_bounds__(kBlock) augment_dual_residual_kernel(double *dres, const X *xp, const X *pgp,
                                                                        const double *var_rescale, double t, int n,
                                                                        GridRed g) {
    double v[2] = {0.0, 0.0};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double r = ((double)xp[i] - (double)pgp[i]) / t * var_rescale[i];
        if (!isfinite(r)) r = copysign(INFINITY, r);
        double dr = dres[i];
        if (fabs(r) > fabs(dr)) dr = r;
        dres[i] = dr;
        v[0] += dr * dr;
        v[1] = fmax(v[1], abs_or_inf(dr));
    }
    grid_reduce<1, 1>(g, v);
}

// Dual objective correction for cone slots: the cone dual slack is
// c_eff - A'y (reference pdhcg_cone_common_kernels.cu:23-40), stored into
// dslack there; out = [sum over cone slots of (c_eff - A'y - slack_box) x].
template <class X>
__global__ void __launch_bounds__(kBlock) cone_slack_correction_kernel(const X *c, const X *qx, const X *aty,
                                                                        X *dslack, const X *xp, const char *mask,
                                                                        int n, GridRed g) {
    double v[1] = {0.0};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        if (mask[i]) {
            const double cs = (double)c[i] + (qx ? (double)qx[i] : 0.0) - (double)aty[i];
            v[0] += (cs - (double)dslack[i]) * (double)xp[i];
            dslack[i] = (X)cs;
        }
    grid_reduce<1, 0>(g, v);
}

// x_bar = 2 x_p - x_k, inner count, optional Halpern (reference :664-668, kernel :577-589).
template <class X>
__global__ void __launch_bounds__(kBlock) bb_final_kernel(X *x, const X *xp, X *xr, const X *x0, int n, int halpern,
                                                           double w, double rho, DevScalars *ds) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const double xi = (double)x[i];
        const double ref = 2.0 * (double)xp[i] - xi;
        xr[i] = (X)ref;
        if (halpern) {
            const double refl = rho * ref + (1.0 - rho) * xi;
            x[i] = (X)(w * refl + (1.0 - w) * (double)x0[i]);
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) ds->inner_total += ds->bb_iter - 1;
}

// ---------------------------------------------------------------------------
// Low-rank product R'DRx with R dense row-major (k x n)
// ---------------------------------------------------------------------------

constexpr int kRxChunk = 2048;  // columns of R per workgroup (x chunk staged in LDS)
constexpr int kRxMaxRank = 256;

// Per-workgroup partial sums of R x over a chunk of kRxChunk columns: the x
// chunk is staged once in LDS, then each wave streams whole contiguous row
// segments R[j, chunk] (coalesced, four loads in flight per lane) and reduces
// them with one butterfly per (row, chunk).  The last workgroup folds the
// partials in workgroup order, applies D and writes u = D R x (FP64).
// D: dkind 1 diagonal, 2 dense (symmetric part taken from the upper triangle
// of the row-major array, as the reference's lower-fill Dsymv on the
// column-major view does, reference pdhg_core_op.cu:272-303).
template <class V, class X, bool Gated>
__global__ void __launch_bounds__(kBlock) rx_dense_kernel(const V *R, const X *x, int n, int k, double *partials,
                                                           unsigned *counter, double *u, const double *ddiag,
                                                           const double *ddense, int dkind, const int *gate) {
    if (Gated && *gate == 0) return;
    __shared__ double xs[kRxChunk];
    __shared__ double rx[kRxMaxRank];
    __shared__ int last;
    const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
    const long long c0 = (long long)blockIdx.x * kRxChunk;
    const int len = (int)((n - c0) < kRxChunk ? (n - c0) : kRxChunk);
    for (int t = threadIdx.x; t < len; t += kBlock) xs[t] = (double)x[c0 + t];
    __syncthreads();
    for (int j = wave; j < k; j += kBlock / kWave) {
        const V *row = R + (size_t)j * n + c0;
        double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
        int t = lane;
        for (; t + 3 * kWave < len; t += 4 * kWave) {
            s0 += (double)row[t] * xs[t];
            s1 += (double)row[t + kWave] * xs[t + kWave];
            s2 += (double)row[t + 2 * kWave] * xs[t + 2 * kWave];
            s3 += (double)row[t + 3 * kWave] * xs[t + 3 * kWave];
        }
        for (; t < len; t += kWave) s0 += (double)row[t] * xs[t];
        double s = wave_sum((s0 + s1) + (s2 + s3));
        if (lane == 0) partials[(size_t)j * gridDim.x + blockIdx.x] = s;
    }
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) last = (atomicAdd(counter, 1u) == gridDim.x - 1) ? 1 : 0;
    __syncthreads();
    if (!last) return;
    __threadfence();
    // fold: one wave per row j, lanes stride over the workgroups (coalesced)
    for (int j = wave; j < k; j += kBlock / kWave) {
        const double *P = partials + (size_t)j * gridDim.x;
        double a = 0.0;
        for (int b = lane; b < (int)gridDim.x; b += kWave) a += coherent_load(P + b);
        a = wave_sum(a);
        if (lane == 0) rx[j] = a;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < k; i += blockDim.x) {
        double r;
        if (dkind == 1) r = rx[i] * ddiag[i];
        else if (dkind == 2) {
            r = 0.0;
            for (int j = 0; j < k; ++j) {
                const int a = i < j ? i : j, b = i < j ? j : i;
                r += ddense[(size_t)a * k + b] * rx[j];
            }
        } else r = rx[i];
        u[i] = r;
    }
    if (threadIdx.x == 0) *counter = 0u;
}

// qx = R'u (+ qx, or + qdiag .* x)  with R dense row-major, one thread per column.
// Gated launches (inside the BB loop) are a separate instantiation so that
// profiles separate them from the always-running ones.
template <class V, class X, bool Gated>
__global__ void __launch_bounds__(kBlock) rtu_dense_kernel(const V *R, const double *u, int n, int k, X *qx,
                                                            const X *qdiag, const X *x, int accumulate,
                                                            const int *gate) {
    if (Gated && *gate == 0) return;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double s = 0.0;
        for (int j = 0; j < k; ++j) s += (double)R[(size_t)j * n + i] * u[j];
        if (qdiag) s = s + (double)qdiag[i] * (double)x[i];
        else if (accumulate) s = s + (double)qx[i];
        qx[i] = (X)s;
    }
}

// u = D rx for a CSR-stored R (rx from an SpMV), single block.
template <class X>
__global__ void apply_middle_kernel(const X *rx, double *u, int k, const double *ddiag, const double *ddense, int dkind,
                                    const int *gate) {
    if (gate && *gate == 0) return;
    for (int i = threadIdx.x; i < k; i += blockDim.x) {
        double r;
        if (dkind == 1) r = (double)rx[i] * ddiag[i];
        else if (dkind == 2) {
            r = 0.0;
            for (int j = 0; j < k; ++j) {
                const int a = i < j ? i : j, b = i < j ? j : i;
                r += ddense[(size_t)a * k + b] * (double)rx[j];
            }
        } else r = (double)rx[i];
        u[i] = r;
    }
}

// ---------------------------------------------------------------------------
// KKT residuals (reference pdhg_core_op.cu:1037-1294, kernels :638-752)
// ---------------------------------------------------------------------------

// Rows: primal residual (v - P(v)) D_r and dual-objective row terms.
// out = [sum pres^2, sum row terms, max |pres|]
template <class X>
__global__ void __launch_bounds__(kBlock) residual_rows_kernel(const X *ax, const X *g, const X *cl, const X *cu,
                                                                const X *yp, const double *con_rescale,
                                                                const X *cl_fin, const X *cu_fin, int m, GridRed gr) {
    double v[3] = {0.0, 0.0, 0.0};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gridDim.x * blockDim.x) {
        const double val = (double)ax[i] + (double)g[i];
        const double pj = fmax((double)cl[i], fmin(val, (double)cu[i]));
        const double pres = (val - pj) * con_rescale[i];
        const double y = (double)yp[i];
        v[0] += pres * pres;
        v[1] += fmax(y, 0.0) * (double)cl_fin[i] + fmin(y, 0.0) * (double)cu_fin[i] - (double)g[i] * y;
        v[2] = fmax(v[2], abs_or_inf(pres));
    }
    grid_reduce<2, 1>(gr, v);
}

// Columns: dual residual, dual slack and objective terms.  qp = 0 uses the
// stored LP dual slack; qp = 1 recomputes it from the projected gradient.
// out = [sum dres^2, c'x, x'Qx, slack'x, sum (Qx D_c)^2, sum (A'y D_c)^2,
//        max|dres|, max|Qx D_c|, max|A'y D_c|]
template <class X>
__global__ void __launch_bounds__(kBlock) residual_cols_kernel(const X *c, const X *aty, const X *qx, const X *xp,
                                                                const X *lb, const X *ub, X *dslack,
                                                                const double *var_rescale, double tau, int qp, int n,
                                                                double *dres_out, GridRed gr) {
    double v[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const double x = (double)xp[i], a = (double)aty[i], ci = (double)c[i], dc = var_rescale[i];
        double dres, ds, q = 0.0;
        if (qp) {
            q = (double)qx[i];
            const double grd = q + ci - a;
            const double t = x - tau * grd;
            const double pj = fmax((double)lb[i], fmin((double)ub[i], t));
            ds = (pj - t) / tau;
            dres = (grd - ds) * dc;
            dslack[i] = (X)ds;
        } else {
            ds = (double)dslack[i];
            dres = (ci - a - ds) * dc;
        }
        if (dres_out) dres_out[i] = dres;  // cones adjust it before the norms are taken
        v[0] += dres * dres;
        v[1] += ci * x;
        v[2] += x * q;
        v[3] += ds * x;
        const double qs = q * dc, as = a * dc;
        v[4] += qs * qs;
        v[5] += as * as;
        v[6] = fmax(v[6], abs_or_inf(dres));
        v[7] = fmax(v[7], abs_or_inf(qs));
        v[8] = fmax(v[8], abs_or_inf(as));
    }
    grid_reduce<6, 3>(gr, v);
}

// ---------------------------------------------------------------------------
// Infeasibility rays (reference pdhg_core_op.cu:1296-1429, kernels :754-859)
// ---------------------------------------------------------------------------

// Projects the ray estimates; out = [max|dx|, max|dy|]
template <class X>
__global__ void __launch_bounds__(kBlock) ray_project_kernel(X *dx, const X *lb, const X *ub, X *dy, const X *cl,
                                                              const X *cu, int n, int m, GridRed gr) {
    double v[2] = {0.0, 0.0};
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < (long long)n + m;
         i += (long long)gridDim.x * blockDim.x) {
        if (i < n) {
            double d = (double)dx[i];
            if (isfinite((double)lb[i])) d = fmax(d, 0.0);
            if (isfinite((double)ub[i])) d = fmin(d, 0.0);
            dx[i] = (X)d;
            v[0] = fmax(v[0], abs_or_inf(d));
        } else {
            const long long j = i - n;
            double d = (double)dy[j];
            if (!isfinite((double)cl[j])) d = fmin(d, 0.0);
            if (!isfinite((double)cu[j])) d = fmax(d, 0.0);
            dy[j] = (X)d;
            v[1] = fmax(v[1], abs_or_inf(d));
        }
    }
    grid_reduce<0, 2>(gr, v);
}

// out = [sum dual-ray row terms, max primal infeasibility]
template <class X>
__global__ void __launch_bounds__(kBlock) ray_rows_kernel(const X *dy, const X *ax, const X *g, const X *cl,
                                                           const X *cu, const X *cl_fin, const X *cu_fin,
                                                           const double *con_rescale, int m, GridRed gr) {
    double v[2] = {0.0, 0.0};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gridDim.x * blockDim.x) {
        const double d = (double)dy[i];
        v[0] += fmax(d, 0.0) * (double)cl_fin[i] + fmin(d, 0.0) * (double)cu_fin[i] - (double)g[i] * d;
        const double pv = (double)ax[i];
        const double inf = (fmax(0.0, -pv) * (double)isfinite((double)cl[i]) +
                            fmax(0.0, pv) * (double)isfinite((double)cu[i])) * con_rescale[i];
        v[1] = fmax(v[1], abs_or_inf(inf));
    }
    grid_reduce<1, 1>(gr, v);
}

// out = [c'dx, sum dual-slack ray terms, max dual infeasibility]
template <class X>
__global__ void __launch_bounds__(kBlock) ray_cols_kernel(const X *c, const X *dx, const X *aty, const X *lb,
                                                           const X *ub, const X *lb_fin, const X *ub_fin,
                                                           const double *var_rescale, int n, GridRed gr) {
    double v[3] = {0.0, 0.0, 0.0};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        v[0] += (double)c[i] * (double)dx[i];
        const double na = -(double)aty[i];
        v[1] += fmax(na, 0.0) * (double)lb_fin[i] + fmin(na, 0.0) * (double)ub_fin[i];
        const double inf = (fmax(0.0, na) * (double)!isfinite((double)lb[i]) -
                            fmin(0.0, na) * (double)!isfinite((double)ub[i])) * var_rescale[i];
        v[2] = fmax(v[2], abs_or_inf(inf));
    }
    grid_reduce<2, 1>(gr, v);
}

// out = [max |a|]
template <class X>
__global__ void __launch_bounds__(kBlock) inf_norm_kernel(const X *a, int n, GridRed gr) {
    double v[1] = {0.0};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        v[0] = fmax(v[0], abs_or_inf((double)a[i]));
    grid_reduce<0, 1>(gr, v);
}

// ---------------------------------------------------------------------------
// Result (reference pdhg_core_op.cu:847-857,1444-1456; kernels :308-327,861-894)
// ---------------------------------------------------------------------------
template <class X>
__global__ void __launch_bounds__(kBlock) result_kernel(const X *xp, const X *yp, const X *c, const X *qx,
                                                         const X *aty, const X *lb, const X *ub,
                                                         const double *var_rescale, const double *con_rescale,
                                                         double ovr, double cbr, int n, int m, double *x_out,
                                                         double *y_out, double *rc_out) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < (long long)n + m;
         i += (long long)gridDim.x * blockDim.x) {
        if (i < n) {
            double gr = (double)c[i];
            if (qx) gr += (double)qx[i];
            double rc = (gr - (double)aty[i]) * var_rescale[i] / ovr;
            if (!isfinite((double)lb[i])) rc = fmin(rc, 0.0);
            if (!isfinite((double)ub[i])) rc = fmax(rc, 0.0);
            rc_out[i] = rc;
            x_out[i] = (double)xp[i] / var_rescale[i] / cbr;
        } else {
            const long long j = i - n;
            y_out[j] = (double)yp[j] / con_rescale[j] / ovr;
        }
    }
}

// ---------------------------------------------------------------------------
// Power iterations for ||A||_2 and the extreme eigenvalues of a sparse Q
// (reference pdhg_core_op.cu:1498-1788), device resident with a convergence
// flag (pi_active) so the host only polls between chunks.
// ---------------------------------------------------------------------------

// sigma^2 = ne'e, ne -= sigma^2 e, residual test; on continue ne += sigma^2 e
// and inv = 1/||ne||.  out of the reduction: [ne'e] then [|ne|^2].
template <class X>
__global__ void __launch_bounds__(kBlock) pi_dot_kernel(const X *ne, const X *e, int m, GridRed gr, DevScalars *ds) {
    if (ds->pi_active == 0) return;
    double v[1] = {0.0};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gridDim.x * blockDim.x)
        v[0] += (double)ne[i] * (double)e[i];
    if (grid_reduce<1, 0>(gr, v) && threadIdx.x == 0) ds->pi_value = v[0];
}

template <class X>
__global__ void __launch_bounds__(kBlock) pi_residual_kernel(X *ne, const X *e, int m, double tol, GridRed gr,
                                                              DevScalars *ds) {
    if (ds->pi_active == 0) return;
    const double s2 = ds->pi_value;
    double v[1] = {0.0};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gridDim.x * blockDim.x) {
        const double t = (double)ne[i] + (-s2) * (double)e[i];
        ne[i] = (X)t;
        v[0] += t * t;
    }
    if (grid_reduce<1, 0>(gr, v) && threadIdx.x == 0) {
        if (sqrt(v[0]) < tol) ds->pi_active = 0;
    }
}

template <class X>
__global__ void __launch_bounds__(kBlock) pi_restore_kernel(X *ne, const X *e, int m, GridRed gr, DevScalars *ds) {
    if (ds->pi_active == 0) return;
    const double s2 = ds->pi_value;
    double v[1] = {0.0};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gridDim.x * blockDim.x) {
        const double t = (double)ne[i] + s2 * (double)e[i];
        ne[i] = (X)t;
        v[0] += t * t;
    }
    if (grid_reduce<1, 0>(gr, v) && threadIdx.x == 0) {
        const double nrm = sqrt(v[0]);
        ds->pi_inv = 1.0 / sqrt(nrm * nrm);
        ds->pi_iter += 1;
    }
}

// Eigenvalue power iteration pieces: inv = 1/||v||.
template <class X>
__global__ void __launch_bounds__(kBlock) pi_norm_kernel(const X *a, int n, GridRed gr, DevScalars *ds) {
    if (ds->pi_active == 0) return;
    double v[1] = {0.0};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        v[0] += (double)a[i] * (double)a[i];
    if (grid_reduce<1, 0>(gr, v) && threadIdx.x == 0) ds->pi_inv = 1.0 / sqrt(v[0]);
}

// shifted = -av + shift*v (shift = 0 means shifted = av); lambda = v'shifted;
// stop when |lambda - old| < tol after the first iteration; else v = shifted.
template <class X>
__global__ void __launch_bounds__(kBlock) pi_rayleigh_kernel(const X *v_in, X *av, double shift, int use_shift, int n,
                                                              double tol, GridRed gr, DevScalars *ds) {
    if (ds->pi_active == 0) return;
    double v[1] = {0.0};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double t = (double)av[i];
        if (use_shift) {
            t = -t + shift * (double)v_in[i];
            av[i] = (X)t;
        }
        v[0] += (double)v_in[i] * t;
    }
    if (grid_reduce<1, 0>(gr, v) && threadIdx.x == 0) {
        const double old = ds->pi_value;
        ds->pi_old = old;
        ds->pi_value = v[0];
        const int it = ds->pi_iter;
        ds->pi_iter = it + 1;
        if (it > 0 && fabs(v[0] - old) < tol) ds->pi_active = 0;
    }
}

}  // namespace pdhcg
