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
    // When set, sigma and w come from device memory: dyn = (tau, sigma, k0)
    // with w = (k0 + pos + 1) / (k0 + pos + 2), so that one recorded launch
    // serves every window of iterations (graph replay).
    const double *dyn = nullptr;
    int pos = 0;
    template <class Acc> __device__ void operator()(int r, Acc s) const {
        double sigma = this->sigma, w = this->w;
        if (dyn) {
            sigma = dyn[1];
            const double k = dyn[2] + pos;
            w = (k + 1.0) / (k + 2.0);
        }
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
    const double *dyn = nullptr;  // as in EpiDualStep: tau = dyn[0]
    int pos = 0;
    template <class Acc> __device__ void operator()(int r, Acc s) const {
        double tau = this->tau, w = this->w;
        if (dyn) {
            tau = dyn[0];
            const double k = dyn[2] + pos;
            w = (k + 1.0) / (k + 2.0);
        }
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

// out = (double)in on two vectors at once (the FP32 iterates handed to the
// FP64 evaluation of the mixed-precision solver).
template <class X>
__global__ void __launch_bounds__(kBlock) widen2_kernel(const X *a, double *a_out, int na, const X *b, double *b_out,
                                                         int nb) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < (long long)na + nb;
         i += (long long)gridDim.x * blockDim.x) {
        if (i < na) a_out[i] = (double)a[i];
        else b_out[i - na] = (double)b[i - na];
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
    grid_partials<1, 0>(g, v);
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
    grid_partials<2, 0>(g, v);
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
                                                             int n, GridRed g) {
    double v[1] = {0.0};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double mm = diag_h[i] + inv_tau;
        if (mm <= 0.0) mm = 1.0;
        m_diag[i] = (X)mm;
        m_inv[i] = (X)(1.0 / mm);
        v[0] += fabs(mm);
    }
    grid_partials<1, 0>(g, v);
}

struct TailTolScale {  // tol_scale = sqrt(sum |M| / n)
    int n;
    __device__ void operator()(const double *r, DevScalars *ds) const {
        ds->tol_scale = n > 0 ? sqrt(r[0] / (double)n) : 1.0;
    }
};

// First projected gradient step from x_k (reference :525-559 with kernels :511-534,591-615).
template <class X>
__global__ void __launch_bounds__(kBlock) bb_init_kernel(const X *aty, X *grad, X *dirn, const X *x, X *xp,
                                                          const X *c, const X *qx, const X *lb, const X *ub,
                                                          const X *m_inv, const X *m_diag, double alpha0, int n,
                                                          GridRed g, const DevScalars *ds, int defer) {
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
    grid_partials<1, 0>(g, v);
}

// Stopping test after a step (reference :575-597): iteration count, |s| or
// sqrt(s'Ms), and the active flag.  first: the step of bb_init (iteration 1).
struct TailBBStep {
    int first, precond, limit;
    double tol;
    __device__ void operator()(const double *r, DevScalars *ds) const {
        const int it = first ? 1 : ds->bb_iter + 1;
        ds->bb_iter = it;
        if (precond) ds->bb_stms = r[0];
        ds->bb_nrm = sqrt(r[0]);
        ds->bb_active = (it < limit && !(ds->bb_nrm <= tol)) ? 1 : 0;
    }
};

// g_new = Qx + c - A'y + (x_p - x)/tau; tmp = s'(g_new - g); alpha
// (reference :599-650 with kernels :536-555,392-400).
template <class X>
__global__ void __launch_bounds__(kBlock) bb_gradient_kernel(const X *xp, const X *x, const X *c, const X *aty,
                                                              const X *qx, X *grad, const X *dirn, double inv_tau,
                                                              int n, GridRed g, const DevScalars *ds) {
    if (ds->bb_active == 0) return;
    double v[1] = {0.0};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const double last = (double)grad[i];
        const double cur = (double)qx[i] + (double)c[i] - (double)aty[i] + inv_tau * ((double)xp[i] - (double)x[i]);
        const double dg = (double)(X)(cur - last);
        grad[i] = (X)cur;
        v[0] += (double)dirn[i] * dg;
    }
    grid_partials<1, 0>(g, v);
}

struct TailBBAlpha {  // alpha = s'Ms / s'dg or |s|^2 / s'dg
    int precond;
    __device__ void operator()(const double *r, DevScalars *ds) const {
        ds->bb_tmp = r[0];
        ds->bb_alpha = precond ? ds->bb_stms / r[0] : (ds->bb_nrm * ds->bb_nrm) / r[0];
    }
};

// x_p <- P(x_p - alpha [M^-1] g), s = step, next stopping test
// (reference kernels :557-575,617-636 and loop head :575-597).
template <class X>
__global__ void __launch_bounds__(kBlock) bb_direction_kernel(X *xp, const X *grad, X *dirn, const X *lb, const X *ub,
                                                               const X *m_inv, const X *m_diag, int n, GridRed g,
                                                               const DevScalars *ds, int defer) {
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
    grid_partials<1, 0>(g, v);
}

// With cones: after the cone projection of x_p the step is s = x_p - base
// (base = x_k for the first step, the pre-step snapshot afterwards) and the
// stopping test uses that s (reference :561-566,652-657 and :575-597).
template <class X>
__global__ void __launch_bounds__(kBlock) bb_cone_step_kernel(const X *xp, const X *base, X *dirn, const X *m_diag,
                                                               int n, int first, GridRed g, const DevScalars *ds) {
    if (!first && ds->bb_active == 0) return;
    const bool pre = m_diag != nullptr;
    double v[1] = {0.0};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const double s = (double)(X)((double)xp[i] - (double)base[i]);
        dirn[i] = (X)s;
        v[0] += pre ? s * ((double)m_diag[i] * s) : s * s;
    }
    grid_partials<1, 0>(g, v);
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
    grid_partials<1, 1>(g, v);
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
    grid_partials<1, 0>(g, v);
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

constexpr int kRxLaneCols = 4;                      // columns of R per lane
constexpr int kRxWaveCols = kWave * kRxLaneCols;    // columns of R per wave (one chunk)
constexpr int kRxMaxRank = 256;

inline int rx_chunks(int n) { return (n + kRxWaveCols - 1) / kRxWaveCols; }

// R is read once per pass: non-temporal loads keep it from displacing the
// reused vectors in the caches.
template <class V> __device__ __forceinline__ double stream_load(const V *p) {
    return (double)__builtin_nontemporal_load(p);
}
template <class V> struct Vec2;
template <> struct Vec2<double> {
    typedef double type __attribute__((ext_vector_type(2)));
};
template <> struct Vec2<float> {
    typedef float type __attribute__((ext_vector_type(2)));
};
// Two adjacent elements; p must be aligned to 2 elements.
template <class V> __device__ __forceinline__ void stream_load2(const V *p, double (&o)[2]) {
    typedef typename Vec2<V>::type T;
    const T t = __builtin_nontemporal_load(reinterpret_cast<const T *>(p));
    o[0] = (double)t[0];
    o[1] = (double)t[1];
}

// Dense R storage: column tiles of kRxWaveCols columns x k rows, each tile
// row-major and contiguous, the last tile zero padded.  One wave's share of
// R x (a tile) is one contiguous block, and R'u reads 64-column row pieces
// kRxWaveCols elements apart.
__host__ __device__ __forceinline__ size_t rx_tile_index(int j, long long i, int k) {
    return ((size_t)(i / kRxWaveCols) * k + j) * kRxWaveCols + (size_t)(i % kRxWaveCols);
}

// Sums four per-lane values s[0..3] over the wave with a transposed
// butterfly: after the xor-32 and xor-16 exchanges each lane holds one row's
// partial, so only 7 exchanges are needed instead of 24.  Returns the total of
// row (lane / 16) in lanes 0, 16, 32 and 48.  Fixed combination order.
__device__ __forceinline__ double wave_sum4(const double (&s)[4], int lane) {
    const bool hi = lane >= 32, mid = (lane & 16) != 0;
    const double k0 = hi ? s[2] : s[0], k1 = hi ? s[3] : s[1];
    const double g0 = hi ? s[0] : s[2], g1 = hi ? s[1] : s[3];
    const double t0 = k0 + __shfl_xor(g0, 32, kWave);
    const double t1 = k1 + __shfl_xor(g1, 32, kWave);
    double u = (mid ? t1 : t0) + __shfl_xor(mid ? t0 : t1, 16, kWave);
#pragma unroll
    for (int off = 8; off > 0; off >>= 1) u += __shfl_xor(u, off, kWave);
    return u;
}

// Column of R x handled by lane `lane` in slot q: two adjacent columns per
// 16-byte load, pairs 128 columns apart (2 lane + 128 (q / 2) + q % 2).
__device__ __forceinline__ int rx_lane_col(int lane, int q) { return 2 * lane + (q / 2) * (kRxWaveCols / 2) + (q % 2); }
static_assert(kRxLaneCols == 4, "two column pairs per lane");

// RU rows of a tile over this lane's kRxLaneCols columns: all RU * 2 pair
// loads are issued before the multiply-adds (one register block).
template <int RU, class V>
__device__ __forceinline__ void rx_rows(const V *T, int j0, int lane, const double (&xv)[kRxLaneCols], double (&s)[RU]) {
    double r[RU][kRxLaneCols];
#pragma unroll
    for (int a = 0; a < RU; ++a)
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            double pr[2];
            stream_load2(T + (size_t)(j0 + a) * kRxWaveCols + rx_lane_col(lane, 2 * h), pr);
            r[a][2 * h] = pr[0];
            r[a][2 * h + 1] = pr[1];
        }
#pragma unroll
    for (int a = 0; a < RU; ++a) {
        double t = 0.0;
#pragma unroll
        for (int q = 0; q < kRxLaneCols; ++q) t += r[a][q] * xv[q];
        s[a] = t;
    }
}

// Dense tiles of a CSR R whose rows have strictly increasing columns: entry
// e (row found by a binary search of rp) goes to its tile slot, converted to
// the storage type V; slots without an entry stay zero.
template <class V>
__global__ void r_tiles_kernel(const int *rp, const int *ci, const double *v, int k, long long nnz, V *Rd) {
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < nnz;
         e += (long long)gridDim.x * blockDim.x) {
        int lo = 0, hi = k;  // rp[lo] <= e < rp[hi]
        while (hi - lo > 1) {
            const int mid = lo + (hi - lo) / 2;
            if (rp[mid] <= e) lo = mid;
            else hi = mid;
        }
        Rd[rx_tile_index(lo, ci[e], k)] = (V)v[e];
    }
}

// Pass 1 of u = D R x: each wave owns one column tile (kRxWaveCols columns,
// x for them held in registers), streams its rows eight at a time (32 loads
// in flight per lane, one contiguous block per wave) and writes one partial
// sum per (row, tile): partials[j * nchunks + chunk].  Padding columns of the
// last tile are zero in R and take x = 0.
template <class V, class X, bool Gated>
__global__ void __launch_bounds__(kBlock) rx_partial_kernel(const V *R, const X *x, int n, int k, double *partials,
                                                             int nchunks, const int *gate) {
    if (Gated && *gate == 0) return;
    const int lane = threadIdx.x % kWave;
    const int chunk = blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave;
    if (chunk >= nchunks) return;
    const long long base = (long long)chunk * kRxWaveCols;
    const V *T = R + (size_t)chunk * k * kRxWaveCols;
    double xv[kRxLaneCols];
#pragma unroll
    for (int q = 0; q < kRxLaneCols; ++q) {
        const long long c = base + rx_lane_col(lane, q);
        xv[q] = c < n ? (double)x[c] : 0.0;
    }
    auto emit = [&](const double (&s4)[4], int j0) {
        const double tot = wave_sum4(s4, lane);
        const int r = lane >> 4;
        if ((lane & 15) == 0 && j0 + r < k) partials[(size_t)(j0 + r) * nchunks + chunk] = tot;
    };
    int j = 0;
    for (; j + 8 <= k; j += 8) {
        double s[8];
        rx_rows<8>(T, j, lane, xv, s);
        emit({s[0], s[1], s[2], s[3]}, j);
        emit({s[4], s[5], s[6], s[7]}, j + 4);
    }
    for (; j < k; j += 4) {
        double s4[4] = {0.0, 0.0, 0.0, 0.0};
        for (int a = 0; a < 4 && j + a < k; ++a) {
            double s1[1];
            rx_rows<1>(T, j + a, lane, xv, s1);
            s4[a] = s1[0];
        }
        emit(s4, j);
    }
}

// Pass 2: workgroup j folds row j's partials in chunk order into rx[j]; the
// last workgroup to finish applies D and writes u = D R x (FP64).
// D: dkind 1 diagonal, 2 dense (symmetric part taken from the upper triangle
// of the row-major array, as the reference's lower-fill Dsymv on the
// column-major view does, reference pdhg_core_op.cu:272-303).
template <bool Gated>
__global__ void __launch_bounds__(kBlock) rx_fold_kernel(const double *partials, int nchunks, int k, double *rx,
                                                          unsigned *counter, double *u, const double *ddiag,
                                                          const double *ddense, int dkind, const int *gate) {
    if (Gated && *gate == 0) return;
    __shared__ double lds[16];
    __shared__ double rxs[kRxMaxRank];
    __shared__ int last;
    const int j = blockIdx.x;
    const double *P = partials + (size_t)j * nchunks;
    double a[4] = {0.0, 0.0, 0.0, 0.0};  // four loads in flight, fixed order
    int c = threadIdx.x;
    for (; c + 3 * kBlock < nchunks; c += 4 * kBlock)
#pragma unroll
        for (int q = 0; q < 4; ++q) a[q] += P[c + q * kBlock];
    for (; c < nchunks; c += kBlock) a[0] += P[c];
    double v[1] = {(a[0] + a[1]) + (a[2] + a[3])};
    block_reduce<1, 0>(v, lds);
    if (threadIdx.x == 0) {
        rx[j] = v[0];
        __threadfence();
        last = (atomicAdd(counter, 1u) == gridDim.x - 1) ? 1 : 0;
    }
    __syncthreads();
    if (!last) return;
    __threadfence();
    for (int i = threadIdx.x; i < k; i += blockDim.x) rxs[i] = coherent_load(rx + i);
    __syncthreads();
    for (int i = threadIdx.x; i < k; i += blockDim.x) {
        double r;
        if (dkind == 1) r = rxs[i] * ddiag[i];
        else if (dkind == 2) {
            r = 0.0;
            for (int jj = 0; jj < k; ++jj) {
                const int a = i < jj ? i : jj, b = i < jj ? jj : i;
                r += ddense[(size_t)a * k + b] * rxs[jj];
            }
        } else r = rxs[i];
        u[i] = r;
    }
    if (threadIdx.x == 0) *counter = 0u;
}

// Adjacent columns i0 .. i0 + RtuCols<V> - 1 of R'u, one 16-byte load per row
// of R (two FP64 or four FP32 columns), each column's sum over the rows in
// order j = 0 .. k - 1.  Columns past n lie in the zero padding of the last
// tile.  On MI355X this pattern streams the tiles at about 6.4 TB/s in FP64,
// against 6.2 TB/s for one column per lane; four FP64 columns per lane (32
// bytes) need too many registers for ten rows in flight (4.0 TB/s).  The
// loaded values stay in the storage type until the multiply-add.
template <class V> struct RtuCols {
    static constexpr int value = 16 / (int)sizeof(V);
};
template <class V>
__device__ __forceinline__ void rtu_cols(const V *R, const double *u, long long i0, int k,
                                         double (&s)[RtuCols<V>::value]) {
    constexpr int RC = RtuCols<V>::value;
    typedef V T __attribute__((ext_vector_type(RC)));
    constexpr int JU = 10;  // rows of R loaded before the multiply-adds
    const V *col = R + rx_tile_index(0, i0, k);  // row j of these columns at col[j * kRxWaveCols]
#pragma unroll
    for (int q = 0; q < RC; ++q) s[q] = 0.0;
    int j = 0;
    for (; j + JU <= k; j += JU) {
        T r[JU];
#pragma unroll
        for (int a = 0; a < JU; ++a)
            r[a] = __builtin_nontemporal_load(reinterpret_cast<const T *>(col + (size_t)(j + a) * kRxWaveCols));
#pragma unroll
        for (int a = 0; a < JU; ++a)
#pragma unroll
            for (int q = 0; q < RC; ++q) s[q] += (double)r[a][q] * u[j + a];
    }
    for (; j < k; ++j) {
        const T r = __builtin_nontemporal_load(reinterpret_cast<const T *>(col + (size_t)j * kRxWaveCols));
#pragma unroll
        for (int q = 0; q < RC; ++q) s[q] += (double)r[q] * u[j];
    }
}

// qx = R'u (+ qx, or + qdiag .* x)  with R dense in column tiles,
// RtuCols<V> adjacent columns per thread (grid over ceil(n / RtuCols<V>) threads).
// Gated launches (inside the BB loop) are a separate instantiation so that
// profiles separate them from the always-running ones.
template <class V, class X, bool Gated>
__global__ void __launch_bounds__(kBlock) rtu_dense_kernel(const V *R, const double *u, int n, int k, X *qx,
                                                            const X *qdiag, const X *x, int accumulate,
                                                            const int *gate) {
    if (Gated && *gate == 0) return;
    constexpr int RC = RtuCols<V>::value;
    const long long groups = ((long long)n + RC - 1) / RC;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < groups;
         t += (long long)gridDim.x * blockDim.x) {
        const long long i0 = RC * t;
        // the epilogue's operands are loaded before the pass over R, so that
        // their latency overlaps it
        double ea[RC], eb[RC];
#pragma unroll
        for (int q = 0; q < RC; ++q) {
            const long long i = i0 + q;
            ea[q] = eb[q] = 0.0;
            if (i < n && qdiag) {
                ea[q] = (double)qdiag[i];
                eb[q] = (double)x[i];
            } else if (i < n && accumulate) {
                ea[q] = (double)qx[i];
            }
        }
        double s[RC];
        rtu_cols(R, u, i0, k, s);
#pragma unroll
        for (int q = 0; q < RC; ++q) {
            const long long i = i0 + q;
            if (i >= n) break;
            double v = s[q];
            if (qdiag) v = v + ea[q] * eb[q];
            else if (accumulate) v = v + ea[q];
            qx[i] = (X)v;
        }
    }
}

// BB body on the dense low-rank path: the R'u pass fused with the gradient
// step of bb_gradient_kernel.  Qx_p = R'u (+ diag(Q) x_p, or + the sparse-Q
// product already in qx_sparse); g = Qx_p + c - A'y + (x_p - x)/tau;
// partial of s'(g - g_old); grad = g.  Qx_p itself is not stored.
// RtuCols<V> adjacent columns per thread, as rtu_dense_kernel.
template <class V, class X>
__global__ void __launch_bounds__(kBlock) rtu_grad_kernel(const V *R, const double *u, int n, int k,
                                                           const X *qx_sparse, const X *qdiag, const X *xp, const X *x,
                                                           const X *c, const X *aty, X *grad, const X *dirn,
                                                           double inv_tau, GridRed g, const DevScalars *ds) {
    if (ds->bb_active == 0) return;
    constexpr int RC = RtuCols<V>::value;
    double v[1] = {0.0};
    const long long groups = ((long long)n + RC - 1) / RC;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < groups;
         t += (long long)gridDim.x * blockDim.x) {
        const long long i0 = RC * t;
        // the epilogue's operands are loaded before the pass over R, so that
        // their latency overlaps it
        double exp_[RC], eq[RC], ec[RC], ex[RC], eg[RC], ed[RC];
#pragma unroll
        for (int q = 0; q < RC; ++q) {
            const long long i = i0 + q;
            exp_[q] = eq[q] = ec[q] = ex[q] = eg[q] = ed[q] = 0.0;
            if (i < n) {
                exp_[q] = (double)xp[i];
                eq[q] = qdiag ? (double)qdiag[i] : qx_sparse ? (double)qx_sparse[i] : 0.0;
                ec[q] = (double)c[i] - (double)aty[i];
                ex[q] = (double)x[i];
                eg[q] = (double)grad[i];
                ed[q] = (double)dirn[i];
            }
        }
        double s[RC];
        rtu_cols(R, u, i0, k, s);
#pragma unroll
        for (int q = 0; q < RC; ++q) {
            const long long i = i0 + q;
            if (i >= n) break;
            double sq = s[q];
            const double xpi = exp_[q];
            if (qdiag) sq = sq + eq[q] * xpi;
            else if (qx_sparse) sq = sq + eq[q];
            const double qv = (double)(X)sq;  // the stored product's rounding, as bb_gradient_kernel sees it
            const double cur = qv + ec[q] + inv_tau * (xpi - ex[q]);
            const double dg = (double)(X)(cur - eg[q]);
            grad[i] = (X)cur;
            v[0] += ed[q] * dg;
        }
    }
    grid_partials<1, 0>(g, v);
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
    grid_partials<2, 1>(gr, v);
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
    grid_partials<6, 3>(gr, v);
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
    grid_partials<0, 2>(gr, v);
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
    grid_partials<1, 1>(gr, v);
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
    grid_partials<2, 1>(gr, v);
}

// out = [max |a|]
template <class X>
__global__ void __launch_bounds__(kBlock) inf_norm_kernel(const X *a, int n, GridRed gr) {
    double v[1] = {0.0};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        v[0] = fmax(v[0], abs_or_inf((double)a[i]));
    grid_partials<0, 1>(gr, v);
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
    grid_partials<1, 0>(gr, v);
}

struct TailPiValue {
    __device__ void operator()(const double *r, DevScalars *ds) const { ds->pi_value = r[0]; }
};

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
    grid_partials<1, 0>(gr, v);
}

struct TailPiResidual {  // stop when the eigen-residual is below tol
    double tol;
    __device__ void operator()(const double *r, DevScalars *ds) const {
        if (sqrt(r[0]) < tol) ds->pi_active = 0;
    }
};

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
    grid_partials<1, 0>(gr, v);
}

struct TailPiRestore {
    __device__ void operator()(const double *r, DevScalars *ds) const {
        const double nrm = sqrt(r[0]);
        ds->pi_inv = 1.0 / sqrt(nrm * nrm);
        ds->pi_iter += 1;
    }
};

// Eigenvalue power iteration pieces: inv = 1/||v||.
template <class X>
__global__ void __launch_bounds__(kBlock) pi_norm_kernel(const X *a, int n, GridRed gr, DevScalars *ds) {
    if (ds->pi_active == 0) return;
    double v[1] = {0.0};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        v[0] += (double)a[i] * (double)a[i];
    grid_partials<1, 0>(gr, v);
}

struct TailPiNorm {
    __device__ void operator()(const double *r, DevScalars *ds) const { ds->pi_inv = 1.0 / sqrt(r[0]); }
};

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
    grid_partials<1, 0>(gr, v);
}

struct TailPiRayleigh {  // stop when the Rayleigh quotient settles
    double tol;
    __device__ void operator()(const double *r, DevScalars *ds) const {
        const double old = ds->pi_value;
        ds->pi_old = old;
        ds->pi_value = r[0];
        const int it = ds->pi_iter;
        ds->pi_iter = it + 1;
        if (it > 0 && fabs(r[0] - old) < tol) ds->pi_active = 0;
    }
};

}  // namespace pdhcg
