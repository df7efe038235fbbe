// Batched solve: many independent problems in one persistent launch, one
// 256-thread workgroup per problem running the whole restarted Halpern PDHG
// loop of reference src/solver.cu:121-169 on the device (evaluation cadence,
// KKT residuals, restarts with the PID primal weight, the BB inner solve,
// affine cones, termination).  The arithmetic follows the single-problem
// solver step for step (reference src/pdhg_core_op.cu); every vector lives in
// HBM in a per-problem slab and every reduction is a fixed-order workgroup
// reduction, so a problem's trajectory does not depend on the rest of the
// batch.  Scalars are held by every thread (each thread computes the same
// value from the same broadcast reductions), so the control flow is uniform.
// The once-per-iteration helpers are not inlined: each gets its own register
// allocation, which keeps the kernel at 248 VGPRs without spills, i.e. two
// waves per SIMD with 512 threads per problem (fully inlined at 256 threads
// it needed 294 registers and ran one problem per CU).
//
// Supported per problem: box bounds, rows [A; F] in CSR, objective of type
// LP, diagonal Q, or dense low rank R'DR (k <= 64) plus a diagonal part, and
// affine cones (SOC / RSOC up to 64 vector slots, exponential, power) on the
// rows.  Variable cones and off-diagonal sparse Q are handled by the
// single-problem solver only.
#pragma once

#include <cfloat>
#include <type_traits>

#include "dev_common.hpp"
#include "dev_cones.hpp"

namespace pdhcg {
namespace batch {

// Threads per problem: the launch picks 512 (two waves per SIMD at 248
// VGPRs) or 256 by timing (batch.hip); device loops stride by the template
// parameter BT, a compile-time constant.
constexpr int kBT = 512, kBTSmall = 256;
constexpr int kBW = kBT / kWave;  // waves of the largest workgroup (reduction slots)
constexpr int kMaxRank = 256;  // low-rank terms up to this rank (R x, u live in the slab)

enum QKind { kQNone = 0, kQDiag = 1, kQLowRank = 2 };

// Loop parameters shared by the batch (pdhg_parameters_t fields).
struct Params {
    int ef, iteration_limit, norm_inf, inner_limit, precond, stall_iters;
    double handover;  // FP32 phase: KKT error below which the FP64 kernel takes over
    double eps_opt, eps_feas, eps_inf, rho;
    double inner_tol0, inner_min_tol;
    double i_smooth, k_p, k_i, k_d, suff, nec, art;
};

// One problem: sizes, pool offsets and scaled-problem scalars.
struct Inst {
    int n, m, k, qkind, dkind, has_affine, ncones, cone0;
    int nlong;                     // SOC/RSOC cones above kMaxThreadCone slots: the last nlong cones
    int a_rp, a_ci, at_rp, at_ci;  // int-pool offsets of [A;F] and its transpose
    int a_long, at_long;           // 1: a wavefront per row, 0: a thread per row
    long long a_v, at_v;           // value-pool offsets of the matrix values
    long long vec;                 // slab offset of the problem's n- and m-vectors
    long long wvec;                // FP64 slab offset of its k-vectors and per-cone values
    long long rdense, dmat;        // value-pool offsets of R (k x n row-major) and D
    double cbr, ovr, obj_const, obj_norm, bound_norm, step, pw0, nonconvexity;
};

struct Cone {
    int type, start, vdim;
    double alpha;
};

// Per-problem results (the iterates stay in the slab).
struct Out {
    int status, iters, inner;
    double pobj, dobj, abs_pr, abs_du, rel_pr, rel_du, gap, rel_gap;
    double max_pray, max_dray, pray_lin_obj, dray_obj;
};

// Slab layout: n-vectors, then m-vectors (storage type of the kernel); the
// FP64 slab: k-vectors, then per-cone values.
enum NVec { vC, vLB, vUB, vVR, vQD, vDH, vX0, vX, vXP, vXR, vATY, vDS, vGR, vDIR, vQX, vPB, vDX, vMD, vMI, vLBF, vUBF,
            kNVec };
enum MVec { wG, wCL, wCU, wCLF, wCUF, wCR, wY0, wY, wYP, wYR, wAX, wDY, wPRES, wPSL, wAPR, wAMEM, wAFD, kMVec };
constexpr int kKVec = 2;      // rx, rx buffer
constexpr int kConeVals = 3;  // projection warm start, residual warm start, complementarity

inline long long slab_size(int n, int m) { return (long long)kNVec * n + (long long)kMVec * m; }
inline long long wslab_size(int k, int ncones) { return (long long)kKVec * k + (long long)kConeVals * ncones; }

__device__ inline int print_frequency(int it) {
    int step = 10;
    long long thr = 1000;
    while (it >= thr) {
        step *= 10;
        thr *= 10;
    }
    return step;
}

__device__ inline double clampd(double v, double lo, double hi) { return fmax(lo, fmin(v, hi)); }

// Workgroup-wide reductions; every thread receives the result (fixed order).
template <int BT> struct Red {
    double *lds;  // 2 * kBW doubles
    __device__ double sum(double v) const {
        v = wave_sum(v);
        const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
        if (lane == 0) lds[w] = v;
        __syncthreads();
        double s = lds[0];
        for (int i = 1; i < BT / kWave; ++i) s += lds[i];
        __syncthreads();
        return s;
    }
    __device__ double max(double v) const {
        for (int off = kWave / 2; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, kWave));
        const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
        if (lane == 0) lds[w] = v;
        __syncthreads();
        double s = lds[0];
        for (int i = 1; i < BT / kWave; ++i) s = fmax(s, lds[i]);
        __syncthreads();
        return s;
    }
    // L2 or max norm of the values f(i), i < count
    template <class F> __device__ double norm(int count, bool inf, F f) const {
        double a = 0.0;
        for (int i = threadIdx.x; i < count; i += BT) {
            const double v = f(i);
            a = inf ? fmax(a, fabs(v)) : a + v * v;
        }
        return inf ? max(a) : sqrt(sum(a));
    }
    template <class F> __device__ double dot(int count, F f) const {
        double a = 0.0;
        for (int i = threadIdx.x; i < count; i += BT) a += f(i);
        return sum(a);
    }
};

// y = M x for a CSR matrix of `rows` rows.  With one workgroup per problem
// and two waves per SIMD there is little occupancy to hide memory latency,
// so each lane keeps four independent partial sums (four loads in flight)
// and adds them in a fixed order.
template <class T, int BT>
__device__ inline void spmv(const int *rp, const int *ci, const T *v, const T *x, T *y, int rows,
                            int wave_rows) {
    if (wave_rows) {
        const int lane = threadIdx.x % kWave;
        for (int r = threadIdx.x / kWave; r < rows; r += BT / kWave) {
            T s0 = 0, s1 = 0, s2 = 0, s3 = 0;
            const int end = rp[r + 1];
            int e = rp[r] + lane;
            for (; e + 3 * kWave < end; e += 4 * kWave) {
                s0 += v[e] * x[ci[e]];
                s1 += v[e + kWave] * x[ci[e + kWave]];
                s2 += v[e + 2 * kWave] * x[ci[e + 2 * kWave]];
                s3 += v[e + 3 * kWave] * x[ci[e + 3 * kWave]];
            }
            for (; e < end; e += kWave) s0 += v[e] * x[ci[e]];
            const T s = wave_sum((s0 + s1) + (s2 + s3));
            if (lane == 0) y[r] = s;
        }
    } else {
        for (int r = threadIdx.x; r < rows; r += BT) {
            T s0 = 0, s1 = 0, s2 = 0, s3 = 0;
            const int end = rp[r + 1];
            int e = rp[r];
            for (; e + 3 < end; e += 4) {
                s0 += v[e] * x[ci[e]];
                s1 += v[e + 1] * x[ci[e + 1]];
                s2 += v[e + 2] * x[ci[e + 2]];
                s3 += v[e + 3] * x[ci[e + 3]];
            }
            for (; e < end; ++e) s0 += v[e] * x[ci[e]];
            y[r] = (s0 + s1) + (s2 + s3);
        }
    }
    __syncthreads();
}

// Everything one problem's loop needs, as device pointers.  T is the
// storage and product type of the iterates and matrix values (double, or
// float for the FP32 phase of the mixed batch); the k-vectors, per-cone
// values, D and every reduction and scalar stay FP64.
template <class T, int BT> struct Ctx {
    const Params *P;
    Inst I;
    const int *arp, *aci, *atrp, *atci;
    const T *av, *atv, *R;
    const double *D;
    T *V;       // n- and m-vectors of the slab
    double *W;  // k-vectors and per-cone values (FP64)
    const Cone *cones;
    Red<BT> red;
    __device__ T *nv(int which) const { return V + (long long)which * I.n; }
    __device__ T *mv(int which) const { return V + (long long)kNVec * I.n + (long long)which * I.m; }
    __device__ double *kv(int which) const { return W + (long long)which * I.k; }
    __device__ double *cv(int which) const { return W + (long long)kKVec * I.k + (long long)which * I.ncones; }
    __device__ void A(const T *x, T *y) const { spmv<T, BT>(arp, aci, av, x, y, I.m, I.a_long); }
    __device__ void At(const T *y, T *x) const { spmv<T, BT>(atrp, atci, atv, y, x, I.n, I.at_long); }

    // qx = Q x (reference src/pdhg_core_op.cu:305-413): diagonal part plus R'DR x
    __device__ __noinline__ void obj_product(const T *x, T *qx) const {
        const int n = I.n, k = I.k;
        if (I.qkind == kQNone) return;
        const T *qd = nv(vQD);
        if (I.qkind == kQDiag) {
            for (int i = threadIdx.x; i < n; i += BT) qx[i] = qd[i] * x[i];
            __syncthreads();
            return;
        }
        double *rx = kv(0), *rb = kv(1);
        const int lane = threadIdx.x % kWave;
        for (int r = threadIdx.x / kWave; r < k; r += BT / kWave) {
            const T *row = R + (long long)r * n;
            T s0 = 0, s1 = 0, s2 = 0, s3 = 0;  // four loads in flight, fixed order
            int j = lane;
            for (; j + 3 * kWave < n; j += 4 * kWave) {
                s0 += row[j] * x[j];
                s1 += row[j + kWave] * x[j + kWave];
                s2 += row[j + 2 * kWave] * x[j + 2 * kWave];
                s3 += row[j + 3 * kWave] * x[j + 3 * kWave];
            }
            for (; j < n; j += kWave) s0 += row[j] * x[j];
            const double s = (double)wave_sum((s0 + s1) + (s2 + s3));
            if (lane == 0) rx[r] = s;
        }
        __syncthreads();
        // u = D rx: diagonal, or the symmetric part of a dense D from its
        // upper triangle (reference src/pdhg_core_op.cu:272-303)
        if (I.dkind == 1) {
            for (int r = threadIdx.x; r < k; r += BT) rb[r] = rx[r] * D[r];
        } else if (I.dkind == 2) {
            for (int i = threadIdx.x; i < k; i += BT) {
                double a = 0.0;
                for (int j = 0; j < k; ++j) {
                    const int lo = i < j ? i : j, hi = i < j ? j : i;
                    a += D[(long long)lo * k + hi] * rx[j];
                }
                rb[i] = a;
            }
        } else {
            for (int r = threadIdx.x; r < k; r += BT) rb[r] = rx[r];
        }
        __syncthreads();
        for (int j = threadIdx.x; j < n; j += BT) {
            T s0 = 0, s1 = 0, s2 = 0, s3 = 0;  // four loads in flight, fixed order
            int r = 0;
            for (; r + 3 < k; r += 4) {
                s0 += R[(long long)r * n + j] * (T)rb[r];
                s1 += R[(long long)(r + 1) * n + j] * (T)rb[r + 1];
                s2 += R[(long long)(r + 2) * n + j] * (T)rb[r + 2];
                s3 += R[(long long)(r + 3) * n + j] * (T)rb[r + 3];
            }
            for (; r < k; ++r) s0 += R[(long long)r * n + j] * (T)rb[r];
            qx[j] = qd[j] * x[j] + ((s0 + s1) + (s2 + s3));
        }
        __syncthreads();
    }

    // Projection of the m-vector v onto the affine cones, scaling affd
    // (reference src/cone_dispatch.cu:38-57 on the row space).
    __device__ __noinline__ void project_affine(T *v, double *warm) const {
        const T *d = mv(wAFD);
        const int nshort = I.ncones - I.nlong;
        for (int c = threadIdx.x; c < nshort; c += BT) {
            const Cone cn = cones[c];
            const int s = cn.start;
            if (cn.type == CONE_STANDARD_SOC || cn.type == CONE_ROTATED_SOC) {
                double pv[kMaxThreadCone + 2];
                for (int q = 0; q < cn.vdim + 2; ++q) pv[q] = v[s + q];
                if constexpr (std::is_same<T, double>::value) {
                    soc_project_block(pv, cn.vdim, d + s, cn.type == CONE_ROTATED_SOC, warm + c);
                } else {
                    double pd[kMaxThreadCone + 2];
                    for (int q = 0; q < cn.vdim + 2; ++q) pd[q] = d[s + q];
                    soc_project_block(pv, cn.vdim, pd, cn.type == CONE_ROTATED_SOC, warm + c);
                }
                for (int q = 0; q < cn.vdim + 2; ++q) v[s + q] = (T)pv[q];
            } else if (cn.type == CONE_EXPONENTIAL) {
                double o[3];
                exp_project_point(v[s], v[s + 1], v[s + 2], d[s], d[s + 1], d[s + 2], o);
                v[s] = (T)o[0];
                v[s + 1] = (T)o[1];
                v[s + 2] = (T)o[2];
            } else {
                double o[3];
                const double d0 = d[s], d1 = d[s + 1], d2 = d[s + 2];
                power_project_point(v[s] / d0, v[s + 1] / d1, v[s + 2] / d2, d0 * d0, d1 * d1, d2 * d2, cn.alpha, o);
                v[s] = (T)(o[0] * d0);
                v[s + 1] = (T)(o[1] * d1);
                v[s + 2] = (T)(o[2] * d2);
            }
        }
        __syncthreads();
        // long second-order cones: the whole workgroup on one cone at a time
        for (int c = nshort; c < I.ncones; ++c) {
            const Cone cn = cones[c];
            T *pv = v + cn.start;
            const T *pd = d + cn.start;
            auto P = [&](int m) { return (double)pv[m]; };
            auto E = [&](int m) { return (double)pd[m]; };
            auto Set = [&](int m, double val) { pv[m] = (T)val; };
            auto sum = [&](double x) { return red.sum(x); };
            if (cn.type == CONE_ROTATED_SOC) soc_project_coop<true>(P, E, Set, cn.vdim, warm + c, sum, false);
            else soc_project_coop<false>(P, E, Set, cn.vdim, warm + c, sum, false);
            __syncthreads();
        }
    }
};

// A problem that leaves the FP32 phase of the mixed batch before
// terminating (its iterates stopped improving at the FP32 rounding floor).
constexpr int kStalled = 100;
constexpr int kSatStall = 1;  // FP32 inner solves at the inner limit before the hand-over

// Loop state (identical in every thread).  best_err / best_iter track the
// KKT error of the FP32 phase (a problem whose error has not halved for
// max(1000, half its iterations) leaves that phase as kStalled), sat the
// FP32 inner solves in a row that reached the inner iteration limit, and
// eval_now asks for a KKT evaluation at the next pass.
struct State {
    int total, inner_count, inner_total, is_major, do_restart, reason, best_iter, sat, eval_now, pad;
    double best_err;
    double pw, pw_err_sum, pw_last_err, best_pw, best_gap;
    double fpe, init_fpe, last_trial_fpe, inner_tol, cached_inv_tau, tol_scale;
    double abs_pr, abs_du, rel_pr, rel_du, pobj, dobj, gap, rel_gap;
    double max_pray, max_dray, pray_lin_obj, dray_obj;
};

// Projected BB inner solve of the primal prox (reference src/pdhg_core_op.cu:495-669).
template <class T, int BT> __device__ __noinline__ void bb_prox(const Ctx<T, BT> &C, State &S, double tau) {
    const Params &P = *C.P;
    const int n = C.I.n;
    T *x = C.nv(vX), *xp = C.nv(vXP), *xr = C.nv(vXR), *grad = C.nv(vGR), *dirn = C.nv(vDIR), *qx = C.nv(vQX),
           *pb = C.nv(vPB), *md = C.nv(vMD), *mi = C.nv(vMI);
    const T *c = C.nv(vC), *lb = C.nv(vLB), *ub = C.nv(vUB), *aty = C.nv(vATY), *dh = C.nv(vDH);
    const double inv = 1.0 / tau;
    double a0 = 1.0 / inv;
    const bool pre = P.precond != 0;
    if (pre && S.cached_inv_tau != inv) {
        double sacc = 0.0;
        for (int i = threadIdx.x; i < n; i += BT) {
            double mm = dh[i] + inv;
            if (mm <= 0.0) mm = 1.0;
            md[i] = mm;
            mi[i] = 1.0 / mm;
            sacc += fabs(mm);
        }
        const double sum = C.red.sum(sacc);
        S.cached_inv_tau = inv;
        S.tol_scale = n > 0 ? sqrt(sum / (double)n) : 1.0;
    }
    if (pre) a0 *= S.tol_scale * S.tol_scale;
    C.obj_product(x, qx);
    for (int i = threadIdx.x; i < n; i += BT) {
        const double g = qx[i] + c[i] - aty[i];
        double nx = pre ? x[i] - a0 * mi[i] * g : x[i] - a0 * g;
        nx = clampd(nx, lb[i], ub[i]);
        xp[i] = nx;
        grad[i] = g;
        dirn[i] = nx - x[i];
    }
    __syncthreads();
    int it = 1;
    while (it < P.inner_limit) {
        double stms = 0.0, nrm;
        if (pre) {
            stms = C.red.dot(n, [&](int i) { return dirn[i] * (md[i] * dirn[i]); });
            nrm = sqrt(stms);
        } else {
            nrm = sqrt(C.red.dot(n, [&](int i) { return dirn[i] * dirn[i]; }));
        }
        if (nrm <= S.inner_tol) break;
        C.obj_product(xp, qx);
        double tacc = 0.0;
        for (int i = threadIdx.x; i < n; i += BT) {
            const double last = grad[i];
            const double cur = qx[i] + c[i] - aty[i] + inv * (xp[i] - x[i]);
            pb[i] = cur - last;
            grad[i] = cur;
            tacc += dirn[i] * pb[i];
        }
        const double tmp = C.red.sum(tacc);
        const double alpha = pre ? stms / tmp : (nrm * nrm) / tmp;
        for (int i = threadIdx.x; i < n; i += BT) {
            const double cur = xp[i];
            double nx = pre ? cur - alpha * mi[i] * grad[i] : cur - alpha * grad[i];
            nx = clampd(nx, lb[i], ub[i]);
            dirn[i] = nx - cur;
            xp[i] = nx;
        }
        __syncthreads();
        it++;
    }
    for (int i = threadIdx.x; i < n; i += BT) xr[i] = 2.0 * xp[i] - x[i];
    __syncthreads();
    S.inner_total += it - 1;
    if (!std::is_same<T, double>::value) S.sat = (P.inner_limit > 1 && it >= P.inner_limit) ? S.sat + 1 : 0;
}

// One PDHG step (reference src/pdhg_core_op.cu:705-830).
template <class T, int BT> __device__ __noinline__ void pdhg_step(const Ctx<T, BT> &C, State &S) {
    const Inst &I = C.I;
    const int n = I.n, m = I.m;
    double tau = I.step / S.pw;
    if (I.nonconvexity < 0) {
        tau = fmax(tau, -1.01 * fmin(0.0, I.nonconvexity));
        tau /= 100;
    }
    const double sigma = I.step * S.pw;
    T *x = C.nv(vX), *xp = C.nv(vXP), *xr = C.nv(vXR), *aty = C.nv(vATY), *ds = C.nv(vDS);
    const T *c = C.nv(vC), *lb = C.nv(vLB), *ub = C.nv(vUB), *qd = C.nv(vQD);
    C.At(C.mv(wY), aty);
    const int tc = S.total;
    const bool store = S.is_major || ((tc + 2) % print_frequency(tc + 2)) == 0;
    if (I.qkind == kQNone) {
        for (int i = threadIdx.x; i < n; i += BT) {
            const double xi = x[i];
            const double t = xi - tau * (c[i] - aty[i]);
            const double pj = clampd(t, lb[i], ub[i]);
            xr[i] = 2.0 * pj - xi;
            if (store) {
                xp[i] = pj;
                ds[i] = (pj - t) / tau;
            }
        }
        __syncthreads();
    } else if (I.qkind == kQDiag) {
        for (int i = threadIdx.x; i < n; i += BT) {
            const double xi = x[i];
            const double t = (xi - tau * (c[i] - aty[i])) / (1.0 + tau * qd[i]);
            const double pj = clampd(t, lb[i], ub[i]);
            xr[i] = 2.0 * pj - xi;
            if (store) xp[i] = pj;
        }
        __syncthreads();
    } else {
        bb_prox(C, S, tau);
    }
    S.inner_total++;
    T *ax = C.mv(wAX), *y = C.mv(wY), *yp = C.mv(wYP), *yr = C.mv(wYR), *apr = C.mv(wAPR);
    const T *g = C.mv(wG), *cl = C.mv(wCL), *cu = C.mv(wCU);
    C.A(xr, ax);
    if (m == 0) return;
    if (I.has_affine) {  // reference src/pdhg_core_op.cu:809-829
        for (int i = threadIdx.x; i < m; i += BT) apr[i] = clampd(ax[i] + g[i] - y[i] / sigma, cl[i], cu[i]);
        __syncthreads();
        C.project_affine(apr, C.cv(0));
    }
    for (int i = threadIdx.x; i < m; i += BT) {
        const double cur = y[i];
        const double v = ax[i] + g[i];
        const double pj = I.has_affine ? apr[i] : clampd(v - cur / sigma, cl[i], cu[i]);
        const double nx = cur - sigma * v + sigma * pj;
        if (store) yp[i] = nx;
        yr[i] = 2.0 * nx - cur;
    }
    __syncthreads();
}

// reference src/pdhg_core_op.cu:832-845
template <class T, int BT> __device__ inline void halpern(const Ctx<T, BT> &C, const State &S) {
    const double rho = C.P->rho;
    const double wgt = (double)(S.inner_count + 1) / (S.inner_count + 2);
    T *x = C.nv(vX), *y = C.mv(wY);
    const T *xr = C.nv(vXR), *x0 = C.nv(vX0), *yr = C.mv(wYR), *y0 = C.mv(wY0);
    for (int i = threadIdx.x; i < C.I.n; i += BT) {
        const double r = rho * xr[i] + (1.0 - rho) * x[i];
        x[i] = wgt * r + (1.0 - wgt) * x0[i];
    }
    for (int i = threadIdx.x; i < C.I.m; i += BT) {
        const double r = rho * yr[i] + (1.0 - rho) * y[i];
        y[i] = wgt * r + (1.0 - wgt) * y0[i];
    }
    __syncthreads();
}

// reference src/pdhg_core_op.cu:973-1035
template <class T, int BT> __device__ __noinline__ void fixed_point_error(const Ctx<T, BT> &C, State &S) {
    const int n = C.I.n, m = C.I.m;
    T *dx = C.nv(vDX), *dy = C.mv(wDY), *aty = C.nv(vATY);
    const T *xr = C.nv(vXR), *x = C.nv(vX), *yr = C.mv(wYR), *y = C.mv(wY);
    for (int i = threadIdx.x; i < n; i += BT) dx[i] = xr[i] - x[i];
    for (int i = threadIdx.x; i < m; i += BT) dy[i] = yr[i] - y[i];
    __syncthreads();
    C.At(dy, aty);
    const double dn = sqrt(C.red.dot(m, [&](int i) { return dy[i] * dy[i]; }));
    const double pn = sqrt(C.red.dot(n, [&](int i) { return dx[i] * dx[i]; }));
    const double mov = pn * pn * S.pw + dn * dn / S.pw;
    const double cross = C.red.dot(n, [&](int i) { return aty[i] * dx[i]; });
    S.fpe = sqrt(mov + 2 * C.I.step * cross);
    if (C.I.qkind == kQLowRank)
        S.inner_tol = fmin(S.inner_tol, fmax(0.0005 * pn / C.I.step * S.pw, C.P->inner_min_tol));
}

// reference src/pdhg_core_op.cu:859-931
template <class T, int BT> __device__ __noinline__ void restart(const Ctx<T, BT> &C, State &S) {
    const Params &P = *C.P;
    const int n = C.I.n, m = C.I.m;
    T *x0 = C.nv(vX0), *x = C.nv(vX), *y0 = C.mv(wY0), *y = C.mv(wY);
    const T *xp = C.nv(vXP), *yp = C.mv(wYP);
    const double pd = sqrt(C.red.dot(n, [&](int i) { return (xp[i] - x0[i]) * (xp[i] - x0[i]); }));
    const double dd = sqrt(C.red.dot(m, [&](int i) { return (yp[i] - y0[i]) * (yp[i] - y0[i]); }));
    const double ratio = S.rel_du / S.rel_pr;
    if (pd > 1e-16 && dd > 1e-16 && pd < 1e12 && dd < 1e12 && ratio > 1e-8 && ratio < 1e8) {
        const double err = log(dd) - log(pd) - log(S.pw);
        S.pw_err_sum *= P.i_smooth;
        S.pw_err_sum += err;
        const double derr = err - S.pw_last_err;
        S.pw *= exp(P.k_p * err + P.k_i * S.pw_err_sum + P.k_d * derr);
        S.pw_last_err = err;
    } else {
        S.pw = S.best_pw;
        S.pw_err_sum = 0.0;
        S.pw_last_err = 0.0;
    }
    const double gapv = fabs(log10(S.rel_du / S.rel_pr));
    if (gapv < S.best_gap) {
        S.best_gap = gapv;
        S.best_pw = S.pw;
    }
    for (int i = threadIdx.x; i < n; i += BT) x0[i] = x[i] = xp[i];
    for (int i = threadIdx.x; i < m; i += BT) y0[i] = y[i] = yp[i];
    __syncthreads();
    S.inner_count = 0;
    S.last_trial_fpe = INFINITY;
}

// reference src/utils.cu:291-322
template <class T, int BT> __device__ inline bool should_restart(const Ctx<T, BT> &C, State &S) {
    const Params &P = *C.P;
    bool r = false;
    if (S.total == P.ef) r = true;
    else if (S.total > P.ef) {
        if (S.fpe <= P.suff * S.init_fpe) r = true;
        if (S.fpe <= P.nec * S.init_fpe && S.fpe > S.last_trial_fpe) r = true;
        if (S.inner_count >= P.art * S.total) r = true;
    }
    S.last_trial_fpe = S.fpe;
    return r;
}

// KKT residuals (reference src/pdhg_core_op.cu:1037-1294).
template <class T, int BT> __device__ __noinline__ void residual(const Ctx<T, BT> &C, State &S) {
    const Inst &I = C.I;
    const int n = I.n, m = I.m;
    const bool linf = C.P->norm_inf != 0;
    const double tau = I.step / S.pw;
    T *ax = C.mv(wAX), *aty = C.nv(vATY), *qx = C.nv(vQX), *ds = C.nv(vDS), *pres = C.mv(wPRES),
           *psl = C.mv(wPSL), *amem = C.mv(wAMEM);
    const T *xp = C.nv(vXP), *yp = C.mv(wYP), *g = C.mv(wG), *cl = C.mv(wCL), *cu = C.mv(wCU),
                 *clf = C.mv(wCLF), *cuf = C.mv(wCUF), *cr = C.mv(wCR), *c = C.nv(vC), *lb = C.nv(vLB),
                 *ub = C.nv(vUB), *vr = C.nv(vVR), *afd = C.mv(wAFD);
    C.A(xp, ax);
    C.At(yp, aty);
    C.obj_product(xp, qx);
    for (int i = threadIdx.x; i < m; i += BT) {
        const double v = ax[i] + g[i];
        const double pj = clampd(v, cl[i], cu[i]);
        pres[i] = (v - pj) * cr[i];
        psl[i] = fmax(yp[i], 0.0) * clf[i] + fmin(yp[i], 0.0) * cuf[i] - g[i] * yp[i];
    }
    const bool lp = I.qkind == kQNone;
    double dres_acc = 0.0;
    for (int j = threadIdx.x; j < n; j += BT) {
        double dres;
        if (lp) {
            dres = (c[j] - aty[j] - ds[j]) * vr[j];
        } else {
            const double gr = qx[j] + c[j] - aty[j];
            const double t = xp[j] - tau * gr;
            const double pj = fmax(lb[j], fmin(ub[j], t));
            const double dsj = (pj - t) / tau;
            dres = (gr - dsj) * vr[j];
            ds[j] = dsj;
        }
        dres_acc = linf ? fmax(dres_acc, fabs(dres)) : dres_acc + dres * dres;
    }
    __syncthreads();
    double abs_du = linf ? C.red.max(dres_acc) : sqrt(C.red.sum(dres_acc));
    double amem_norm = 0.0, acomp_norm = 0.0;
    if (I.has_affine) {  // reference src/pdhg_core_op.cu:1125-1129,114-201
        double *acomp = C.cv(2);
        for (int i = threadIdx.x; i < m; i += BT) {
            pres[i] = clampd(ax[i] + g[i], cl[i], cu[i]);
            amem[i] = 0.0;
        }
        __syncthreads();
        C.project_affine(pres, C.cv(1));
        for (int k = threadIdx.x; k < I.ncones; k += BT) {
            const Cone cn = C.cones[k];
            const int L = (cn.type == CONE_STANDARD_SOC || cn.type == CONE_ROTATED_SOC) ? cn.vdim + 2 : 3;
            double dotv = 0.0;
            for (int i = cn.start; i < cn.start + L; ++i) {
                amem[i] = -yp[i];
                dotv += yp[i] * (ax[i] + g[i]);
            }
            acomp[k] = fabs(dotv) / I.cbr;
        }
        __syncthreads();
        C.project_affine(amem, C.cv(1));
        for (int i = threadIdx.x; i < m; i += BT) {
            pres[i] = (ax[i] + g[i] - pres[i]) * cr[i];
            amem[i] *= afd[i];
        }
        __syncthreads();
        amem_norm = C.red.norm(m, linf, [&](int i) { return amem[i]; });
        acomp_norm = C.red.norm(I.ncones, linf, [&](int i) { return acomp[i]; });
    }
    S.abs_pr = C.red.norm(m, linf, [&](int i) { return pres[i]; }) / I.cbr;
    const double lin_pr = S.abs_pr;
    if (I.has_affine)
        abs_du = linf ? fmax(fmax(abs_du, amem_norm), acomp_norm)
                      : sqrt(abs_du * abs_du + amem_norm * amem_norm + acomp_norm * acomp_norm);
    S.abs_du = abs_du / I.ovr;
    const double half_xqx = lp ? 0.0 : 0.5 * C.red.dot(n, [&](int j) { return xp[j] * qx[j]; });
    S.pobj = (C.red.dot(n, [&](int j) { return c[j] * xp[j]; }) + half_xqx) / (I.cbr * I.ovr) + I.obj_const;
    const double base = C.red.dot(n, [&](int j) { return ds[j] * xp[j]; });
    const double sps = C.red.dot(m, [&](int i) { return psl[i]; });
    S.dobj = (base + sps - half_xqx) / (I.cbr * I.ovr) + I.obj_const;
    S.rel_pr = lin_pr / (1.0 + I.bound_norm);
    double den;
    if (lp) den = 1.0 + I.obj_norm;
    else {
        const double qn = C.red.norm(n, linf, [&](int j) { return qx[j] * vr[j]; });
        const double an = C.red.norm(n, linf, [&](int j) { return aty[j] * vr[j]; });
        den = 1.0 + fmax(I.obj_norm, fmax(qn / I.ovr, an / I.ovr));
    }
    S.rel_du = S.abs_du / den;
    S.gap = fabs(S.pobj - S.dobj);
    S.rel_gap = S.gap / (1.0 + fabs(S.pobj) + fabs(S.dobj));
}

// Infeasibility rays (reference src/pdhg_core_op.cu:1296-1429).
template <class T, int BT> __device__ __noinline__ void infeasibility(const Ctx<T, BT> &C, State &S) {
    const Inst &I = C.I;
    const int n = I.n, m = I.m;
    T *dx = C.nv(vDX), *dy = C.mv(wDY), *ax = C.mv(wAX), *aty = C.nv(vATY), *qx = C.nv(vQX);
    const T *lb = C.nv(vLB), *ub = C.nv(vUB), *cl = C.mv(wCL), *cu = C.mv(wCU), *c = C.nv(vC),
                 *clf = C.mv(wCLF), *cuf = C.mv(wCUF), *g = C.mv(wG), *lbf = C.nv(vLBF), *ubf = C.nv(vUBF),
                 *cr = C.mv(wCR), *vr = C.nv(vVR);
    for (int j = threadIdx.x; j < n; j += BT) {
        if (isfinite(lb[j])) dx[j] = fmax(dx[j], 0.0);
        if (isfinite(ub[j])) dx[j] = fmin(dx[j], 0.0);
    }
    for (int i = threadIdx.x; i < m; i += BT) {
        if (!isfinite(cl[i])) dy[i] = fmin(dy[i], 0.0);
        if (!isfinite(cu[i])) dy[i] = fmax(dy[i], 0.0);
    }
    __syncthreads();
    const double prn = C.red.norm(n, true, [&](int j) { return dx[j]; });
    if (prn > 0.0) {
        const double sc = 1.0 / prn;
        for (int j = threadIdx.x; j < n; j += BT) dx[j] *= sc;
        __syncthreads();
    }
    const double drn = C.red.norm(m, true, [&](int i) { return dy[i]; });
    C.A(dx, ax);
    C.At(dy, aty);
    S.pray_lin_obj = C.red.dot(n, [&](int j) { return c[j] * dx[j]; }) / (I.cbr * I.ovr);
    const double sps = C.red.dot(m, [&](int i) {
        return fmax(dy[i], 0.0) * clf[i] + fmin(dy[i], 0.0) * cuf[i] - g[i] * dy[i];
    });
    const double sds = C.red.dot(n, [&](int j) { return fmax(-aty[j], 0.0) * lbf[j] + fmin(-aty[j], 0.0) * ubf[j]; });
    S.dray_obj = (sps + sds) / (I.cbr * I.ovr);
    S.max_pray = C.red.norm(m, true, [&](int i) {
        const double pv = ax[i];
        return (fmax(0.0, -pv) * (double)isfinite(cl[i]) + fmax(0.0, pv) * (double)isfinite(cu[i])) * cr[i];
    });
    if (I.qkind != kQNone) {
        C.obj_product(dx, qx);
        const double qn = C.red.norm(n, true, [&](int j) { return qx[j]; });
        S.max_pray = fmax(S.max_pray, qn / I.ovr);
    }
    const double dsn = C.red.norm(n, true, [&](int j) {
        const double dv = -aty[j];
        return (fmax(0.0, dv) * (double)!isfinite(lb[j]) - fmin(0.0, dv) * (double)!isfinite(ub[j])) * vr[j];
    });
    S.max_dray = dsn;
    const double sc = fmax(drn, dsn);
    if (sc > 0.0) {
        S.max_dray /= sc;
        S.dray_obj /= sc;
    } else {
        S.max_dray = 0.0;
        S.dray_obj = 0.0;
    }
}

// reference src/utils.cu:219-289 (the batch applies no wall-clock limit)
template <class T, int BT> __device__ inline void check_termination(const Ctx<T, BT> &C, State &S) {
    const Params &P = *C.P;
    if (S.rel_du < P.eps_feas && S.rel_pr < P.eps_feas && S.rel_gap < P.eps_opt) {
        S.reason = TERMINATION_REASON_OPTIMAL;
        return;
    }
    if (C.I.has_affine) {
        if (S.total >= P.iteration_limit) S.reason = TERMINATION_REASON_ITERATION_LIMIT;
        return;
    }
    if (S.dray_obj > 0.0 && S.max_dray / S.dray_obj <= P.eps_inf) {
        S.reason = TERMINATION_REASON_PRIMAL_INFEASIBLE;
        return;
    }
    if (S.pray_lin_obj < 0.0 && S.max_pray / (-S.pray_lin_obj) <= P.eps_inf) {
        S.reason = TERMINATION_REASON_DUAL_INFEASIBLE;
        return;
    }
    if (S.total >= P.iteration_limit) S.reason = TERMINATION_REASON_ITERATION_LIMIT;
}

// The persistent launch: workgroup b advances problem b by up to `budget`
// passes of the main loop from its saved state (every loop is bounded), and
// a problem that has terminated leaves A'y and Q x of its final point in the
// slab for the reduced costs and returns at once on later launches.
//
// T = float is the FP32 phase of the mixed batch: iterates, matrix values and
// products in FP32 (vpool, slabs), D, the k-vectors and the cone values in
// FP64 (dpool, wslabs), and a problem whose KKT error stops improving leaves
// with kStalled; the host then widens the iterates into the FP64 slabs and
// the FP64 kernel re-evaluates every problem before it terminates.
template <class T, int BT>
__global__ void __launch_bounds__(BT) batch_solve_kernel(const Params *P, const Inst *insts, const int *ipool,
                                                          const T *vpool, const double *dpool, T *slabs,
                                                          double *wslabs, const Cone *cones, State *states, Out *out,
                                                          int budget) {
    __shared__ double lds[2 * kBW];
    State S = states[blockIdx.x];
    if (S.reason != TERMINATION_REASON_UNSPECIFIED) return;
    constexpr bool fp32 = !std::is_same<T, double>::value;
    Ctx<T, BT> C;
    C.P = P;
    C.I = insts[blockIdx.x];
    C.arp = ipool + C.I.a_rp;
    C.aci = ipool + C.I.a_ci;
    C.atrp = ipool + C.I.at_rp;
    C.atci = ipool + C.I.at_ci;
    C.av = vpool + C.I.a_v;
    C.atv = vpool + C.I.at_v;
    C.R = vpool + C.I.rdense;
    C.D = dpool + C.I.dmat;
    C.V = slabs + C.I.vec;
    C.W = wslabs + C.I.wvec;
    C.cones = cones + C.I.cone0;
    C.red.lds = lds;
    const int ef = P->ef;
    for (int pass = 0; pass < budget; ++pass) {
        if (!(S.total < P->iteration_limit)) {
            S.reason = TERMINATION_REASON_ITERATION_LIMIT;
            residual(C, S);
            break;
        }
        const int tc = S.total;
        if (S.is_major || tc == 0 || S.eval_now || tc % print_frequency(tc) == 0) {
            S.eval_now = 0;
            residual(C, S);
            if (!C.I.has_affine && S.is_major && tc < 3 * ef) infeasibility(C, S);
            check_termination(C, S);
            if (S.reason != TERMINATION_REASON_UNSPECIFIED) break;
            if (fp32) {  // near the tolerance or at the FP32 floor: hand the problem to the FP64 kernel
                const double err = fmax(S.rel_pr, fmax(S.rel_du, S.rel_gap));
                if (err < P->handover) {
                    S.reason = kStalled;
                    break;
                }
                if (err < 0.5 * S.best_err) {
                    S.best_err = err;
                    S.best_iter = tc;
                } else if (tc - S.best_iter >= max(P->stall_iters, tc / 2)) {
                    S.reason = kStalled;
                    break;
                }
            }
        }
        if (fp32 && S.sat >= kSatStall) {
            S.reason = kStalled;
            break;
        }
        if (S.is_major || tc == 0) {
            S.do_restart = should_restart(C, S);
            if (S.do_restart) restart(C, S);
        }
        S.is_major = ((tc + 1) % ef) == 0;
        pdhg_step(C, S);
        if (S.is_major || S.do_restart) {
            fixed_point_error(C, S);
            if (S.do_restart) {
                S.init_fpe = S.fpe;
                S.do_restart = 0;
            }
        }
        halpern(C, S);
        S.inner_count++;
        S.total++;
    }
    if (S.reason != TERMINATION_REASON_UNSPECIFIED && !fp32) {
        C.At(C.mv(wYP), C.nv(vATY));
        C.obj_product(C.nv(vXP), C.nv(vQX));
    }
    if (threadIdx.x == 0) {
        states[blockIdx.x] = S;
        Out o;
        o.status = S.reason;
        o.iters = S.total;
        o.inner = S.inner_total;
        o.pobj = S.pobj;
        o.dobj = S.dobj;
        o.abs_pr = S.abs_pr;
        o.abs_du = S.abs_du;
        o.rel_pr = S.rel_pr;
        o.rel_du = S.rel_du;
        o.gap = S.gap;
        o.rel_gap = S.rel_gap;
        o.max_pray = S.max_pray;
        o.max_dray = S.max_dray;
        o.pray_lin_obj = S.pray_lin_obj;
        o.dray_obj = S.dray_obj;
        out[blockIdx.x] = o;
    }
}

}  // namespace batch
}  // namespace pdhcg
