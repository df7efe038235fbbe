// Host side of the batched solve (pdhcg_solve_batch): each problem is checked,
// scaled on the host exactly as solve_qp_problem scales it, and packed into
// pooled device arrays; the device loop of dev_batch.hpp then advances every
// unfinished problem by a window of iterations per launch until all have
// terminated (or the batch's time limit is reached between windows).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

#include "../../include/pdhcg.h"
#include "../../include/pdhcg_amd.h"
#include "dev_batch.hpp"
#include "host_scale.h"

namespace pdhcg {
namespace batch {
namespace {

// Loop state of one problem, persisted between launches (POD, same layout
// on host and device).
struct HostState {
    int total, inner_count, inner_total, is_major, do_restart, reason, best_iter, sat, eval_now, pad;
    double best_err;
    double pw, pw_err_sum, pw_last_err, best_pw, best_gap;
    double fpe, init_fpe, last_trial_fpe, inner_tol, cached_inv_tau, tol_scale;
    double abs_pr, abs_du, rel_pr, rel_du, pobj, dobj, gap, rel_gap;
    double max_pray, max_dray, pray_lin_obj, dray_obj;
};
static_assert(sizeof(HostState) == sizeof(State), "host and device loop state differ");

inline int grid_of(long long n) { return (int)std::max<long long>(1, std::min<long long>((n + 255) / 256, 8192)); }

__global__ void narrow_kernel(const double *a, float *b, long long n) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        b[i] = (float)a[i];
}

// The iterate and work vectors of problem blockIdx.x widened from the FP32
// slab into the FP64 slab; the problem data (c, bounds, scalings, the
// Jacobi diagonal, g, row bounds) keep their FP64 values.
__global__ void widen_iterates_kernel(const Inst *insts, const float *s32, double *s64) {
    const Inst I = insts[blockIdx.x];
    const int nv[] = {vX0, vX, vXP, vXR, vATY, vDS, vGR, vDIR, vQX, vPB, vDX, vMD, vMI};
    const int mv[] = {wY0, wY, wYP, wYR, wAX, wDY, wPRES, wPSL, wAPR, wAMEM};
    for (int w : nv)
        for (int i = threadIdx.x; i < I.n; i += blockDim.x) {
            const long long o = I.vec + (long long)w * I.n + i;
            s64[o] = (double)s32[o];
        }
    for (int w : mv)
        for (int i = threadIdx.x; i < I.m; i += blockDim.x) {
            const long long o = I.vec + (long long)kNVec * I.n + (long long)w * I.m + i;
            s64[o] = (double)s32[o];
        }
}

// Every problem that left the FP32 phase (terminated there, stalled or at
// the iteration limit) is evaluated again by the FP64 kernel at its next
// pass, which decides its status; the Jacobi metric is rebuilt from the
// FP64 diagonal.
__global__ void resume_fp64_kernel(State *states, int count) {
    for (int b = blockIdx.x * blockDim.x + threadIdx.x; b < count; b += gridDim.x * blockDim.x) {
        State &S = states[b];
        S.reason = TERMINATION_REASON_UNSPECIFIED;
        S.eval_now = 1;
        S.cached_inv_tau = -1.0;
    }
}

// sigma_max(A) by the power iteration on A A' the single-problem solver uses
// (reference src/pdhg_core_op.cu:1673-1788): start rand_r(1234) in [-1, 1].
double host_max_sv(const HostCSR &A, const HostCSR &At, int maxit, double tol) {
    const int m = A.rows, n = A.cols;
    std::vector<double> e((size_t)m), ne((size_t)m), dp((size_t)n);
    unsigned int seed = 1234;
    for (int i = 0; i < m; ++i) e[i] = ((double)rand_r(&seed) / (double)RAND_MAX) * 2.0 - 1.0;
    auto mv = [](const HostCSR &M, const std::vector<double> &x, std::vector<double> &y) {
        for (int r = 0; r < M.rows; ++r) {
            double s = 0.0;
            for (int q = M.rp[r]; q < M.rp[(size_t)r + 1]; ++q) s += M.v[q] * x[M.ci[q]];
            y[r] = s;
        }
    };
    auto nrm = [](const std::vector<double> &v) {
        double s = 0.0;
        for (double a : v) s += a * a;
        return std::sqrt(s);
    };
    double s2 = 1.0, inv = 1.0 / nrm(e);
    for (double &a : e) a *= inv;
    for (int it = 0; it < maxit; ++it) {
        mv(At, e, dp);
        mv(A, dp, ne);
        s2 = 0.0;
        for (int i = 0; i < m; ++i) s2 += ne[i] * e[i];
        for (int i = 0; i < m; ++i) ne[i] += -s2 * e[i];
        if (nrm(ne) < tol) break;
        for (int i = 0; i < m; ++i) ne[i] += s2 * e[i];
        inv = 1.0 / nrm(ne);
        for (double &a : ne) a *= inv;
        std::swap(e, ne);
    }
    return std::sqrt(s2);
}

struct Packed {
    Inst inst;
    ScaledProblem sp;
    std::vector<int> ai;     // A rp, A ci, At rp, At ci
    std::vector<double> av;  // A v, At v, R dense, D
    std::vector<double> slab;   // n- and m-vectors
    std::vector<double> wslab;  // k-vectors and per-cone values
    std::vector<Cone> cones;
    HostState st;
};

// Problems the batch takes; a message names what is not supported.
const char *unsupported(const qp_problem_t *p) {
    if (!p) return "problem is NULL";
    if (p->cones.num_cones > 0) return "variable cone blocks (use solve_qp_problem)";
    if (p->num_quadratic_constraints > 0) return "quadratic constraints";
    if (p->num_rank_lowrank_obj > kMaxRank) return "a low-rank term of rank above 256";
    for (int c = 0; c < p->affine_cones.num_cones; ++c) {
        const cone_type_t ct = p->affine_cones.type[c];
        if (ct == CONE_PSD) return "an affine PSD cone";
    }
    return nullptr;
}

// Scales and packs one problem; returns false (with *why set) for an
// objective the batch does not handle.
bool pack(const qp_problem_t *prob, const pdhg_parameters_t &P, Packed &k, const char **why) {
    k.sp = scale_problem(P, *prob);
    ScaledProblem &sp = k.sp;
    if (sp.qtype == PDHCG_SPARSE_Q) {
        *why = "a sparse Q (use solve_qp_problem)";
        return false;
    }
    if (sp.qtype == PDHCG_LOW_RANK_PLUS_SPARSE_Q)
        for (int i = 0; i < sp.n; ++i)
            for (int q = sp.Q.rp[i]; q < sp.Q.rp[(size_t)i + 1]; ++q)
                if (sp.Q.ci[q] != i) {
                    *why = "a sparse Q with off-diagonal entries (use solve_qp_problem)";
                    return false;
                }
    const int n = sp.n, m = sp.m, r = sp.k;
    Inst &I = k.inst;
    I.n = n;
    I.m = m;
    I.k = 0;
    I.cbr = sp.cbr;
    I.ovr = sp.ovr;
    I.obj_const = sp.obj_const;
    I.obj_norm = sp.obj_norm;
    I.bound_norm = sp.bound_norm;
    I.dkind = (int)sp.dkind;
    I.nonconvexity = 0.0;
    std::vector<double> qd((size_t)n, 0.0), dh((size_t)n, 0.0);
    if (sp.qtype == PDHCG_NON_Q) I.qkind = kQNone;
    else if (sp.qtype == PDHCG_DIAG_Q) {
        I.qkind = kQDiag;
        qd = sp.qdiag;
        double mn = 0.0;
        for (double q : qd) mn = std::fmin(mn, q);
        I.nonconvexity = mn;
    } else {
        I.qkind = kQLowRank;
        I.k = r;
        if (sp.qtype == PDHCG_LOW_RANK_PLUS_SPARSE_Q)
            for (int i = 0; i < n; ++i)
                for (int q = sp.Q.rp[i]; q < sp.Q.rp[(size_t)i + 1]; ++q) qd[i] += sp.Q.v[q];
    }
    // pools: [A rp | A ci | At rp | At ci] and [A v | At v | R | D]
    const HostCSR At = transpose(sp.A);
    I.a_rp = 0;
    I.a_ci = m + 1;
    I.at_rp = I.a_ci + sp.A.nnz();
    I.at_ci = I.at_rp + n + 1;
    k.ai.insert(k.ai.end(), sp.A.rp.begin(), sp.A.rp.end());
    k.ai.insert(k.ai.end(), sp.A.ci.begin(), sp.A.ci.end());
    k.ai.insert(k.ai.end(), At.rp.begin(), At.rp.end());
    k.ai.insert(k.ai.end(), At.ci.begin(), At.ci.end());
    I.a_long = sp.A.nnz() >= 16LL * std::max(m, 1);
    I.at_long = At.nnz() >= 16LL * std::max(n, 1);
    I.a_v = 0;
    I.at_v = sp.A.nnz();
    k.av.insert(k.av.end(), sp.A.v.begin(), sp.A.v.end());
    k.av.insert(k.av.end(), At.v.begin(), At.v.end());
    I.rdense = (long long)k.av.size();
    if (I.qkind == kQLowRank) {
        std::vector<double> Rd((size_t)r * n, 0.0);
        for (int j = 0; j < r; ++j)
            for (int q = sp.R.rp[j]; q < sp.R.rp[(size_t)j + 1]; ++q) Rd[(size_t)j * n + sp.R.ci[q]] += sp.R.v[q];
        // Jacobi diagonal: diag(Q) + diag(R'DR) (reference src/solver_state.cu:280-370)
        dh = qd;
        for (int i = 0; i < n; ++i) {
            double acc = 0.0;
            if (sp.dkind == MiddleKind::Diag) {
                for (int j = 0; j < r; ++j) acc += sp.ddiag[j] * Rd[(size_t)j * n + i] * Rd[(size_t)j * n + i];
            } else if (sp.dkind == MiddleKind::Dense) {
                for (int a = 0; a < r; ++a)
                    for (int b = 0; b < r; ++b)
                        acc += Rd[(size_t)a * n + i] * Rd[(size_t)b * n + i] * sp.ddense[(size_t)a * r + b];
            } else {
                for (int j = 0; j < r; ++j) acc += Rd[(size_t)j * n + i] * Rd[(size_t)j * n + i];
            }
            dh[i] += acc;
        }
        k.av.insert(k.av.end(), Rd.begin(), Rd.end());
    }
    I.dmat = (long long)k.av.size();
    if (sp.dkind == MiddleKind::Diag) k.av.insert(k.av.end(), sp.ddiag.begin(), sp.ddiag.end());
    if (sp.dkind == MiddleKind::Dense) k.av.insert(k.av.end(), sp.ddense.begin(), sp.ddense.end());
    // affine cones on the rows
    const cone_blocks_t *ac = sp.affine_cones;
    I.ncones = (ac && m > 0) ? ac->num_cones : 0;
    I.has_affine = I.ncones > 0;
    // cones in input order, long second-order cones moved to the end (the
    // workgroup projects those one at a time); the per-cone values follow
    // this order
    I.nlong = 0;
    for (int pass = 0; pass < 2; ++pass)
        for (int c = 0; c < I.ncones; ++c) {
            const bool lng = (ac->type[c] == CONE_STANDARD_SOC || ac->type[c] == CONE_ROTATED_SOC) &&
                             ac->v_dim[c] > kMaxThreadCone;
            if (lng != (pass == 1)) continue;
            I.nlong += lng ? 1 : 0;
            k.cones.push_back(Cone{(int)ac->type[c], ac->start_idx[c], ac->v_dim[c],
                                   ac->power_alpha ? ac->power_alpha[c] : 0.0});
        }
    // step size and primal weight (reference src/pdhg_core_op.cu:933-971)
    if (sp.A.nnz() == 0) I.step = 1.0;
    else {
        const double sv = host_max_sv(sp.A, At, P.sv_max_iter, P.sv_tol);
        I.step = sv < 1e-12 ? 1.0 : 0.998 / sv;
    }
    I.pw0 = P.bound_objective_rescaling ? 1.0 : (sp.obj_norm + 1.0) / (sp.bound_norm + 1.0);
    // the slab (FP64), starting point projected onto the box
    k.slab.assign((size_t)slab_size(n, m), 0.0);
    k.wslab.assign((size_t)wslab_size(I.k, I.ncones), 0.0);
    auto N = [&](int w) { return k.slab.data() + (size_t)w * n; };
    auto M = [&](int w) { return k.slab.data() + (size_t)kNVec * n + (size_t)w * m; };
    for (int i = 0; i < n; ++i) {
        N(vC)[i] = sp.c[i];
        N(vLB)[i] = sp.lb[i];
        N(vUB)[i] = sp.ub[i];
        N(vVR)[i] = sp.var_rescale[i];
        N(vQD)[i] = qd[i];
        N(vDH)[i] = dh[i];
        N(vLBF)[i] = std::isfinite(sp.lb[i]) ? sp.lb[i] : 0.0;
        N(vUBF)[i] = std::isfinite(sp.ub[i]) ? sp.ub[i] : 0.0;
        double x0 = sp.primal_start.empty() ? 0.0 : sp.primal_start[i] * sp.var_rescale[i] * sp.cbr;
        x0 = std::fmax(sp.lb[i], std::fmin(x0, sp.ub[i]));
        N(vX0)[i] = N(vX)[i] = N(vXP)[i] = N(vXR)[i] = x0;
    }
    for (int i = 0; i < m; ++i) {
        M(wG)[i] = sp.g.empty() ? 0.0 : sp.g[i];
        M(wCL)[i] = sp.cl[i];
        M(wCU)[i] = sp.cu[i];
        M(wCLF)[i] = std::isfinite(sp.cl[i]) ? sp.cl[i] : 0.0;
        M(wCUF)[i] = std::isfinite(sp.cu[i]) ? sp.cu[i] : 0.0;
        M(wCR)[i] = sp.con_rescale[i];
        M(wAFD)[i] = 1.0 / sp.con_rescale[i];
        const double y0 = sp.dual_start.empty() ? 0.0 : sp.dual_start[i] * sp.con_rescale[i] * sp.ovr;
        M(wY0)[i] = M(wY)[i] = M(wYP)[i] = y0;
    }
    HostState &S = k.st;
    std::memset(&S, 0, sizeof(S));
    S.pw = S.best_pw = I.pw0;
    S.best_gap = INFINITY;
    S.last_trial_fpe = INFINITY;
    S.best_err = INFINITY;
    S.inner_tol = P.inner_solver_parameters.initial_tolerance > 0 ? P.inner_solver_parameters.initial_tolerance : 1e-3;
    S.cached_inv_tau = -1.0;
    S.tol_scale = 1.0;
    S.reason = TERMINATION_REASON_UNSPECIFIED;
    return true;
}

// Device copy of h; on a failed copy the allocation is released before the
// error propagates.
template <class T> T *upload(const T *h, size_t count) {
    T *d = nullptr;
    HIP_OK(hipMalloc(&d, std::max<size_t>(count, 1) * sizeof(T)));
    if (count) {
        const hipError_t e = hipMemcpy(d, h, count * sizeof(T), hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            (void)hipFree(d);
            HIP_OK(e);
        }
    }
    return d;
}
template <class T> T *upload(const std::vector<T> &h) { return upload(h.data(), h.size()); }

// Host buffer of `count` elements left uninitialised (every element is
// written before it is read).
template <class T> std::unique_ptr<T[]> raw_buffer(size_t count) {
    return std::unique_ptr<T[]>(new T[std::max<size_t>(count, 1)]);
}

// Runs f(i) for i < count on up to 16 host threads.
template <class F> void parallel_for(int count, F f) {
    const int nt = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::atomic<int> next{0};
    auto work = [&]() {
        for (int i = next++; i < count; i = next++) f(i);
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < nt; ++t) pool.emplace_back(work);
    work();
    for (std::thread &t : pool) t.join();
}

// Launch shapes chosen by the probe, per batch shape, for the process.
struct ShapeKey {
    int mixed, count;
    long long n, m, nnz;
    bool operator<(const ShapeKey &o) const {
        return std::tie(mixed, count, n, m, nnz) < std::tie(o.mixed, o.count, o.n, o.m, o.nnz);
    }
};
struct ShapeChoice {
    int threads, row_len;
};
std::mutex g_shape_mu;
std::map<ShapeKey, ShapeChoice> g_shapes;

bool lookup_shape(const ShapeKey &k, ShapeChoice &out) {
    std::lock_guard<std::mutex> lock(g_shape_mu);
    auto it = g_shapes.find(k);
    if (it == g_shapes.end()) return false;
    out = it->second;
    return true;
}

void store_shape(const ShapeKey &k, const ShapeChoice &c) {
    std::lock_guard<std::mutex> lock(g_shape_mu);
    g_shapes[k] = c;
}

}  // namespace
}  // namespace batch
}  // namespace pdhcg

using namespace pdhcg;
using namespace pdhcg::batch;

extern "C" int pdhcg_solve_batch(const qp_problem_t *const *problems, int count, const pdhg_parameters_t *params,
                                 pdhcg_result_t **results) {
    if (count < 0 || (count > 0 && (!problems || !results))) {
        std::fprintf(stderr, "[interface] pdhcg_solve_batch: invalid arguments.\n");
        return -1;
    }
    for (int i = 0; i < count; ++i) results[i] = nullptr;
    if (count == 0) return 0;
    pdhg_parameters_t P;
    if (params) P = *params;
    else set_default_parameters(&P);
    std::vector<void *> dev;
    try {
        const auto t0 = std::chrono::steady_clock::now();
        // scale and pack the problems on all host cores (they are independent)
        std::vector<Packed> all((size_t)count);
        std::vector<const char *> why((size_t)count, nullptr);
        std::vector<std::string> err((size_t)count);
        parallel_for(count, [&](int i) {
            try {
                why[i] = unsupported(problems[i]);
                if (!why[i]) pack(problems[i], P, all[i], &why[i]);
            } catch (const std::exception &e) {
                err[i] = e.what();
            }
        });
        for (int i = 0; i < count; ++i)
            if (!err[i].empty()) fail("problem %d: %s", i, err[i].c_str());
        // largest problems first, so they start in the first round of workgroups
        std::vector<int> idx;  // problem index of each packed entry
        for (int i = 0; i < count; ++i) {
            if (why[i]) std::fprintf(stderr, "[interface] pdhcg_solve_batch: problem %d has %s.\n", i, why[i]);
            else idx.push_back(i);
        }
        auto work_of = [&](int i) { return (double)all[i].inst.n * (1.0 + all[i].inst.k) + all[i].inst.m; };
        std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return work_of(a) > work_of(b); });
        std::vector<Packed> pk;
        pk.reserve(idx.size());
        for (int i : idx) pk.push_back(std::move(all[i]));
        const int B = (int)pk.size();
        if (B == 0) return 0;
        // pool the problems: offsets first, then every problem copied into
        // its slices of the pools in parallel
        std::vector<Inst> insts((size_t)B);
        std::vector<HostState> states((size_t)B);
        std::vector<long long> ioff((size_t)B + 1, 0), voff((size_t)B + 1, 0), coff((size_t)B + 1, 0),
            soff((size_t)B + 1, 0), woff((size_t)B + 1, 0);
        for (int b = 0; b < B; ++b) {
            ioff[b + 1] = ioff[b] + (long long)pk[b].ai.size();
            voff[b + 1] = voff[b] + (long long)pk[b].av.size();
            coff[b + 1] = coff[b] + (long long)pk[b].cones.size();
            soff[b + 1] = soff[b] + (long long)pk[b].slab.size();
            woff[b + 1] = woff[b] + (long long)pk[b].wslab.size();
        }
        if (ioff[B] > (long long)INT32_MAX) fail("the batch's index pool exceeds 2^31 entries");
        auto ipool = raw_buffer<int>((size_t)ioff[B]);
        auto vpool = raw_buffer<double>((size_t)voff[B]);
        auto cones = raw_buffer<Cone>((size_t)coff[B]);
        auto slabs = raw_buffer<double>((size_t)soff[B]);
        auto wslabs = raw_buffer<double>((size_t)woff[B]);
        std::atomic<bool> pack_error{false};
        parallel_for(B, [&](int b) {
            Packed &q = pk[b];
            Inst I = q.inst;
            const int io = (int)ioff[b];
            const long long vo = voff[b];
            I.a_rp += io;
            I.a_ci += io;
            I.at_rp += io;
            I.at_ci += io;
            I.a_v += vo;
            I.at_v += vo;
            I.rdense += vo;
            I.dmat += vo;
            I.cone0 = (int)coff[b];
            I.vec = soff[b];
            I.wvec = woff[b];
            std::copy(q.ai.begin(), q.ai.end(), ipool.get() + ioff[b]);
            std::copy(q.av.begin(), q.av.end(), vpool.get() + vo);
            std::copy(q.cones.begin(), q.cones.end(), cones.get() + coff[b]);
            std::copy(q.slab.begin(), q.slab.end(), slabs.get() + soff[b]);
            std::copy(q.wslab.begin(), q.wslab.end(), wslabs.get() + woff[b]);
            // every offset the kernel derives from I stays inside this
            // problem's share of the pools (a failure is a packing bug)
            const long long nnz = (long long)q.inst.at_rp - q.inst.a_ci;
            const int *rp = ipool.get() + I.a_rp, *trp = ipool.get() + I.at_rp;
            bool ok = I.a_ci == I.a_rp + I.m + 1 && (long long)I.at_rp == I.a_ci + nnz &&
                      I.at_ci == I.at_rp + I.n + 1 && I.at_ci + nnz <= ioff[b + 1] && rp[0] == 0 &&
                      rp[I.m] == nnz && trp[0] == 0 && trp[I.n] == nnz && I.at_v == I.a_v + nnz &&
                      I.rdense == I.at_v + nnz && I.dmat >= I.rdense && I.dmat <= voff[b + 1] &&
                      I.vec + slab_size(I.n, I.m) <= soff[b + 1] && I.wvec + wslab_size(I.k, I.ncones) <= woff[b + 1] &&
                      I.cone0 + I.ncones <= coff[b + 1];
            for (long long e = 0; ok && e < nnz; ++e) {
                const int c = ipool[(size_t)I.a_ci + e], r = ipool[(size_t)I.at_ci + e];
                ok = c >= 0 && c < I.n && r >= 0 && r < I.m;
            }
            for (int c = 0; ok && c < I.ncones; ++c) {
                const Cone &cn = cones[(size_t)I.cone0 + c];
                const int len = cn.type == CONE_EXPONENTIAL || cn.type == CONE_POWER ? 3 : cn.vdim + 2;
                ok = cn.start >= 0 && (long long)cn.start + len <= I.m &&
                     (c >= I.ncones - I.nlong || cn.type == CONE_EXPONENTIAL || cn.type == CONE_POWER ||
                      cn.vdim <= kMaxThreadCone);
            }
            if (!ok) pack_error.store(true);
            insts[b] = I;
            states[b] = q.st;
            // the pools hold the data now
            std::vector<int>().swap(q.ai);
            std::vector<double>().swap(q.av);
            std::vector<double>().swap(q.slab);
            std::vector<double>().swap(q.wslab);
        });
        if (pack_error.load()) fail("packing produced an offset outside a problem's pools (internal error)");
        Params dp;
        dp.ef = std::max(1, P.termination_evaluation_frequency);
        dp.iteration_limit = P.termination_criteria.iteration_limit;
        dp.norm_inf = P.optimality_norm == NORM_TYPE_L_INF;
        dp.inner_limit = std::max(1, P.inner_solver_parameters.iteration_limit);
        dp.precond = P.diag_jacobi_precond ? 1 : 0;
        dp.eps_opt = P.termination_criteria.eps_optimal_relative;
        dp.eps_feas = P.termination_criteria.eps_feasible_relative;
        dp.eps_inf = P.termination_criteria.eps_infeasible;
        dp.rho = P.reflection_coefficient;
        dp.inner_tol0 = P.inner_solver_parameters.initial_tolerance > 0 ? P.inner_solver_parameters.initial_tolerance
                                                                         : 1e-3;
        dp.inner_min_tol = P.inner_solver_parameters.min_tolerance > 0 ? P.inner_solver_parameters.min_tolerance : 1e-9;
        dp.i_smooth = P.restart_params.i_smooth;
        dp.k_p = P.restart_params.k_p;
        dp.k_i = P.restart_params.k_i;
        dp.k_d = P.restart_params.k_d;
        dp.suff = P.restart_params.sufficient_reduction_for_restart;
        dp.nec = P.restart_params.necessary_reduction_for_restart;
        dp.art = P.restart_params.artificial_restart_threshold;
        {  // iterations without a halving of the KKT error before a problem leaves the FP32 phase
            const char *e = std::getenv("PDHCG_MIXED_STALL");
            dp.stall_iters = e ? std::atoi(e) : std::max(1000, 5 * dp.ef);
        }
        // the FP32 phase hands a problem over once its KKT error is within a
        // decade of the tolerance: FP32 iterates take the bulk of the
        // iterations, FP64 ones the last decade (many small problems have
        // their FP32 floor there, and iterating on at the floor is waste)
        dp.handover = 10.0 * std::fmax(dp.eps_opt, dp.eps_feas);
        const bool mixed = pdhcg_get_precision() == PDHCG_PRECISION_MIXED;
        auto track = [&](void *p) {
            dev.push_back(p);
            return p;
        };
        Params *d_params = (Params *)track(upload(std::vector<Params>{dp}));
        Inst *d_insts = (Inst *)track(upload(insts));
        int *d_ipool = (int *)track(upload(ipool.get(), (size_t)ioff[B]));
        double *d_vpool = (double *)track(upload(vpool.get(), (size_t)voff[B]));
        double *d_slabs = (double *)track(upload(slabs.get(), (size_t)soff[B]));
        double *d_wslabs = (double *)track(upload(wslabs.get(), (size_t)woff[B]));
        Cone *d_cones = (Cone *)track(upload(cones.get(), (size_t)coff[B]));
        State *d_states = (State *)track(upload(states));
        Out *d_out = (Out *)track(upload(std::vector<Out>((size_t)B)));
        // stream owned by a guard, so it is destroyed on every exit
        struct StreamGuard {
            hipStream_t s = nullptr;
            ~StreamGuard() {
                if (s) (void)hipStreamDestroy(s);
            }
        } sg;
        HIP_OK(hipStreamCreateWithFlags(&sg.s, hipStreamNonBlocking));
        hipStream_t st = sg.s;
        // mixed precision: FP32 copies of the matrix values and of the slabs
        float *d_vpool32 = nullptr, *d_slabs32 = nullptr;
        if (mixed) {
            HIP_OK(hipMalloc(&d_vpool32, std::max<size_t>((size_t)voff[B], 1) * sizeof(float)));
            dev.push_back(d_vpool32);
            HIP_OK(hipMalloc(&d_slabs32, std::max<size_t>((size_t)soff[B], 1) * sizeof(float)));
            dev.push_back(d_slabs32);
            hipLaunchKernelGGL(narrow_kernel, dim3(grid_of(voff[B])), dim3(256), 0, st, (const double *)d_vpool,
                               d_vpool32, voff[B]);
            hipLaunchKernelGGL(narrow_kernel, dim3(grid_of(soff[B])), dim3(256), 0, st, (const double *)d_slabs,
                               d_slabs32, soff[B]);
            HIP_OK(hipGetLastError());
        }
        const auto t_dev = std::chrono::steady_clock::now();
        int windows = 0;
        // windows of iterations; every problem that has terminated returns at once
        const int window = std::max(1, std::min(dp.ef * 16, 2000));
        std::vector<Out> out((size_t)B);
        bool time_up = false, cancelled = false;
        // runs windows until no problem is running (or time / cancel); `budget`
        // passes per window; `phase` names the kernel in the progress lines
        auto run_phase = [&](auto launch, const char *phase, int max_windows) {
            for (int w = 0; w < max_windows; ++w) {
                const auto tw = std::chrono::steady_clock::now();
                launch();
                ++windows;
                HIP_OK(hipGetLastError());
                HIP_OK(hipMemcpyAsync(out.data(), d_out, out.size() * sizeof(Out), hipMemcpyDeviceToHost, st));
                HIP_OK(hipStreamSynchronize(st));
                int running = 0;
                for (const Out &o : out) running += o.status == TERMINATION_REASON_UNSPECIFIED ? 1 : 0;
                if (P.verbose >= 1)  // the tail: how many problems each window leaves running
                    std::fprintf(stderr, "[batch] %s window %d (%d passes): %.3fs, %d of %d problems still running\n",
                                 phase, windows, window,
                                 std::chrono::duration<double>(std::chrono::steady_clock::now() - tw).count(), running,
                                 B);
                if (running == 0) return;
                if (g_pdhcg_cancel) {  // user interrupt, checked between windows
                    cancelled = true;
                    return;
                }
                const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
                if (el >= P.termination_criteria.time_sec_limit) {
                    time_up = true;
                    return;
                }
            }
        };
        // launches of the main phase's kernel at either workgroup size
        auto launch32 = [&](int bt, int budget) {
            if (bt == kBTSmall)
                hipLaunchKernelGGL((batch_solve_kernel<float, kBTSmall>), dim3(B), dim3(kBTSmall), 0, st, d_params,
                                   d_insts, d_ipool, (const float *)d_vpool32, (const double *)d_vpool, d_slabs32,
                                   d_wslabs, d_cones, d_states, d_out, budget);
            else
                hipLaunchKernelGGL((batch_solve_kernel<float, kBT>), dim3(B), dim3(kBT), 0, st, d_params, d_insts,
                                   d_ipool, (const float *)d_vpool32, (const double *)d_vpool, d_slabs32, d_wslabs,
                                   d_cones, d_states, d_out, budget);
        };
        auto launch64 = [&](int bt, int budget) {
            if (bt == kBTSmall)
                hipLaunchKernelGGL((batch_solve_kernel<double, kBTSmall>), dim3(B), dim3(kBTSmall), 0, st, d_params,
                                   d_insts, d_ipool, (const double *)d_vpool, (const double *)d_vpool, d_slabs,
                                   d_wslabs, d_cones, d_states, d_out, budget);
            else
                hipLaunchKernelGGL((batch_solve_kernel<double, kBT>), dim3(B), dim3(kBT), 0, st, d_params, d_insts,
                                   d_ipool, (const double *)d_vpool, (const double *)d_vpool, d_slabs, d_wslabs,
                                   d_cones, d_states, d_out, budget);
        };
        // Launch-shape autotuning.  Each candidate runs the same short probe
        // from the same saved state (loop states, slabs and FP64 slabs are
        // copied device-to-device and restored after each), and the batch
        // runs with the fastest, kept only when 5% faster than the default:
        //   threads per problem, 512 or 256;
        //   then, when some problem's A or A' has rows of 4 to 63 entries on
        //   average (either form could win), the row length from which a
        //   product takes a wavefront per row instead of a thread per row:
        //   16 (default), 4 or 64.
        // PDHCG_BATCH_THREADS=256|512 and PDHCG_BATCH_ROWS=4|16|64 pin them.
        // The choice is kept per process for the batch's shape (precision,
        // problem count, total n, m and nnz(A)), so a family solved again
        // (cold, warm, at another tolerance) is not probed again.
        int bt = kBT, row_len = 16;
        const char *et = std::getenv("PDHCG_BATCH_THREADS");
        const char *er = std::getenv("PDHCG_BATCH_ROWS");
        if (et) bt = std::atoi(et) == kBTSmall ? kBTSmall : kBT;
        if (er && (std::atoi(er) == 4 || std::atoi(er) == 64)) row_len = std::atoi(er);
        auto set_rows = [&](int len) {  // the per-problem row forms for a threshold, on the device
            for (Inst &I : insts) {
                const long long nnz = (long long)I.at_rp - I.a_ci;
                I.a_long = nnz >= (long long)len * std::max(I.m, 1);
                I.at_long = nnz >= (long long)len * std::max(I.n, 1);
            }
            HIP_OK(hipMemcpyAsync(d_insts, insts.data(), (size_t)B * sizeof(Inst), hipMemcpyHostToDevice, st));
        };
        bool ambiguous = false;
        for (const Inst &I : insts) {
            const long long nnz = (long long)I.at_rp - I.a_ci;
            for (int rows : {I.m, I.n})
                if (rows > 0 && nnz >= 4LL * rows && nnz < 64LL * rows) ambiguous = true;
        }
        long long tot_n = 0, tot_m = 0, tot_nnz = 0;
        for (const Inst &I : insts) {
            tot_n += I.n;
            tot_m += I.m;
            tot_nnz += (long long)I.at_rp - I.a_ci;
        }
        const ShapeKey key{mixed ? 1 : 0, B, tot_n, tot_m, tot_nnz};
        ShapeChoice known{};
        const bool have = lookup_shape(key, known);
        if (have && !et) bt = known.threads;
        if (have && !er) row_len = known.row_len;
        if (row_len != 16) set_rows(row_len);
        const bool tune_bt = !et && !have && window > 1, tune_rows = !er && !have && ambiguous && window > 1;
        if (tune_bt || tune_rows) {
            const int probe = std::min(window, 12);
            const size_t sbytes = (size_t)soff[B] * (mixed ? sizeof(float) : sizeof(double));
            void *slab_src = mixed ? (void *)d_slabs32 : (void *)d_slabs;
            void *s_copy = nullptr, *w_copy = nullptr, *st_copy = nullptr;
            HIP_OK(hipMalloc(&s_copy, std::max<size_t>(sbytes, 1)));
            dev.push_back(s_copy);
            HIP_OK(hipMalloc(&w_copy, std::max<size_t>((size_t)woff[B], 1) * sizeof(double)));
            dev.push_back(w_copy);
            HIP_OK(hipMalloc(&st_copy, (size_t)B * sizeof(State)));
            dev.push_back(st_copy);
            auto copy_all = [&](bool save) {
                auto cp = [&](void *a, void *b, size_t n) {
                    if (n) HIP_OK(hipMemcpyAsync(save ? b : a, save ? a : b, n, hipMemcpyDeviceToDevice, st));
                };
                cp(slab_src, s_copy, sbytes);
                cp(d_wslabs, w_copy, (size_t)woff[B] * sizeof(double));
                cp(d_states, st_copy, (size_t)B * sizeof(State));
            };
            struct EventPair {  // destroyed on every exit, exceptions included
                hipEvent_t a = nullptr, b = nullptr;
                ~EventPair() {
                    if (a) (void)hipEventDestroy(a);
                    if (b) (void)hipEventDestroy(b);
                }
            } ev;
            HIP_OK(hipEventCreate(&ev.a));
            HIP_OK(hipEventCreate(&ev.b));
            hipEvent_t e0 = ev.a, e1 = ev.b;
            auto timed = [&](int threads) {  // one probe from the saved state, restored afterwards
                float ms = 0.0f;
                HIP_OK(hipEventRecord(e0, st));
                if (mixed) launch32(threads, probe);
                else launch64(threads, probe);
                HIP_OK(hipEventRecord(e1, st));
                HIP_OK(hipEventSynchronize(e1));
                HIP_OK(hipGetLastError());
                HIP_OK(hipEventElapsedTime(&ms, e0, e1));
                copy_all(false);
                return ms;
            };
            copy_all(true);
            if (tune_bt) {
                const float m512 = timed(kBT), m256 = timed(kBTSmall);
                bt = m256 < 0.95f * m512 ? kBTSmall : kBT;
                if (P.verbose >= 1)
                    std::fprintf(stderr,
                                 "[batch] threads per problem: %d (probe of %d passes: 512 -> %.1f ms, 256 -> %.1f ms)\n",
                                 bt, probe, m512, m256);
            }
            if (tune_rows) {
                const int lens[3] = {16, 4, 64};
                float ms[3];
                for (int c = 0; c < 3; ++c) {
                    set_rows(lens[c]);
                    ms[c] = timed(bt);
                }
                int best = 0;
                for (int c = 1; c < 3; ++c)
                    if (ms[c] < 0.95f * ms[0] && ms[c] < ms[best]) best = c;
                row_len = lens[best];
                set_rows(row_len);
                if (P.verbose >= 1)
                    std::fprintf(stderr,
                                 "[batch] wavefront rows from %d entries (probe: 16 -> %.1f ms, 4 -> %.1f ms, 64 -> %.1f ms)\n",
                                 row_len, ms[0], ms[1], ms[2]);
            }
            HIP_OK(hipStreamSynchronize(st));
            if (!et && !er) store_shape(key, ShapeChoice{bt, row_len});  // only fully probed choices
        }
        if (mixed) {
            run_phase([&] { launch32(bt, window); }, "FP32", INT_MAX);
            // FP32 -> FP64: the iterates widened exactly into the FP64 slabs
            // (the FP64 problem data stay), and every problem that has not
            // hit a limit evaluated again in FP64 before it may terminate
            hipLaunchKernelGGL(widen_iterates_kernel, dim3(B), dim3(256), 0, st, (const Inst *)d_insts,
                               (const float *)d_slabs32, d_slabs);
            hipLaunchKernelGGL(resume_fp64_kernel, dim3(grid_of(B)), dim3(256), 0, st, d_states, B);
            HIP_OK(hipGetLastError());
            if (P.verbose >= 1) {
                int stalled = 0;
                for (const Out &o : out) stalled += o.status == kStalled ? 1 : 0;
                std::fprintf(stderr, "[batch] FP32 phase done: %d of %d problems left it at the FP32 floor\n",
                             stalled, B);
            }
        }
        // FP64; after a timed-out or cancelled FP32 phase, one pass that
        // evaluates every problem in FP64
        const bool stop = time_up || cancelled;
        const bool was_time_up = time_up, was_cancelled = cancelled;
        run_phase([&] { launch64(bt, stop ? 1 : window); }, "FP64", stop ? 1 : INT_MAX);
        if (stop) {
            time_up = was_time_up;
            cancelled = was_cancelled;
        }
        const auto t_end = std::chrono::steady_clock::now();
        const double elapsed = std::chrono::duration<double>(t_end - t0).count();
        if (P.verbose >= 1)
            std::fprintf(stderr, "[batch] %d problems: host scaling and packing %.3fs, device %.3fs in %d windows\n", B,
                         std::chrono::duration<double>(t_dev - t0).count(),
                         std::chrono::duration<double>(t_end - t_dev).count(), windows);
        HIP_OK(hipMemcpy(slabs.get(), d_slabs, (size_t)soff[B] * sizeof(double), hipMemcpyDeviceToHost));
        for (int b = 0; b < B; ++b) {
            const Inst &I = insts[b];
            const ScaledProblem &sp = pk[b].sp;
            const Out &o = out[b];
            const double *V = slabs.get() + I.vec;
            const double *xp = V + (size_t)vXP * I.n, *aty = V + (size_t)vATY * I.n, *qx = V + (size_t)vQX * I.n;
            const double *yp = V + (size_t)kNVec * I.n + (size_t)wYP * I.m;
            pdhcg_result_t *r = ccalloc<pdhcg_result_t>(1);
            r->primal_solution = cmalloc<double>(I.n);
            r->dual_solution = cmalloc<double>(I.m);
            r->reduced_cost = cmalloc<double>(I.n);
            for (int j = 0; j < I.n; ++j) {  // reference src/pdhg_core_op.cu:1431-1496
                double gr = sp.c[j];
                if (I.qkind != kQNone) gr += qx[j];
                double rc = (gr - aty[j]) * sp.var_rescale[j] / sp.ovr;
                if (!std::isfinite(sp.lb[j])) rc = std::fmin(rc, 0.0);
                if (!std::isfinite(sp.ub[j])) rc = std::fmax(rc, 0.0);
                r->reduced_cost[j] = rc;
                r->primal_solution[j] = xp[j] / sp.var_rescale[j] / sp.cbr;
            }
            for (int i = 0; i < I.m; ++i) r->dual_solution[i] = yp[i] / sp.con_rescale[i] / sp.ovr;
            const qp_problem_t *orig = problems[idx[b]];
            r->num_variables = orig->num_variables;
            r->num_constraints = orig->num_constraints;
            r->num_nonzeros = orig->constraint_matrix_num_nonzeros;
            r->total_count = o.iters;
            r->total_inner_count = o.inner;
            r->rescaling_time_sec = sp.rescale_time;
            r->cumulative_time_sec = elapsed;
            r->relative_primal_residual = o.rel_pr;
            r->relative_dual_residual = o.rel_du;
            r->absolute_primal_residual = o.abs_pr;
            r->absolute_dual_residual = o.abs_du;
            r->primal_objective_value = o.pobj;
            r->dual_objective_value = o.dobj;
            r->objective_gap = o.gap;
            r->relative_objective_gap = o.rel_gap;
            r->max_primal_ray_infeasibility = o.max_pray;
            r->max_dual_ray_infeasibility = o.max_dray;
            r->primal_ray_linear_objective = o.pray_lin_obj;
            r->dual_ray_objective = o.dray_obj;
            r->termination_reason = o.status != TERMINATION_REASON_UNSPECIFIED
                                        ? (termination_reason_t)o.status
                                        : cancelled ? TERMINATION_REASON_USER_INTERRUPT
                                        : time_up   ? TERMINATION_REASON_TIME_LIMIT
                                                    : TERMINATION_REASON_UNSPECIFIED;
            results[idx[b]] = r;
        }
        for (void *p : dev) (void)hipFree(p);
        return 0;
    } catch (const std::exception &e) {
        for (void *p : dev) (void)hipFree(p);
        std::fprintf(stderr, "[interface] pdhcg_solve_batch: %s\n", e.what());
        return -1;
    }
}
