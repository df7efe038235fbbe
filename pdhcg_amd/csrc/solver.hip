// Restarted Halpern PDHG with the projected Barzilai-Borwein inner solve on
// one MI355X (gfx950).  The host drives the outer loop exactly as reference
// src/solver.cu:121-169 does (evaluation cadence, restarts, fixed-point error,
// Halpern weights), while every vector operation runs in the HIP kernels of
// dev_spmv.hpp / dev_kernels.hpp.  Host synchronisation happens only where the
// host must decide: KKT evaluations, restarts, fixed-point errors, and
// between chunks of the device-resident BB and power iterations.
//
// Template parameters: V is the storage type of matrix values, X of vectors.
// Solver<double, double> is the plain FP64 form; Solver<float, float> keeps
// matrix values and vector traffic in FP32 (FP32 SpMV accumulation) with all
// reductions, scalars, restart and KKT logic in FP64.
#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <type_traits>
#include <utility>
#include <vector>

#include "../../include/pdhcg_amd.h"
#include "dev_cones.hpp"
#include "dev_sections.hpp"
#include "dev_kernels.hpp"
#include "dev_timer.hpp"
#include "host_scale.h"

namespace pdhcg {

// Cancellation request checked at every KKT evaluation (reference src/pdhcg.c:31,
// src/utils.cu:252-256).
volatile int g_pdhcg_cancel = 0;

// Problems with at least this many stored matrix entries (A, Q and R) are
// scaled with the entry-wise passes on the device (scale.hip).
constexpr long long kDeviceScaleMinEntries = 1LL << 20;

namespace {

int print_frequency(int iter) {  // reference src/utils.cu:535-546
    int step = 10;
    long long threshold = 1000;
    while (iter >= threshold) {
        step *= 10;
        threshold *= 10;
    }
    return step;
}

// Prologue of a recorded window of iterations: writes the window's scalars
// (primal step, dual step, Halpern count at the window start) to device
// memory, where the step epilogues of the window read them.
__global__ void graph_dyn_kernel(double *dyn, double tau, double sigma, double k0) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        dyn[0] = tau;
        dyn[1] = sigma;
        dyn[2] = k0;
    }
}

// Graph replay is off by default: measured on MI355X, the launch-by-launch
// loop already runs at the device's dispatch rate (about 6.5 us per SpMV
// launch, the host running ahead), and a replayed window was 3-7% slower.
int g_graph_window = -1;  // -1: read PDHCG_GRAPH once

bool graph_window_default() {
    if (g_graph_window < 0) {
        const char *e = std::getenv("PDHCG_GRAPH");
        g_graph_window = (e && std::strcmp(e, "1") == 0) ? 1 : 0;
    }
    return g_graph_window != 0;
}

template <class T> T *dalloc(size_t count) {
    T *p = nullptr;
    HIP_OK(hipMalloc(&p, (count ? count : 1) * sizeof(T)));
    return p;
}

template <class T, class S> T *dupload(const std::vector<S> &h, size_t count) {
    T *p = dalloc<T>(count);
    if (count && std::is_same<T, S>::value) {
        HIP_OK(hipMemcpy(p, h.data(), count * sizeof(T), hipMemcpyHostToDevice));
    } else if (count) {
        std::vector<T> tmp(count);
        for (size_t i = 0; i < count; ++i) tmp[i] = (T)h[i];
        HIP_OK(hipMemcpy(p, tmp.data(), count * sizeof(T), hipMemcpyHostToDevice));
    }
    return p;
}

template <class T> T *dzeros(size_t count) {
    T *p = dalloc<T>(count);
    HIP_OK(hipMemset(p, 0, (count ? count : 1) * sizeof(T)));
    return p;
}

// Bump allocation from a caller-provided device region (the cone set of the
// allocation-free entry point).  With base == nullptr it only measures: the
// returned offsets are never dereferenced and no copies are issued.
struct DevArena {
    char *base = nullptr;
    size_t cap = 0, used = 0;
    bool measuring() const { return base == nullptr; }
    template <class T> T *take(size_t count) {
        const size_t bytes = ((std::max<size_t>(count, 1) * sizeof(T)) + 255) & ~(size_t)255;
        if (used + bytes > cap) fail("device scratch of %zu bytes exhausted (%zu more needed)", cap, used + bytes - cap);
        T *p = reinterpret_cast<T *>(base + used);
        used += bytes;
        return p;
    }
};

template <class V> BinnedCsr<V> upload_csr(const HostCSR &h) {
    return make_binned<V>(h.rows, h.cols, h.rp, h.ci, h.v);
}

struct RunStats {
    int total_count = 0;
    int inner_total = 0;
};

// The variable cone blocks on the device: one block list per type, PSD
// blocks grouped by order, and a slot mask (reference
// internal/internal_types.h:106-127, dispatch src/cone_dispatch.cu).
// Every PSD block of a cone set: one launch of psd_project_kernel, largest
// orders first.
struct PsdSet {
    int count = 0;
    size_t lds = 0;             // dynamic LDS bytes per workgroup (the largest need)
    PsdBlock *blocks = nullptr;
    double *work = nullptr;     // psd_work_doubles(order, mode) per block
};

// LDS a workgroup may allocate on the current device, less the PSD kernel's
// static arrays and a margin.
inline size_t psd_lds_budget() {
    int dev = 0, lds = 0;
    HIP_OK(hipGetDevice(&dev));
    HIP_OK(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    return lds > 1024 ? (size_t)lds - 1024 : 0;
}

// Launches the PSD projection of every block; `site` selects which kept
// eigenvector set warm-starts it (0: the iterate, 1: the dual residual).
// Dynamic LDS above the default limit is requested once per instance.
template <class X> void psd_launch(const PsdSet &g, X *v, int site, const int *gate, hipStream_t st) {
    static bool raised[2] = {false, false};
    if (!g.count) return;
    // PDHCG_PSD_THREADS=512|1024 picks the workgroup size (default 1024)
    static const int threads = [] {
        const char *e = std::getenv("PDHCG_PSD_THREADS");
        return e && std::atoi(e) == 512 ? kPsdThreads : kPsdThreadsLarge;
    }();
    auto go = [&](auto kernel, int which, int nt) {
        if (!raised[which] && g.lds > 0) {
            HIP_OK(hipFuncSetAttribute((const void *)kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)psd_lds_budget()));
            raised[which] = true;
        }
        hipLaunchKernelGGL(kernel, dim3(g.count), dim3(nt), g.lds, st, v, (const PsdBlock *)g.blocks, g.work, site,
                           gate);
    };
    if (threads == kPsdThreadsLarge) go(psd_project_kernel<X, kPsdThreadsLarge>, 1, kPsdThreadsLarge);
    else go(psd_project_kernel<X, kPsdThreads>, 0, kPsdThreads);
}

class DevConeSet {
  public:
    bool psd = false, power = false;
    ConeList rsoc, soc, expc, powc;
    ConeList lrsoc, lsoc;                        // SOC/RSOC blocks above kMaxThreadCone vector slots
    ConeList powall;                             // every power block, for the primal violation
    SectionList sec;                             // blocks with fixed slots (any non-PSD type)
    SectionList lsec;                            // SOC/RSOC blocks with fixed slots above kMaxThreadCone
    double *lsec_scratch = nullptr;              // 4 (v_dim + 2) values per long section block
    long long *lsec_off = nullptr;
    char *fixed = nullptr;                       // per-slot fixed flags, when sections exist
    PsdSet pset;                                 // every PSD block
    char *mask = nullptr;
    int total = 0;                               // all blocks, in input order:
    int *all_start = nullptr, *all_len = nullptr;  // first slot and slot count

    DevConeSet() = default;
    DevConeSet(const DevConeSet &) = delete;
    DevConeSet &operator=(const DevConeSet &) = delete;
    ~DevConeSet() { release(); }

    // arena: place every array in that region instead of hipMalloc (nothing
    // is freed by release() then)
    void build(const cone_blocks_t *b, int n, DevArena *arena = nullptr) {
        arena_ = arena;
        struct HostList {
            std::vector<int> start, vdim;
            std::vector<double> alpha;
        } lists[6];  // rotated SOC, SOC, exponential, power, long rotated SOC, long SOC
        std::vector<char> hmask((size_t)n, 0);
        std::vector<std::pair<int, int>> ps;  // (order, start)
        std::vector<int> hs, hl;
        if (b->is_fixed && b->fixed_mask_size < n)
            fail("the fixed-slot mask has %d entries; %d variables", b->fixed_mask_size, n);
        std::vector<int> sec_start, sec_vdim, sec_type, pall_start, lsec_start, lsec_vdim, lsec_type;
        std::vector<double> sec_alpha, pall_alpha;
        std::vector<long long> lsec_off_h;
        long long lsec_total = 0;
        for (int k = 0; k < b->num_cones; ++k) {
            const int st = b->start_idx[k], vd = b->v_dim[k], len = cone_block_slots(b, k);
            hs.push_back(st);
            hl.push_back(len);
            const cone_type_t t = b->type[k];
            if (st < 0 || len < 0 || (long long)st + len > n) fail("cone block %d lies outside the variables", k);
            for (int i = st; i < st + len; ++i) hmask[i] = t == CONE_PSD ? kMaskPsd : kMaskCone;
            if (t == CONE_PSD) {
                psd = true;
                ps.emplace_back(vd, st);
                continue;
            }
            if (t == CONE_POWER) {
                power = true;
                pall_start.push_back(st);
                pall_alpha.push_back(b->power_alpha ? b->power_alpha[k] : 0.0);
            }
            bool has_fixed = false;
            for (int i = st; b->is_fixed && i < st + len; ++i) has_fixed |= b->is_fixed[i] != 0;
            if (has_fixed) {
                if ((t == CONE_STANDARD_SOC || t == CONE_ROTATED_SOC) && vd > kMaxThreadCone) {
                    lsec_start.push_back(st);
                    lsec_vdim.push_back(vd);
                    lsec_type.push_back((int)t);
                    lsec_off_h.push_back(lsec_total);
                    lsec_total += 4LL * (vd + 2);
                    continue;
                }
                sec_start.push_back(st);
                sec_vdim.push_back(vd);
                sec_type.push_back((int)t);
                sec_alpha.push_back(t == CONE_POWER && b->power_alpha ? b->power_alpha[k] : 0.0);
                continue;
            }
            const bool long_block = (t == CONE_STANDARD_SOC || t == CONE_ROTATED_SOC) && vd > kMaxThreadCone;
            const int li = t == CONE_ROTATED_SOC ? (long_block ? 4 : 0)
                           : t == CONE_STANDARD_SOC ? (long_block ? 5 : 1)
                           : t == CONE_EXPONENTIAL ? 2
                                                   : 3;
            lists[li].start.push_back(st);
            lists[li].vdim.push_back(vd);
            lists[li].alpha.push_back(t == CONE_POWER && b->power_alpha ? b->power_alpha[k] : 0.0);
        }
        ConeList *dst[6] = {&rsoc, &soc, &expc, &powc, &lrsoc, &lsoc};
        for (int li = 0; li < 6; ++li) {
            const HostList &h = lists[li];
            ConeList &L = *dst[li];
            L.count = (int)h.start.size();
            if (!L.count) continue;
            L.start = up<int>(h.start, h.start.size());
            L.vdim = up<int>(h.vdim, h.vdim.size());
            L.alpha = up<double>(h.alpha, h.alpha.size());
            L.warm = zeros<double>(h.start.size());
            L.rwarm = zeros<double>(h.start.size());
        }
        if (!pall_start.empty()) {
            powall.count = (int)pall_start.size();
            powall.start = up<int>(pall_start, pall_start.size());
            powall.alpha = up<double>(pall_alpha, pall_alpha.size());
        }
        if (!sec_start.empty()) {
            sec.count = (int)sec_start.size();
            sec.start = up<int>(sec_start, sec_start.size());
            sec.vdim = up<int>(sec_vdim, sec_vdim.size());
            sec.type = up<int>(sec_type, sec_type.size());
            sec.alpha = up<double>(sec_alpha, sec_alpha.size());
        }
        if (!lsec_start.empty()) {
            lsec.count = (int)lsec_start.size();
            lsec.start = up<int>(lsec_start, lsec_start.size());
            lsec.vdim = up<int>(lsec_vdim, lsec_vdim.size());
            lsec.type = up<int>(lsec_type, lsec_type.size());
            std::vector<double> zero_alpha(lsec_start.size(), 0.0);
            lsec.alpha = up<double>(zero_alpha, zero_alpha.size());
            lsec_off = up<long long>(lsec_off_h, lsec_off_h.size());
            lsec_scratch = raw<double>((size_t)lsec_total);
        }
        if (!sec_start.empty() || !lsec_start.empty()) {
            std::vector<char> hf(b->is_fixed, b->is_fixed + n);
            fixed = up<char>(hf, hf.size());
        }
        // PSD blocks, largest orders first (stable in input order otherwise)
        std::stable_sort(ps.begin(), ps.end(),
                         [](const std::pair<int, int> &x, const std::pair<int, int> &y) { return x.first > y.first; });
        if (!ps.empty()) {
            const size_t budget = psd_lds_budget();
            // PDHCG_PSD_MODE=1|2|3 moves V' (1), A (2) or both (3) to the work
            // buffer on orders whose layout would fit LDS (the tests run every
            // mode)
            const char *fm = std::getenv("PDHCG_PSD_MODE");
            std::vector<PsdBlock> hb(ps.size());
            long long work = 0;
            for (size_t i = 0; i < ps.size(); ++i) {
                PsdBlock &pb = hb[i];
                pb.order = ps[i].first;
                pb.start = ps[i].second;
                auto fits = [&](int md) { return psd_lds_bytes(pb.order, md) <= budget; };
                pb.mode = fits(0) ? 0 : (fits(1) ? 1 : (fits(2) ? 2 : 3));
                if (fm) {
                    const int f = std::atoi(fm);
                    if (f >= 1 && f <= 3) pb.mode = fits(f) ? f : 3;
                }
                pb.pad = 0;
                pb.work = work;
                work += (long long)psd_work_doubles(pb.order, pb.mode);
                pset.lds = std::max(pset.lds, psd_lds_bytes(pb.order, pb.mode));
            }
            pset.count = (int)hb.size();
            pset.blocks = up<PsdBlock>(hb, hb.size());
            pset.work = raw<double>((size_t)work);
            if (!measuring()) {
                hipLaunchKernelGGL(psd_identity_kernel, dim3(pset.count), dim3(kBlock), 0, nullptr,
                                   (const PsdBlock *)pset.blocks, pset.work);
                HIP_OK(hipGetLastError());
                HIP_OK(hipStreamSynchronize(nullptr));
            }
        }
        mask = up<char>(hmask, hmask.size());
        total = (int)hs.size();
        all_start = up<int>(hs, hs.size());
        all_len = up<int>(hl, hl.size());
    }

    // Projects v onto the scaled cones D K block by block (reference
    // src/cone_dispatch.cu:38-57).  residual_warm selects the warm starts of
    // the residual evaluation instead of those of the iteration.
    // The PSD launch (at most one workgroup per PSD block, so most CUs are
    // free) runs on a second stream beside the other cone lists: the solver
    // enables this; the standalone cone entry points keep one stream.
    void enable_overlap() {
        if (side_ || !pset.count) return;
        HIP_OK(hipStreamCreateWithFlags(&side_, hipStreamNonBlocking));
        HIP_OK(hipEventCreateWithFlags(&fork_, hipEventDisableTiming));
        HIP_OK(hipEventCreateWithFlags(&join_, hipEventDisableTiming));
    }

    template <class X>
    void project(X *v, const double *d, bool residual_warm, const int *gate, hipStream_t st) const {
        const int rw = residual_warm ? 1 : 0;
        const bool overlap = side_ && pset.count;
        if (overlap) {
            // the PSD slots are disjoint from every other list's slots
            HIP_OK(hipEventRecord(fork_, st));
            HIP_OK(hipStreamWaitEvent(side_, fork_, 0));
            psd_launch<X>(pset, v, 0, gate, side_);
            HIP_OK(hipEventRecord(join_, side_));
        }
        if (rsoc.count)
            hipLaunchKernelGGL((cone_project_kernel<X, CONE_ROTATED_SOC>), blocks(rsoc.count), dim3(kBlock), 0, st, v,
                               d, rsoc, rw, gate);
        if (soc.count)
            hipLaunchKernelGGL((cone_project_kernel<X, CONE_STANDARD_SOC>), blocks(soc.count), dim3(kBlock), 0, st, v,
                               d, soc, rw, gate);
        if (expc.count)
            hipLaunchKernelGGL((cone_project_kernel<X, CONE_EXPONENTIAL>), blocks(expc.count), dim3(kBlock), 0, st, v,
                               d, expc, rw, gate);
        if (powc.count)
            hipLaunchKernelGGL((cone_project_kernel<X, CONE_POWER>), blocks(powc.count), dim3(kBlock), 0, st, v, d,
                               powc, rw, gate);
        if (lrsoc.count)
            hipLaunchKernelGGL((soc_long_project_kernel<X, true>), dim3(lrsoc.count), dim3(kBlock), 0, st, v, d,
                               (const X *)nullptr, 0.0, lrsoc, rw, gate);
        if (lsoc.count)
            hipLaunchKernelGGL((soc_long_project_kernel<X, false>), dim3(lsoc.count), dim3(kBlock), 0, st, v, d,
                               (const X *)nullptr, 0.0, lsoc, rw, gate);
        if (sec.count)
            hipLaunchKernelGGL(section_project_kernel<X>, blocks(sec.count), dim3(kBlock), 0, st, v, d, (const X *)nullptr,
                               0.0, fixed, sec, gate);
        if (lsec.count)
            hipLaunchKernelGGL(section_long_project_kernel<X>, dim3(lsec.count), dim3(kBlock), 0, st, v, d,
                               (const X *)nullptr, 0.0, fixed, lsec, lsec_scratch, lsec_off, gate);
        if (overlap) HIP_OK(hipStreamWaitEvent(st, join_, 0));
        else psd_launch<X>(pset, v, 0, gate, st);
    }

    // Cone step of the diagonal-Q prox: projection onto diag(d) K in the
    // metric diag(1 + tau q) (PSD blocks excluded; they take the BB path).
    template <class X> void project_diag_q(X *v, const double *d, const X *q, double tau, hipStream_t st) const {
        if (rsoc.count)
            hipLaunchKernelGGL((cone_project_diag_q_kernel<X, CONE_ROTATED_SOC>), blocks(rsoc.count), dim3(kBlock), 0,
                               st, v, d, q, tau, rsoc);
        if (soc.count)
            hipLaunchKernelGGL((cone_project_diag_q_kernel<X, CONE_STANDARD_SOC>), blocks(soc.count), dim3(kBlock), 0,
                               st, v, d, q, tau, soc);
        if (expc.count)
            hipLaunchKernelGGL((cone_project_diag_q_kernel<X, CONE_EXPONENTIAL>), blocks(expc.count), dim3(kBlock), 0,
                               st, v, d, q, tau, expc);
        if (powc.count)
            hipLaunchKernelGGL((cone_project_diag_q_kernel<X, CONE_POWER>), blocks(powc.count), dim3(kBlock), 0, st, v,
                               d, q, tau, powc);
        if (lrsoc.count)
            hipLaunchKernelGGL((soc_long_project_kernel<X, true>), dim3(lrsoc.count), dim3(kBlock), 0, st, v, d, q,
                               tau, lrsoc, 0, (const int *)nullptr);
        if (lsoc.count)
            hipLaunchKernelGGL((soc_long_project_kernel<X, false>), dim3(lsoc.count), dim3(kBlock), 0, st, v, d, q,
                               tau, lsoc, 0, (const int *)nullptr);
        if (sec.count)
            hipLaunchKernelGGL(section_project_kernel<X>, blocks(sec.count), dim3(kBlock), 0, st, v, d, q, tau, fixed,
                               sec, (const int *)nullptr);
        if (lsec.count)
            hipLaunchKernelGGL(section_long_project_kernel<X>, dim3(lsec.count), dim3(kBlock), 0, st, v, d, q, tau,
                               fixed, lsec, lsec_scratch, lsec_off, (const int *)nullptr);
    }

    // Dual-cone residuals (rc - P_{K*}(rc)) d of rc = c + qx - aty (qx may be
    // null) written over dres on every cone slot; scratch (n entries) holds the
    // PSD projections (reference src/cone_dispatch.cu:88-115).  Blocks with
    // fixed slots get the prox residual of their section at the point x.
    template <class X>
    void dual_residual(double *dres, X *scratch, const X *x, const X *c, const X *qx, const X *aty, const double *d,
                       int n, hipStream_t st) const {
        if (sec.count)
            hipLaunchKernelGGL(section_residual_kernel<X>, blocks(sec.count), dim3(kBlock), 0, st, dres, x, c, qx, aty,
                               d, fixed, sec);
        if (lsec.count)
            hipLaunchKernelGGL(section_long_residual_kernel<X>, dim3(lsec.count), dim3(kBlock), 0, st, dres, x, c, qx,
                               aty, d, fixed, lsec, lsec_scratch, lsec_off);
        if (rsoc.count)
            hipLaunchKernelGGL((cone_dual_residual_kernel<X, CONE_ROTATED_SOC>), blocks(rsoc.count), dim3(kBlock), 0,
                               st, dres, c, qx, aty, d, rsoc);
        if (soc.count)
            hipLaunchKernelGGL((cone_dual_residual_kernel<X, CONE_STANDARD_SOC>), blocks(soc.count), dim3(kBlock), 0,
                               st, dres, c, qx, aty, d, soc);
        if (expc.count)
            hipLaunchKernelGGL((cone_dual_residual_kernel<X, CONE_EXPONENTIAL>), blocks(expc.count), dim3(kBlock), 0,
                               st, dres, c, qx, aty, d, expc);
        if (powc.count)
            hipLaunchKernelGGL((cone_dual_residual_kernel<X, CONE_POWER>), blocks(powc.count), dim3(kBlock), 0, st,
                               dres, c, qx, aty, d, powc);
        if (lrsoc.count)
            hipLaunchKernelGGL((soc_long_dual_residual_kernel<X, true>), dim3(lrsoc.count), dim3(kBlock), 0, st, dres,
                               c, qx, aty, d, lrsoc);
        if (lsoc.count)
            hipLaunchKernelGGL((soc_long_dual_residual_kernel<X, false>), dim3(lsoc.count), dim3(kBlock), 0, st, dres,
                               c, qx, aty, d, lsoc);
        if (psd) {
            hipLaunchKernelGGL(psd_rc_kernel<X>, dim3(blocks_for(n)), dim3(kBlock), 0, st, scratch, dres, c, qx, aty,
                               mask, n);
            psd_launch<X>(pset, scratch, 1, nullptr, st);
            hipLaunchKernelGGL(psd_dres_kernel<X>, dim3(blocks_for(n)), dim3(kBlock), 0, st, dres, scratch, d, mask,
                               n);
        }
    }

    void release() {
        if (side_) {
            (void)hipStreamSynchronize(side_);
            (void)hipStreamDestroy(side_);
            (void)hipEventDestroy(fork_);
            (void)hipEventDestroy(join_);
            side_ = nullptr;
            fork_ = join_ = nullptr;
        }
        if (arena_) {  // arena memory belongs to the caller
            *this = DevConeSet::Cleared();
            return;
        }
        for (ConeList *L : {&soc, &rsoc, &expc, &powc, &lsoc, &lrsoc, &powall}) {
            void *lp[] = {L->start, L->vdim, L->alpha, L->warm, L->rwarm};
            for (void *p : lp)
                if (p) (void)hipFree(p);
            *L = ConeList();
        }
        for (void *p : {(void *)sec.start, (void *)sec.vdim, (void *)sec.type, (void *)sec.alpha, (void *)fixed,
                        (void *)lsec.start, (void *)lsec.vdim, (void *)lsec.type, (void *)lsec.alpha,
                        (void *)lsec_scratch, (void *)lsec_off})
            if (p) (void)hipFree(p);
        sec = SectionList();
        lsec = SectionList();
        lsec_scratch = nullptr;
        lsec_off = nullptr;
        fixed = nullptr;
        if (pset.blocks) (void)hipFree(pset.blocks);
        if (pset.work) (void)hipFree(pset.work);
        pset = PsdSet();
        for (int *p : {all_start, all_len})
            if (p) (void)hipFree(p);
        if (mask) (void)hipFree(mask);
        mask = nullptr;
        all_start = all_len = nullptr;
        total = 0;
    }

  private:
    static dim3 blocks(int count) { return dim3((unsigned)((count + kBlock - 1) / kBlock)); }

    DevArena *arena_ = nullptr;
    hipStream_t side_ = nullptr;  // the PSD stream of enable_overlap()
    hipEvent_t fork_ = nullptr, join_ = nullptr;
    struct Cleared {};
    DevConeSet &operator=(Cleared) {
        psd = power = false;
        rsoc = soc = expc = powc = lrsoc = lsoc = powall = ConeList();
        sec = lsec = SectionList();
        lsec_scratch = nullptr;
        lsec_off = nullptr;
        fixed = mask = nullptr;
        pset = PsdSet();
        total = 0;
        all_start = all_len = nullptr;
        arena_ = nullptr;
        return *this;
    }
    template <class T, class S> T *up(const std::vector<S> &h, size_t count) {
        if (!arena_) return dupload<T>(h, count);
        T *p = arena_->take<T>(count);
        if (count && !arena_->measuring()) {
            std::vector<T> tmp(count);
            for (size_t i = 0; i < count; ++i) tmp[i] = (T)h[i];
            HIP_OK(hipMemcpy(p, tmp.data(), count * sizeof(T), hipMemcpyHostToDevice));
        }
        return p;
    }
    template <class T> T *zeros(size_t count) {
        if (!arena_) return dzeros<T>(count);
        T *p = arena_->take<T>(count);
        if (!arena_->measuring()) HIP_OK(hipMemset(p, 0, std::max<size_t>(count, 1) * sizeof(T)));
        return p;
    }
    template <class T> T *raw(size_t count) { return arena_ ? arena_->take<T>(count) : dalloc<T>(count); }
    bool measuring() const { return arena_ && arena_->measuring(); }
};

}  // namespace

class SolverBase {
  public:
    virtual ~SolverBase() = default;
    virtual int run(int iterations) = 0;
    virtual pdhcg_result_t *result(const qp_problem_t *original) = 0;
    virtual void counts(int *total, int *inner) = 0;
    virtual void synchronize() = 0;
    virtual double setup_seconds() const = 0;
    virtual KernelTimer &timer() = 0;
    virtual void graph_stats(int *windows, int *iterations) const = 0;
};

template <class V, class X> class Solver final : public SolverBase {
    template <class, class> friend class Solver;

  public:
    // eval_only: an FP64 evaluation copy of the problem for the mixed-precision
    // solver (KKT residuals, infeasibility rays, result); it shares the
    // owner's stream and runs no iterations of its own.
    Solver(const ScaledProblem &sp, const pdhg_parameters_t &params, hipStream_t shared = nullptr,
           bool eval_only = false);
    ~Solver() override;
    int run(int iterations) override;
    pdhcg_result_t *result(const qp_problem_t *original) override;
    void counts(int *total, int *inner) override {
        if (promoted_) return eval_->counts(total, inner);
        read_scalars();
        *total = total_count_;
        *inner = host_inner_ + hs_->inner_total;
    }
    void synchronize() override { HIP_OK(hipStreamSynchronize(st_)); }
    double setup_seconds() const override { return setup_seconds_; }
    KernelTimer &timer() override { return timer_; }
    void graph_stats(int *windows, int *iterations) const override {
        *windows = graph_windows_;
        *iterations = graph_iters_;
    }

  private:
    using Acc = typename std::conditional<std::is_same<V, double>::value || std::is_same<X, double>::value, double,
                                          float>::type;

    // ---- setup helpers
    void upload(const ScaledProblem &sp);
    void tune_spmv();
    float time_launches(const std::function<void()> &launch);
    int shift_A_ = 0, shift_At_ = 0;  // autotuned lane shifts of the A and A' products
    void setup_cones(const ScaledProblem &sp, std::vector<double> &xs);
    void init_inner(const ScaledProblem &sp);
    double power_sv();
    double power_eig(bool minimum, double shift);

    // ---- device operations
    GridRed red(int offset = 0) const { return GridRed{red_partials_, ds_->red + offset}; }
    // Folds the partials of the reducing launch just enqueued (grid of
    // `blocks`) into ds_->red[offset...] and runs its tail on the device.
    template <int NS, int NM, class Tail = NoTail>
    void fold(int blocks, int offset = 0, Tail tail = Tail(), const int *gate = nullptr) {
        hipLaunchKernelGGL((grid_fold_kernel<NS, NM, Tail, DevScalars>), dim3(1), dim3(kBlock), 0, st_, red(offset),
                           blocks, tail, ds_, gate);
    }
    void read_scalars() {
        HIP_OK(hipMemcpyAsync(hs_, ds_, sizeof(DevScalars), hipMemcpyDeviceToHost, st_));
        HIP_OK(hipStreamSynchronize(st_));
    }
    int grid(long long n) const { return blocks_for(n); }
    // one thread per element, no cap (kernels without a grid reduction)
    static int full_grid(long long n) { return (int)std::max<long long>(1, (n + kBlock - 1) / kBlock); }
    // LP and diagonal-Q primal steps are closed form; the rest use BB
    bool closed_prox() const { return qtype_ == PDHCG_NON_Q || (qtype_ == PDHCG_DIAG_Q && !has_psd_); }
    void q_product(const X *x, X *qx, const int *gate);
    void project_cones(X *v, bool residual_warm, const int *gate) {
        cones_.project(v, var_rescale_, residual_warm, gate, st_);
    }
    void bb_prox(double tau, bool fuse_halpern, double w);
    void pdhg_step(bool fuse_halpern);
    void halpern(double w);
    void fixed_point_error();
    void restart();
    bool should_restart();
    void residual();
    void infeasibility();
    void check_termination();

    pdhg_parameters_t P_;
    hipStream_t st_ = nullptr;
    bool own_stream_ = true, eval_only_ = false;
    // Mixed precision: the termination test, the infeasibility rays and the
    // returned point are evaluated in FP64 on FP64 problem data, from the
    // FP32 iterates widened exactly (null in the FP64 solver).
    Solver<double, double> *eval_ = nullptr;
    void eval_residual();
    void eval_infeasibility();
    // FP32 iterates stop improving at a floor set by their rounding (a
    // relative KKT error near 1e-6 on problems whose |A||x| is large next to
    // the right-hand sides).  When the FP64-evaluated error has not halved
    // for stall_iters_ iterations nor for half of the iterations so far, the
    // state moves to the evaluation copy,
    // which finishes the solve with FP64 iterates (promoted_).  The inner
    // BB solve meets the same floor sooner: its tolerance follows the primal
    // movement down to 1e-9, below what FP32 steps can resolve, and every
    // inner solve then runs to the inner iteration limit.  kSatPromote such
    // solves in a row also hand over to FP64.
    static constexpr int kSatPromote = 4;
    bool promoted_ = false;
    double best_err_ = INFINITY;
    int best_err_iter_ = 0, stall_iters_ = 1000, sat_run_ = 0;
    bool stalled(int tc);
    void promote();
    int n_ = 0, m_ = 0, k_ = 0;
    quad_obj_type_t qtype_ = PDHCG_NON_Q;
    bool is_lp_ = true;
    double cbr_ = 1, ovr_ = 1, obj_const_ = 0, obj_norm_ = 0, bound_norm_ = 0, rescale_time_ = 0;
    double setup_seconds_ = 0;
    // setup phases (name, seconds), printed at verbose >= 1
    std::vector<std::pair<const char *, double>> laps_;
    std::chrono::steady_clock::time_point lap_t_;
    void lap(const char *name) {
        const auto now = std::chrono::steady_clock::now();
        laps_.emplace_back(name, std::chrono::duration<double>(now - lap_t_).count());
        lap_t_ = now;
    }

    BinnedCsr<V> A_, At_, Q_, Rc_, Rtc_;
    PanelSpmv<V, Acc> pnA_, pnAt_;  // column-panel forms, built when x is larger than an L2

    // y = A x and A'y through the panel form when it is built, with the epilogue
    template <class Epi> void apply_A(const X *x, Epi epi, hipStream_t s, const int *gate = nullptr) {
        if (pnA_.on) pnA_.launch(x, epi, s, gate);
        else spmv_launch<V, X, Acc>(A_, x, epi, s, gate);
    }
    template <class Epi> void apply_At(const X *y, Epi epi, hipStream_t s, const int *gate = nullptr) {
        if (pnAt_.on) pnAt_.launch(y, epi, s, gate);
        else spmv_launch<V, X, Acc>(At_, y, epi, s, gate);
    }
    bool q_sparse_ = false;          // Q stored as CSR (has off-diagonal entries)
    X *qsdiag_ = nullptr;            // Q stored as its diagonal
    X *qdiag_ = nullptr;             // DIAG_Q prox diagonal (diag + 1e-12)
    V *Rd_ = nullptr;                // dense R in column tiles (rx_tile_index), zero padded
    bool r_dense_ = false;
    int dkind_ = 0;
    double *ddiag_ = nullptr, *ddense_ = nullptr, *u_ = nullptr, *rx_partials_ = nullptr, *rx_sums_ = nullptr;
    unsigned *rx_counter_ = nullptr;
    int rx_chunks_ = 0;
    X *rxv_ = nullptr;
    int rx_blocks_ = 0;

    X *lb_ = nullptr, *ub_ = nullptr, *c_ = nullptr, *cl_ = nullptr, *cu_ = nullptr, *g_ = nullptr;
    X *cl_fin_ = nullptr, *cu_fin_ = nullptr, *lb_fin_ = nullptr, *ub_fin_ = nullptr;
    double *con_rescale_ = nullptr, *var_rescale_ = nullptr;
    X *x0_ = nullptr, *x_ = nullptr, *xp_ = nullptr, *xr_ = nullptr;
    X *y0_ = nullptr, *y_ = nullptr, *yp_ = nullptr, *yr_ = nullptr;
    X *aty_ = nullptr, *ax_ = nullptr, *dslack_ = nullptr, *dx_ = nullptr, *dy_ = nullptr, *qx_ = nullptr;

    // variable cone blocks, one list per type (reference internal/internal_types.h:106-127)
    bool has_cones_ = false, has_psd_ = false, has_power_ = false;
    DevConeSet cones_;
    // affine cones on the rows of [A; F] (reference src/solver_state.cu:621-638)
    bool has_affine_ = false;
    DevConeSet acones_;
    double *aff_d_ = nullptr, *acomp_ = nullptr;  // 1 / row scaling; per-block complementarity
    X *aproj_ = nullptr, *apres_ = nullptr, *amem_ = nullptr;
    X *snap_ = nullptr, *pgp_ = nullptr;
    double *dres_ = nullptr;

    // BB
    bool bb_ = false, precond_ = false;
    X *grad_ = nullptr, *dirn_ = nullptr, *mdiag_ = nullptr, *minv_ = nullptr;
    double *diag_h_ = nullptr;
    double cached_inv_tau_ = -1.0, inner_tol_ = 1e-3, inner_min_tol_ = 1e-9;
    int inner_limit_ = 1000;
    int host_inner_ = 0;  // +1 per outer iteration (reference pdhg_core_op.cu:762)

    // reductions and device scalars
    double *red_partials_ = nullptr;
    DevScalars *ds_ = nullptr, *hs_ = nullptr;
    int *bb_poll_ = nullptr;   // pinned: two (bb_active, bb_iter) poll slots
    hipEvent_t bb_ev_[2] = {nullptr, nullptr};
    int bb_pred_ = 2;          // BB bodies the previous inner solve ran

    // Recorded windows of plain iterations (closed-form prox, no cones): runs
    // of iterations with no evaluation, restart or fixed-point error are
    // replayed from an instantiated hipGraph holding their SpMV launches.
    // One graph per (length, store pattern); the window's scalars are set by
    // the graph's prologue node before each replay.
    struct GraphWindow {
        int len = 0;
        std::vector<char> store;
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        hipGraphNode_t prologue = nullptr;
        hipEvent_t done = nullptr;
        long long used = 0;
    };
    static constexpr int kMinGraphWindow = 4, kMaxGraphWindow = 256, kMaxGraphs = 8;
    std::vector<GraphWindow> graphs_;
    double *dyn_ = nullptr;
    bool graph_on_ = false;
    int graph_windows_ = 0, graph_iters_ = 0;
    long long graph_clock_ = 0;
    bool plain_iteration(int t) const;
    double primal_tau() const;
    GraphWindow &graph_for(int len, const std::vector<char> &store);
    void run_window(int len);
    void release_graph(GraphWindow &g);

    // outer-loop state (reference internal/internal_types.h:137-236)
    double q_norm_ = 0, nonconvexity_ = 0;
    double step_ = 0, pw_ = 1, pw_err_sum_ = 0, pw_last_err_ = 0, best_pw_ = 1, best_gap_ = INFINITY;
    double fpe_ = 0, init_fpe_ = 0, last_trial_fpe_ = INFINITY;
    int total_count_ = 0, inner_count_ = 0;
    bool is_major_ = false, do_restart_ = false;
    double abs_pr_ = 0, rel_pr_ = 0, abs_du_ = 0, rel_du_ = 0, pobj_ = 0, dobj_ = 0, gap_ = 0, rel_gap_ = 0;
    double max_pray_ = 0, max_dray_ = 0, pray_lin_obj_ = 0, dray_obj_ = 0;
    termination_reason_t reason_ = TERMINATION_REASON_UNSPECIFIED;
    double cum_time_ = 0;
    std::chrono::steady_clock::time_point t_start_;
    KernelTimer timer_;
};

template <class V, class X>
Solver<V, X>::Solver(const ScaledProblem &sp, const pdhg_parameters_t &params, hipStream_t shared, bool eval_only)
    : P_(params) {
    const auto t0 = std::chrono::steady_clock::now();
    lap_t_ = t0;
    eval_only_ = eval_only;
    if (shared) {
        st_ = shared;
        own_stream_ = false;
    } else HIP_OK(hipStreamCreateWithFlags(&st_, hipStreamNonBlocking));
    n_ = sp.n;
    m_ = sp.m;
    k_ = sp.k;
    qtype_ = sp.qtype;
    is_lp_ = qtype_ == PDHCG_NON_Q;
    cbr_ = sp.cbr;
    ovr_ = sp.ovr;
    obj_const_ = sp.obj_const;
    obj_norm_ = sp.obj_norm;
    bound_norm_ = sp.bound_norm;
    rescale_time_ = sp.rescale_time;
    red_partials_ = dzeros<double>(kRedPartials);
    ds_ = dzeros<DevScalars>(1);
    HIP_OK(hipHostMalloc((void **)&hs_, sizeof(DevScalars), hipHostMallocDefault));
    std::memset(hs_, 0, sizeof(DevScalars));
    HIP_OK(hipHostMalloc((void **)&bb_poll_, 4 * sizeof(int), hipHostMallocDefault));
    for (hipEvent_t &e : bb_ev_) HIP_OK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    upload(sp);
    if (eval_only_) {
        lap("vectors and cones");
        HIP_OK(hipStreamSynchronize(st_));
        return;
    }
    if constexpr (std::is_same<X, float>::value) {
        eval_ = new Solver<double, double>(sp, params, st_, true);
        // PDHCG_MIXED_STALL=N: iterations without a halving of the FP64 KKT
        // error before the FP64 finish (0: never)
        const char *e = std::getenv("PDHCG_MIXED_STALL");
        stall_iters_ = e ? std::atoi(e) : std::max(1000, 5 * P_.termination_evaluation_frequency);
        lap("FP64 evaluation copy");
    }

    // quadratic term information (reference src/solver_state.cu:380-420)
    if (qtype_ == PDHCG_SPARSE_Q) {
        q_norm_ = power_eig(false, 0.0);
        nonconvexity_ = power_eig(true, q_norm_);
    } else if (qtype_ == PDHCG_DIAG_Q) {
        double mx = 0.0, mn = 0.0;
        for (double q : sp.qdiag) {
            mx = std::fmax(mx, std::fabs(q));
            mn = std::fmin(mn, q);
        }
        q_norm_ = mx;
        nonconvexity_ = mn;
    }
    init_inner(sp);
    if (nonconvexity_ < 0) inner_limit_ = 1;
    lap("Q estimates and inner solver");
    // step size and primal weight (reference src/pdhg_core_op.cu:933-971)
    if (A_.nnz == 0) step_ = 1.0;
    else {
        const double sv = power_sv();
        step_ = sv < 1e-12 ? 1.0 : 0.998 / sv;
    }
    pw_ = P_.bound_objective_rescaling ? 1.0 : (obj_norm_ + 1.0) / (bound_norm_ + 1.0);
    best_pw_ = pw_;
    graph_on_ = graph_window_default() && closed_prox() && !has_cones_ && !has_affine_;
    if (graph_on_) dyn_ = dzeros<double>(4);
    HIP_OK(hipStreamSynchronize(st_));
    lap("step size");
    t_start_ = std::chrono::steady_clock::now();
    setup_seconds_ = std::chrono::duration<double>(t_start_ - t0).count();
    if (P_.verbose >= 1) {
        std::fprintf(stderr, "[setup] host scaling %.3fs, device setup %.3fs:", sp.rescale_time, setup_seconds_);
        for (const auto &l : laps_) std::fprintf(stderr, " %s %.3fs;", l.first, l.second);
        std::fprintf(stderr, "\n");
    }
}

template <class V, class X> void Solver<V, X>::upload(const ScaledProblem &sp) {
    A_ = upload_csr<V>(sp.A);
    // panels pay off when the gathered vector is larger than one XCD's 4 MiB
    // L2 (measured on an 8 MB x: 145 -> 109 us; on a 4 MB y they lose,
    // 99 -> 123 us); PDHCG_SPMV_PANELS=0 turns them off, =1 forces them
    const char *pe = std::getenv("PDHCG_SPMV_PANELS");
    const int panel_mode = pe ? std::atoi(pe) : -1;
    const size_t kPanelMinBytes = (size_t)4 << 20;
    auto want_panels = [&](int cols, int nnz, int rows) {
        if (panel_mode == 0 || eval_only_) return false;
        if (panel_mode == 1) return rows > 0;
        return (size_t)cols * sizeof(X) > kPanelMinBytes && nnz >= 4 * rows;
    };
    lap("A binned");
    if (want_panels(n_, sp.A.nnz(), m_)) pnA_.build(m_, n_, sp.A.rp, sp.A.ci, sp.A.v);
    lap("A column panels");
    {
        // the device scaling hands over the transpose it built
        HostCSR own;
        if (sp.At.rp.empty()) own = transpose(sp.A);
        const HostCSR &t = sp.At.rp.empty() ? own : sp.At;
        lap("A transposed");
        At_ = upload_csr<V>(t);
        if (want_panels(m_, t.nnz(), n_)) pnAt_.build(n_, m_, t.rp, t.ci, t.v);
        lap("A' binned");
    }
    if (qtype_ == PDHCG_SPARSE_Q || qtype_ == PDHCG_LOW_RANK_PLUS_SPARSE_Q) {
        bool diag_only = true;
        for (int i = 0; i < n_ && diag_only; ++i)
            for (int e = sp.Q.rp[i]; e < sp.Q.rp[(size_t)i + 1]; ++e)
                if (sp.Q.ci[e] != i) {
                    diag_only = false;
                    break;
                }
        if (diag_only) {
            std::vector<double> d((size_t)n_, 0.0);
            for (int i = 0; i < n_; ++i)
                for (int e = sp.Q.rp[i]; e < sp.Q.rp[(size_t)i + 1]; ++e) d[i] += sp.Q.v[e];
            qsdiag_ = dupload<X>(d, (size_t)n_);
        } else {
            q_sparse_ = true;
            Q_ = upload_csr<V>(sp.Q);
        }
    }
    if (qtype_ == PDHCG_DIAG_Q) qdiag_ = dupload<X>(sp.qdiag, (size_t)n_);
    if (qtype_ == PDHCG_LOW_RANK_Q || qtype_ == PDHCG_LOW_RANK_PLUS_SPARSE_Q) {
        // the dense tile pass keeps one LDS slot per row of R, so ranks above
        // kRxMaxRank take the CSR products, which have no rank limit
        const long long dense_elems = (long long)k_ * n_;
        r_dense_ = dense_elems <= 4LL * sp.R.nnz() + 4096 && k_ <= kRxMaxRank;
        if (r_dense_) {
            rx_chunks_ = rx_chunks(n_);
            const size_t tiles = (size_t)rx_chunks_ * kRxWaveCols * k_;
            // rows with strictly increasing columns (the usual case) are
            // scattered into the tiles on the device; duplicate entries are
            // summed in entry order on the host
            bool unique = true;
            for (int j = 0; j < k_ && unique; ++j)
                for (int e = sp.R.rp[j] + 1; e < sp.R.rp[(size_t)j + 1]; ++e)
                    if (sp.R.ci[e] <= sp.R.ci[e - 1]) {
                        unique = false;
                        break;
                    }
            const long long rnnz = sp.R.nnz();
            if (unique && rnnz > 0) {
                Rd_ = dzeros<V>(tiles);
                int *drp = dupload<int>(sp.R.rp, sp.R.rp.size()), *dci = dupload<int>(sp.R.ci, (size_t)rnnz);
                double *dv = dupload<double>(sp.R.v, (size_t)rnnz);
                hipLaunchKernelGGL(r_tiles_kernel<V>, dim3(full_grid(rnnz)), dim3(kBlock), 0, nullptr,
                                   (const int *)drp, (const int *)dci, (const double *)dv, k_, rnnz, Rd_);
                HIP_OK(hipGetLastError());
                HIP_OK(hipDeviceSynchronize());
                for (void *q : {(void *)drp, (void *)dci, (void *)dv}) HIP_OK(hipFree(q));
            } else {
                std::vector<double> Rd(tiles, 0.0);
                for (int j = 0; j < k_; ++j)
                    for (int e = sp.R.rp[j]; e < sp.R.rp[(size_t)j + 1]; ++e)
                        Rd[rx_tile_index(j, sp.R.ci[e], k_)] += sp.R.v[e];
                Rd_ = dupload<V>(Rd, Rd.size());
            }
            rx_blocks_ = std::max(1, (rx_chunks_ + kBlock / kWave - 1) / (kBlock / kWave));
            rx_partials_ = dzeros<double>((size_t)rx_chunks_ * k_);
            rx_sums_ = dzeros<double>((size_t)k_);
            rx_counter_ = dzeros<unsigned>(1);
        } else {
            Rc_ = upload_csr<V>(sp.R);
            HostCSR rt = transpose(sp.R);
            Rtc_ = upload_csr<V>(rt);
            rxv_ = dzeros<X>((size_t)k_);
        }
        dkind_ = (int)sp.dkind;
        if (sp.dkind == MiddleKind::Diag) ddiag_ = dupload<double>(sp.ddiag, sp.ddiag.size());
        if (sp.dkind == MiddleKind::Dense) ddense_ = dupload<double>(sp.ddense, sp.ddense.size());
        u_ = dzeros<double>((size_t)k_);
    }
    lap("Q and R");
    lb_ = dupload<X>(sp.lb, n_);
    ub_ = dupload<X>(sp.ub, n_);
    c_ = dupload<X>(sp.c, n_);
    cl_ = dupload<X>(sp.cl, m_);
    cu_ = dupload<X>(sp.cu, m_);
    g_ = dupload<X>(sp.g, m_);
    auto finite_or_zero = [](const std::vector<double> &v) {
        std::vector<double> r(v.size());
        for (size_t i = 0; i < v.size(); ++i) r[i] = std::isfinite(v[i]) ? v[i] : 0.0;
        return r;
    };
    cl_fin_ = dupload<X>(finite_or_zero(sp.cl), m_);
    cu_fin_ = dupload<X>(finite_or_zero(sp.cu), m_);
    lb_fin_ = dupload<X>(finite_or_zero(sp.lb), n_);
    ub_fin_ = dupload<X>(finite_or_zero(sp.ub), n_);
    con_rescale_ = dupload<double>(sp.con_rescale, m_);
    var_rescale_ = dupload<double>(sp.var_rescale, n_);

    // starting point (reference src/solver_state.cu:809-829,976-995)
    std::vector<double> xs((size_t)n_, 0.0), ys((size_t)m_, 0.0);
    if (!sp.primal_start.empty())
        for (int i = 0; i < n_; ++i) xs[i] = sp.primal_start[i] * sp.var_rescale[i] * sp.cbr;
    if (!sp.dual_start.empty())
        for (int i = 0; i < m_; ++i) ys[i] = sp.dual_start[i] * sp.con_rescale[i] * sp.ovr;
    setup_cones(sp, xs);
    if (sp.affine_cones && sp.affine_cones->num_cones > 0 && m_ > 0) {
        has_affine_ = true;
        acones_.build(sp.affine_cones, m_);
        std::vector<double> inv((size_t)m_);
        for (int i = 0; i < m_; ++i) inv[i] = 1.0 / sp.con_rescale[i];
        aff_d_ = dupload<double>(inv, inv.size());
        acomp_ = dzeros<double>((size_t)acones_.total);
        aproj_ = dzeros<X>(m_);
        apres_ = dzeros<X>(m_);
        amem_ = dzeros<X>(m_);
    }
    x0_ = dupload<X>(xs, n_);
    if (has_cones_) project_cones(x0_, false, nullptr);
    hipLaunchKernelGGL(clamp_kernel<X>, dim3(grid(n_)), dim3(kBlock), 0, st_, x0_, lb_, ub_, n_);
    x_ = dalloc<X>(n_);
    xp_ = dalloc<X>(n_);
    xr_ = dalloc<X>(n_);
    for (X *dst : {x_, xp_, xr_})
        HIP_OK(hipMemcpyAsync(dst, x0_, (size_t)n_ * sizeof(X), hipMemcpyDeviceToDevice, st_));
    HIP_OK(hipStreamSynchronize(st_));
    y0_ = dupload<X>(ys, m_);
    y_ = dupload<X>(ys, m_);
    yp_ = dupload<X>(ys, m_);
    yr_ = dzeros<X>(m_);
    aty_ = dzeros<X>(n_);
    ax_ = dzeros<X>(m_);
    dslack_ = dzeros<X>(n_);
    dx_ = dzeros<X>(n_);
    dy_ = dzeros<X>(m_);
    qx_ = dzeros<X>(n_);
    if (!eval_only_) {
        lap("vectors and cones");
        tune_spmv();
        lap("SpMV tuning");
    }
}

// Mean duration (ms) of a launch sequence on the solver stream: one
// untimed warm-up, then five timed repetitions.
template <class V, class X> float Solver<V, X>::time_launches(const std::function<void()> &launch) {
    hipEvent_t a, b;
    HIP_OK(hipEventCreate(&a));
    HIP_OK(hipEventCreate(&b));
    launch();
    HIP_OK(hipEventRecord(a, st_));
    for (int r = 0; r < 5; ++r) launch();
    HIP_OK(hipEventRecord(b, st_));
    HIP_OK(hipEventSynchronize(b));
    float ms = 0.0f;
    HIP_OK(hipEventElapsedTime(&ms, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return ms / 5.0f;
}

// SpMV autotuner: times the lane shifts -1, 0, +1 of the A and A' products
// on the problem's own matrices (the shift moves every row class up or down
// one power of two of lanes; the stored layout is the same) and keeps the
// fastest, preferring the unshifted form unless another is 2% faster.  When
// the column-panel form of a product was built, it is kept only if it beats
// the best plain form (unless PDHCG_SPMV_PANELS=1 forces panels).
// PDHCG_SPMV_SHIFT forces a shift.
template <class V, class X> void Solver<V, X>::tune_spmv() {
    const char *forced = std::getenv("PDHCG_SPMV_SHIFT");
    const char *pe = std::getenv("PDHCG_SPMV_PANELS");
    const bool panels_forced = pe && std::atoi(pe) == 1;
    X *xs = dzeros<X>((size_t)std::max(n_, m_));
    X *ys = dzeros<X>((size_t)std::max(n_, m_));
    auto tune = [&](BinnedCsr<V> &M, PanelSpmv<V, Acc> &pn) {
        int best = 0;
        if (forced) best = std::atoi(forced);
        else if (M.total_blocks > 0) {
            float best_ms = time_launches([&] { spmv_launch<V, X, Acc>(M, xs, EpiStore<X>{ys}, st_); });
            for (int sh : {-1, 1}) {
                set_lane_shift(M, sh);
                const float ms = time_launches([&] { spmv_launch<V, X, Acc>(M, xs, EpiStore<X>{ys}, st_); });
                if (ms < 0.98f * best_ms) {
                    best_ms = ms;
                    best = sh;
                }
            }
            if (pn.on && !forced && !panels_forced) {
                const float pms = time_launches([&] { pn.launch(xs, EpiStore<X>{ys}, st_, nullptr); });
                if (!(pms < best_ms)) pn.release();
            }
        }
        set_lane_shift(M, best);
        return best;
    };
    shift_A_ = tune(A_, pnA_);
    shift_At_ = tune(At_, pnAt_);
    HIP_OK(hipStreamSynchronize(st_));
    (void)hipFree(xs);
    (void)hipFree(ys);
    if (P_.verbose >= 1)
        std::fprintf(stderr, "[spmv] lane shift A %d%s, A' %d%s\n", shift_A_, pnA_.on ? " (column panels)" : "", shift_At_,
                    pnAt_.on ? " (column panels)" : "");
}

// Initial auxiliary values of SOC / RSOC blocks written into the
// (unprojected) starting point (reference src/solver_state.cu:578-686), and
// the device cone set with its work vectors.
template <class V, class X> void Solver<V, X>::setup_cones(const ScaledProblem &sp, std::vector<double> &xs) {
    const cone_blocks_t *b = sp.cones;
    if (!b || b->num_cones <= 0) return;
    cones_.build(b, n_);
    cones_.enable_overlap();
    // fixed slots keep their start value (reference src/solver_state.cu:650-651,673)
    auto free_slot = [&](int i) { return !(b->is_fixed && b->is_fixed[i]); };
    for (int k = 0; k < b->num_cones; ++k) {
        const int aux = b->start_idx[k] + b->v_dim[k];
        if (b->type[k] == CONE_STANDARD_SOC) {
            if (free_slot(aux)) xs[aux] = -kInvSqrt2 * sp.cbr * sp.var_rescale[aux];
            if (free_slot(aux + 1)) xs[(size_t)aux + 1] = kInvSqrt2 * sp.cbr * sp.var_rescale[(size_t)aux + 1];
        } else if (b->type[k] == CONE_ROTATED_SOC) {
            if (free_slot(aux + 1)) xs[(size_t)aux + 1] = sp.cbr * sp.var_rescale[(size_t)aux + 1];
        }
    }
    has_cones_ = true;
    has_psd_ = cones_.psd;
    has_power_ = cones_.power;
    pgp_ = dzeros<X>(n_);
    dres_ = dzeros<double>(n_);
}

// Jacobi metric diag(Q) + diag(R'DR) from the scaled host data
// (reference src/solver_state.cu:280-370, kernels pdhcg_kernels.cu:407-495).
template <class V, class X> void Solver<V, X>::init_inner(const ScaledProblem &sp) {
    int lim = P_.inner_solver_parameters.iteration_limit;
    double t0 = P_.inner_solver_parameters.initial_tolerance, tmin = P_.inner_solver_parameters.min_tolerance;
    if (lim < 1) {
        std::fprintf(stderr, "Warning: inner_iter_limit (%d) < 1. Resetting to 1.\n", lim);
        lim = 1;
    }
    if (t0 <= 0.0) t0 = 1e-3;
    if (tmin <= 0.0) tmin = 1e-9;
    inner_limit_ = lim;
    // the diagonal-Q prox is closed form unless a PSD block needs the BB solve
    // (reference src/pdhg_core_op.cu:740-741)
    if (qtype_ == PDHCG_NON_Q || (qtype_ == PDHCG_DIAG_Q && !has_psd_)) return;
    bb_ = true;
    inner_tol_ = t0;
    inner_min_tol_ = tmin;
    grad_ = dzeros<X>(n_);
    dirn_ = dzeros<X>(n_);
    if (has_cones_) snap_ = dzeros<X>(n_);
    precond_ = P_.diag_jacobi_precond && !has_psd_;
    if (!precond_) return;
    std::vector<double> dh((size_t)n_, 0.0);
    if (qtype_ == PDHCG_SPARSE_Q || qtype_ == PDHCG_LOW_RANK_PLUS_SPARSE_Q)
        for (int i = 0; i < n_; ++i)
            for (int e = sp.Q.rp[i]; e < sp.Q.rp[(size_t)i + 1]; ++e)
                if (sp.Q.ci[e] == i) dh[i] += sp.Q.v[e];
    if ((qtype_ == PDHCG_LOW_RANK_Q || qtype_ == PDHCG_LOW_RANK_PLUS_SPARSE_Q) && sp.dkind == MiddleKind::Dense) {
        const HostCSR rt = transpose(sp.R);
        for (int i = 0; i < n_; ++i) {
            double acc = 0.0;
            const int e0 = rt.rp[i], e1 = rt.rp[(size_t)i + 1];
            for (int a = e0; a < e1; ++a) {
                const double *Drow = sp.ddense.data() + (size_t)rt.ci[a] * k_;
                for (int b = e0; b < e1; ++b) acc += rt.v[a] * rt.v[b] * Drow[rt.ci[b]];
            }
            dh[i] += acc;
        }
    } else if (qtype_ == PDHCG_LOW_RANK_Q || qtype_ == PDHCG_LOW_RANK_PLUS_SPARSE_Q) {
        // diag(R'DR) for diagonal (or no) D, accumulated row by row of R: each
        // column receives its terms in increasing row order, as a pass over
        // the transpose would add them, without building the transpose
        std::vector<double> acc((size_t)n_, 0.0);
        const bool dg = sp.dkind == MiddleKind::Diag;
        for (int j = 0; j < k_; ++j)
            for (int e = sp.R.rp[j]; e < sp.R.rp[(size_t)j + 1]; ++e) {
                const double v = sp.R.v[e];
                acc[sp.R.ci[e]] += dg ? sp.ddiag[j] * v * v : v * v;
            }
        for (int i = 0; i < n_; ++i) dh[i] += acc[i];
    }
    diag_h_ = dupload<double>(dh, n_);
    mdiag_ = dzeros<X>(n_);
    minv_ = dzeros<X>(n_);
}

template <class V, class X> Solver<V, X>::~Solver() {
    if (st_) (void)hipStreamSynchronize(st_);
    delete eval_;
    for (GraphWindow &g : graphs_) release_graph(g);
    if (dyn_) (void)hipFree(dyn_);
    free_binned(A_);
    free_binned(At_);
    free_binned(Q_);
    free_binned(Rc_);
    free_binned(Rtc_);
    pnA_.release();
    pnAt_.release();
    void *ptrs[] = {qsdiag_, qdiag_, Rd_,     ddiag_,  ddense_,      u_,          rx_partials_, rx_counter_, rx_sums_,
                    rxv_,    lb_,    ub_,     c_,      cl_,          cu_,         g_,           cl_fin_,
                    cu_fin_, lb_fin_, ub_fin_, con_rescale_, var_rescale_, x0_,     x_,           xp_,
                    xr_,     y0_,    y_,      yp_,     yr_,          aty_,        ax_,          dslack_,
                    dx_,     dy_,    qx_,     grad_,   dirn_,        mdiag_,      minv_,        diag_h_,
                    red_partials_, ds_, snap_, pgp_, dres_, aff_d_, acomp_, aproj_, apres_, amem_};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    cones_.release();
    acones_.release();
    if (hs_) (void)hipHostFree(hs_);
    if (bb_poll_) (void)hipHostFree(bb_poll_);
    for (hipEvent_t e : bb_ev_)
        if (e) (void)hipEventDestroy(e);
    if (st_ && own_stream_) (void)hipStreamDestroy(st_);
}

// ||A||_2 by power iteration on A A' (reference src/pdhg_core_op.cu:1673-1788).
template <class V, class X> double Solver<V, X>::power_sv() {
    std::vector<double> e((size_t)m_);
    unsigned int seed = 1234;
    for (int i = 0; i < m_; ++i) e[i] = ((double)rand_r(&seed) / (double)RAND_MAX) * 2.0 - 1.0;
    double nrm = 0.0;
    for (double v : e) nrm += v * v;
    nrm = std::sqrt(nrm);
    const double inv = 1.0 / std::sqrt(nrm * nrm);
    for (double &v : e) v *= inv;
    X *buf[2] = {dupload<X>(e, m_), dzeros<X>(m_)};
    X *dp = dzeros<X>(n_);
    std::memset(hs_, 0, sizeof(DevScalars));
    hs_->pi_active = 1;
    hs_->pi_value = 1.0;
    HIP_OK(hipMemcpyAsync(ds_, hs_, sizeof(DevScalars), hipMemcpyHostToDevice, st_));
    const int *gate = &ds_->pi_active;
    int it = 0, chunk = 8;
    while (it < P_.sv_max_iter) {
        const int todo = std::min(chunk, P_.sv_max_iter - it);
        for (int c = 0; c < todo; ++c, ++it) {
            X *ev = buf[it % 2], *ne = buf[(it + 1) % 2];
            apply_At(ev, EpiStore<X>{dp}, st_, gate);
            apply_A(dp, EpiStore<X>{ne}, st_, gate);
            hipLaunchKernelGGL(pi_dot_kernel<X>, dim3(grid(m_)), dim3(kBlock), 0, st_, ne, ev, m_, red(), ds_);
            fold<1, 0>(grid(m_), 0, TailPiValue{}, gate);
            hipLaunchKernelGGL(pi_residual_kernel<X>, dim3(grid(m_)), dim3(kBlock), 0, st_, ne, ev, m_, P_.sv_tol,
                               red(), ds_);
            fold<1, 0>(grid(m_), 0, TailPiResidual{P_.sv_tol}, gate);
            hipLaunchKernelGGL(pi_restore_kernel<X>, dim3(grid(m_)), dim3(kBlock), 0, st_, ne, ev, m_, red(), ds_);
            fold<1, 0>(grid(m_), 0, TailPiRestore{}, gate);
            hipLaunchKernelGGL(scale_dev_kernel<X>, dim3(grid(m_)), dim3(kBlock), 0, st_, ne, &ds_->pi_inv, m_, gate);
        }
        HIP_OK(hipGetLastError());
        read_scalars();
        if (!hs_->pi_active) break;
        chunk = std::min(chunk * 2, 64);
    }
    read_scalars();
    const double s2 = hs_->pi_value;
    (void)hipFree(buf[0]);
    (void)hipFree(buf[1]);
    (void)hipFree(dp);
    return std::sqrt(s2);
}

// Largest eigenvalue (minimum = false) or lambda_max - largest eigenvalue of
// (lambda_max I - Q) (minimum = true) of the sparse Q
// (reference src/pdhg_core_op.cu:1498-1671).
template <class V, class X> double Solver<V, X>::power_eig(bool minimum, double shift) {
    std::vector<double> v0((size_t)n_);
    unsigned int seed = 1234;
    for (int i = 0; i < n_; ++i) v0[i] = (double)rand_r(&seed) / RAND_MAX;
    X *v = dupload<X>(v0, n_), *av = dzeros<X>(n_);
    std::memset(hs_, 0, sizeof(DevScalars));
    hs_->pi_active = 1;
    HIP_OK(hipMemcpyAsync(ds_, hs_, sizeof(DevScalars), hipMemcpyHostToDevice, st_));
    const int *gate = &ds_->pi_active;
    int it = 0, chunk = 8;
    while (it < P_.sv_max_iter) {
        const int todo = std::min(chunk, P_.sv_max_iter - it);
        for (int c = 0; c < todo; ++c, ++it) {
            hipLaunchKernelGGL(pi_norm_kernel<X>, dim3(grid(n_)), dim3(kBlock), 0, st_, v, n_, red(), ds_);
            fold<1, 0>(grid(n_), 0, TailPiNorm{}, gate);
            hipLaunchKernelGGL(scale_dev_kernel<X>, dim3(grid(n_)), dim3(kBlock), 0, st_, v, &ds_->pi_inv, n_, gate);
            spmv_launch<V, X, Acc>(Q_, v, EpiStore<X>{av}, st_, gate);
            hipLaunchKernelGGL(pi_rayleigh_kernel<X>, dim3(grid(n_)), dim3(kBlock), 0, st_, v, av, shift,
                               minimum ? 1 : 0, n_, P_.sv_tol, red(), ds_);
            fold<1, 0>(grid(n_), 0, TailPiRayleigh{P_.sv_tol}, gate);
            hipLaunchKernelGGL(copy_gated_kernel<X>, dim3(grid(n_)), dim3(kBlock), 0, st_, v, av, n_, gate);
        }
        HIP_OK(hipGetLastError());
        read_scalars();
        if (!hs_->pi_active) break;
        chunk = std::min(chunk * 2, 64);
    }
    read_scalars();
    const double val = hs_->pi_value;
    (void)hipFree(v);
    (void)hipFree(av);
    return minimum ? shift - val : val;
}

// Q x by objective type (reference src/pdhg_core_op.cu:305-413).
template <class V, class X> void Solver<V, X>::q_product(const X *x, X *qx, const int *gate) {
    const int nb = grid(n_);
    switch (qtype_) {
    case PDHCG_NON_Q:
        return;
    case PDHCG_DIAG_Q:
        hipLaunchKernelGGL(mul_kernel<X>, dim3(nb), dim3(kBlock), 0, st_, qdiag_, x, qx, n_, gate);
        return;
    case PDHCG_SPARSE_Q:
        if (q_sparse_) spmv_launch<V, X, Acc>(Q_, x, EpiStore<X>{qx}, st_, gate);
        else hipLaunchKernelGGL(mul_kernel<X>, dim3(nb), dim3(kBlock), 0, st_, qsdiag_, x, qx, n_, gate);
        return;
    case PDHCG_LOW_RANK_Q:
    case PDHCG_LOW_RANK_PLUS_SPARSE_Q: {
        const bool plus = qtype_ == PDHCG_LOW_RANK_PLUS_SPARSE_Q;
        if (plus && q_sparse_) spmv_launch<V, X, Acc>(Q_, x, EpiStore<X>{qx}, st_, gate);
        if (r_dense_) {
            const X *qd = (plus && !q_sparse_) ? qsdiag_ : nullptr;
            const int acc = (plus && q_sparse_) ? 1 : 0;
            if (gate) {
                hipLaunchKernelGGL((rx_partial_kernel<V, X, true>), dim3(rx_blocks_), dim3(kBlock), 0, st_, Rd_, x, n_,
                                   k_, rx_partials_, rx_chunks_, gate);
                hipLaunchKernelGGL(rx_fold_kernel<true>, dim3(k_), dim3(kBlock), 0, st_, rx_partials_, rx_chunks_, k_,
                                   rx_sums_, rx_counter_, u_, ddiag_, ddense_, dkind_, gate);
                hipLaunchKernelGGL((rtu_dense_kernel<V, X, true>), dim3(full_grid(((long long)n_ + RtuCols<V>::value - 1) / RtuCols<V>::value)), dim3(kBlock), 0, st_, Rd_, u_,
                                   n_, k_, qx, qd, x, acc, gate);
            } else {
                const int t0 = timer_.start(kTimedRx, st_);
                hipLaunchKernelGGL((rx_partial_kernel<V, X, false>), dim3(rx_blocks_), dim3(kBlock), 0, st_, Rd_, x,
                                   n_, k_, rx_partials_, rx_chunks_, gate);
                timer_.stop(kTimedRx, t0, st_);
                hipLaunchKernelGGL(rx_fold_kernel<false>, dim3(k_), dim3(kBlock), 0, st_, rx_partials_, rx_chunks_, k_,
                                   rx_sums_, rx_counter_, u_, ddiag_, ddense_, dkind_, gate);
                const int t1 = timer_.start(kTimedRtu, st_);
                hipLaunchKernelGGL((rtu_dense_kernel<V, X, false>), dim3(full_grid(((long long)n_ + RtuCols<V>::value - 1) / RtuCols<V>::value)), dim3(kBlock), 0, st_, Rd_,
                                   u_, n_, k_, qx, qd, x, acc, gate);
                timer_.stop(kTimedRtu, t1, st_);
            }
        } else {
            spmv_launch<V, X, Acc>(Rc_, x, EpiStore<X>{rxv_}, st_, gate);
            hipLaunchKernelGGL(apply_middle_kernel<X>, dim3(1), dim3(kBlock), 0, st_, rxv_, u_, k_, ddiag_, ddense_,
                               dkind_, gate);
            if (plus && !q_sparse_)
                hipLaunchKernelGGL(mul_kernel<X>, dim3(nb), dim3(kBlock), 0, st_, qsdiag_, x, qx, n_, gate);
            if (plus) spmv_launch<V, double, double>(Rtc_, u_, EpiAccumulate<X>{qx}, st_, gate);
            else spmv_launch<V, double, double>(Rtc_, u_, EpiStore<X>{qx}, st_, gate);
        }
        return;
    }
    default:
        fail("unknown quadratic objective type");
    }
}

// Projected BB inner solve of the primal prox (reference src/pdhg_core_op.cu:495-669).
template <class V, class X> void Solver<V, X>::bb_prox(double tau, bool fuse_halpern, double w) {
    const double inv = 1.0 / tau;
    const double alpha0 = 1.0 / inv;
    const int nb = grid(n_);
    if (precond_ && cached_inv_tau_ != inv) {
        hipLaunchKernelGGL(bb_refresh_kernel<X>, dim3(nb), dim3(kBlock), 0, st_, diag_h_, inv, mdiag_, minv_, n_,
                           red());
        fold<1, 0>(nb, 0, TailTolScale{n_});
        cached_inv_tau_ = inv;
    }
    const X *mi = precond_ ? minv_ : nullptr;
    const X *md = precond_ ? mdiag_ : nullptr;
    q_product(x_, qx_, nullptr);
    // With cones every box step is followed by the cone projection and the
    // step is re-measured from the projected point (reference :561-566,652-657).
    const int defer = has_cones_ ? 1 : 0;
    const int pre = precond_ ? 1 : 0;
    hipLaunchKernelGGL(bb_init_kernel<X>, dim3(nb), dim3(kBlock), 0, st_, aty_, grad_, dirn_, x_, xp_, c_, qx_, lb_,
                       ub_, mi, md, alpha0, n_, red(), ds_, defer);
    if (has_cones_) {
        project_cones(xp_, false, nullptr);
        hipLaunchKernelGGL(bb_cone_step_kernel<X>, dim3(nb), dim3(kBlock), 0, st_, xp_, x_, dirn_, md, n_, 1, red(),
                           ds_);
    }
    fold<1, 0>(nb, 0, TailBBStep{1, pre, inner_limit_, inner_tol_});
    const int *gate = &ds_->bb_active;
    const bool plus = qtype_ == PDHCG_LOW_RANK_PLUS_SPARSE_Q;
    const bool fused_body = r_dense_ && (qtype_ == PDHCG_LOW_RANK_Q || plus);
    const int max_bodies = inner_limit_ - 1;
    const int nb_grad = fused_body ? grid(((long long)n_ + RtuCols<V>::value - 1) / RtuCols<V>::value) : nb;  // rtu_grad: RtuCols<V> columns per thread
    auto enqueue_body = [&]() {
        if (fused_body) {
            // Q x_p with the gradient step folded into the R'u pass
            if (plus && q_sparse_) spmv_launch<V, X, Acc>(Q_, xp_, EpiStore<X>{qx_}, st_, gate);
            hipLaunchKernelGGL((rx_partial_kernel<V, X, true>), dim3(rx_blocks_), dim3(kBlock), 0, st_, Rd_, xp_,
                               n_, k_, rx_partials_, rx_chunks_, gate);
            hipLaunchKernelGGL(rx_fold_kernel<true>, dim3(k_), dim3(kBlock), 0, st_, rx_partials_, rx_chunks_, k_,
                               rx_sums_, rx_counter_, u_, ddiag_, ddense_, dkind_, gate);
            hipLaunchKernelGGL((rtu_grad_kernel<V, X>), dim3(nb_grad), dim3(kBlock), 0, st_, Rd_, u_, n_, k_,
                               (plus && q_sparse_) ? qx_ : nullptr, (plus && !q_sparse_) ? qsdiag_ : nullptr,
                               xp_, x_, c_, aty_, grad_, dirn_, inv, red(), ds_);
        } else {
            q_product(xp_, qx_, gate);
            hipLaunchKernelGGL(bb_gradient_kernel<X>, dim3(nb), dim3(kBlock), 0, st_, xp_, x_, c_, aty_, qx_,
                               grad_, dirn_, inv, n_, red(), ds_);
        }
        fold<1, 0>(nb_grad, 0, TailBBAlpha{pre}, gate);
        if (has_cones_)
            hipLaunchKernelGGL(copy_gated_kernel<X>, dim3(nb), dim3(kBlock), 0, st_, snap_, xp_, n_, gate);
        hipLaunchKernelGGL(bb_direction_kernel<X>, dim3(nb), dim3(kBlock), 0, st_, xp_, grad_, dirn_, lb_, ub_,
                           mi, md, n_, red(), ds_, defer);
        if (has_cones_) {
            project_cones(xp_, false, gate);
            hipLaunchKernelGGL(bb_cone_step_kernel<X>, dim3(nb), dim3(kBlock), 0, st_, xp_, snap_, dirn_, md, n_,
                               0, red(), ds_);
        }
        fold<1, 0>(nb, 0, TailBBStep{0, pre, inner_limit_, inner_tol_}, gate);
    };
    // The loop stops on the device (bb_active); the host learns of it through
    // polls of (bb_active, bb_iter) copied after a chunk of bodies.  One more
    // chunk is always queued behind the polled one, so the stream never
    // drains while the host waits: when the polled flag says the loop ended,
    // the queued chunk runs gated (empty launches).  The first chunk is the
    // body count of the previous solve.
    int issued = 0, inflight = 0, head = 0, chunk = std::max(1, bb_pred_), last_iter = 0;
    auto enqueue_chunk = [&](int count) {
        count = std::min(count, max_bodies - issued);
        for (int c = 0; c < count; ++c) enqueue_body();
        issued += count;
        HIP_OK(hipGetLastError());
        const int slot = (head + inflight) % 2;
        HIP_OK(hipMemcpyAsync(bb_poll_ + 2 * slot, &ds_->bb_active, 2 * sizeof(int), hipMemcpyDeviceToHost, st_));
        HIP_OK(hipEventRecord(bb_ev_[slot], st_));
        ++inflight;
    };
    while (issued < max_bodies || inflight > 0) {
        while (issued < max_bodies && inflight < 2) {
            enqueue_chunk(chunk);
            chunk = inflight == 1 ? 1 : std::min(chunk * 2, 16);
        }
        HIP_OK(hipEventSynchronize(bb_ev_[head]));
        const int active = bb_poll_[2 * head], iter = bb_poll_[2 * head + 1];
        last_iter = iter;
        head ^= 1;
        --inflight;
        if (!active) {
            bb_pred_ = std::max(1, iter - 1);
            break;
        }
        if (inflight == 0) chunk = std::max(chunk, 2);
    }
    if (eval_) sat_run_ = (inner_limit_ > 1 && last_iter >= inner_limit_) ? sat_run_ + 1 : 0;
    hipLaunchKernelGGL(bb_final_kernel<X>, dim3(nb), dim3(kBlock), 0, st_, x_, xp_, xr_, x0_, n_,
                       fuse_halpern ? 1 : 0, w, P_.reflection_coefficient, ds_);
}

// One PDHG step (reference src/pdhg_core_op.cu:705-830).
template <class V, class X> double Solver<V, X>::primal_tau() const {
    double tau = step_ / pw_;
    if (nonconvexity_ < 0) {
        tau = std::fmax(tau, -1.01 * std::fmin(0.0, nonconvexity_));
        tau /= 100;
    }
    return tau;
}

// Iteration t runs no evaluation, restart or fixed-point error, given that the
// previous one did not end a major cycle (run() below).
template <class V, class X> bool Solver<V, X>::plain_iteration(int t) const {
    const int ef = P_.termination_evaluation_frequency;
    return t > 0 && t % ef != 0 && t % print_frequency(t) != 0 && (t + 1) % ef != 0;
}

template <class V, class X> void Solver<V, X>::release_graph(GraphWindow &g) {
    if (g.done) (void)hipEventSynchronize(g.done);
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
    if (g.graph) (void)hipGraphDestroy(g.graph);
    if (g.done) (void)hipEventDestroy(g.done);
    g = GraphWindow();
}

// The instantiated graph of `len` plain iterations with the given store
// pattern: a prologue kernel node, then the iterations' launches exactly as
// pdhg_step enqueues them (fused Halpern), recorded by stream capture.
template <class V, class X>
typename Solver<V, X>::GraphWindow &Solver<V, X>::graph_for(int len, const std::vector<char> &store) {
    for (GraphWindow &g : graphs_)
        if (g.len == len && g.store == store) {
            g.used = ++graph_clock_;
            return g;
        }
    if ((int)graphs_.size() >= kMaxGraphs) {
        auto lru = std::min_element(graphs_.begin(), graphs_.end(),
                                    [](const GraphWindow &a, const GraphWindow &b) { return a.used < b.used; });
        release_graph(*lru);
        graphs_.erase(lru);
    }
    const double rho = P_.reflection_coefficient;
    hipGraph_t body = nullptr;
    HIP_OK(hipStreamBeginCapture(st_, hipStreamCaptureModeThreadLocal));
    try {
        for (int p = 0; p < len; ++p) {
            EpiPrimalStep<X> epi{x_, xp_, xr_, dslack_, c_, lb_, ub_, x0_, qtype_ == PDHCG_DIAG_Q ? qdiag_ : nullptr,
                                 0.0, 0.0, rho, (int)store[p], 1, dyn_, p};
            apply_At(y_, epi, st_);
            EpiDualStep<X> depi{y_, yp_, yr_, g_, cl_, cu_, y0_, 0.0, 0.0, rho, (int)store[p], 1, dyn_, p};
            apply_A(xr_, depi, st_);
        }
    } catch (...) {
        (void)hipStreamEndCapture(st_, &body);
        if (body) (void)hipGraphDestroy(body);
        throw;
    }
    HIP_OK(hipStreamEndCapture(st_, &body));
    GraphWindow g;
    g.len = len;
    g.store = store;
    HIP_OK(hipGraphCreate(&g.graph, 0));
    double *d = dyn_, zero = 0.0;
    void *args[] = {&d, &zero, &zero, &zero};
    hipKernelNodeParams kp{};
    kp.func = reinterpret_cast<void *>(graph_dyn_kernel);
    kp.gridDim = dim3(1);
    kp.blockDim = dim3(kWave);
    kp.sharedMemBytes = 0;
    kp.kernelParams = args;
    kp.extra = nullptr;
    HIP_OK(hipGraphAddKernelNode(&g.prologue, g.graph, nullptr, 0, &kp));
    hipGraphNode_t child = nullptr;
    HIP_OK(hipGraphAddChildGraphNode(&child, g.graph, &g.prologue, 1, body));
    HIP_OK(hipGraphDestroy(body));
    HIP_OK(hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0));
    HIP_OK(hipEventCreateWithFlags(&g.done, hipEventDisableTiming));
    g.used = ++graph_clock_;
    graphs_.push_back(g);
    return graphs_.back();
}

// Replays `len` plain iterations from total_count_ (state as after len
// passes of run()'s loop body).
template <class V, class X> void Solver<V, X>::run_window(int len) {
    const int tc = total_count_;
    std::vector<char> store((size_t)len);
    for (int p = 0; p < len; ++p) store[p] = ((tc + p + 2) % print_frequency(tc + p + 2)) == 0 ? 1 : 0;
    GraphWindow &g = graph_for(len, store);
    double *d = dyn_, tau = primal_tau(), sigma = step_ * pw_, k0 = (double)inner_count_;
    void *args[] = {&d, &tau, &sigma, &k0};
    hipKernelNodeParams kp{};
    kp.func = reinterpret_cast<void *>(graph_dyn_kernel);
    kp.gridDim = dim3(1);
    kp.blockDim = dim3(kWave);
    kp.sharedMemBytes = 0;
    kp.kernelParams = args;
    kp.extra = nullptr;
    HIP_OK(hipEventSynchronize(g.done));  // the previous replay of this graph has finished
    HIP_OK(hipGraphExecKernelNodeSetParams(g.exec, g.prologue, &kp));
    HIP_OK(hipGraphLaunch(g.exec, st_));
    HIP_OK(hipEventRecord(g.done, st_));
    host_inner_ += len;
    inner_count_ += len;
    total_count_ += len;
    graph_windows_ += 1;
    graph_iters_ += len;
}

template <class V, class X> void Solver<V, X>::pdhg_step(bool fuse_halpern) {
    double tau = primal_tau();
    const double sigma = step_ * pw_;
    const int tc = total_count_;
    const int store = (is_major_ || ((tc + 2) % print_frequency(tc + 2)) == 0) ? 1 : 0;
    const double w = (double)(inner_count_ + 1) / (inner_count_ + 2);
    const double rho = P_.reflection_coefficient;
    if (closed_prox()) {
        // cones force the stored form and replace the reflection on cone slots
        // after the projection (reference :428,730-735), so no Halpern fusion
        const bool cones = has_cones_;
        EpiPrimalStep<X> epi{x_, xp_, xr_, dslack_, c_, lb_, ub_, x0_, qtype_ == PDHCG_DIAG_Q ? qdiag_ : nullptr,
                             tau, w, rho, cones ? 1 : store, fuse_halpern ? 1 : 0};
        const int t = timer_.start(kTimedPrimalStep, st_);
        apply_At(y_, epi, st_);
        timer_.stop(kTimedPrimalStep, t, st_);
        if (cones) {
            if (qtype_ == PDHCG_DIAG_Q) cones_.project_diag_q(xp_, var_rescale_, qdiag_, tau, st_);
            else project_cones(xp_, false, nullptr);
            hipLaunchKernelGGL(cone_reflect_kernel<X>, dim3(grid(n_)), dim3(kBlock), 0, st_, xr_, xp_, x_, cones_.mask,
                               n_);
        }
    } else {
        const int t = timer_.start(kTimedPrimalStep, st_);
        apply_At(y_, EpiStore<X>{aty_}, st_);
        timer_.stop(kTimedPrimalStep, t, st_);
        bb_prox(tau, fuse_halpern, w);
    }
    host_inner_ += 1;
    if (has_affine_) {
        // the dual prox point is projected onto the affine cones between the
        // SpMV and the update (reference src/pdhg_core_op.cu:809-829)
        EpiDualPrepare<X> prep{ax_, aproj_, y_, g_, cl_, cu_, sigma};
        const int t = timer_.start(kTimedDualStep, st_);
        apply_A(xr_, prep, st_);
        timer_.stop(kTimedDualStep, t, st_);
        acones_.project(aproj_, aff_d_, false, nullptr, st_);
        hipLaunchKernelGGL(affine_dual_finish_kernel<X>, dim3(grid(m_)), dim3(kBlock), 0, st_, y_, yp_, yr_, ax_,
                           aproj_, sigma, store, m_);
        return;
    }
    EpiDualStep<X> depi{y_, yp_, yr_, g_, cl_, cu_, y0_, sigma, w, rho, store, fuse_halpern ? 1 : 0};
    const int t = timer_.start(kTimedDualStep, st_);
    apply_A(xr_, depi, st_);
    timer_.stop(kTimedDualStep, t, st_);
}

template <class V, class X> void Solver<V, X>::halpern(double w) {
    hipLaunchKernelGGL(halpern_kernel<X>, dim3(grid((long long)n_ + m_)), dim3(kBlock), 0, st_, x0_, x_, xr_, y0_, y_,
                       yr_, n_, m_, w, P_.reflection_coefficient);
}

// reference src/pdhg_core_op.cu:973-1035
template <class V, class X> void Solver<V, X>::fixed_point_error() {
    hipLaunchKernelGGL(delta_norms_kernel<X>, dim3(grid((long long)n_ + m_)), dim3(kBlock), 0, st_, xr_, x_, dx_, yr_,
                       y_, dy_, n_, m_, red());
    fold<2, 0>(grid((long long)n_ + m_));
    read_scalars();
    const double pn2 = hs_->red[0], dn2 = hs_->red[1];
    apply_At(dy_, EpiStore<X>{aty_}, st_);
    hipLaunchKernelGGL(dot_kernel<X>, dim3(grid(n_)), dim3(kBlock), 0, st_, aty_, dx_, n_, red(), nullptr);
    fold<1, 0>(grid(n_));
    read_scalars();
    const double cross = hs_->red[0];
    const double pn = std::sqrt(pn2), dn = std::sqrt(dn2);
    const double movement = pn * pn * pw_ + dn * dn / pw_;
    fpe_ = std::sqrt(movement + 2 * step_ * cross);
    if (!is_lp_ && bb_) inner_tol_ = std::fmin(inner_tol_, std::fmax(0.0005 * pn / step_ * pw_, inner_min_tol_));
}

// reference src/pdhg_core_op.cu:859-931
template <class V, class X> void Solver<V, X>::restart() {
    const restart_parameters_t &rp = P_.restart_params;
    hipLaunchKernelGGL(delta_norms_kernel<X>, dim3(grid((long long)n_ + m_)), dim3(kBlock), 0, st_, xp_, x0_, dx_, yp_,
                       y0_, dy_, n_, m_, red());
    fold<2, 0>(grid((long long)n_ + m_));
    read_scalars();
    const double pd = std::sqrt(hs_->red[0]), dd = std::sqrt(hs_->red[1]);
    const double ratio = rel_du_ / rel_pr_;
    if (pd > 1e-16 && dd > 1e-16 && pd < 1e12 && dd < 1e12 && ratio > 1e-8 && ratio < 1e8) {
        const double err = std::log(dd) - std::log(pd) - std::log(pw_);
        pw_err_sum_ *= rp.i_smooth;
        pw_err_sum_ += err;
        const double derr = err - pw_last_err_;
        pw_ *= std::exp(rp.k_p * err + rp.k_i * pw_err_sum_ + rp.k_d * derr);
        pw_last_err_ = err;
    } else {
        pw_ = best_pw_;
        pw_err_sum_ = 0.0;
        pw_last_err_ = 0.0;
    }
    const double gapv = std::fabs(std::log10(rel_du_ / rel_pr_));
    if (gapv < best_gap_) {
        best_gap_ = gapv;
        best_pw_ = pw_;
    }
    HIP_OK(hipMemcpyAsync(x0_, xp_, (size_t)n_ * sizeof(X), hipMemcpyDeviceToDevice, st_));
    HIP_OK(hipMemcpyAsync(x_, xp_, (size_t)n_ * sizeof(X), hipMemcpyDeviceToDevice, st_));
    HIP_OK(hipMemcpyAsync(y0_, yp_, (size_t)m_ * sizeof(X), hipMemcpyDeviceToDevice, st_));
    HIP_OK(hipMemcpyAsync(y_, yp_, (size_t)m_ * sizeof(X), hipMemcpyDeviceToDevice, st_));
    inner_count_ = 0;
    last_trial_fpe_ = INFINITY;
}

// reference src/utils.cu:291-322
template <class V, class X> bool Solver<V, X>::should_restart() {
    const restart_parameters_t &rp = P_.restart_params;
    const int ef = P_.termination_evaluation_frequency;
    bool r = false;
    if (total_count_ == ef) r = true;
    else if (total_count_ > ef) {
        if (fpe_ <= rp.sufficient_reduction_for_restart * init_fpe_) r = true;
        if (fpe_ <= rp.necessary_reduction_for_restart * init_fpe_ && fpe_ > last_trial_fpe_) r = true;
        if (inner_count_ >= rp.artificial_restart_threshold * total_count_) r = true;
    }
    last_trial_fpe_ = fpe_;
    return r;
}

// reference src/pdhg_core_op.cu:1037-1294
// FP64 evaluation of the FP32 iterates (mixed precision): x_bar, y_bar and,
// for an LP, the dual slack of the last primal step are widened exactly into
// the evaluation copy, whose FP64 products and reductions give the residuals.
template <class V, class X> void Solver<V, X>::eval_residual() {
    Solver<double, double> &E = *eval_;
    hipLaunchKernelGGL(widen2_kernel<X>, dim3(full_grid((long long)n_ + m_)), dim3(kBlock), 0, st_, xp_, E.xp_, n_,
                       yp_, E.yp_, m_);
    if (is_lp_)
        hipLaunchKernelGGL(widen2_kernel<X>, dim3(full_grid(n_)), dim3(kBlock), 0, st_, dslack_, E.dslack_, n_,
                           (const X *)nullptr, (double *)nullptr, 0);
    HIP_OK(hipGetLastError());
    E.step_ = step_;
    E.pw_ = pw_;
    E.residual();
    abs_pr_ = E.abs_pr_;
    rel_pr_ = E.rel_pr_;
    abs_du_ = E.abs_du_;
    rel_du_ = E.rel_du_;
    pobj_ = E.pobj_;
    dobj_ = E.dobj_;
    gap_ = E.gap_;
    rel_gap_ = E.rel_gap_;
}

template <class V, class X> void Solver<V, X>::eval_infeasibility() {
    Solver<double, double> &E = *eval_;
    hipLaunchKernelGGL(widen2_kernel<X>, dim3(full_grid((long long)n_ + m_)), dim3(kBlock), 0, st_, dx_, E.dx_, n_,
                       dy_, E.dy_, m_);
    HIP_OK(hipGetLastError());
    E.infeasibility();
    pray_lin_obj_ = E.pray_lin_obj_;
    dray_obj_ = E.dray_obj_;
    max_pray_ = E.max_pray_;
    max_dray_ = E.max_dray_;
}

template <class V, class X> bool Solver<V, X>::stalled(int tc) {
    if (stall_iters_ <= 0) return false;
    const double err = std::fmax(rel_pr_, std::fmax(rel_du_, rel_gap_));
    if (err < 0.5 * best_err_) {
        best_err_ = err;
        best_err_iter_ = tc;
        return false;
    }
    // a plateau between restarts is not a stall: the error must also have
    // stood still for half of the run so far
    return tc - best_err_iter_ >= std::max(stall_iters_, tc / 2);
}

// Hands the loop state to the FP64 evaluation copy: the iterates widened
// exactly, the inner-solve buffers, and every scalar of the outer loop, so
// that it continues the same iteration in FP64.
template <class V, class X> void Solver<V, X>::promote() {
    Solver<double, double> &E = *eval_;
    auto widen = [&](const X *a, double *ao, int na, const X *b, double *bo, int nb) {
        hipLaunchKernelGGL(widen2_kernel<X>, dim3(full_grid((long long)na + nb)), dim3(kBlock), 0, st_, a, ao, na, b,
                           bo, nb);
    };
    widen(x_, E.x_, n_, y_, E.y_, m_);
    widen(xp_, E.xp_, n_, yp_, E.yp_, m_);
    widen(xr_, E.xr_, n_, yr_, E.yr_, m_);
    widen(x0_, E.x0_, n_, y0_, E.y0_, m_);
    widen(dx_, E.dx_, n_, dy_, E.dy_, m_);
    widen(dslack_, E.dslack_, n_, (const X *)nullptr, (double *)nullptr, 0);
    HIP_OK(hipGetLastError());
    if (bb_) {
        E.bb_ = true;
        E.precond_ = precond_;
        E.grad_ = dzeros<double>(n_);
        E.dirn_ = dzeros<double>(n_);
        if (has_cones_) E.snap_ = dzeros<double>(n_);
        if (precond_) {
            E.diag_h_ = dalloc<double>(n_);
            HIP_OK(hipMemcpyAsync(E.diag_h_, diag_h_, (size_t)n_ * sizeof(double), hipMemcpyDeviceToDevice, st_));
            E.mdiag_ = dzeros<double>(n_);
            E.minv_ = dzeros<double>(n_);
        }
    }
    E.inner_tol_ = inner_tol_;
    E.inner_min_tol_ = inner_min_tol_;
    E.inner_limit_ = inner_limit_;
    E.cached_inv_tau_ = -1.0;
    E.bb_pred_ = bb_pred_;
    E.q_norm_ = q_norm_;
    E.nonconvexity_ = nonconvexity_;
    E.step_ = step_;
    E.pw_ = pw_;
    E.pw_err_sum_ = pw_err_sum_;
    E.pw_last_err_ = pw_last_err_;
    E.best_pw_ = best_pw_;
    E.best_gap_ = best_gap_;
    E.fpe_ = fpe_;
    E.init_fpe_ = init_fpe_;
    E.last_trial_fpe_ = last_trial_fpe_;
    E.total_count_ = total_count_;
    E.inner_count_ = inner_count_;
    E.is_major_ = is_major_;
    E.do_restart_ = do_restart_;
    E.max_pray_ = max_pray_;
    E.max_dray_ = max_dray_;
    E.pray_lin_obj_ = pray_lin_obj_;
    E.dray_obj_ = dray_obj_;
    E.reason_ = reason_;
    E.cum_time_ = cum_time_;
    E.t_start_ = t_start_;
    read_scalars();
    E.host_inner_ = host_inner_ + hs_->inner_total;
    E.graph_on_ = false;
    HIP_OK(hipStreamSynchronize(st_));
    promoted_ = true;
}

template <class V, class X> void Solver<V, X>::residual() {
    if (eval_) return eval_residual();
    const bool linf = P_.optimality_norm == NORM_TYPE_L_INF;
    const double tau = step_ / pw_;
    apply_A(xp_, EpiStore<X>{ax_}, st_);
    apply_At(yp_, EpiStore<X>{aty_}, st_);
    q_product(xp_, qx_, nullptr);
    const GridRed r0 = red();
    if (has_affine_) {
        // red[1] row terms, [20..23] rows (sum pres^2, sum amem^2, max pres, max amem), [24..25] complementarity
        hipLaunchKernelGGL(affine_rows_prepare_kernel<X>, dim3(grid(m_)), dim3(kBlock), 0, st_, ax_, g_, cl_, cu_,
                           yp_, cl_fin_, cu_fin_, apres_, amem_, m_, r0);
        fold<1, 0>(grid(m_), 1);
        acones_.project(apres_, aff_d_, true, nullptr, st_);
        const int waves = kBlock / kWave;
        hipLaunchKernelGGL(affine_complementarity_kernel<X>, dim3((acones_.total + waves - 1) / waves), dim3(kBlock), 0,
                           st_, ax_, g_, yp_, amem_, acones_.all_start, acones_.all_len, acones_.total, cbr_, acomp_);
        acones_.project(amem_, aff_d_, true, nullptr, st_);
        hipLaunchKernelGGL(affine_rows_finish_kernel<X>, dim3(grid(m_)), dim3(kBlock), 0, st_, ax_, g_, apres_, amem_,
                           con_rescale_, aff_d_, m_, r0);
        fold<2, 2>(grid(m_), 20);
        hipLaunchKernelGGL(sq_max_kernel, dim3(grid(acones_.total)), dim3(kBlock), 0, st_, acomp_, acones_.total, r0);
        fold<1, 1>(grid(acones_.total), 24);
    } else {
        hipLaunchKernelGGL(residual_rows_kernel<X>, dim3(grid(m_)), dim3(kBlock), 0, st_, ax_, g_, cl_, cu_, yp_,
                           con_rescale_, cl_fin_, cu_fin_, m_, r0);
        fold<2, 1>(grid(m_), 0);
    }
    hipLaunchKernelGGL(residual_cols_kernel<X>, dim3(grid(n_)), dim3(kBlock), 0, st_, c_, aty_, qx_, xp_, lb_, ub_,
                       dslack_, var_rescale_, tau, is_lp_ ? 0 : 1, n_, has_cones_ ? dres_ : nullptr, r0);
    fold<6, 3>(grid(n_), 3);
    if (has_cones_) {
        // red[12..13] augmented dual residual, [14..17] power violation, [18] cone slack correction
        const X *qx = is_lp_ ? nullptr : qx_;
        cones_.dual_residual(dres_, pgp_, xp_, c_, qx, aty_, var_rescale_, n_, st_);
        // projected-gradient mapping (reference src/pdhg_core_op.cu:67-90)
        double t = step_ / pw_;
        if (!(t > 0.0) || !std::isfinite(t)) t = 1.0;
        hipLaunchKernelGGL(projected_gradient_point_kernel<X>, dim3(grid(n_)), dim3(kBlock), 0, st_, pgp_, xp_, c_, qx,
                           aty_, lb_, ub_, t, n_);
        project_cones(pgp_, true, nullptr);
        hipLaunchKernelGGL(augment_dual_residual_kernel<X>, dim3(grid(n_)), dim3(kBlock), 0, st_, dres_, xp_, pgp_,
                           var_rescale_, t, n_, r0);
        fold<1, 1>(grid(n_), 12);
        if (has_power_) {
            hipLaunchKernelGGL(power_violation_kernel<X>, dim3(grid(cones_.powall.count)), dim3(kBlock), 0, st_, xp_,
                               var_rescale_, cones_.powall, cbr_, r0);
            fold<2, 2>(grid(cones_.powall.count), 14);
        }
        hipLaunchKernelGGL(cone_slack_correction_kernel<X>, dim3(grid(n_)), dim3(kBlock), 0, st_, c_, qx, aty_,
                           dslack_, xp_, cones_.mask, n_, r0);
        fold<1, 0>(grid(n_), 18);
    }
    HIP_OK(hipGetLastError());
    read_scalars();
    const double *R = hs_->red;
    // rows: [sum pres^2, row terms, max|pres|]; cols: [sum dres^2, c'x, x'Qx, slack'x, (Qx)^2, (A'y)^2, max dres, max Qx, max A'y]
    const double pres2 = has_affine_ ? R[20] : R[0], pres_max = has_affine_ ? R[22] : R[2];
    abs_pr_ = (linf ? pres_max : std::sqrt(pres2)) / cbr_;
    const double lin_pr = abs_pr_;
    double pv_rel = 0.0;
    if (has_power_) {  // reference src/pdhg_core_op.cu:214-270
        const double pv_abs = (linf ? R[16] : std::sqrt(R[14])) / cbr_;
        pv_rel = linf ? R[17] : std::sqrt(R[15]);
        abs_pr_ = linf ? std::fmax(abs_pr_, pv_abs) : std::hypot(abs_pr_, pv_abs);
    }
    const double dres2 = has_cones_ ? R[12] : R[3 + 0], dres_max = has_cones_ ? R[13] : R[3 + 6];
    if (has_affine_)  // reference src/pdhg_core_op.cu:1166-1189
        abs_du_ = linf ? std::fmax(std::fmax(dres_max, R[23]), R[25]) : std::sqrt(dres2 + R[21] + R[24]);
    else abs_du_ = linf ? dres_max : std::sqrt(dres2);
    abs_du_ /= ovr_;
    const double half_xqx = qtype_ == PDHCG_NON_Q ? 0.0 : 0.5 * R[3 + 2];
    const double slack_x = R[3 + 3] + (has_cones_ ? R[18] : 0.0);
    pobj_ = (R[3 + 1] + half_xqx) / (cbr_ * ovr_) + obj_const_;
    dobj_ = (slack_x + R[1] - half_xqx) / (cbr_ * ovr_) + obj_const_;
    rel_pr_ = lin_pr / (1.0 + bound_norm_);
    if (has_power_) rel_pr_ = linf ? std::fmax(rel_pr_, pv_rel) : std::hypot(rel_pr_, pv_rel);
    double den;
    if (is_lp_) den = 1.0 + obj_norm_;
    else {
        const double qn = linf ? R[3 + 7] : std::sqrt(R[3 + 4]);
        const double an = linf ? R[3 + 8] : std::sqrt(R[3 + 5]);
        den = 1.0 + std::fmax(obj_norm_, std::fmax(qn / ovr_, an / ovr_));
    }
    rel_du_ = abs_du_ / den;
    gap_ = std::fabs(pobj_ - dobj_);
    rel_gap_ = gap_ / (1.0 + std::fabs(pobj_) + std::fabs(dobj_));
}

// reference src/pdhg_core_op.cu:1296-1429
template <class V, class X> void Solver<V, X>::infeasibility() {
    if (eval_) return eval_infeasibility();
    hipLaunchKernelGGL(ray_project_kernel<X>, dim3(grid((long long)n_ + m_)), dim3(kBlock), 0, st_, dx_, lb_, ub_, dy_,
                       cl_, cu_, n_, m_, red());
    fold<0, 2>(grid((long long)n_ + m_));
    read_scalars();
    const double prn = hs_->red[0], drn = hs_->red[1];
    if (prn > 0.0) hipLaunchKernelGGL(scale_kernel<X>, dim3(grid(n_)), dim3(kBlock), 0, st_, dx_, 1.0 / prn, n_);
    apply_A(dx_, EpiStore<X>{ax_}, st_);
    apply_At(dy_, EpiStore<X>{aty_}, st_);
    const GridRed r0 = red();
    hipLaunchKernelGGL(ray_rows_kernel<X>, dim3(grid(m_)), dim3(kBlock), 0, st_, dy_, ax_, g_, cl_, cu_, cl_fin_,
                       cu_fin_, con_rescale_, m_, r0);
    fold<1, 1>(grid(m_), 0);
    hipLaunchKernelGGL(ray_cols_kernel<X>, dim3(grid(n_)), dim3(kBlock), 0, st_, c_, dx_, aty_, lb_, ub_, lb_fin_,
                       ub_fin_, var_rescale_, n_, r0);
    fold<2, 1>(grid(n_), 2);
    double qn = 0.0;
    const bool has_q = !is_lp_ && qtype_ != PDHCG_NON_Q;
    if (has_q) {
        q_product(dx_, qx_, nullptr);
        hipLaunchKernelGGL(inf_norm_kernel<X>, dim3(grid(n_)), dim3(kBlock), 0, st_, qx_, n_, r0);
        fold<0, 1>(grid(n_), 5);
    }
    HIP_OK(hipGetLastError());
    read_scalars();
    const double *R = hs_->red;
    pray_lin_obj_ = R[2] / (cbr_ * ovr_);
    dray_obj_ = (R[0] + R[3]) / (cbr_ * ovr_);
    max_pray_ = R[1];
    if (has_q) {
        qn = R[5];
        max_pray_ = std::fmax(max_pray_, qn / ovr_);
    }
    const double dsn = R[4];
    max_dray_ = dsn;
    const double sc = std::fmax(drn, dsn);
    if (sc > 0.0) {
        max_dray_ /= sc;
        dray_obj_ /= sc;
    } else {
        max_dray_ = 0.0;
        dray_obj_ = 0.0;
    }
}

// reference src/utils.cu:219-289
template <class V, class X> void Solver<V, X>::check_termination() {
    const termination_criteria_t &c = P_.termination_criteria;
    if (g_pdhcg_cancel) {
        reason_ = TERMINATION_REASON_USER_INTERRUPT;
        return;
    }
    if (rel_du_ < c.eps_feasible_relative && rel_pr_ < c.eps_feasible_relative && rel_gap_ < c.eps_optimal_relative) {
        reason_ = TERMINATION_REASON_OPTIMAL;
        return;
    }
    if (has_cones_) {  // ray certificates are box-only (reference src/utils.cu:263-278)
        if (total_count_ >= c.iteration_limit) reason_ = TERMINATION_REASON_ITERATION_LIMIT;
        else if (cum_time_ >= c.time_sec_limit) reason_ = TERMINATION_REASON_TIME_LIMIT;
        return;
    }
    if (dray_obj_ > 0.0 && max_dray_ / dray_obj_ <= c.eps_infeasible) {
        reason_ = TERMINATION_REASON_PRIMAL_INFEASIBLE;
        return;
    }
    if (pray_lin_obj_ < 0.0 && max_pray_ / (-pray_lin_obj_) <= c.eps_infeasible) {
        reason_ = TERMINATION_REASON_DUAL_INFEASIBLE;
        return;
    }
    if (total_count_ >= c.iteration_limit) {
        reason_ = TERMINATION_REASON_ITERATION_LIMIT;
        return;
    }
    if (cum_time_ >= c.time_sec_limit) reason_ = TERMINATION_REASON_TIME_LIMIT;
}

// Main loop body, reference src/solver.cu:121-169.
template <class V, class X> int Solver<V, X>::run(int iterations) {
    if (promoted_) return eval_->run(iterations);
    const int ef = P_.termination_evaluation_frequency;
    for (int done = 0; done < iterations; ++done) {
        if (reason_ != TERMINATION_REASON_UNSPECIFIED) break;
        if (!(total_count_ < P_.termination_criteria.iteration_limit)) {
            reason_ = TERMINATION_REASON_ITERATION_LIMIT;
            residual();
            break;
        }
        if (graph_on_ && !timer_.enabled && !is_major_ && !do_restart_ && plain_iteration(total_count_)) {
            const int lim = std::min(iterations - done, P_.termination_criteria.iteration_limit - total_count_);
            int len = 1;
            while (len < lim && len < kMaxGraphWindow && plain_iteration(total_count_ + len)) ++len;
            if (len >= kMinGraphWindow) {
                run_window(len);
                done += len - 1;
                continue;
            }
        }
        const int tc = total_count_;
        if (eval_ && stall_iters_ > 0 && sat_run_ >= kSatPromote) {
            promote();
            if (P_.verbose >= 1)
                std::fprintf(stderr, "[mixed] %d FP32 inner solves in a row reached the inner limit; FP64 iterates "
                             "from iteration %d\n", sat_run_, tc);
            return eval_->run(iterations - done);
        }
        if (is_major_ || tc == 0 || tc % print_frequency(tc) == 0) {
            residual();
            if (!has_cones_ && !has_affine_ && is_major_ && tc < 3 * ef) infeasibility();
            cum_time_ = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start_).count();
            check_termination();
            if (P_.verbose >= 2 && tc % print_frequency(tc) == 0)
                std::printf("%6d %7d %.1e | %8.1e  %8.1e | %.1e %.1e %.1e | %.1e %.1e %.1e \n", tc,
                            host_inner_ + hs_->inner_total, cum_time_, pobj_, dobj_, abs_pr_, abs_du_, gap_, rel_pr_,
                            rel_du_, rel_gap_);
            if (reason_ != TERMINATION_REASON_UNSPECIFIED) break;
            if (eval_ && stalled(tc)) {
                promote();
                if (P_.verbose >= 1)
                    std::fprintf(stderr, "[mixed] FP32 iterates stalled at a KKT error of %.1e; FP64 iterates from "
                                 "iteration %d\n", best_err_, tc);
                return eval_->run(iterations - done);
            }
        }
        if (is_major_ || tc == 0) {
            do_restart_ = should_restart();
            if (do_restart_) restart();
        }
        is_major_ = ((tc + 1) % ef) == 0;
        const bool need_fpe = is_major_ || do_restart_;
        const double w = (double)(inner_count_ + 1) / (inner_count_ + 2);
        // the closed-form prox with cones rewrites x_bar after the step, so
        // its Halpern update cannot ride in the SpMV epilogue
        const bool fuse = !need_fpe && !(has_cones_ && closed_prox()) && !has_affine_;
        pdhg_step(fuse);
        if (need_fpe) {
            fixed_point_error();
            if (do_restart_) {
                init_fpe_ = fpe_;
                do_restart_ = false;
            }
        }
        if (!fuse) halpern(w);
        inner_count_++;
        total_count_++;
    }
    HIP_OK(hipGetLastError());
    return (int)reason_;
}

// reference src/pdhg_core_op.cu:1431-1496
template <class V, class X> pdhcg_result_t *Solver<V, X>::result(const qp_problem_t *orig) {
    if (promoted_) return eval_->result(orig);
    if (eval_) {
        // the returned point, reduced costs and residuals come from the FP64
        // evaluation copy; counters, times and status are this solver's
        Solver<double, double> &E = *eval_;
        hipLaunchKernelGGL(widen2_kernel<X>, dim3(full_grid((long long)n_ + m_)), dim3(kBlock), 0, st_, xp_, E.xp_,
                           n_, yp_, E.yp_, m_);
        HIP_OK(hipGetLastError());
        pdhcg_result_t *r = E.result(orig);
        read_scalars();
        r->total_count = total_count_;
        r->total_inner_count = host_inner_ + hs_->inner_total;
        r->rescaling_time_sec = rescale_time_;
        r->cumulative_time_sec = cum_time_;
        r->termination_reason = reason_;
        return r;
    }
    apply_At(yp_, EpiStore<X>{aty_}, st_);
    q_product(xp_, qx_, nullptr);
    double *xo = dalloc<double>(n_), *yo = dalloc<double>(m_), *rco = dalloc<double>(n_);
    hipLaunchKernelGGL(result_kernel<X>, dim3(grid((long long)n_ + m_)), dim3(kBlock), 0, st_, xp_, yp_, c_,
                       is_lp_ ? (const X *)nullptr : (const X *)qx_, aty_, lb_, ub_, var_rescale_, con_rescale_, ovr_,
                       cbr_, n_, m_, xo, yo, rco);
    HIP_OK(hipGetLastError());
    pdhcg_result_t *r = ccalloc<pdhcg_result_t>(1);
    r->primal_solution = cmalloc<double>(n_);
    r->dual_solution = cmalloc<double>(m_);
    r->reduced_cost = cmalloc<double>(n_);
    HIP_OK(hipMemcpyAsync(r->primal_solution, xo, (size_t)n_ * sizeof(double), hipMemcpyDeviceToHost, st_));
    HIP_OK(hipMemcpyAsync(r->dual_solution, yo, (size_t)m_ * sizeof(double), hipMemcpyDeviceToHost, st_));
    HIP_OK(hipMemcpyAsync(r->reduced_cost, rco, (size_t)n_ * sizeof(double), hipMemcpyDeviceToHost, st_));
    read_scalars();
    (void)hipFree(xo);
    (void)hipFree(yo);
    (void)hipFree(rco);
    r->num_variables = orig->num_variables;
    r->num_constraints = orig->num_constraints;
    r->num_nonzeros = orig->constraint_matrix_num_nonzeros;
    r->total_count = total_count_;
    r->total_inner_count = host_inner_ + hs_->inner_total;
    r->rescaling_time_sec = rescale_time_;
    r->cumulative_time_sec = cum_time_;
    r->relative_primal_residual = rel_pr_;
    r->relative_dual_residual = rel_du_;
    r->absolute_primal_residual = abs_pr_;
    r->absolute_dual_residual = abs_du_;
    r->primal_objective_value = pobj_;
    r->dual_objective_value = dobj_;
    r->objective_gap = gap_;
    r->relative_objective_gap = rel_gap_;
    r->max_primal_ray_infeasibility = max_pray_;
    r->max_dual_ray_infeasibility = max_dray_;
    r->primal_ray_linear_objective = pray_lin_obj_;
    r->dual_ray_objective = dray_obj_;
    r->termination_reason = reason_;
    return r;
}

template class Solver<double, double>;
template class Solver<float, float>;

}  // namespace pdhcg

using namespace pdhcg;

namespace {
int g_default_precision = -1;  // -1: read PDHCG_PRECISION once

int default_precision() {
    if (g_default_precision < 0) {
        const char *e = std::getenv("PDHCG_PRECISION");
        g_default_precision = (e && (std::strcmp(e, "mixed") == 0 || std::strcmp(e, "fp32") == 0)) ? PDHCG_PRECISION_MIXED
                                                                                                  : PDHCG_PRECISION_FP64;
    }
    return g_default_precision;
}

const char *reason_name(int r) {
    static const char *names[] = {"UNSPECIFIED",  "OPTIMAL",         "PRIMAL_INFEASIBLE", "DUAL_INFEASIBLE",
                                  "INFEASIBLE_OR_UNBOUNDED", "TIME_LIMIT", "ITERATION_LIMIT", "USER_INTERRUPT",
                                  "FEAS_POLISH_SUCCESS"};
    return (r >= 0 && r < 9) ? names[r] : "UNKNOWN";
}

void final_log(const pdhcg_result_t *r, const pdhg_parameters_t &p) {  // reference src/utils.cu:487-510
    if (p.verbose < 1) return;
    std::printf("Solution Summary\n");
    std::printf("  Status             : %s\n", reason_name(r->termination_reason));
    std::printf("  Solve time         : %.3g sec\n", r->cumulative_time_sec);
    std::printf("  Iterations         : %d\n", r->total_count);
    if (r->total_inner_count > 0) std::printf("  Inner Iterations   : %d\n", r->total_inner_count);
    std::printf("  Primal objective   : %.10g\n", r->primal_objective_value);
    std::printf("  Dual objective     : %.10g\n", r->dual_objective_value);
    std::printf("  Objective gap      : %.3e\n", r->relative_objective_gap);
    std::printf("  Primal infeas      : %.3e\n", r->relative_primal_residual);
    std::printf("  Dual infeas        : %.3e\n", r->relative_dual_residual);
}
}  // namespace

struct pdhcg_solver {
    std::unique_ptr<SolverBase> impl;
    const qp_problem_t *problem;   // what the device solves (the lifted copy for a QCQP)
    const qp_problem_t *original;  // the caller's problem
    qp_problem_t *lifted = nullptr;
    pdhg_parameters_t params;
    ~pdhcg_solver() {
        impl.reset();
        if (lifted) qp_problem_free(lifted);
    }
};

extern "C" {

void pdhcg_set_precision(int precision) {
    g_default_precision = precision == PDHCG_PRECISION_MIXED ? PDHCG_PRECISION_MIXED : PDHCG_PRECISION_FP64;
}

int pdhcg_get_precision(void) { return default_precision(); }

void pdhcg_request_cancel(int flag) { g_pdhcg_cancel = flag; }

pdhcg_solver_t *pdhcg_solver_create(const qp_problem_t *prob, const pdhg_parameters_t *params, int precision) {
    if (!prob) {
        std::fprintf(stderr, "[interface] pdhcg_solver_create: invalid arguments.\n");
        return nullptr;
    }
    pdhcg_solver *s = nullptr;
    try {
        s = new pdhcg_solver();
        s->problem = s->original = prob;
        if (params) s->params = *params;
        else fill_default_parameters(&s->params);
        // quadratic constraints are lifted into cone blocks first
        // (reference src/solver.cu:44-66)
        if (prob->num_quadratic_constraints > 0) {
            s->lifted = lift_quadratic_constraints(prob, s->params.default_cone_type);
            if (!s->lifted) {
                std::fprintf(stderr, "Error: QCQP -> SOCP transformation failed; cannot solve.\n");
                delete s;
                return nullptr;
            }
            if (s->params.verbose >= 1)
                std::fprintf(stderr, "[QCQP] %d quadratic constraint(s) reformulated as %d %s SOC block(s); extended "
                             "problem: %d vars, %d rows, %d nnz.\n", prob->num_quadratic_constraints,
                             s->lifted->cones.num_cones,
                             s->params.default_cone_type == CONE_STANDARD_SOC ? "standard" : "rotated",
                             s->lifted->num_variables, s->lifted->num_constraints,
                             s->lifted->constraint_matrix_num_nonzeros);
            s->problem = s->lifted;
        }
        // the entry-wise scaling passes and the transpose of A run on the
        // device for problems of at least kDeviceScaleMinEntries stored
        // entries (PDHCG_DEVICE_SCALING=0 keeps them on the host, =1 forces
        // the device); both give the same scaled problem bit for bit
        const qp_problem_t &q = *s->problem;
        const long long entries = (long long)q.constraint_matrix_num_nonzeros +
                                  q.objective_sparse_matrix_num_nonzeros + q.objective_lowrank_matrix_num_nonzeros;
        const char *ds = std::getenv("PDHCG_DEVICE_SCALING");
        const bool on_device = ds ? std::atoi(ds) != 0 : entries >= kDeviceScaleMinEntries;
        std::unique_ptr<MatrixPasses> passes = on_device ? make_device_passes() : nullptr;
        ScaledProblem sp = scale_problem(s->params, q, passes.get());
        passes.reset();
        if (precision == PDHCG_PRECISION_MIXED) s->impl.reset(new Solver<float, float>(sp, s->params));
        else s->impl.reset(new Solver<double, double>(sp, s->params));
        return s;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "[interface] solver setup failed: %s\n", e.what());
        delete s;
        return nullptr;
    }
}

int pdhcg_solver_run(pdhcg_solver_t *s, int iterations) {
    if (!s) return -1;
    try {
        return s->impl->run(iterations);
    } catch (const std::exception &e) {
        std::fprintf(stderr, "[interface] solver step failed: %s\n", e.what());
        return -1;
    }
}

int pdhcg_solver_counts(pdhcg_solver_t *s, int *total, int *inner) {
    if (!s) return -1;
    try {
        s->impl->counts(total, inner);
        return 0;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "[interface] %s\n", e.what());
        return -1;
    }
}

int pdhcg_solver_synchronize(pdhcg_solver_t *s) {
    if (!s) return -1;
    try {
        s->impl->synchronize();
        return 0;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "[interface] %s\n", e.what());
        return -1;
    }
}

pdhcg_result_t *pdhcg_solver_result(pdhcg_solver_t *s) {
    if (!s) return nullptr;
    try {
        pdhcg_result_t *r = s->impl->result(s->problem);
        if (s->lifted) restore_result_dimensions(r, s->original);  // reference src/solver.cu:199
        return r;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "[interface] result extraction failed: %s\n", e.what());
        return nullptr;
    }
}

int pdhcg_solver_timing(pdhcg_solver_t *s, int enable) {
    if (!s) return -1;
    try {
        s->impl->synchronize();
        if (enable < 0) throw std::runtime_error("timing period must be >= 0");
        s->impl->timer().reset();
        s->impl->timer().enabled = enable != 0;
        s->impl->timer().period = enable > 0 ? enable : 1;
        return 0;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "[interface] %s\n", e.what());
        return -1;
    }
}

void pdhcg_set_graph_window(int enable) { g_graph_window = enable ? 1 : 0; }

int pdhcg_get_graph_window(void) { return graph_window_default() ? 1 : 0; }

int pdhcg_solver_graph_stats(pdhcg_solver_t *s, int *windows, int *iterations) {
    if (!s || !windows || !iterations) return -1;
    s->impl->graph_stats(windows, iterations);
    return 0;
}

int pdhcg_solver_kernel_time(pdhcg_solver_t *s, int kernel, double *total_ms, int *launches) {
    if (!s || kernel < 0 || kernel >= kNumTimed || !total_ms || !launches) return -1;
    try {
        s->impl->synchronize();
        s->impl->timer().totals(kernel, total_ms, launches);
        return 0;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "[interface] %s\n", e.what());
        return -1;
    }
}

void pdhcg_solver_free(pdhcg_solver_t *s) {
    if (!s) return;
    try {
        delete s;
    } catch (...) {
    }
}

// Standalone cone operations on host vectors (one device round trip each).
namespace {
// Device buffers and the stream of one standalone cone call, released on
// every exit (hipFree waits for the device, so kernels still queued after
// an error finish first).
struct CallBuffers {
    hipStream_t st = nullptr;
    std::vector<void *> held;
    double *keep(double *p) {
        held.push_back(p);
        return p;
    }
    ~CallBuffers() {
        for (void *p : held)
            if (p) (void)hipFree(p);
        if (st) (void)hipStreamDestroy(st);
    }
};

// count > 1 (projections only): `count` vectors of n slots, one after
// another in in/out, projected in order through one cone set.
int cone_entry(const char *name, const cone_blocks_t *cones, const double *d, const double *in, double *out, int n,
               bool dual, int count = 1) {
    if (!cones || !in || !out || n < 0 || count < 1 || (dual && count != 1) || cones->num_cones < 0 ||
        (cones->num_cones > 0 && !cones->start_idx)) {
        std::fprintf(stderr, "[interface] %s: invalid arguments.\n", name);
        return -1;
    }
    try {
        CallBuffers cb;
        HIP_OK(hipStreamCreateWithFlags(&cb.st, hipStreamNonBlocking));
        hipStream_t st = cb.st;
        DevConeSet set;
        set.build(cones, n);
        if (dual && (set.sec.count || set.lsec.count))
            throw std::runtime_error("blocks with fixed slots have no dual-cone residual without a primal point");
        std::vector<double> dh((size_t)n, 1.0);
        if (d) std::copy(d, d + n, dh.begin());
        double *dd = cb.keep(dupload<double>(dh, n));
        double *v = cb.keep(dalloc<double>((size_t)n));
        double *zero = cb.keep(dzeros<double>(n));
        if (dual) {
            HIP_OK(hipMemcpy(v, in, (size_t)n * sizeof(double), hipMemcpyHostToDevice));
            double *res = cb.keep(dupload<double>(std::vector<double>(out, out + n), n));
            double *psd_buf = cb.keep(dzeros<double>(n));
            set.dual_residual<double>(res, psd_buf, zero, v, nullptr, zero, dd, n, st);
            HIP_OK(hipGetLastError());
            HIP_OK(hipMemcpyAsync(out, res, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, st));
        } else {
            for (int i = 0; i < count; ++i) {
                HIP_OK(hipMemcpyAsync(v, in + (size_t)i * n, (size_t)n * sizeof(double), hipMemcpyHostToDevice, st));
                set.project<double>(v, dd, false, nullptr, st);
                HIP_OK(hipGetLastError());
                HIP_OK(hipMemcpyAsync(out + (size_t)i * n, v, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, st));
            }
        }
        HIP_OK(hipStreamSynchronize(st));
        return 0;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "[interface] %s: %s\n", name, e.what());
        return -1;
    }
}
}  // namespace

int pdhcg_project_cones(const cone_blocks_t *cones, const double *d, double *x, int n) {
    return cone_entry("pdhcg_project_cones", cones, d, x, x, n, false);
}

int pdhcg_project_cones_seq(const cone_blocks_t *cones, const double *d, double *x, int n, int count) {
    return cone_entry("pdhcg_project_cones_seq", cones, d, x, x, n, false, count);
}

// ---- allocation-free variant of pdhcg_project_cones ----
//
// Every device buffer of the call (the point, the scaling, the per-type block
// lists, their warm starts, the PSD work and long-section scratch) is carved
// from `scratch`, a fixed array of kNoAllocScratchBytes in each device's
// global memory for the life of the process.  The host reaches it only
// through the address hipGetSymbolAddress gives on that device (kept per
// device in g_scratch_ptr); the array's name is never passed to a kernel.
namespace {
constexpr size_t kNoAllocScratchBytes = (size_t)32 << 20;
constexpr int kNoAllocMaxDevices = 64;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];
std::mutex g_scratch_mu;                        // one call at a time per process
void *g_scratch_ptr[kNoAllocMaxDevices] = {};  // device address of `scratch`, per device

__global__ void fill_kernel(double *v, double value, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) v[i] = value;
}
}  // namespace

int pdhcg_project_cones_noalloc(const cone_blocks_t *cones, const double *d, double *x, int n) {
    const char *name = "pdhcg_project_cones_noalloc";
    if (!cones || !x || n < 0 || cones->num_cones < 0 || (cones->num_cones > 0 && !cones->start_idx)) {
        std::fprintf(stderr, "[interface] %s: invalid arguments.\n", name);
        return -1;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    try {
        int dev = 0;
        HIP_OK(hipGetDevice(&dev));
        if (dev < 0 || dev >= kNoAllocMaxDevices) fail("device %d is beyond the %d the scratch table covers", dev,
                                                       kNoAllocMaxDevices);
        if (!g_scratch_ptr[dev]) {
            void *ptr = nullptr;
            const hipError_t rc = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
            if (rc != hipSuccess || !ptr)
                fail("hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(rc));
            g_scratch_ptr[dev] = ptr;
        }
        // size: the point, the scaling and the cone set, measured before anything is written
        DevArena need;
        need.cap = ~(size_t)0;
        need.take<double>((size_t)n);
        need.take<double>((size_t)n);
        {
            DevConeSet probe;
            probe.build(cones, n, &need);
        }
        if (need.used > kNoAllocScratchBytes)
            fail("needs %zu bytes of device scratch; the fixed scratch holds %zu (n = %d, %d cone blocks)", need.used,
                 kNoAllocScratchBytes, n, cones->num_cones);
        DevArena arena;
        arena.base = static_cast<char *>(g_scratch_ptr[dev]);
        arena.cap = kNoAllocScratchBytes;
        double *v = arena.take<double>((size_t)n), *dd = arena.take<double>((size_t)n);
        if (n > 0) {
            HIP_OK(hipMemcpy(v, x, (size_t)n * sizeof(double), hipMemcpyHostToDevice));
            if (d) HIP_OK(hipMemcpy(dd, d, (size_t)n * sizeof(double), hipMemcpyHostToDevice));
            else hipLaunchKernelGGL(fill_kernel, dim3(blocks_for(n)), dim3(kBlock), 0, 0, dd, 1.0, n);
        }
        DevConeSet set;
        set.build(cones, n, &arena);
        set.project<double>(v, dd, false, nullptr, 0);
        HIP_OK(hipGetLastError());
        if (n > 0) HIP_OK(hipMemcpy(x, v, (size_t)n * sizeof(double), hipMemcpyDeviceToHost));
        HIP_OK(hipDeviceSynchronize());
        return 0;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "[interface] %s: %s\n", name, e.what());
        return -1;
    }
}

int pdhcg_cone_dual_residual(const cone_blocks_t *cones, const double *d, const double *rc, double *out, int n) {
    return cone_entry("pdhcg_cone_dual_residual", cones, d, rc, out, n, true);
}

pdhcg_result_t *solve_qp_problem(const qp_problem_t *prob, const pdhg_parameters_t *params) {
    if (!prob) {
        std::fprintf(stderr, "[interface] solve_qp_problem: invalid arguments.\n");
        return nullptr;
    }
    if (validate_fixed_sections(prob) != 0) return nullptr;
    pdhg_parameters_t local;
    if (params) local = *params;
    else fill_default_parameters(&local);
    pdhcg_solver_t *s = pdhcg_solver_create(prob, &local, default_precision());
    if (!s) {
        std::fprintf(stderr, "[interface] optimize returned NULL.\n");
        return nullptr;
    }
    int rc = 0;
    while (rc == 0) rc = pdhcg_solver_run(s, 1 << 30);
    pdhcg_result_t *r = rc < 0 ? nullptr : pdhcg_solver_result(s);
    pdhcg_solver_free(s);
    if (!r) {
        std::fprintf(stderr, "[interface] optimize returned NULL.\n");
        return nullptr;
    }
    final_log(r, local);
    return r;
}

}  // extern "C"
