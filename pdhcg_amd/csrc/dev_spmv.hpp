// Row-length-binned CSR sparse matrix-vector product for gfx950.
//
// Rows are sorted (stably) into classes by length.  A class assigns L lanes
// of a 64-wide wavefront to each row (L = 1, 2, ..., 64, shifted up or down
// one power of two by the setup autotuner), so short rows are
// packed many to a wave and long rows get a whole wave; rows longer than
// kLongRow get a full 256-thread workgroup with an LDS reduction.  One launch
// covers every class: the workgroup index selects the class through a small
// offset table.  The matrix is stored in that slot order (row pointers
// contiguous per class, plus the original row id of each slot), so a row's
// extent is read coalesced and independently of its output index, instead
// of through a dependent row-id lookup.  Each row's sum is produced in a
// fixed order (lane-strided partial sums, then an xor-butterfly), so the
// result is deterministic.  The per-row result is handed to an epilogue
// functor, which lets the PDHG proximal updates run inside the SpMV without
// writing the product to HBM.
#pragma once

#include <algorithm>
#include <type_traits>
#include <vector>

#include "dev_common.hpp"

namespace pdhcg {

constexpr int kSpmvClasses = 8;      // L = 1,2,4,8,16,32,64, and workgroup rows
constexpr int kLongRow = 2048;       // rows longer than this get a workgroup

// Lanes per row for a row of `len` entries: about four entries per lane.
inline int spmv_class_of(int len) {
    if (len > kLongRow) return 7;
    int lanes = 1, c = 0;
    while (lanes < 64 && lanes * 4 < len) {
        lanes <<= 1;
        ++c;
    }
    return c;
}

struct SpmvClassTable {
    int class_start[kSpmvClasses + 1];  // first slot of each class
    int block_start[kSpmvClasses + 1];  // first workgroup of each class
    int shift;                          // class c < 7 runs 2^clamp(c + shift, 0, 6) lanes per row
};

// Lanes (as a power of two) of class c < 7 under a lane shift.
__host__ __device__ inline int spmv_lane_log(int c, int shift) {
    const int l = c + shift;
    return l < 0 ? 0 : (l > 6 ? 6 : l);
}

// Workgroup offsets of the classes for a lane shift (counts = rows per class).
inline void spmv_set_blocks(SpmvClassTable &tab, int shift) {
    tab.shift = shift;
    tab.block_start[0] = 0;
    for (int c = 0; c < kSpmvClasses; ++c) {
        const long long count = tab.class_start[c + 1] - tab.class_start[c];
        long long blocks;
        if (c == 7) blocks = count;
        else blocks = (count * (1LL << spmv_lane_log(c, shift)) + kBlock - 1) / kBlock;
        tab.block_start[c + 1] = tab.block_start[c] + (int)blocks;
    }
}

// A CSR matrix in slot order: slot s holds original row row[s], whose
// entries are ci/v[rp[s] .. rp[s + 1]).
template <class V> struct BinnedCsr {
    int rows = 0, cols = 0, nnz = 0, total_blocks = 0;
    int *rp = nullptr, *row = nullptr, *ci = nullptr;
    V *v = nullptr;
    SpmvClassTable tab = {};
};

template <class T> T *upload_vec(const std::vector<T> &h) {
    T *d = nullptr;
    HIP_OK(hipMalloc(&d, std::max<size_t>(h.size(), 1) * sizeof(T)));
    if (!h.empty()) HIP_OK(hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
    return d;
}

// Builds the slot order (stable within a class) from a host CSR
// (rows x cols) and uploads the matrix in that order.
template <class V>
BinnedCsr<V> make_binned(int rows, int cols, const std::vector<int> &rp, const std::vector<int> &ci,
                         const std::vector<double> &val) {
    BinnedCsr<V> b;
    b.rows = rows;
    b.cols = cols;
    std::vector<int> count(kSpmvClasses, 0), cls((size_t)rows);
    for (int r = 0; r < rows; ++r) {
        cls[r] = spmv_class_of(rp[(size_t)r + 1] - rp[r]);
        count[cls[r]]++;
    }
    b.tab.class_start[0] = 0;
    for (int c = 0; c < kSpmvClasses; ++c) b.tab.class_start[c + 1] = b.tab.class_start[c] + count[c];
    std::vector<int> order((size_t)rows), cursor(b.tab.class_start, b.tab.class_start + kSpmvClasses);
    for (int r = 0; r < rows; ++r) order[cursor[cls[r]]++] = r;
    spmv_set_blocks(b.tab, 0);
    b.total_blocks = b.tab.block_start[kSpmvClasses];
    std::vector<int> srp((size_t)rows + 1, 0);
    for (int s = 0; s < rows; ++s) srp[(size_t)s + 1] = srp[s] + (rp[(size_t)order[s] + 1] - rp[order[s]]);
    b.nnz = srp[rows];
    std::vector<int> sci((size_t)b.nnz);
    std::vector<V> sv((size_t)b.nnz);
    for (int s = 0; s < rows; ++s) {
        const int r = order[s];
        int o = srp[s];
        for (int e = rp[r]; e < rp[(size_t)r + 1]; ++e, ++o) {
            sci[o] = ci[e];
            sv[o] = (V)val[e];
        }
    }
    b.rp = upload_vec(srp);
    b.row = upload_vec(order);
    b.ci = upload_vec(sci);
    b.v = upload_vec(sv);
    return b;
}

template <class V> void set_lane_shift(BinnedCsr<V> &b, int shift) {
    spmv_set_blocks(b.tab, shift);
    b.total_blocks = b.tab.block_start[kSpmvClasses];
}

template <class V> void free_binned(BinnedCsr<V> &b) {
    if (b.rp) (void)hipFree(b.rp);
    if (b.row) (void)hipFree(b.row);
    if (b.ci) (void)hipFree(b.ci);
    if (b.v) (void)hipFree(b.v);
    b = BinnedCsr<V>();
}

// Sum of val[e] * x[ci[e]] over e = e, e + step, ... < e1, added in that
// order.  Four entries are taken per round: their indices and values are
// loaded first and the four gathers of x issued together, instead of one
// dependent index-then-gather chain per entry.  (Non-temporal loads for the
// matrix stream were measured slower here: A'y went from 114 to 153 us.)
template <class V, class X, class Acc>
__device__ __forceinline__ Acc row_dot(const V *__restrict__ val, const int *__restrict__ ci, const X *__restrict__ x,
                                       int e, int e1, int step, Acc acc) {
    for (; e + 3 * step < e1; e += 4 * step) {
        int c[4];
        Acc v[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            c[q] = ci[e + q * step];
            v[q] = (Acc)val[e + q * step];
        }
        Acc g[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) g[q] = (Acc)x[c[q]];
#pragma unroll
        for (int q = 0; q < 4; ++q) acc += v[q] * g[q];
    }
    for (; e < e1; e += step) acc += (Acc)val[e] * (Acc)x[ci[e]];
    return acc;
}

// Workgroup `b` of a binned product (0 <= b < M.total_blocks): calls
// out(row, sum) once per row it covers, from exactly one thread.
template <class V, class X, class Acc, class Out>
__device__ __forceinline__ void binned_block(const BinnedCsr<V> &M, const X *__restrict__ x, int b, Out out) {
    const SpmvClassTable &tab = M.tab;
    int c = 0;
#pragma unroll
    for (int k = 1; k < kSpmvClasses; ++k) c += (b >= tab.block_start[k]) ? 1 : 0;
    const int local_block = b - tab.block_start[c];
    if (c < 7) {
        const int ll = spmv_lane_log(c, tab.shift);
        const int lanes = 1 << ll;
        const int t = local_block * kBlock + threadIdx.x;
        const int slot = t >> ll;
        const int lane = t & (lanes - 1);
        const int count = tab.class_start[c + 1] - tab.class_start[c];
        const bool active = slot < count;
        int row = 0;
        Acc acc = Acc(0);
        if (active) {
            const int s = tab.class_start[c] + slot;
            row = M.row[s];
            acc = row_dot<V, X, Acc>(M.v, M.ci, x, M.rp[s] + lane, M.rp[s + 1], lanes, acc);
        }
        if (lanes > 1) acc = group_sum(acc, lanes);
        if (active && lane == 0) out(row, acc);
    } else {
        __shared__ Acc part[kBlock / kWave];
        const int s = tab.class_start[7] + local_block;
        const int row = M.row[s];
        Acc acc = row_dot<V, X, Acc>(M.v, M.ci, x, M.rp[s] + threadIdx.x, M.rp[s + 1], kBlock, Acc(0));
        acc = wave_sum(acc);
        if ((threadIdx.x % kWave) == 0) part[threadIdx.x / kWave] = acc;
        __syncthreads();
        if (threadIdx.x == 0) {
            Acc sum = part[0];
            for (int w = 1; w < kBlock / kWave; ++w) sum += part[w];
            out(row, sum);
        }
    }
}

// y-side epilogue contract: `epi(row, sum)` is called once per row by exactly
// one thread.  `gate` (may be null) skips the whole launch when *gate == 0.
template <class V, class X, class Acc, class Epi>
__global__ void __launch_bounds__(kBlock) spmv_binned_kernel(BinnedCsr<V> M, const X *__restrict__ x, Epi epi,
                                                              const int *gate) {
    if (gate && *gate == 0) return;
    binned_block<V, X, Acc>(M, x, blockIdx.x, epi);
}

template <class V, class X, class Acc, class Epi>
void spmv_launch(const BinnedCsr<V> &M, const X *x, Epi epi, hipStream_t s, const int *gate = nullptr) {
    if (M.total_blocks == 0) return;
    hipLaunchKernelGGL((spmv_binned_kernel<V, X, Acc, Epi>), dim3(M.total_blocks), dim3(kBlock), 0, s, M, x, epi,
                       gate);
    HIP_OK(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Column-panel SpMV for gathers larger than one XCD's L2.
//
// The columns are split into kPanels contiguous panels and the matrix into one
// binned CSR per panel (all rows, global column indices).  Workgroup b is
// dispatched to XCD b % kXcds, so sending workgroup b to panel
// (b % kXcds) % kPanels makes each XCD gather from one panel's slice of x,
// which stays in its L2.  Each panel writes its per-row partial sums; a
// combine kernel adds a row's kPanels partials in panel order and calls the
// epilogue.
// ---------------------------------------------------------------------------
constexpr int kXcds = 8;
constexpr int kPanels = 4;
constexpr int kXcdsPerPanel = kXcds / kPanels;
static_assert(kXcds % kPanels == 0, "the panels must divide the XCDs evenly");

template <class V, class Acc> struct PanelSet {
    BinnedCsr<V> M[kPanels];
    Acc *partial;  // [kPanels][rows]
    int rows;
};

template <class Acc> struct PanelStore {
    Acc *out;
    __device__ void operator()(int r, Acc s) const { out[r] = s; }
};

template <class V, class X, class Acc>
__global__ void __launch_bounds__(kBlock) spmv_panel_kernel(PanelSet<V, Acc> S, const X *__restrict__ x,
                                                             const int *gate) {
    if (gate && *gate == 0) return;
    const int xcd = blockIdx.x % kXcds;
    const int p = xcd % kPanels;
    const int b = (blockIdx.x / kXcds) * kXcdsPerPanel + xcd / kPanels;
    if (b >= S.M[p].total_blocks) return;
    binned_block<V, X, Acc>(S.M[p], x, b, PanelStore<Acc>{S.partial + (size_t)p * S.rows});
}

template <class Acc, class Epi>
__global__ void __launch_bounds__(kBlock) spmv_panel_combine_kernel(const Acc *__restrict__ partial, int rows, Epi epi,
                                                                     const int *gate) {
    if (gate && *gate == 0) return;
    for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < rows; r += gridDim.x * blockDim.x) {
        Acc s = partial[r];
#pragma unroll
        for (int p = 1; p < kPanels; ++p) s += partial[(size_t)p * rows + r];
        epi(r, s);
    }
}

// Host side: the panel matrices of one operator.
template <class V, class Acc> struct PanelSpmv {
    bool on = false;
    PanelSet<V, Acc> set = {};
    int max_blocks = 0;

    // rp/ci/v: host CSR of the operator (rows x cols).
    void build(int nrows, int ncols, const std::vector<int> &rp, const std::vector<int> &ci,
               const std::vector<double> &v) {
        // panel p holds columns [lo_p, lo_{p+1}) with lo_p the first column
        // whose col * kPanels / ncols is p; one pass counts, one scatters
        int lo[kPanels + 1];
        for (int p = 0; p <= kPanels; ++p)
            lo[p] = p == kPanels ? ncols : (int)(((long long)p * std::max(ncols, 1) + kPanels - 1) / kPanels);
        auto panel_of = [&](int col) {
            int p = 0;
            while (p + 1 < kPanels && col >= lo[p + 1]) ++p;
            return p;
        };
        std::vector<std::vector<int>> prp(kPanels, std::vector<int>((size_t)nrows + 1, 0));
        for (int r = 0; r < nrows; ++r)
            for (int e = rp[r]; e < rp[(size_t)r + 1]; ++e) prp[panel_of(ci[e])][(size_t)r + 1]++;
        for (int p = 0; p < kPanels; ++p)
            for (int r = 0; r < nrows; ++r) prp[p][(size_t)r + 1] += prp[p][r];
        std::vector<std::vector<int>> pci(kPanels), fill(kPanels);
        std::vector<std::vector<double>> pv(kPanels);
        for (int p = 0; p < kPanels; ++p) {
            pci[p].resize((size_t)prp[p][nrows]);
            pv[p].resize((size_t)prp[p][nrows]);
            fill[p].assign(prp[p].begin(), prp[p].end() - 1);
        }
        for (int r = 0; r < nrows; ++r)
            for (int e = rp[r]; e < rp[(size_t)r + 1]; ++e) {
                const int p = panel_of(ci[e]);
                const int o = fill[p][r]++;
                pci[p][o] = ci[e];
                pv[p][o] = v[e];
            }
        for (int p = 0; p < kPanels; ++p) {
            set.M[p] = make_binned<V>(nrows, ncols, prp[p], pci[p], pv[p]);
            max_blocks = std::max(max_blocks, set.M[p].total_blocks);
        }
        HIP_OK(hipMalloc(&set.partial, (size_t)kPanels * std::max(nrows, 1) * sizeof(Acc)));
        set.rows = nrows;
        on = true;
    }

    void release() {
        for (int p = 0; p < kPanels; ++p) free_binned(set.M[p]);
        if (set.partial) (void)hipFree(set.partial);
        set = PanelSet<V, Acc>();
        max_blocks = 0;
        on = false;
    }

    template <class X, class Epi> void launch(const X *x, Epi epi, hipStream_t s, const int *gate) const {
        if (max_blocks > 0) {
            const int rounds = (max_blocks + kXcdsPerPanel - 1) / kXcdsPerPanel;
            hipLaunchKernelGGL((spmv_panel_kernel<V, X, Acc>), dim3(kXcds * rounds), dim3(kBlock), 0, s, set, x, gate);
        }
        const int blocks = std::max(1, std::min((set.rows + kBlock - 1) / kBlock, 65535));
        hipLaunchKernelGGL((spmv_panel_combine_kernel<Acc, Epi>), dim3(blocks), dim3(kBlock), 0, s,
                           (const Acc *)set.partial, set.rows, epi, gate);
        HIP_OK(hipGetLastError());
    }
};

// Plain epilogues.
template <class X> struct EpiStore {
    X *y;
    template <class Acc> __device__ void operator()(int r, Acc s) const { y[r] = (X)s; }
};
template <class X> struct EpiAccumulate {
    X *y;  // y += M x  (beta = 1)
    template <class Acc> __device__ void operator()(int r, Acc s) const { y[r] = (X)(s + (Acc)y[r]); }
};

}  // namespace pdhcg
