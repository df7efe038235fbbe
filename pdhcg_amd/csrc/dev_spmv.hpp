// Row-length-binned CSR sparse matrix-vector product for gfx950.
//
// Rows are sorted (stably) into classes by length.  A class assigns L lanes
// of a 64-wide wavefront to each row (L = 1, 2, ..., 64), so short rows are
// packed many to a wave and long rows get a whole wave; rows longer than
// kLongRow get a full 256-thread workgroup with an LDS reduction.  One launch
// covers every class: the workgroup index selects the class through a small
// offset table.  Each row's sum is produced in a fixed order (lane-strided
// partial sums, then an xor-butterfly), so the result is deterministic.
// The per-row result is handed to an epilogue functor, which lets the PDHG
// proximal updates run inside the SpMV without writing the product to HBM.
#pragma once

#include <vector>

#include "dev_common.hpp"

namespace pdhcg {

constexpr int kSpmvClasses = 8;      // L = 1,2,4,8,16,32,64, and workgroup rows
constexpr int kLongRow = 2048;       // rows longer than this get a workgroup

struct SpmvPlan {
    int rows = 0;
    int *d_order = nullptr;                       // row ids grouped by class
    int class_start[kSpmvClasses + 1] = {0};      // offsets into d_order
    int block_start[kSpmvClasses + 1] = {0};      // first workgroup of each class
    int total_blocks = 0;
};

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

// Builds the class lists on the host and uploads them (stable in row order).
inline SpmvPlan make_spmv_plan(int rows, const std::vector<int> &rp) {
    SpmvPlan p;
    p.rows = rows;
    std::vector<int> count(kSpmvClasses, 0), cls((size_t)rows);
    for (int r = 0; r < rows; ++r) {
        cls[r] = spmv_class_of(rp[(size_t)r + 1] - rp[r]);
        count[cls[r]]++;
    }
    p.class_start[0] = 0;
    for (int c = 0; c < kSpmvClasses; ++c) p.class_start[c + 1] = p.class_start[c] + count[c];
    std::vector<int> order((size_t)(rows > 0 ? rows : 1)), cursor(p.class_start, p.class_start + kSpmvClasses);
    for (int r = 0; r < rows; ++r) order[cursor[cls[r]]++] = r;
    p.block_start[0] = 0;
    for (int c = 0; c < kSpmvClasses; ++c) {
        long long blocks;
        if (c == 7) blocks = count[c];
        else {
            const long long lanes = 1LL << c;
            blocks = (count[c] * lanes + kBlock - 1) / kBlock;
        }
        p.block_start[c + 1] = p.block_start[c] + (int)blocks;
    }
    p.total_blocks = p.block_start[kSpmvClasses];
    HIP_OK(hipMalloc(&p.d_order, order.size() * sizeof(int)));
    HIP_OK(hipMemcpy(p.d_order, order.data(), order.size() * sizeof(int), hipMemcpyHostToDevice));
    return p;
}

inline void free_spmv_plan(SpmvPlan &p) {
    if (p.d_order) (void)hipFree(p.d_order);
    p.d_order = nullptr;
}

struct SpmvClassTable {
    int class_start[kSpmvClasses + 1];
    int block_start[kSpmvClasses + 1];
};

template <class V> struct DevCsr {
    int rows = 0, cols = 0, nnz = 0;
    int *rp = nullptr, *ci = nullptr;
    V *v = nullptr;
};

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

// y-side epilogue contract: `epi(row, sum)` is called once per row by exactly
// one thread.  `gate` (may be null) skips the whole launch when *gate == 0.
template <class V, class X, class Acc, class Epi>
__global__ void __launch_bounds__(kBlock) spmv_binned_kernel(DevCsr<V> M, const X *__restrict__ x,
                                                              const int *__restrict__ order, SpmvClassTable tab,
                                                              Epi epi, const int *gate) {
    if (gate && *gate == 0) return;
    const int b = blockIdx.x;
    int c = 0;
#pragma unroll
    for (int k = 1; k < kSpmvClasses; ++k) c += (b >= tab.block_start[k]) ? 1 : 0;
    const int local_block = b - tab.block_start[c];
    const int *rp = M.rp;
    const int *ci = M.ci;
    const V *val = M.v;
    if (c < 7) {
        const int lanes = 1 << c;
        const int t = local_block * kBlock + threadIdx.x;
        const int slot = t >> c;
        const int lane = t & (lanes - 1);
        const int count = tab.class_start[c + 1] - tab.class_start[c];
        const bool active = slot < count;
        int row = 0;
        Acc acc = Acc(0);
        if (active) {
            row = order[tab.class_start[c] + slot];
            const int e0 = rp[row], e1 = rp[row + 1];
            acc = row_dot<V, X, Acc>(val, ci, x, e0 + lane, e1, lanes, acc);
        }
        if (lanes > 1) acc = group_sum(acc, lanes);
        if (active && lane == 0) epi(row, acc);
    } else {
        __shared__ Acc part[kBlock / kWave];
        const int row = order[tab.class_start[7] + local_block];
        const int e0 = rp[row], e1 = rp[row + 1];
        Acc acc = Acc(0);
        acc = row_dot<V, X, Acc>(val, ci, x, e0 + threadIdx.x, e1, kBlock, acc);
        acc = wave_sum(acc);
        if ((threadIdx.x % kWave) == 0) part[threadIdx.x / kWave] = acc;
        __syncthreads();
        if (threadIdx.x == 0) {
            Acc s = part[0];
            for (int w = 1; w < kBlock / kWave; ++w) s += part[w];
            epi(row, s);
        }
    }
}

template <class V, class X, class Acc, class Epi>
void spmv_launch(const DevCsr<V> &M, const SpmvPlan &plan, const X *x, Epi epi, hipStream_t s,
                 const int *gate = nullptr) {
    if (plan.total_blocks == 0) return;
    SpmvClassTable tab;
    for (int i = 0; i <= kSpmvClasses; ++i) {
        tab.class_start[i] = plan.class_start[i];
        tab.block_start[i] = plan.block_start[i];
    }
    hipLaunchKernelGGL((spmv_binned_kernel<V, X, Acc, Epi>), dim3(plan.total_blocks), dim3(kBlock), 0, s, M, x,
                       plan.d_order, tab, epi, gate);
    HIP_OK(hipGetLastError());
}

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
