// Device-side building blocks for gfx950 (CDNA4, wave64): error checking,
// wave/block reductions and a deterministic grid reduction.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "host_common.h"

#define HIP_OK(call)                                                                                   \
    do {                                                                                               \
        hipError_t err_ = (call);                                                                      \
        if (err_ != hipSuccess)                                                                        \
            ::pdhcg::fail("HIP error %s at %s:%d: %s", hipGetErrorName(err_), __FILE__, __LINE__,      \
                          hipGetErrorString(err_));                                                    \
    } while (0)

namespace pdhcg {

constexpr int kWave = 64;
constexpr int kBlock = 256;            // threads per block of the elementwise/reduction kernels
constexpr int kMaxRedBlocks = 4096;    // upper bound on blocks of any reducing launch
constexpr int kMaxRedValues = 16;      // values reduced by one launch
constexpr size_t kRedPartials = (size_t)kMaxRedValues * kMaxRedBlocks;

template <class T> __device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

// Sum over `width` consecutive lanes (width a power of two <= 64); every lane
// of the group receives the group total.
template <class T> __device__ __forceinline__ T group_sum(T v, int width) {
    for (int off = width >> 1; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

// Combines two partial values: the first NS of a value pack are sums, the
// remaining ones are maxima.
template <int NS> __device__ __forceinline__ double combine(int i, double a, double b) {
    return i < NS ? a + b : fmax(a, b);
}

// Reduction request: the output receives NS sums followed by NM maxima (FP64).
// A reducing kernel writes one partial pack per workgroup; the fold kernel
// launched after it on the same stream combines them (no atomics, no fences:
// the kernel boundary orders the writes before the reads).
struct GridRed {
    double *partials;   // kRedPartials: [kMaxRedValues][kMaxRedBlocks]
    double *out;
};

// Load of a value written by another workgroup of the same launch: coherent
// at device scope (bypasses a possibly stale per-CU cache line) but, unlike a
// volatile access, free to be pipelined with other loads.
__device__ __forceinline__ double coherent_load(const double *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Block reduction of a value pack into thread 0 (fixed combination order).
template <int NS, int NM> __device__ __forceinline__ void block_reduce(double (&v)[NS + NM], double *lds) {
    constexpr int NV = NS + NM;
    const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave, waves = blockDim.x / kWave;
#pragma unroll
    for (int i = 0; i < NV; ++i)
#pragma unroll
        for (int off = kWave / 2; off > 0; off >>= 1) v[i] = combine<NS>(i, v[i], __shfl_xor(v[i], off, kWave));
    if (lane == 0)
#pragma unroll
        for (int i = 0; i < NV; ++i) lds[i * 16 + wave] = v[i];
    __syncthreads();
    if (threadIdx.x == 0)
#pragma unroll
        for (int i = 0; i < NV; ++i) {
            double a = lds[i * 16];
            for (int w = 1; w < waves; ++w) a = combine<NS>(i, a, lds[i * 16 + w]);
            v[i] = a;
        }
    __syncthreads();
}

// Stores this workgroup's partial pack (thread 0 writes after a block
// reduction in a fixed order).  Launch grid_fold_kernel with the same grid
// size afterwards.
template <int NS, int NM> __device__ void grid_partials(const GridRed &g, double (&v)[NS + NM]) {
    constexpr int NV = NS + NM;
    static_assert(NV <= kMaxRedValues, "too many reduced values");
    __shared__ double lds[NV * 16];
    block_reduce<NS, NM>(v, lds);
    if (threadIdx.x == 0)
#pragma unroll
        for (int i = 0; i < NV; ++i) g.partials[(size_t)i * kMaxRedBlocks + blockIdx.x] = v[i];
}

// No follow-up work after a reduction.
struct NoTail {
    template <class S> __device__ void operator()(const double *, S *) const {}
};

// One workgroup folds the partial packs of `nblocks` workgroups in a fixed
// order (thread t takes blocks t, t + 256, ... into four accumulators, so four
// loads are in flight), writes g.out and hands the values to
// tail(values, scalars) in thread 0.
// gate: skip everything when *gate == 0 (device-resident loops).
template <int NS, int NM, class Tail, class S>
__global__ void __launch_bounds__(kBlock) grid_fold_kernel(GridRed g, int nblocks, Tail tail, S *scalars,
                                                            const int *gate) {
    if (gate && *gate == 0) return;
    constexpr int NV = NS + NM;
    __shared__ double lds[NV * 16];
    double t[NV];
#pragma unroll
    for (int i = 0; i < NV; ++i) {
        const double id = i < NS ? 0.0 : -INFINITY;
        const double *P = g.partials + (size_t)i * kMaxRedBlocks;
        double a[4] = {id, id, id, id};
        for (int b0 = threadIdx.x; b0 < nblocks; b0 += 4 * kBlock)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int b = b0 + q * kBlock;
                a[q] = combine<NS>(i, a[q], b < nblocks ? P[b] : id);
            }
        t[i] = combine<NS>(i, combine<NS>(i, a[0], a[1]), combine<NS>(i, a[2], a[3]));
    }
    block_reduce<NS, NM>(t, lds);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < NV; ++i) g.out[i] = t[i];
        tail(t, scalars);
    }
}

inline int blocks_for(long long n, int per_block = kBlock, int cap = kMaxRedBlocks) {
    long long b = (n + per_block - 1) / per_block;
    if (b < 1) b = 1;
    if (b > cap) b = cap;
    return (int)b;
}

// |x| for the infinity norm; a non-finite entry makes the norm infinite
// (reference src/utils.cu:563-564).
__device__ __forceinline__ double abs_or_inf(double x) { return isfinite(x) ? fabs(x) : INFINITY; }

}  // namespace pdhcg
