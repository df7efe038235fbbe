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
constexpr int kMaxRedBlocks = 2048;    // upper bound on blocks of any reducing launch
constexpr int kMaxRedValues = 16;      // values reduced by one launch

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
struct GridRed {
    double *partials;   // [kMaxRedValues][kMaxRedBlocks]
    unsigned *counter;  // arrival counter, zero between launches
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

// Grid reduction with a fixed combination order: every block stores its
// partials, the last block to arrive folds them in block order, writes g.out
// and returns true (in all of its threads); the other blocks return false.
// The result does not depend on block scheduling.  Maxima start from -inf,
// so callers reduce magnitudes (>= 0) or explicit values.
template <int NS, int NM> __device__ bool grid_reduce(const GridRed &g, double (&v)[NS + NM]) {
    constexpr int NV = NS + NM;
    static_assert(NV <= kMaxRedValues, "too many reduced values");
    __shared__ double lds[NV * 16];
    __shared__ int last;
    block_reduce<NS, NM>(v, lds);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < NV; ++i) g.partials[(size_t)i * kMaxRedBlocks + blockIdx.x] = v[i];
        __threadfence();
        unsigned prev = atomicAdd(g.counter, 1u);
        last = (prev == gridDim.x - 1) ? 1 : 0;
    }
    __syncthreads();
    if (!last) return false;
    __threadfence();
    double t[NV];
#pragma unroll
    for (int i = 0; i < NV; ++i) t[i] = i < NS ? 0.0 : -INFINITY;
    for (int b = threadIdx.x; b < (int)gridDim.x; b += blockDim.x)
#pragma unroll
        for (int i = 0; i < NV; ++i)
            t[i] = combine<NS>(i, t[i], coherent_load(g.partials + (size_t)i * kMaxRedBlocks + b));
    block_reduce<NS, NM>(t, lds);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < NV; ++i) g.out[i] = t[i];
        *g.counter = 0u;
#pragma unroll
        for (int i = 0; i < NV; ++i) v[i] = t[i];
    }
    __syncthreads();
    return true;
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
