// Learned warm start: a 3-layer MLP (in -> H -> H -> out, GELU) that maps the
// parameters of a problem family to a primal/dual starting point.  Training
// and inference run on the device: BF16 matrix-core products with FP32
// accumulation whose epilogues apply the elementwise work (bias, GELU, loss,
// GELU derivative, bias-gradient column sums), and one Adam pass over the
// flat FP32 parameter array.
//
// Matrix product: C[M x N] = op(A)[M x K] op(B)[K x N], BF16 operands, op =
// identity or transpose of a row-major array; M, N, K multiples of 16 (the
// host pads).  A workgroup of four waves owns a 64 x 64 tile of C and walks K
// in steps of 128 (few, large steps: the products here are small and each
// step costs a memory round trip): each step's A and B tiles go global ->
// registers (16-byte loads along the contiguous dimension) -> LDS
// (double-buffered, one barrier per step), and each wave multiplies its
// 32 x 32 quarter as 2 x 2 blocks of gfx950's v_mfma_f32_16x16x32_bf16 (K = 32
// per instruction, twice the K of the CDNA3 16x16x16 form).  Per lane, the A
// operand holds op(A)[r][k + 8g .. k + 8g + 7] for r = lane % 16,
// g = lane / 16, the B operand op(B)[k + 8g .. + 7][c] for c = lane % 16, and
// the 16 x 16 result holds C[4g + i][lane % 16] in register i.  An operand
// whose storage is contiguous along k is kept k-major in LDS (one 16-byte
// read per operand); the others are kept in their storage order and read as
// eight 2-byte values.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

#include "../../include/pdhcg_mlp.h"
#include "dev_common.hpp"

namespace pdhcg {
namespace mlp {

typedef unsigned short bf16;
typedef short s8 __attribute__((ext_vector_type(8)));
typedef __bf16 b8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));

// float -> bf16 with round to nearest even (NaN kept quiet)
__host__ __device__ inline bf16 to_bf16(float f) {
    unsigned u;
    memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (bf16)((u >> 16) | 0x40);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (bf16)(u >> 16);
}

constexpr int kBM = 64, kBN = 64, kBK = 128;
constexpr int kPitchK = kBK + 8;     // k-major LDS rows (272 bytes: 16 rows on distinct banks)
constexpr int kPitchMN = kBM + 8;    // storage-order LDS rows (144 bytes)
constexpr int kTileElems = kBK * kPitchMN > 64 * kPitchK ? kBK * kPitchMN : 64 * kPitchK;
constexpr int kLoads = kBM * kBK / 8 / 256;  // 16-byte loads per thread per operand tile

__device__ inline float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ inline float gelu_grad(float z) {
    return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) + z * 0.39894228040143268f * expf(-0.5f * z * z);
}

// Epilogue of a product, applied to v = (op(A) op(B))[r][c]:
enum Epi {
    kStore,      // C = v
    kBiasGelu,   // z = v + bias[c]; C = z; H = bf16(GELU(z))
    kBias,       // C = v + bias[c]
    kBiasMse,    // y = v + bias[c]; g = scale (y - T[r][c]) for c < out, else 0; H = bf16(g);
                 // column sums of g and the sum of (y - T)^2 per tile
    kGeluBack,   // g = v GELU'(Z[r][c]); H = bf16(g); column sums of g per tile
};

struct GemmArgs {
    int M, N, K;
    const bf16 *A;
    int lda;
    const bf16 *B;
    int ldb;
    float *C;        // M x N (ldc = N) for kStore / kBiasGelu / kBias
    const float *bias;
    bf16 *H;         // M x N BF16 output
    const float *Z;  // M x N (kGeluBack)
    const float *T;  // M x out targets (kBiasMse)
    int out;
    float scale;
    float *colpart;  // [row tile][N] column sums
    double *lossp;   // [tile] sums of squared errors
};

// One K-step of an operand tile: 256 threads x kLoads x 8 BF16 values.  KMAJ:
// storage is contiguous along k (64 rows of kBK k-values); otherwise storage
// is contiguous along the M / N index (kBK k-rows of 64 values).  Chunk
// e = t + 256 u of the tile is row e / (kBK / 8), 8-value column e % (kBK / 8)
// (KMAJ), or k-row e / 8, column e % 8.
struct TileRegs {
    uint4 v[kLoads];
};
template <bool KMAJ>
__device__ inline void load_tile(TileRegs &r, const bf16 *P, int ld, int i0, int imax, int k0, int K) {
#pragma unroll
    for (int u = 0; u < kLoads; ++u) {
        const int e = threadIdx.x + 256 * u;
        int i, k;
        if (KMAJ) {
            i = i0 + e / (kBK / 8);
            k = k0 + (e % (kBK / 8)) * 8;
        } else {
            k = k0 + e / 8;
            i = i0 + (e % 8) * 8;
        }
        r.v[u] = uint4{0u, 0u, 0u, 0u};
        if (i < imax && k < K)
            r.v[u] = *reinterpret_cast<const uint4 *>(P + (KMAJ ? (size_t)i * ld + k : (size_t)k * ld + i));
    }
}
template <bool KMAJ> __device__ inline void store_tile(bf16 *S, const TileRegs &r) {
#pragma unroll
    for (int u = 0; u < kLoads; ++u) {
        const int e = threadIdx.x + 256 * u;
        bf16 *d = KMAJ ? S + (e / (kBK / 8)) * kPitchK + (e % (kBK / 8)) * 8 : S + (e / 8) * kPitchMN + (e % 8) * 8;
        *reinterpret_cast<uint4 *>(d) = r.v[u];
    }
}
// MFMA operand (8 k-values) for tile-local index i (row of op(A) / column of
// op(B)) and k-group base kb.
template <bool KMAJ> __device__ inline s8 read_op(const bf16 *S, int i, int kb) {
    if (KMAJ) return *reinterpret_cast<const s8 *>(S + i * kPitchK + kb);
    s8 v;
#pragma unroll
    for (int q = 0; q < 8; ++q) v[q] = (short)S[(kb + q) * kPitchMN + i];
    return v;
}

// Grid: (ceil(N / 64), ceil(M / 64)); 256 threads.
template <bool TA, bool TB, int EPI> __global__ void __launch_bounds__(256) gemm_kernel(GemmArgs g) {
    constexpr bool AK = !TA;  // op(A) = A (M x K storage): contiguous along k
    constexpr bool BK = TB;   // op(B) = B' (N x K storage): contiguous along k
    __shared__ __align__(16) bf16 sA[2][kTileElems];
    __shared__ __align__(16) bf16 sB[2][kTileElems];
    __shared__ float scol[2][kBN];
    __shared__ double sloss[4];
    const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
    const int l16 = lane % 16, kg = 4 * (lane / 16), kop = 8 * (lane / 16);
    const int r0 = blockIdx.y * kBM, c0 = blockIdx.x * kBN;
    const int wr = (wave / 2) * 32, wc = (wave % 2) * 32;
    const int nk = (g.K + kBK - 1) / kBK;
    f4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};

    TileRegs ra, rb;
    load_tile<AK>(ra, g.A, g.lda, r0, g.M, 0, g.K);
    load_tile<BK>(rb, g.B, g.ldb, c0, g.N, 0, g.K);
    store_tile<AK>(sA[0], ra);
    store_tile<BK>(sB[0], rb);
    __syncthreads();
    for (int kt = 0; kt < nk; ++kt) {
        const int cur = kt & 1;
        if (kt + 1 < nk) {
            load_tile<AK>(ra, g.A, g.lda, r0, g.M, (kt + 1) * kBK, g.K);
            load_tile<BK>(rb, g.B, g.ldb, c0, g.N, (kt + 1) * kBK, g.K);
        }
#pragma unroll
        for (int kk = 0; kk < kBK; kk += 32) {
            s8 a[2], b[2];
#pragma unroll
            for (int i = 0; i < 2; ++i) a[i] = read_op<AK>(sA[cur], wr + 16 * i + l16, kk + kop);
#pragma unroll
            for (int j = 0; j < 2; ++j) b[j] = read_op<BK>(sB[cur], wc + 16 * j + l16, kk + kop);
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(b8, a[i]),
                                                                        __builtin_bit_cast(b8, b[j]), acc[i][j], 0, 0, 0);
        }
        if (kt + 1 < nk) {
            store_tile<AK>(sA[cur ^ 1], ra);
            store_tile<BK>(sB[cur ^ 1], rb);
        }
        __syncthreads();
    }

    // epilogue
    constexpr bool kReduce = EPI == kBiasMse || EPI == kGeluBack;
    float csum[2] = {0.f, 0.f};
    double lsum = 0.0;
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        const int c = c0 + wc + 16 * j + l16;
        if (c0 + wc + 16 * j >= g.N) continue;
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            if (r0 + wr + 16 * i >= g.M) continue;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int r = r0 + wr + 16 * i + kg + q;
                const size_t o = (size_t)r * g.N + c;
                const float v = acc[i][j][q];
                if (EPI == kStore) {
                    g.C[o] = v;
                } else if (EPI == kBiasGelu) {
                    const float z = v + g.bias[c];
                    g.C[o] = z;
                    g.H[o] = to_bf16(gelu(z));
                } else if (EPI == kBias) {
                    g.C[o] = v + g.bias[c];
                } else if (EPI == kBiasMse) {
                    float gr = 0.0f;
                    if (c < g.out) {
                        const float d = v + g.bias[c] - g.T[(size_t)r * g.out + c];
                        lsum += (double)d * d;
                        gr = g.scale * d;
                    }
                    g.H[o] = to_bf16(gr);
                    csum[j] += gr;
                } else {
                    const float gr = v * gelu_grad(g.Z[o]);
                    g.H[o] = to_bf16(gr);
                    csum[j] += gr;
                }
            }
        }
    }
    if (kReduce) {
        // column sums: over the 4 row groups of the wave, then the 2 waves of a column half
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            csum[j] += __shfl_xor(csum[j], 16, kWave);
            csum[j] += __shfl_xor(csum[j], 32, kWave);
        }
        if (lane < 16) {
            scol[wave / 2][wc + l16] = csum[0];
            scol[wave / 2][wc + 16 + l16] = csum[1];
        }
        if (EPI == kBiasMse) {
            lsum = wave_sum(lsum);
            if (lane == 0) sloss[wave] = lsum;
        }
        __syncthreads();
        if (threadIdx.x < kBN && c0 + (int)threadIdx.x < g.N)
            g.colpart[(size_t)blockIdx.y * g.N + c0 + threadIdx.x] = scol[0][threadIdx.x] + scol[1][threadIdx.x];
        if (EPI == kBiasMse && threadIdx.x == 0)
            g.lossp[blockIdx.y * gridDim.x + blockIdx.x] = ((sloss[0] + sloss[1]) + sloss[2]) + sloss[3];
    }
}

__global__ void to_bf16_kernel(const float *x, bf16 *y, long long n) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        y[i] = to_bf16(x[i]);
}

// Minibatch assembly from the device-resident samples: row r of the batch is
// sample idx[r] for r < valid and zero for valid <= r < rows; X (count x in)
// goes into the padded BF16 input (rows x in_p), T (count x out), when given,
// into the targets (rows x out).
__global__ void gather_kernel(const float *X, const float *T, const int *idx, int rows, int valid, int in, int in_p,
                              int out, bf16 *xb, float *t) {
    const long long nx = (long long)rows * in_p, nt = T ? (long long)rows * out : 0;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < nx + nt;
         i += (long long)gridDim.x * blockDim.x) {
        if (i < nx) {
            const int r = (int)(i / in_p), c = (int)(i % in_p);
            xb[i] = (r < valid && c < in) ? to_bf16(X[(size_t)idx[r] * in + c]) : (bf16)0;
        } else {
            const long long j = i - nx;
            const int r = (int)(j / out), c = (int)(j % out);
            t[j] = r < valid ? T[(size_t)idx[r] * out + c] : 0.0f;
        }
    }
}

// Bias gradients from the per-row-tile column sums (fixed order), and the
// loss from the per-tile squared-error sums (blockIdx.y == 3).
struct FoldArgs {
    const float *colpart[3];
    float *gb[3];
    int cols[3];
    int tiles;           // row tiles of the batch
    const double *lossp;
    int nloss;
    double scale;
    double *loss;
};
__global__ void fold_kernel(FoldArgs f) {
    const int s = blockIdx.y;
    if (s < 3) {
        const int c = blockIdx.x * blockDim.x + threadIdx.x;
        if (c >= f.cols[s]) return;
        float acc = 0.0f;
        for (int t = 0; t < f.tiles; ++t) acc += f.colpart[s][(size_t)t * f.cols[s] + c];
        f.gb[s][c] = acc;
        return;
    }
    if (blockIdx.x != 0) return;
    double acc = 0.0;
    for (int i = threadIdx.x; i < f.nloss; i += blockDim.x) acc += f.lossp[i];
    __shared__ double lds[16];
    double v[1] = {acc};
    block_reduce<1, 0>(v, lds);
    if (threadIdx.x == 0) *f.loss = v[0] * f.scale;
}

// Adam (Kingma & Ba 2015) on the flat FP32 parameters; refreshes the BF16 copy.
__global__ void adam_kernel(float *w, const float *g, float *m1, float *m2, long long n, float lr, float b1, float b2,
                            float c1, float c2, float eps, bf16 *wb) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float a = b1 * m1[i] + (1.0f - b1) * gi, b = b2 * m2[i] + (1.0f - b2) * gi * gi;
        m1[i] = a;
        m2[i] = b;
        const float wi = w[i] - lr * (a / c1) / (sqrtf(b / c2) + eps);
        w[i] = wi;
        wb[i] = to_bf16(wi);
    }
}

inline int pad16(int v) { return (v + 15) / 16 * 16; }
inline int ew_grid(long long n) { return (int)std::max<long long>(1, std::min<long long>((n + 255) / 256, 4096)); }

template <class T> T *dnew(size_t n) {
    T *p = nullptr;
    HIP_OK(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)));
    HIP_OK(hipMemset(p, 0, std::max<size_t>(n, 1) * sizeof(T)));
    return p;
}

// Layer l: weights at offset w (in x out, row-major, padded sizes), bias at b.
struct Layer {
    int in, out;
    size_t w, b;
};

struct Net {
    int in, hidden, out, in_p, out_p;
    Layer L[3];
    size_t np = 0;                      // flat parameter count (padded)
    float *P = nullptr, *G = nullptr, *M1 = nullptr, *M2 = nullptr;
    bf16 *PB = nullptr;                 // BF16 copy of P
    int step = 0;
    hipStream_t st = nullptr;
    // activations of the current minibatch (cap rows)
    int cap = 0;
    bf16 *xb = nullptr, *h1b = nullptr, *h2b = nullptr, *dyb = nullptr, *dz2b = nullptr, *dz1b = nullptr;
    float *z1 = nullptr, *z2 = nullptr, *y = nullptr, *t = nullptr;
    float *colpart[3] = {nullptr, nullptr, nullptr};
    double *lossp = nullptr, *loss = nullptr;  // tile partials; per-minibatch losses (kMaxSteps)
    // device-resident samples
    size_t data_cap = 0;
    float *dX = nullptr, *dT = nullptr;
    int *idx = nullptr;
};

constexpr int kMaxSteps = 1 << 16;  // minibatches per epoch

template <bool TA, bool TB, int EPI> void launch(hipStream_t st, const GemmArgs &a) {
    dim3 grid((unsigned)((a.N + kBN - 1) / kBN), (unsigned)((a.M + kBM - 1) / kBM));
    hipLaunchKernelGGL((gemm_kernel<TA, TB, EPI>), grid, dim3(256), 0, st, a);
    HIP_OK(hipGetLastError());
}

GemmArgs gemm_args(int M, int N, int K, const bf16 *A, int lda, const bf16 *B, int ldb) {
    GemmArgs a{};
    a.M = M;
    a.N = N;
    a.K = K;
    a.A = A;
    a.lda = lda;
    a.B = B;
    a.ldb = ldb;
    return a;
}

inline int row_tiles(int rows) { return (rows + kBM - 1) / kBM; }

void free_batch(Net &n) {
    for (void *p : {(void *)n.xb, (void *)n.h1b, (void *)n.h2b, (void *)n.dyb, (void *)n.dz2b, (void *)n.dz1b,
                    (void *)n.z1, (void *)n.z2, (void *)n.y, (void *)n.t, (void *)n.colpart[0], (void *)n.colpart[1],
                    (void *)n.colpart[2], (void *)n.lossp, (void *)n.loss})
        if (p) (void)hipFree(p);
}

void ensure_batch(Net &n, int rows) {
    if (rows <= n.cap) return;
    free_batch(n);
    const size_t r = (size_t)rows, H = (size_t)n.hidden, tiles = (size_t)row_tiles(rows);
    n.xb = dnew<bf16>(r * n.in_p);
    n.z1 = dnew<float>(r * H);
    n.h1b = dnew<bf16>(r * H);
    n.z2 = dnew<float>(r * H);
    n.h2b = dnew<bf16>(r * H);
    n.y = dnew<float>(r * n.out_p);
    n.t = dnew<float>(r * n.out);
    n.dyb = dnew<bf16>(r * n.out_p);
    n.dz2b = dnew<bf16>(r * H);
    n.dz1b = dnew<bf16>(r * H);
    n.colpart[0] = dnew<float>(tiles * H);
    n.colpart[1] = dnew<float>(tiles * H);
    n.colpart[2] = dnew<float>(tiles * n.out_p);
    n.lossp = dnew<double>(tiles * (size_t)((n.out_p + kBN - 1) / kBN));
    n.loss = dnew<double>(kMaxSteps);
    n.cap = rows;
}

// Copies `count` samples (X: count x in; T: count x out, or none) to the
// device and sets idx to the identity.
void stage(Net &n, const float *X, const float *T, int count) {
    if ((size_t)count > n.data_cap) {
        for (void *p : {(void *)n.dX, (void *)n.dT, (void *)n.idx})
            if (p) (void)hipFree(p);
        n.dX = dnew<float>((size_t)count * n.in);
        n.dT = dnew<float>((size_t)count * n.out);
        n.idx = dnew<int>((size_t)count);
        n.data_cap = (size_t)count;
    }
    HIP_OK(hipMemcpyAsync(n.dX, X, (size_t)count * n.in * sizeof(float), hipMemcpyHostToDevice, n.st));
    if (T) HIP_OK(hipMemcpyAsync(n.dT, T, (size_t)count * n.out * sizeof(float), hipMemcpyHostToDevice, n.st));
    std::vector<int> id((size_t)count);
    for (int i = 0; i < count; ++i) id[i] = i;
    HIP_OK(hipMemcpyAsync(n.idx, id.data(), id.size() * sizeof(int), hipMemcpyHostToDevice, n.st));
    HIP_OK(hipStreamSynchronize(n.st));
}

// Minibatch of `rows` rows from samples idx[first .. first + valid).
void gather(Net &n, int first, int rows, int valid, bool targets) {
    const long long work = (long long)rows * n.in_p + (targets ? (long long)rows * n.out : 0);
    hipLaunchKernelGGL(gather_kernel, dim3(ew_grid(work)), dim3(256), 0, n.st, n.dX, targets ? n.dT : nullptr,
                       n.idx + first, rows, valid, n.in, n.in_p, n.out, n.xb, n.t);
    HIP_OK(hipGetLastError());
}

// Forward pass of `rows` (multiple of 16) gathered rows; with `loss` set, the
// last layer's epilogue forms the loss gradient against the gathered targets
// instead of storing the output.
void forward(Net &n, int rows, bool loss) {
    const int H = n.hidden;
    GemmArgs a = gemm_args(rows, H, n.in_p, n.xb, n.in_p, n.PB + n.L[0].w, H);
    a.C = n.z1;
    a.bias = n.P + n.L[0].b;
    a.H = n.h1b;
    launch<false, false, kBiasGelu>(n.st, a);
    a = gemm_args(rows, H, H, n.h1b, H, n.PB + n.L[1].w, H);
    a.C = n.z2;
    a.bias = n.P + n.L[1].b;
    a.H = n.h2b;
    launch<false, false, kBiasGelu>(n.st, a);
    a = gemm_args(rows, n.out_p, H, n.h2b, H, n.PB + n.L[2].w, n.out_p);
    a.bias = n.P + n.L[2].b;
    if (!loss) {
        a.C = n.y;
        launch<false, false, kBias>(n.st, a);
        return;
    }
    a.T = n.t;
    a.out = n.out;
    a.scale = 2.0f / ((float)rows * (float)n.out);
    a.H = n.dyb;
    a.colpart = n.colpart[2];
    a.lossp = n.lossp;
    launch<false, false, kBiasMse>(n.st, a);
}

// Backward pass after forward(n, rows, true): fills G and writes the loss to
// the device slot *loss.
void backward(Net &n, int rows, double *loss) {
    const int H = n.hidden, O = n.out_p;
    // layer 3: gW3 = H2' dY, dZ2 = (dY W3') GELU'(Z2)
    GemmArgs a = gemm_args(H, O, rows, n.h2b, H, n.dyb, O);
    a.C = n.G + n.L[2].w;
    launch<true, false, kStore>(n.st, a);
    a = gemm_args(rows, H, O, n.dyb, O, n.PB + n.L[2].w, O);
    a.Z = n.z2;
    a.H = n.dz2b;
    a.colpart = n.colpart[1];
    launch<false, true, kGeluBack>(n.st, a);
    // layer 2: gW2 = H1' dZ2, dZ1 = (dZ2 W2') GELU'(Z1)
    a = gemm_args(H, H, rows, n.h1b, H, n.dz2b, H);
    a.C = n.G + n.L[1].w;
    launch<true, false, kStore>(n.st, a);
    a = gemm_args(rows, H, H, n.dz2b, H, n.PB + n.L[1].w, H);
    a.Z = n.z1;
    a.H = n.dz1b;
    a.colpart = n.colpart[0];
    launch<false, true, kGeluBack>(n.st, a);
    // layer 1: gW1 = X' dZ1
    a = gemm_args(n.in_p, H, rows, n.xb, n.in_p, n.dz1b, H);
    a.C = n.G + n.L[0].w;
    launch<true, false, kStore>(n.st, a);
    // bias gradients and loss
    FoldArgs f{};
    for (int l = 0; l < 3; ++l) {
        f.colpart[l] = n.colpart[l];
        f.gb[l] = n.G + n.L[l].b;
        f.cols[l] = n.L[l].out;
    }
    f.tiles = row_tiles(rows);
    f.lossp = n.lossp;
    f.nloss = row_tiles(rows) * ((O + kBN - 1) / kBN);
    f.scale = 1.0 / ((double)rows * n.out);
    f.loss = loss;
    hipLaunchKernelGGL(fold_kernel, dim3((unsigned)((std::max(H, O) + 255) / 256), 4), dim3(256), 0, n.st, f);
    HIP_OK(hipGetLastError());
}

void adam_step(Net &n, float lr) {
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    ++n.step;
    const float c1 = 1.0f - std::pow(b1, (float)n.step), c2 = 1.0f - std::pow(b2, (float)n.step);
    hipLaunchKernelGGL(adam_kernel, dim3(ew_grid((long long)n.np)), dim3(256), 0, n.st, n.P, n.G, n.M1, n.M2,
                       (long long)n.np, lr, b1, b2, c1, c2, eps, n.PB);
    HIP_OK(hipGetLastError());
}

void refresh_bf16(Net &n) {
    hipLaunchKernelGGL(to_bf16_kernel, dim3(ew_grid((long long)n.np)), dim3(256), 0, n.st, n.P, n.PB, (long long)n.np);
    HIP_OK(hipGetLastError());
}

}  // namespace mlp
}  // namespace pdhcg

using namespace pdhcg;
using namespace pdhcg::mlp;

struct pdhcg_mlp {
    Net net;
};

namespace {
// Parameter order of get/set and gradient output: W1, b1, W2, b2, W3, b3
// (unpadded; W row-major in x out).
size_t param_count(const Net &n) {
    return (size_t)n.in * n.hidden + n.hidden + (size_t)n.hidden * n.hidden + n.hidden + (size_t)n.hidden * n.out +
           n.out;
}

// Copies between the packed unpadded host layout and the padded flat device array.
void transfer(Net &n, float *host, bool to_host, bool grads) {
    const int dims[3][2] = {{n.in, n.hidden}, {n.hidden, n.hidden}, {n.hidden, n.out}};
    float *dev = grads ? n.G : n.P;
    std::vector<float> flat(n.np, 0.0f);
    if (to_host) {
        HIP_OK(hipMemcpyAsync(flat.data(), dev, n.np * sizeof(float), hipMemcpyDeviceToHost, n.st));
        HIP_OK(hipStreamSynchronize(n.st));
    }
    size_t off = 0;
    for (int l = 0; l < 3; ++l) {
        const Layer &L = n.L[l];
        const int rin = dims[l][0], rout = dims[l][1];
        for (int r = 0; r < rin; ++r)
            for (int c = 0; c < rout; ++c) {
                float &d = flat[L.w + (size_t)r * L.out + c];
                float &h = host[off + (size_t)r * rout + c];
                if (to_host) h = d; else d = h;
            }
        off += (size_t)rin * rout;
        for (int c = 0; c < rout; ++c) {
            float &d = flat[L.b + c];
            float &h = host[off + c];
            if (to_host) h = d; else d = h;
        }
        off += rout;
    }
    if (!to_host) {
        HIP_OK(hipMemcpyAsync(dev, flat.data(), n.np * sizeof(float), hipMemcpyHostToDevice, n.st));
        if (!grads) refresh_bf16(n);
        HIP_OK(hipStreamSynchronize(n.st));
    }
}

template <class F> int guarded(const char *name, F f) {
    try {
        return f();
    } catch (const std::exception &e) {
        std::fprintf(stderr, "[interface] %s: %s\n", name, e.what());
        return -1;
    }
}
}  // namespace

extern "C" {

pdhcg_mlp_t *pdhcg_mlp_create(int in_dim, int hidden, int out_dim, unsigned seed) {
    if (in_dim <= 0 || hidden <= 0 || out_dim <= 0 || hidden % 16 != 0) {
        std::fprintf(stderr, "[interface] pdhcg_mlp_create: invalid sizes (hidden must be a multiple of 16).\n");
        return nullptr;
    }
    pdhcg_mlp_t *m = new pdhcg_mlp_t();
    Net &n = m->net;
    if (guarded("pdhcg_mlp_create", [&] {
            n.in = in_dim;
            n.hidden = hidden;
            n.out = out_dim;
            n.in_p = pad16(in_dim);
            n.out_p = pad16(out_dim);
            HIP_OK(hipStreamCreateWithFlags(&n.st, hipStreamNonBlocking));
            const int ins[3] = {n.in_p, hidden, hidden}, outs[3] = {hidden, hidden, n.out_p};
            size_t off = 0;
            for (int l = 0; l < 3; ++l) {
                // every segment size is a multiple of 16, so every segment is 64-byte aligned
                n.L[l] = Layer{ins[l], outs[l], off, off + (size_t)ins[l] * outs[l]};
                off += (size_t)ins[l] * outs[l] + outs[l];
            }
            n.np = off;
            n.P = dnew<float>(n.np);
            n.G = dnew<float>(n.np);
            n.M1 = dnew<float>(n.np);
            n.M2 = dnew<float>(n.np);
            n.PB = dnew<bf16>(n.np);
            // He-style normal initialisation on the real (unpadded) blocks, zero biases
            std::vector<float> p(n.np, 0.0f);
            std::mt19937_64 rng(seed);
            const int fan_in[3] = {in_dim, hidden, hidden};
            const int rin[3] = {in_dim, hidden, hidden}, rout[3] = {hidden, hidden, out_dim};
            for (int l = 0; l < 3; ++l) {
                std::normal_distribution<float> nd(0.0f, std::sqrt(2.0f / (float)fan_in[l]));
                for (int r = 0; r < rin[l]; ++r)
                    for (int c = 0; c < rout[l]; ++c) p[n.L[l].w + (size_t)r * n.L[l].out + c] = nd(rng);
            }
            HIP_OK(hipMemcpy(n.P, p.data(), n.np * sizeof(float), hipMemcpyHostToDevice));
            refresh_bf16(n);
            HIP_OK(hipStreamSynchronize(n.st));
            return 0;
        }) != 0) {
        pdhcg_mlp_free(m);
        return nullptr;
    }
    return m;
}

int pdhcg_mlp_num_params(const pdhcg_mlp_t *m) { return m ? (int)param_count(m->net) : -1; }

int pdhcg_mlp_get_params(pdhcg_mlp_t *m, float *params) {
    if (!m || !params) return -1;
    return guarded("pdhcg_mlp_get_params", [&] {
        transfer(m->net, params, true, false);
        return 0;
    });
}

int pdhcg_mlp_set_params(pdhcg_mlp_t *m, const float *params) {
    if (!m || !params) return -1;
    return guarded("pdhcg_mlp_set_params", [&] {
        transfer(m->net, const_cast<float *>(params), false, false);
        return 0;
    });
}

int pdhcg_mlp_loss_grad(pdhcg_mlp_t *m, const float *X, const float *T, int rows, double *loss, float *grads) {
    if (!m || !X || !T || rows <= 0 || rows % 16 != 0) {
        std::fprintf(stderr, "[interface] pdhcg_mlp_loss_grad: invalid arguments (rows must be a multiple of 16).\n");
        return -1;
    }
    return guarded("pdhcg_mlp_loss_grad", [&] {
        Net &n = m->net;
        ensure_batch(n, rows);
        stage(n, X, T, rows);
        gather(n, 0, rows, rows, true);
        forward(n, rows, true);
        backward(n, rows, n.loss);
        double l = 0.0;
        HIP_OK(hipMemcpyAsync(&l, n.loss, sizeof(double), hipMemcpyDeviceToHost, n.st));
        HIP_OK(hipStreamSynchronize(n.st));
        if (loss) *loss = l;
        if (grads) transfer(n, grads, true, true);
        return 0;
    });
}

int pdhcg_mlp_train(pdhcg_mlp_t *m, const float *X, const float *T, int count, int epochs, int batch, float lr,
                    unsigned seed, double *final_loss) {
    if (!m || !X || !T || count < batch || batch <= 0 || batch % 16 != 0 || epochs < 0) {
        std::fprintf(stderr, "[interface] pdhcg_mlp_train: invalid arguments (batch a multiple of 16, count >= batch).\n");
        return -1;
    }
    return guarded("pdhcg_mlp_train", [&] {
        Net &n = m->net;
        if (count / batch > kMaxSteps) fail("more than %d minibatches per epoch", kMaxSteps);
        ensure_batch(n, batch);
        stage(n, X, T, count);
        std::vector<int> order((size_t)count);
        for (int i = 0; i < count; ++i) order[i] = i;
        std::mt19937_64 rng(seed);
        std::vector<double> losses;
        double last = 0.0;
        for (int e = 0; e < epochs; ++e) {
            std::shuffle(order.begin(), order.end(), rng);
            HIP_OK(hipMemcpyAsync(n.idx, order.data(), order.size() * sizeof(int), hipMemcpyHostToDevice, n.st));
            int nb = 0;
            for (int s = 0; s + batch <= count; s += batch, ++nb) {
                gather(n, s, batch, batch, true);
                forward(n, batch, true);
                backward(n, batch, n.loss + nb);
                adam_step(n, lr);
            }
            losses.resize((size_t)nb);
            HIP_OK(hipMemcpyAsync(losses.data(), n.loss, (size_t)nb * sizeof(double), hipMemcpyDeviceToHost, n.st));
            HIP_OK(hipStreamSynchronize(n.st));  // also keeps `order` alive until its copy is done
            double sum = 0.0;
            for (double v : losses) sum += v;
            last = nb ? sum / nb : 0.0;
        }
        if (final_loss) *final_loss = last;
        return 0;
    });
}

int pdhcg_mlp_predict(pdhcg_mlp_t *m, const float *X, int count, float *Y) {
    if (!m || !X || !Y || count < 0) return -1;
    return guarded("pdhcg_mlp_predict", [&] {
        Net &n = m->net;
        if (count == 0) return 0;
        const int chunk = 1024;
        ensure_batch(n, chunk);
        stage(n, X, nullptr, count);
        std::vector<float> yb((size_t)chunk * n.out_p);
        for (int s = 0; s < count; s += chunk) {
            const int rows = std::min(chunk, count - s), padded = pad16(rows);
            gather(n, s, padded, rows, false);
            forward(n, padded, false);
            HIP_OK(hipMemcpyAsync(yb.data(), n.y, (size_t)padded * n.out_p * sizeof(float), hipMemcpyDeviceToHost,
                                  n.st));
            HIP_OK(hipStreamSynchronize(n.st));
            for (int r = 0; r < rows; ++r)
                std::memcpy(Y + (size_t)(s + r) * n.out, &yb[(size_t)r * n.out_p], n.out * sizeof(float));
        }
        return 0;
    });
}

// Product C = op(A) op(B) of FP32 host matrices through the BF16 matrix-core
// kernel (operands rounded to BF16, FP32 accumulation); for tests of the kernel.
int pdhcg_mfma_gemm(int ta, int tb, int M, int N, int K, const float *A, const float *B, float *C) {
    if (M <= 0 || N <= 0 || K <= 0 || M % 16 || N % 16 || K % 16 || !A || !B || !C) {
        std::fprintf(stderr, "[interface] pdhcg_mfma_gemm: sizes must be positive multiples of 16.\n");
        return -1;
    }
    return guarded("pdhcg_mfma_gemm", [&] {
        const size_t na = (size_t)M * K, nb = (size_t)K * N, nc = (size_t)M * N;
        // released on every exit, including a HIP error part way through
        std::vector<void *> held;
        struct Release {
            std::vector<void *> &h;
            ~Release() {
                for (void *p : h) (void)hipFree(p);
            }
        } release{held};
        auto hold = [&](auto *p) {
            held.push_back((void *)p);
            return p;
        };
        float *fa = hold(dnew<float>(na)), *fb = hold(dnew<float>(nb)), *fc = hold(dnew<float>(nc));
        bf16 *ba = hold(dnew<bf16>(na)), *bb = hold(dnew<bf16>(nb));
        HIP_OK(hipMemcpy(fa, A, na * sizeof(float), hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(fb, B, nb * sizeof(float), hipMemcpyHostToDevice));
        hipLaunchKernelGGL(to_bf16_kernel, dim3(ew_grid((long long)na)), dim3(256), 0, 0, fa, ba, (long long)na);
        hipLaunchKernelGGL(to_bf16_kernel, dim3(ew_grid((long long)nb)), dim3(256), 0, 0, fb, bb, (long long)nb);
        // row-major storage: op(A) is M x K, so A is K x M when transposed
        GemmArgs a = gemm_args(M, N, K, ba, ta ? M : K, bb, tb ? K : N);
        a.C = fc;
        if (!ta && !tb) launch<false, false, kStore>(0, a);
        else if (!ta && tb) launch<false, true, kStore>(0, a);
        else if (ta && !tb) launch<true, false, kStore>(0, a);
        else launch<true, true, kStore>(0, a);
        HIP_OK(hipMemcpy(C, fc, nc * sizeof(float), hipMemcpyDeviceToHost));
        return 0;
    });
}

void pdhcg_mlp_free(pdhcg_mlp_t *m) {
    if (!m) return;
    Net &n = m->net;
    if (n.st) (void)hipStreamSynchronize(n.st);
    free_batch(n);
    for (void *p : {(void *)n.P, (void *)n.G, (void *)n.M1, (void *)n.M2, (void *)n.PB, (void *)n.dX, (void *)n.dT,
                    (void *)n.idx})
        if (p) (void)hipFree(p);
    if (n.st) (void)hipStreamDestroy(n.st);
    delete m;
}

}  // extern "C"
