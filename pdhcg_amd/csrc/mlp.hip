// Learned warm start: a 3-layer MLP (in -> H -> H -> out, GELU) that maps the
// parameters of a problem family to a primal/dual starting point.  Training
// and inference run on the device with hand-written BF16 MFMA matrix products
// (FP32 accumulation) and FP32 elementwise kernels (bias, GELU, loss, Adam).
//
// Matrix product: C[M x N] (FP32) = op(A)[M x K] op(B)[K x N] with BF16
// operands, op = identity or transpose of a row-major array.  A workgroup of
// four waves owns a 64 x 64 tile of C, each wave a 32 x 32 quarter made of
// 2 x 2 blocks of v_mfma_f32_16x16x16bf16: per lane, A contributes the 4
// values A[r][k0 + 4 (lane / 16) .. + 3] of row r = lane % 16, B the 4 values
// B[k0 + 4 (lane / 16) .. + 3][c] of column c = lane % 16, and the 16 x 16
// result holds C[4 (lane / 16) + i][lane % 16] in lane's i-th register.  M, N
// and K are multiples of 16 (the host pads); tiles past M or N are skipped
// per 16-block.
#include <hip/hip_runtime.h>

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
typedef short s4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

// float -> bf16 with round to nearest even (NaN kept quiet)
__host__ __device__ inline bf16 to_bf16(float f) {
    unsigned u;
    memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (bf16)((u >> 16) | 0x40);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (bf16)(u >> 16);
}
__host__ __device__ inline float from_bf16(bf16 h) {
    const unsigned u = (unsigned)h << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
}

constexpr int kTile = 64;

// Element (r, k) of op(A) and (k, c) of op(B) for row-major storage.
template <bool TA> __device__ inline bf16 a_at(const bf16 *A, int lda, int r, int k) {
    return TA ? A[(size_t)k * lda + r] : A[(size_t)r * lda + k];
}
template <bool TB> __device__ inline bf16 b_at(const bf16 *B, int ldb, int k, int c) {
    return TB ? B[(size_t)c * ldb + k] : B[(size_t)k * ldb + c];
}

// C = op(A) op(B) (accumulate: C += ...).  Grid: (N / 64, M / 64) rounded up.
template <bool TA, bool TB>
__global__ void __launch_bounds__(256) gemm_bf16_kernel(int M, int N, int K, const bf16 *A, int lda, const bf16 *B,
                                                       int ldb, float *C, int ldc, int accumulate) {
    const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
    const int r0 = blockIdx.y * kTile + (wave / 2) * 32, c0 = blockIdx.x * kTile + (wave % 2) * 32;
    const int lr = lane % 16, lk = 4 * (lane / 16);
    f4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};
    const bool rows_ok[2] = {r0 < M, r0 + 16 < M}, cols_ok[2] = {c0 < N, c0 + 16 < N};
    for (int k0 = 0; k0 < K; k0 += 16) {
        s4 a[2], b[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const int r = r0 + 16 * i + lr;
#pragma unroll
            for (int q = 0; q < 4; ++q) a[i][q] = rows_ok[i] ? (short)a_at<TA>(A, lda, r, k0 + lk + q) : (short)0;
        }
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int c = c0 + 16 * j + lr;
#pragma unroll
            for (int q = 0; q < 4; ++q) b[j][q] = cols_ok[j] ? (short)b_at<TB>(B, ldb, k0 + lk + q, c) : (short)0;
        }
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[i], b[j], acc[i][j], 0, 0, 0);
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            if (!rows_ok[i] || !cols_ok[j]) continue;
            const int c = c0 + 16 * j + lr;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int r = r0 + 16 * i + lk + q;
                float *p = C + (size_t)r * ldc + c;
                *p = accumulate ? *p + acc[i][j][q] : acc[i][j][q];
            }
        }
}

__global__ void to_bf16_kernel(const float *x, bf16 *y, long long n) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        y[i] = to_bf16(x[i]);
}

__device__ inline float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ inline float gelu_grad(float z) {
    return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) + z * 0.39894228040143268f * expf(-0.5f * z * z);
}

// Z += bias (row vector); optionally H = GELU(Z) and its BF16 copy.
__global__ void bias_act_kernel(float *Z, const float *bias, int rows, int cols, int act, float *H, bf16 *Hb) {
    const long long n = (long long)rows * cols;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const float z = Z[i] + bias[i % cols];
        Z[i] = z;
        const float h = act ? gelu(z) : z;
        if (H) H[i] = h;
        if (Hb) Hb[i] = to_bf16(h);
    }
}

// dY = 2 (Y - T) / (rows * out) on the first `out` columns (padding: 0);
// loss partial sums per workgroup into part[blockIdx.x].
__global__ void mse_grad_kernel(const float *Y, const float *T, int rows, int cols, int out, float *dY, bf16 *dYb,
                                double *part) {
    const long long n = (long long)rows * cols;
    const float sc = 2.0f / ((float)rows * (float)out);
    double acc = 0.0;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const int c = (int)(i % cols);
        float g = 0.0f;
        if (c < out) {
            const float d = Y[i] - T[(i / cols) * out + c];
            acc += (double)d * d;
            g = sc * d;
        }
        dY[i] = g;
        dYb[i] = to_bf16(g);
    }
    __shared__ double lds[16];
    double v[1] = {acc};
    block_reduce<1, 0>(v, lds);
    if (threadIdx.x == 0) part[blockIdx.x] = v[0];
}

// Sum of the nb loss partials (fixed order) times `scale` into *out.
__global__ void loss_fold_kernel(const double *part, int nb, double scale, double *out) {
    double acc = 0.0;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) acc += part[i];
    __shared__ double lds[16];
    double v[1] = {acc};
    block_reduce<1, 0>(v, lds);
    if (threadIdx.x == 0) *out = v[0] * scale;
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

// dZ = dH * GELU'(Z) (in place on dH), with its BF16 copy
__global__ void gelu_back_kernel(float *dH, const float *Z, long long n, bf16 *dZb) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const float g = dH[i] * gelu_grad(Z[i]);
        dH[i] = g;
        dZb[i] = to_bf16(g);
    }
}

// db[c] = sum_r dZ[r][c] (fixed order)
__global__ void colsum_kernel(const float *dZ, int rows, int cols, float *db) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= cols) return;
    float s = 0.0f;
    for (int r = 0; r < rows; ++r) s += dZ[(size_t)r * cols + c];
    db[c] = s;
}

// Adam (Kingma & Ba 2015) on FP32 master weights; refreshes the BF16 copy.
__global__ void adam_kernel(float *w, const float *g, float *m1, float *m2, long long n, float lr, float b1, float b2,
                            float c1, float c2, float eps, bf16 *wb) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float a = b1 * m1[i] + (1.0f - b1) * gi, b = b2 * m2[i] + (1.0f - b2) * gi * gi;
        m1[i] = a;
        m2[i] = b;
        w[i] -= lr * (a / c1) / (sqrtf(b / c2) + eps);
        if (wb) wb[i] = to_bf16(w[i]);
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

struct Layer {
    int in, out;             // padded sizes
    float *w, *b, *gw, *gb;  // FP32 weights (in x out row-major), bias, gradients
    float *mw, *vw, *mb, *vb;
    bf16 *wb;                // BF16 copy of w
};

struct Net {
    int in, hidden, out, in_p, out_p;
    Layer L[3];
    int step = 0;
    hipStream_t st = nullptr;
    // activations of the current batch (rows = batch)
    int cap = 0;
    bf16 *xb = nullptr, *h1b = nullptr, *h2b = nullptr, *dyb = nullptr, *dz2b = nullptr, *dz1b = nullptr;
    float *z1 = nullptr, *z2 = nullptr, *y = nullptr, *dy = nullptr, *dh2 = nullptr, *dh1 = nullptr, *t = nullptr;
    double *part = nullptr, *loss = nullptr;  // loss partials; per-minibatch losses (kMaxSteps)
    // device-resident samples
    size_t data_cap = 0;
    float *dX = nullptr, *dT = nullptr;
    int *idx = nullptr;
};

constexpr int kMaxSteps = 1 << 16;  // minibatches per epoch

void gemm(hipStream_t st, bool ta, bool tb, int M, int N, int K, const bf16 *A, int lda, const bf16 *B, int ldb,
          float *C, int ldc, int accumulate) {
    dim3 grid((unsigned)((N + kTile - 1) / kTile), (unsigned)((M + kTile - 1) / kTile));
    if (!ta && !tb)
        hipLaunchKernelGGL((gemm_bf16_kernel<false, false>), grid, dim3(256), 0, st, M, N, K, A, lda, B, ldb, C, ldc,
                           accumulate);
    else if (!ta && tb)
        hipLaunchKernelGGL((gemm_bf16_kernel<false, true>), grid, dim3(256), 0, st, M, N, K, A, lda, B, ldb, C, ldc,
                           accumulate);
    else if (ta && !tb)
        hipLaunchKernelGGL((gemm_bf16_kernel<true, false>), grid, dim3(256), 0, st, M, N, K, A, lda, B, ldb, C, ldc,
                           accumulate);
    else
        hipLaunchKernelGGL((gemm_bf16_kernel<true, true>), grid, dim3(256), 0, st, M, N, K, A, lda, B, ldb, C, ldc,
                           accumulate);
    HIP_OK(hipGetLastError());
}

void ensure_batch(Net &n, int rows) {
    if (rows <= n.cap) return;
    for (void *p : {(void *)n.xb, (void *)n.h1b, (void *)n.h2b, (void *)n.dyb, (void *)n.dz2b, (void *)n.dz1b,
                    (void *)n.z1, (void *)n.z2, (void *)n.y, (void *)n.dy, (void *)n.dh2, (void *)n.dh1, (void *)n.t,
                    (void *)n.part, (void *)n.loss})
        if (p) (void)hipFree(p);
    const size_t r = (size_t)rows, H = (size_t)n.hidden;
    n.xb = dnew<bf16>(r * n.in_p);
    n.z1 = dnew<float>(r * H);
    n.h1b = dnew<bf16>(r * H);
    n.z2 = dnew<float>(r * H);
    n.h2b = dnew<bf16>(r * H);
    n.y = dnew<float>(r * n.out_p);
    n.t = dnew<float>(r * n.out);
    n.dy = dnew<float>(r * n.out_p);
    n.dyb = dnew<bf16>(r * n.out_p);
    n.dh2 = dnew<float>(r * H);
    n.dz2b = dnew<bf16>(r * H);
    n.dh1 = dnew<float>(r * H);
    n.dz1b = dnew<bf16>(r * H);
    n.part = dnew<double>(4096);
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
}

// Forward pass of `rows` (multiple of 16) loaded rows: y = W3'GELU(W2'GELU(W1'x + b1) + b2) + b3.
void forward(Net &n, int rows) {
    const int H = n.hidden;
    gemm(n.st, false, false, rows, H, n.in_p, n.xb, n.in_p, n.L[0].wb, H, n.z1, H, 0);
    hipLaunchKernelGGL(bias_act_kernel, dim3(ew_grid((long long)rows * H)), dim3(256), 0, n.st, n.z1, n.L[0].b, rows, H,
                       1, (float *)nullptr, n.h1b);
    gemm(n.st, false, false, rows, H, H, n.h1b, H, n.L[1].wb, H, n.z2, H, 0);
    hipLaunchKernelGGL(bias_act_kernel, dim3(ew_grid((long long)rows * H)), dim3(256), 0, n.st, n.z2, n.L[1].b, rows, H,
                       1, (float *)nullptr, n.h2b);
    gemm(n.st, false, false, rows, n.out_p, H, n.h2b, H, n.L[2].wb, n.out_p, n.y, n.out_p, 0);
    hipLaunchKernelGGL(bias_act_kernel, dim3(ew_grid((long long)rows * n.out_p)), dim3(256), 0, n.st, n.y, n.L[2].b,
                       rows, n.out_p, 0, (float *)nullptr, (bf16 *)nullptr);
}

// Backward pass of the mean squared error against the gathered targets;
// fills the gradients and writes the loss to the device slot *loss.
void backward(Net &n, int rows, double *loss) {
    const int H = n.hidden, O = n.out_p;
    const int g = ew_grid((long long)rows * O);
    hipLaunchKernelGGL(mse_grad_kernel, dim3(g), dim3(256), 0, n.st, n.y, n.t, rows, O, n.out, n.dy, n.dyb, n.part);
    // layer 3: gW3 = H2' dY, gb3 = colsum dY, dH2 = dY W3'
    gemm(n.st, true, false, H, O, rows, n.h2b, H, n.dyb, O, n.L[2].gw, O, 0);
    hipLaunchKernelGGL(colsum_kernel, dim3((O + 255) / 256), dim3(256), 0, n.st, n.dy, rows, O, n.L[2].gb);
    gemm(n.st, false, true, rows, H, O, n.dyb, O, n.L[2].wb, O, n.dh2, H, 0);
    hipLaunchKernelGGL(gelu_back_kernel, dim3(ew_grid((long long)rows * H)), dim3(256), 0, n.st, n.dh2, n.z2,
                       (long long)rows * H, n.dz2b);
    // layer 2
    gemm(n.st, true, false, H, H, rows, n.h1b, H, n.dz2b, H, n.L[1].gw, H, 0);
    hipLaunchKernelGGL(colsum_kernel, dim3((H + 255) / 256), dim3(256), 0, n.st, n.dh2, rows, H, n.L[1].gb);
    gemm(n.st, false, true, rows, H, H, n.dz2b, H, n.L[1].wb, H, n.dh1, H, 0);
    hipLaunchKernelGGL(gelu_back_kernel, dim3(ew_grid((long long)rows * H)), dim3(256), 0, n.st, n.dh1, n.z1,
                       (long long)rows * H, n.dz1b);
    // layer 1
    gemm(n.st, true, false, n.in_p, H, rows, n.xb, n.in_p, n.dz1b, H, n.L[0].gw, H, 0);
    hipLaunchKernelGGL(colsum_kernel, dim3((H + 255) / 256), dim3(256), 0, n.st, n.dh1, rows, H, n.L[0].gb);
    hipLaunchKernelGGL(loss_fold_kernel, dim3(1), dim3(256), 0, n.st, n.part, g, 1.0 / ((double)rows * n.out), loss);
    HIP_OK(hipGetLastError());
}

void adam_step(Net &n, float lr) {
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    ++n.step;
    const float c1 = 1.0f - std::pow(b1, (float)n.step), c2 = 1.0f - std::pow(b2, (float)n.step);
    for (Layer &L : n.L) {
        const long long nw = (long long)L.in * L.out;
        hipLaunchKernelGGL(adam_kernel, dim3(ew_grid(nw)), dim3(256), 0, n.st, L.w, L.gw, L.mw, L.vw, nw, lr, b1, b2, c1,
                           c2, eps, L.wb);
        hipLaunchKernelGGL(adam_kernel, dim3(ew_grid(L.out)), dim3(256), 0, n.st, L.b, L.gb, L.mb, L.vb,
                           (long long)L.out, lr, b1, b2, c1, c2, eps, (bf16 *)nullptr);
    }
    HIP_OK(hipGetLastError());
}

void refresh_bf16(Net &n) {
    for (Layer &L : n.L) {
        const long long nw = (long long)L.in * L.out;
        hipLaunchKernelGGL(to_bf16_kernel, dim3(ew_grid(nw)), dim3(256), 0, n.st, L.w, L.wb, nw);
    }
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

// Copies between the packed unpadded host layout and the padded device arrays.
void transfer(Net &n, float *host, bool to_host, bool grads) {
    const int dims[3][2] = {{n.in, n.hidden}, {n.hidden, n.hidden}, {n.hidden, n.out}};
    size_t off = 0;
    for (int l = 0; l < 3; ++l) {
        Layer &L = n.L[l];
        const int rin = dims[l][0], rout = dims[l][1];
        std::vector<float> w((size_t)L.in * L.out, 0.0f), b((size_t)L.out, 0.0f);
        float *dw = grads ? L.gw : L.w, *db = grads ? L.gb : L.b;
        if (to_host) {
            HIP_OK(hipMemcpy(w.data(), dw, w.size() * sizeof(float), hipMemcpyDeviceToHost));
            HIP_OK(hipMemcpy(b.data(), db, b.size() * sizeof(float), hipMemcpyDeviceToHost));
            for (int r = 0; r < rin; ++r)
                for (int c = 0; c < rout; ++c) host[off + (size_t)r * rout + c] = w[(size_t)r * L.out + c];
            off += (size_t)rin * rout;
            for (int c = 0; c < rout; ++c) host[off + c] = b[c];
            off += rout;
        } else {
            for (int r = 0; r < rin; ++r)
                for (int c = 0; c < rout; ++c) w[(size_t)r * L.out + c] = host[off + (size_t)r * rout + c];
            off += (size_t)rin * rout;
            for (int c = 0; c < rout; ++c) b[c] = host[off + c];
            off += rout;
            HIP_OK(hipMemcpy(dw, w.data(), w.size() * sizeof(float), hipMemcpyHostToDevice));
            HIP_OK(hipMemcpy(db, b.data(), b.size() * sizeof(float), hipMemcpyHostToDevice));
        }
    }
    if (!to_host && !grads) refresh_bf16(n);
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
            const int fan_in[3] = {in_dim, hidden, hidden};
            std::mt19937_64 rng(seed);
            for (int l = 0; l < 3; ++l) {
                Layer &L = n.L[l];
                L.in = ins[l];
                L.out = outs[l];
                const size_t nw = (size_t)L.in * L.out;
                L.w = dnew<float>(nw);
                L.gw = dnew<float>(nw);
                L.mw = dnew<float>(nw);
                L.vw = dnew<float>(nw);
                L.wb = dnew<bf16>(nw);
                L.b = dnew<float>(L.out);
                L.gb = dnew<float>(L.out);
                L.mb = dnew<float>(L.out);
                L.vb = dnew<float>(L.out);
                // He-style normal initialisation on the real (unpadded) block
                std::normal_distribution<float> nd(0.0f, std::sqrt(2.0f / (float)fan_in[l]));
                std::vector<float> w(nw, 0.0f);
                const int rin = l == 0 ? in_dim : hidden, rout = l == 2 ? out_dim : hidden;
                for (int r = 0; r < rin; ++r)
                    for (int c = 0; c < rout; ++c) w[(size_t)r * L.out + c] = nd(rng);
                HIP_OK(hipMemcpy(L.w, w.data(), nw * sizeof(float), hipMemcpyHostToDevice));
            }
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
        HIP_OK(hipStreamSynchronize(m->net.st));
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
        forward(n, rows);
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
                forward(n, batch);
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
        HIP_OK(hipStreamSynchronize(n.st));
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
            forward(n, padded);
            HIP_OK(hipMemcpyAsync(yb.data(), n.y, (size_t)padded * n.out_p * sizeof(float), hipMemcpyDeviceToHost,
                                  n.st));
            HIP_OK(hipStreamSynchronize(n.st));
            for (int r = 0; r < rows; ++r)
                std::memcpy(Y + (size_t)(s + r) * n.out, &yb[(size_t)r * n.out_p], n.out * sizeof(float));
        }
        return 0;
    });
}

// Product C = op(A) op(B) of FP32 host matrices through the BF16 MFMA kernel
// (operands rounded to BF16, FP32 accumulation); for tests of the kernel.
int pdhcg_mfma_gemm(int ta, int tb, int M, int N, int K, const float *A, const float *B, float *C) {
    if (M <= 0 || N <= 0 || K <= 0 || M % 16 || N % 16 || K % 16 || !A || !B || !C) {
        std::fprintf(stderr, "[interface] pdhcg_mfma_gemm: sizes must be positive multiples of 16.\n");
        return -1;
    }
    return guarded("pdhcg_mfma_gemm", [&] {
        const size_t na = (size_t)M * K, nb = (size_t)K * N, nc = (size_t)M * N;
        float *fa = dnew<float>(na), *fb = dnew<float>(nb), *fc = dnew<float>(nc);
        bf16 *ba = dnew<bf16>(na), *bb = dnew<bf16>(nb);
        HIP_OK(hipMemcpy(fa, A, na * sizeof(float), hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(fb, B, nb * sizeof(float), hipMemcpyHostToDevice));
        hipLaunchKernelGGL(to_bf16_kernel, dim3(ew_grid((long long)na)), dim3(256), 0, 0, fa, ba, (long long)na);
        hipLaunchKernelGGL(to_bf16_kernel, dim3(ew_grid((long long)nb)), dim3(256), 0, 0, fb, bb, (long long)nb);
        // row-major storage: op(A) is M x K, so A is K x M when transposed
        gemm(0, ta != 0, tb != 0, M, N, K, ba, ta ? M : K, bb, tb ? K : N, fc, N, 0);
        HIP_OK(hipMemcpy(C, fc, nc * sizeof(float), hipMemcpyDeviceToHost));
        for (void *p : {(void *)fa, (void *)fb, (void *)fc, (void *)ba, (void *)bb}) (void)hipFree(p);
        return 0;
    });
}

void pdhcg_mlp_free(pdhcg_mlp_t *m) {
    if (!m) return;
    Net &n = m->net;
    if (n.st) (void)hipStreamSynchronize(n.st);
    for (Layer &L : n.L)
        for (void *p : {(void *)L.w, (void *)L.b, (void *)L.gw, (void *)L.gb, (void *)L.mw, (void *)L.vw, (void *)L.mb,
                        (void *)L.vb, (void *)L.wb})
            if (p) (void)hipFree(p);
    for (void *p : {(void *)n.xb, (void *)n.h1b, (void *)n.h2b, (void *)n.dyb, (void *)n.dz2b, (void *)n.dz1b,
                    (void *)n.z1, (void *)n.z2, (void *)n.y, (void *)n.dy, (void *)n.dh2, (void *)n.dh1, (void *)n.t,
                    (void *)n.part, (void *)n.loss, (void *)n.dX, (void *)n.dT, (void *)n.idx})
        if (p) (void)hipFree(p);
    if (n.st) (void)hipStreamDestroy(n.st);
    delete m;
}

}  // extern "C"
