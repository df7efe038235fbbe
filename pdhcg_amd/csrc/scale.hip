// Matrix passes of the problem scaling on the device (reference
// src/preconditioner.c:497-652 for the passes, src/solver_state.cu:735-777 for
// the transpose): the stable transpose of [A;F], the row and column maxima
// (Ruiz) and power sums (Pock-Chambolle) of |a|, and the divisions and
// multiplications that apply the factors to A, its transpose, Q and R.  The
// factor vectors themselves (square roots, cone-block merging, bound and
// objective norms) stay on the host in host_scale.cpp, so only m + n doubles
// cross PCIe per pass.
//
// Every entry gets the host pass's arithmetic in the host pass's order: the
// divisions and multiplications are IEEE and per entry; maxima do not depend
// on order; the power sums add a row's (or a column's) terms sequentially in
// storage order, rows in increasing order within a column (the stable
// transpose), as the host loops do.  The scaled matrices are therefore bit for
// bit the host path's.  pow(|a|, 1) is taken as |a| itself (the default
// exponent alpha = 1); other exponents call the device pow, whose last bit may
// differ from the C library's.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <memory>
#include <vector>

#include "dev_common.hpp"
#include "host_scale.h"

namespace pdhcg {
namespace {

// r-th row index of every stored entry of a CSR matrix
// Row of every entry: a binary search of rp per entry (rp[lo] <= e <
// rp[lo + 1]), so a row of any length costs no more than a short one.
__global__ void expand_rows_kernel(const int *rp, int rows, int nnz, int *row_of) {
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < nnz; e += gridDim.x * blockDim.x) {
        int lo = 0, hi = rows;  // rp[lo] <= e < rp[hi]
        while (hi - lo > 1) {
            const int mid = lo + (hi - lo) / 2;
            if (rp[mid] <= e) lo = mid;
            else hi = mid;
        }
        row_of[e] = lo;
    }
}

__global__ void iota_kernel(int *v, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) v[i] = i;
}

__global__ void count_kernel(const int *keys, int n, int *counts) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        atomicAdd(counts + keys[i], 1);
}

// transpose entry p takes source entry src[p]: its row becomes the column
__global__ void gather_transpose_kernel(const int *src, const int *row_of, const double *v, int n, int *tci,
                                        double *tv) {
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
        const int e = src[p];
        tci[p] = row_of[e];
        tv[p] = v[e];
    }
}

// maximum of |v| over each row, one wave per row (order-free)
__global__ void row_absmax_kernel(const int *rp, const double *v, int rows, double *out) {
    const int lane = threadIdx.x % kWave;
    const int waves = gridDim.x * (blockDim.x / kWave);
    for (int r = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave; r < rows; r += waves) {
        double mx = 0.0;
        for (int e = rp[r] + lane; e < rp[r + 1]; e += kWave) mx = fmax(mx, fabs(v[e]));
        for (int off = kWave / 2; off > 0; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off, kWave));
        if (lane == 0) out[r] = mx;
    }
}

// sum of |v|^p over each row, added sequentially in storage order
__global__ void row_powsum_kernel(const int *rp, const double *v, int rows, double p, double *out) {
    for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < rows; r += gridDim.x * blockDim.x) {
        double s = 0.0;
        for (int e = rp[r]; e < rp[r + 1]; ++e) {
            const double a = fabs(v[e]);
            s += p == 1.0 ? a : pow(a, p);
        }
        out[r] = s;
    }
}

// v[e] /= (rs[r] * cs[ci[e]]) over a CSR matrix (rows x cols); rs null: v[e] /= cs[ci[e]]
__global__ void divide_kernel(const int *row_of, const int *ci, double *v, int n, const double *rs,
                              const double *cs) {
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x)
        v[e] = rs ? v[e] / (rs[row_of[e]] * cs[ci[e]]) : v[e] / cs[ci[e]];
}

// the transpose's entries: v[p] /= (cr[ci[p]] * vr[row]) with the operands of
// the product in the order of the source entry's (cr[row of A] * vr[col of A])
__global__ void divide_transpose_kernel(const int *row_of, const int *ci, double *v, int n, const double *cr,
                                        const double *vr) {
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x)
        v[p] = v[p] / (cr[ci[p]] * vr[row_of[p]]);
}

__global__ void multiply_kernel(double *v, long long n, double f) {
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x)
        v[e] *= f;
}

inline int grid_for(long long n) { return (int)std::max<long long>(1, std::min<long long>((n + 255) / 256, 8192)); }

template <class T> T *dev_alloc(size_t count) {
    T *p = nullptr;
    HIP_OK(hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T)));
    return p;
}
template <class T> T *dev_upload(const std::vector<T> &h) {
    T *p = dev_alloc<T>(h.size());
    if (!h.empty()) HIP_OK(hipMemcpy(p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
    return p;
}
template <class T> void dev_download(std::vector<T> &h, const T *p) {
    if (!h.empty()) HIP_OK(hipMemcpy(h.data(), p, h.size() * sizeof(T), hipMemcpyDeviceToHost));
}

// One CSR matrix on the device with the row index of each entry.
struct DevCsr {
    int rows = 0, cols = 0, nnz = 0;
    int *rp = nullptr, *ci = nullptr, *row_of = nullptr;
    double *v = nullptr;
    void load(const HostCSR &h) {
        rows = h.rows;
        cols = h.cols;
        nnz = h.nnz();
        rp = dev_upload(h.rp);
        ci = dev_upload(h.ci);
        v = dev_upload(h.v);
        row_of = dev_alloc<int>((size_t)nnz);
        if (rows > 0 && nnz > 0)
            hipLaunchKernelGGL(expand_rows_kernel, dim3(grid_for(nnz)), dim3(256), 0, 0, rp, rows, nnz, row_of);
        HIP_OK(hipGetLastError());
    }
    void release() {
        for (void *p : {(void *)rp, (void *)ci, (void *)row_of, (void *)v})
            if (p) (void)hipFree(p);
        *this = DevCsr();
    }
};

class DevicePasses final : public MatrixPasses {
  public:
    ~DevicePasses() override {
        for (DevCsr *m : {&A_, &At_, &Q_, &R_}) m->release();
        for (double *p : {cr_, vr_, wr_, wc_})
            if (p) (void)hipFree(p);
    }

    void load(const ScaledProblem &w) override {
        A_.load(w.A);
        Q_.load(w.Q);
        R_.load(w.R);
        transpose_A();
        cr_ = dev_alloc<double>((size_t)A_.rows);
        vr_ = dev_alloc<double>((size_t)A_.cols);
        wr_ = dev_alloc<double>((size_t)A_.rows);
        wc_ = dev_alloc<double>((size_t)A_.cols);
        HIP_OK(hipDeviceSynchronize());
    }

    void abs_max(std::vector<double> &row, std::vector<double> &col) override {
        launch_rows(row_absmax_kernel, A_, wr_);
        launch_rows(row_absmax_kernel, At_, wc_);
        HIP_OK(hipGetLastError());
        dev_download(row, wr_);
        dev_download(col, wc_);
    }

    void pow_sums(double alpha, std::vector<double> &row, std::vector<double> &col) override {
        if (A_.rows > 0)
            hipLaunchKernelGGL(row_powsum_kernel, dim3(grid_for(A_.rows)), dim3(256), 0, 0, A_.rp, A_.v, A_.rows, alpha,
                               wr_);
        if (At_.rows > 0)
            hipLaunchKernelGGL(row_powsum_kernel, dim3(grid_for(At_.rows)), dim3(256), 0, 0, At_.rp, At_.v, At_.rows,
                               2.0 - alpha, wc_);
        HIP_OK(hipGetLastError());
        dev_download(row, wr_);
        dev_download(col, wc_);
    }

    void divide(const std::vector<double> &cr, const std::vector<double> &vr) override {
        HIP_OK(hipMemcpy(cr_, cr.data(), cr.size() * sizeof(double), hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(vr_, vr.data(), vr.size() * sizeof(double), hipMemcpyHostToDevice));
        if (A_.nnz) {
            hipLaunchKernelGGL(divide_kernel, dim3(grid_for(A_.nnz)), dim3(256), 0, 0, A_.row_of, A_.ci, A_.v, A_.nnz,
                               (const double *)cr_, (const double *)vr_);
            hipLaunchKernelGGL(divide_transpose_kernel, dim3(grid_for(At_.nnz)), dim3(256), 0, 0, At_.row_of, At_.ci,
                               At_.v, At_.nnz, (const double *)cr_, (const double *)vr_);
        }
        if (Q_.nnz)
            hipLaunchKernelGGL(divide_kernel, dim3(grid_for(Q_.nnz)), dim3(256), 0, 0, Q_.row_of, Q_.ci, Q_.v, Q_.nnz,
                               (const double *)vr_, (const double *)vr_);
        if (R_.nnz)
            hipLaunchKernelGGL(divide_kernel, dim3(grid_for(R_.nnz)), dim3(256), 0, 0, R_.row_of, R_.ci, R_.v, R_.nnz,
                               (const double *)nullptr, (const double *)vr_);
        HIP_OK(hipGetLastError());
    }

    void multiply_qr(double fq, double fr) override {
        if (Q_.nnz) hipLaunchKernelGGL(multiply_kernel, dim3(grid_for(Q_.nnz)), dim3(256), 0, 0, Q_.v, (long long)Q_.nnz, fq);
        if (R_.nnz) hipLaunchKernelGGL(multiply_kernel, dim3(grid_for(R_.nnz)), dim3(256), 0, 0, R_.v, (long long)R_.nnz, fr);
        HIP_OK(hipGetLastError());
    }

    void finish(ScaledProblem &w) override {
        HIP_OK(hipDeviceSynchronize());
        dev_download(w.A.v, A_.v);
        dev_download(w.Q.v, Q_.v);
        dev_download(w.R.v, R_.v);
        w.At.rows = At_.rows;
        w.At.cols = At_.cols;
        w.At.rp.resize((size_t)At_.rows + 1);
        w.At.ci.resize((size_t)At_.nnz);
        w.At.v.resize((size_t)At_.nnz);
        dev_download(w.At.rp, At_.rp);
        dev_download(w.At.ci, At_.ci);
        dev_download(w.At.v, At_.v);
    }

  private:
    DevCsr A_, At_, Q_, R_;
    double *cr_ = nullptr, *vr_ = nullptr, *wr_ = nullptr, *wc_ = nullptr;

    static void launch_rows(void (*k)(const int *, const double *, int, double *), const DevCsr &M, double *out) {
        if (M.rows == 0) return;
        hipLaunchKernelGGL(k, dim3(grid_for((long long)M.rows * kWave)), dim3(256), 0, 0, (const int *)M.rp,
                           (const double *)M.v, M.rows, out);
    }

    // Stable transpose (rows of the result list source rows in increasing
    // order): a stable radix sort of the entries by column keeps row-major
    // order among equal columns.
    void transpose_A() {
        At_.rows = A_.cols;
        At_.cols = A_.rows;
        At_.nnz = A_.nnz;
        At_.rp = dev_alloc<int>((size_t)At_.rows + 1);
        At_.ci = dev_alloc<int>((size_t)At_.nnz);
        At_.v = dev_alloc<double>((size_t)At_.nnz);
        At_.row_of = dev_alloc<int>((size_t)At_.nnz);
        const int nnz = A_.nnz;
        int *keys_out = dev_alloc<int>((size_t)nnz), *idx_in = dev_alloc<int>((size_t)nnz),
            *idx_out = dev_alloc<int>((size_t)nnz), *counts = dev_alloc<int>((size_t)At_.rows + 1);
        HIP_OK(hipMemset(counts, 0, ((size_t)At_.rows + 1) * sizeof(int)));
        if (nnz > 0) {
            hipLaunchKernelGGL(iota_kernel, dim3(grid_for(nnz)), dim3(256), 0, 0, idx_in, nnz);
            hipLaunchKernelGGL(count_kernel, dim3(grid_for(nnz)), dim3(256), 0, 0, (const int *)A_.ci, nnz, counts);
            HIP_OK(hipGetLastError());
            int bits = 1;
            while (bits < 31 && (1 << bits) < At_.rows) ++bits;
            size_t tmp_bytes = 0;
            HIP_OK(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, A_.ci, keys_out, idx_in, idx_out, nnz, 0, bits));
            void *tmp = dev_alloc<char>(tmp_bytes);
            HIP_OK(hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, A_.ci, keys_out, idx_in, idx_out, nnz, 0, bits));
            (void)hipFree(tmp);
            hipLaunchKernelGGL(gather_transpose_kernel, dim3(grid_for(nnz)), dim3(256), 0, 0, (const int *)idx_out,
                               (const int *)A_.row_of, (const double *)A_.v, nnz, At_.ci, At_.v);
            HIP_OK(hipGetLastError());
        }
        size_t scan_bytes = 0;
        HIP_OK(hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, counts, At_.rp, At_.rows + 1));
        void *scan_tmp = dev_alloc<char>(scan_bytes);
        HIP_OK(hipcub::DeviceScan::ExclusiveSum(scan_tmp, scan_bytes, counts, At_.rp, At_.rows + 1));
        (void)hipFree(scan_tmp);
        if (At_.rows > 0 && nnz > 0)
            hipLaunchKernelGGL(expand_rows_kernel, dim3(grid_for(nnz)), dim3(256), 0, 0, (const int *)At_.rp,
                               At_.rows, nnz, At_.row_of);
        HIP_OK(hipGetLastError());
        HIP_OK(hipDeviceSynchronize());
        for (int *p : {keys_out, idx_in, idx_out, counts}) (void)hipFree(p);
    }
};

}  // namespace

std::unique_ptr<MatrixPasses> make_device_passes() { return std::unique_ptr<MatrixPasses>(new DevicePasses()); }

}  // namespace pdhcg
