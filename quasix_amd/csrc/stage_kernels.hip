// Kernels of the stage entry points that take the reference's intermediate arrays (a given
// P0, the full DF tensor lpq, the Coulomb metric or its square root) instead of building
// them from B_ov: pair rows of lpq, the static CD quadrature, and the elementwise steps of
// the metric-screened W = v^1/2 (I - v^1/2 P0 v^1/2T)^-1 v^1/2T.  All FP64; the GEMMs and
// the factorisation are the MFMA kernels of adjoint_kernels.hip and chol_kernels.hip.
#include "qx_internal.h"

namespace qx {
namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr int kRedBlocks = 1024;

dim3 grid_for(size_t n) {
    size_t b = (n + 255) / 256;
    if (b > 16384) b = 16384;
    if (b < 1) b = 1;
    return dim3((unsigned)b);
}

__global__ void lpq_pair_rows_kernel(const double* __restrict__ lpq, int naux, int nmo,
                                     int npad, double* __restrict__ u, double* __restrict__ v,
                                     int* __restrict__ any_v) {
    const size_t total = (size_t)nmo * nmo * npad;
    bool nz = false;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total;
         i += (size_t)gridDim.x * blockDim.x) {
        const size_t j = i / npad;
        const int p = (int)(i % npad);
        const int n = (int)(j / nmo), m = (int)(j % nmo);
        double a = 0.0, b = 0.0;
        if (p < naux) {
            a = lpq[((size_t)p * nmo + n) * nmo + m];
            b = lpq[((size_t)p * nmo + m) * nmo + n];
        }
        const double d = 0.5 * (a - b);
        u[i] = 0.5 * (a + b);
        v[i] = d;
        nz |= d != 0.0;
    }
    if (nz) any_v[0] = 1;      // every writer stores the same value
}

__global__ void set_int_kernel(int* p, int v) {
    if (blockIdx.x == 0 && threadIdx.x == 0) p[0] = v;
}

// One workgroup per target n; the (f, m) terms are split over the threads and summed in a
// fixed order (thread partials, then a tree over the workgroup).
__global__ __launch_bounds__(256) void sigma_static_cd_kernel(
    const double* __restrict__ wu, const double* __restrict__ wv, int nmo, int nf,
    const double* __restrict__ eps, const double* __restrict__ freqs,
    const double* __restrict__ wts, double eta, double* __restrict__ sigma) {
    __shared__ double red[2][256];
    const int n = blockIdx.x;
    const double en = eps[n];
    const size_t nn = (size_t)nmo * nmo;
    double re = 0.0, im = 0.0;
    for (int idx = threadIdx.x; idx < nmo * nf; idx += blockDim.x) {
        const int f = idx / nmo, m = idx % nmo;
        const size_t j = (size_t)f * nn + (size_t)n * nmo + m;
        const double w = wu[j] - (wv ? wv[j] : 0.0);
        const double s = en > eps[m] ? 1.0 : -1.0;
        const double a = -eps[m], b = -eta * s;           // emo = a + i b
        const double dr = a * a - b * b + freqs[f] * freqs[f], di = 2.0 * a * b;
        const double den = dr * dr + di * di;             // emo / (emo^2 + w^2)
        const double gr = (a * dr + b * di) / den, gi = (b * dr - a * di) / den;
        re += wts[f] * w * gr;
        im += wts[f] * w * gi;
    }
    red[0][threadIdx.x] = re;
    red[1][threadIdx.x] = im;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            red[0][threadIdx.x] += red[0][threadIdx.x + s];
            red[1][threadIdx.x] += red[1][threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        sigma[2 * n] = -red[0][0] / kPi;
        sigma[2 * n + 1] = -red[1][0] / kPi;
    }
}

__global__ void pad_square_kernel(const double* __restrict__ src, size_t src_stride, int n,
                                  int npad, int batch, int diag_pad, double* __restrict__ dst) {
    const size_t per = (size_t)npad * npad, total = per * batch;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total;
         i += (size_t)gridDim.x * blockDim.x) {
        const size_t b = i / per, rc = i % per;
        const int r = (int)(rc / npad), c = (int)(rc % npad);
        double v = 0.0;
        if (src && r < n && c < n) v = src[b * src_stride + (size_t)r * n + c];
        else if (diag_pad && r == c) v = 1.0;
        dst[i] = v;
    }
}

__global__ void eps_sym_kernel(const double* __restrict__ M, int npad, int batch,
                               double* __restrict__ eps) {
    const size_t per = (size_t)npad * npad, total = per * batch;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total;
         i += (size_t)gridDim.x * blockDim.x) {
        const size_t b = i / per, rc = i % per;
        const int r = (int)(rc / npad), c = (int)(rc % npad);
        const double* Mb = M + b * per;
        eps[i] = (r == c ? 1.0 : 0.0) - 0.5 * (Mb[(size_t)r * npad + c] + Mb[(size_t)c * npad + r]);
    }
}

__global__ void sym_compact_kernel(const double* __restrict__ T, int npad, int n, int batch,
                                   double* __restrict__ W) {
    const size_t per = (size_t)n * n, total = per * batch;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total;
         i += (size_t)gridDim.x * blockDim.x) {
        const size_t b = i / per, rc = i % per;
        const int r = (int)(rc / n), c = (int)(rc % n);
        const double* Tb = T + b * (size_t)npad * npad;
        W[i] = 0.5 * (Tb[(size_t)r * npad + c] + Tb[(size_t)c * npad + r]);
    }
}

__global__ void add_identity_kernel(double* __restrict__ a, int npad, int n, int batch,
                                    double alpha) {
    const size_t total = (size_t)n * batch;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total;
         i += (size_t)gridDim.x * blockDim.x) {
        const size_t b = i / n;
        const int r = (int)(i % n);
        a[b * (size_t)npad * npad + (size_t)r * npad + r] += alpha;
    }
}

__global__ void axpby_kernel(double alpha, const double* __restrict__ x, double beta,
                             double* __restrict__ y, size_t count) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < count;
         i += (size_t)gridDim.x * blockDim.x)
        y[i] = (x ? alpha * x[i] : 0.0) + beta * y[i];
}

__device__ double reduce_term(const double* a, const double* b, size_t i, int mode) {
    if (mode == 0) return a[i] * a[i];
    if (mode == 1) return fabs(a[i] - b[i]);
    return fabs(a[i]);
}

// Pass 1: each of kRedBlocks workgroups reduces a strided share into part[blockIdx.x].
__global__ __launch_bounds__(256) void reduce_part_kernel(const double* __restrict__ a,
                                                          const double* __restrict__ b,
                                                          size_t count, int mode,
                                                          double* __restrict__ part) {
    __shared__ double red[256];
    double acc = 0.0;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < count;
         i += (size_t)gridDim.x * blockDim.x) {
        const double t = reduce_term(a, b, i, mode);
        acc = mode == 0 ? acc + t : fmax(acc, t);
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            red[threadIdx.x] = mode == 0 ? red[threadIdx.x] + red[threadIdx.x + s]
                                         : fmax(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// Pass 2: one workgroup combines the partials into out[0].
__global__ __launch_bounds__(256) void reduce_final_kernel(const double* __restrict__ part,
                                                           int nparts, int mode,
                                                           double* __restrict__ out) {
    __shared__ double red[256];
    double acc = 0.0;
    for (int i = threadIdx.x; i < nparts; i += blockDim.x)
        acc = mode == 0 ? acc + part[i] : fmax(acc, part[i]);
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            red[threadIdx.x] = mode == 0 ? red[threadIdx.x] + red[threadIdx.x + s]
                                         : fmax(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0];
}

}  // namespace

hipError_t launch_lpq_pair_rows(const double* lpq, int naux, int nmo, int npad, double* u,
                                double* v, int* any_v, hipStream_t st) {
    hipLaunchKernelGGL(set_int_kernel, dim3(1), dim3(64), 0, st, any_v, 0);
    const size_t total = (size_t)nmo * nmo * npad;
    hipLaunchKernelGGL(lpq_pair_rows_kernel, grid_for(total), dim3(256), 0, st, lpq, naux, nmo,
                       npad, u, v, any_v);
    return hipGetLastError();
}

hipError_t launch_sigma_static_cd(const double* wu, const double* wv, int nmo, int nf,
                                  const double* eps, const double* freqs, const double* wts,
                                  double eta, double* sigma, hipStream_t st) {
    if (nmo <= 0) return hipSuccess;
    hipLaunchKernelGGL(sigma_static_cd_kernel, dim3(nmo), dim3(256), 0, st, wu, wv, nmo, nf, eps,
                       freqs, wts, eta, sigma);
    return hipGetLastError();
}

hipError_t launch_pad_square(const double* src, size_t src_stride, int n, int npad, int batch,
                             bool diag_pad, double* dst, hipStream_t st) {
    hipLaunchKernelGGL(pad_square_kernel, grid_for((size_t)npad * npad * batch), dim3(256), 0,
                       st, src, src_stride, n, npad, batch, diag_pad ? 1 : 0, dst);
    return hipGetLastError();
}

hipError_t launch_eps_sym(const double* M, int npad, int batch, double* eps, hipStream_t st) {
    hipLaunchKernelGGL(eps_sym_kernel, grid_for((size_t)npad * npad * batch), dim3(256), 0, st,
                       M, npad, batch, eps);
    return hipGetLastError();
}

hipError_t launch_sym_compact(const double* T, int npad, int n, int batch, double* W,
                              hipStream_t st) {
    hipLaunchKernelGGL(sym_compact_kernel, grid_for((size_t)n * n * batch), dim3(256), 0, st, T,
                       npad, n, batch, W);
    return hipGetLastError();
}

hipError_t launch_add_identity(double* a, int npad, int n, int batch, double alpha,
                               hipStream_t st) {
    hipLaunchKernelGGL(add_identity_kernel, grid_for((size_t)n * batch), dim3(256), 0, st, a,
                       npad, n, batch, alpha);
    return hipGetLastError();
}

hipError_t launch_reduce(const double* a, const double* b, size_t count, int mode, double* out,
                         hipStream_t st) {
    // out[1 .. kRedBlocks] hold the partials of the first pass
    hipLaunchKernelGGL(reduce_part_kernel, dim3(kRedBlocks), dim3(256), 0, st, a, b, count, mode,
                       out + 1);
    hipLaunchKernelGGL(reduce_final_kernel, dim3(1), dim3(256), 0, st, out + 1, kRedBlocks, mode,
                       out);
    return hipGetLastError();
}

hipError_t launch_axpby(double alpha, const double* x, double beta, double* y, size_t count,
                        hipStream_t st) {
    hipLaunchKernelGGL(axpby_kernel, grid_for(count), dim3(256), 0, st, alpha, x, beta, y, count);
    return hipGetLastError();
}

}  // namespace qx
