// Kernels of the stage entry points that take the reference's intermediate arrays (a given
// P0, the full DF tensor lpq, the Coulomb metric or its square root) instead of building
// them from B_ov: pair rows of lpq, the static CD quadrature, and the elementwise steps of
// the metric-screened W = v^1/2 (I - v^1/2 P0 v^1/2T)^-1 v^1/2T.  All FP64; the GEMMs and
// the factorisation are the MFMA kernels of adjoint_kernels.hip and chol_kernels.hip.
#include "qx_internal.h"

#include <algorithm>

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

// Reductions, one instance per mode.  With the mode as a run-time argument the max |a| form
// returned wrong values on gfx950 (checked element-wise against torch; the generated code
// skipped the |a| term of the three-way selection), so the mode is a template parameter.
// MODE 0: sum a^2; 1: max |a - b|; 2: max |a|.  A NaN term becomes +inf so that the max
// forms (fmax drops NaN) report it.
template <int MODE>
__device__ __forceinline__ double reduce_term(const double* a, const double* b, size_t i) {
    double t;
    if constexpr (MODE == 0) t = a[i] * a[i];
    else if constexpr (MODE == 1) t = fabs(a[i] - b[i]);
    else t = fabs(a[i]);
    return t != t ? INFINITY : t;
}

template <int MODE>
__device__ __forceinline__ double reduce_op(double x, double y) {
    if constexpr (MODE == 0) return x + y;
    else return fmax(x, y);
}

template <int MODE>
__device__ __forceinline__ double block_reduce(double acc, double* red) {
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] = reduce_op<MODE>(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    return red[0];
}

// Pass 1: each of kRedBlocks workgroups reduces a strided share into part[blockIdx.x].
template <int MODE>
__global__ __launch_bounds__(256) void reduce_part_kernel(const double* __restrict__ a,
                                                          const double* __restrict__ b,
                                                          size_t count,
                                                          double* __restrict__ part) {
    __shared__ double red[256];
    double acc = 0.0;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < count;
         i += (size_t)gridDim.x * blockDim.x)
        acc = reduce_op<MODE>(acc, reduce_term<MODE>(a, b, i));
    const double r = block_reduce<MODE>(acc, red);
    if (threadIdx.x == 0) part[blockIdx.x] = r;
}

// Pass 2: one workgroup combines the partials into out[0].
template <int MODE>
__global__ __launch_bounds__(256) void reduce_final_kernel(const double* __restrict__ part,
                                                           int nparts, double* __restrict__ out) {
    __shared__ double red[256];
    double acc = 0.0;
    for (int i = threadIdx.x; i < nparts; i += blockDim.x) acc = reduce_op<MODE>(acc, part[i]);
    const double r = block_reduce<MODE>(acc, red);
    if (threadIdx.x == 0) out[0] = r;
}

template <int MODE>
hipError_t reduce_mode(const double* a, const double* b, size_t count, double* out,
                       hipStream_t st) {
    // out[1 .. kRedBlocks] hold the partials of the first pass
    hipLaunchKernelGGL(reduce_part_kernel<MODE>, dim3(kRedBlocks), dim3(256), 0, st, a, b, count,
                       out + 1);
    hipLaunchKernelGGL(reduce_final_kernel<MODE>, dim3(1), dim3(256), 0, st, out + 1, kRedBlocks,
                       out);
    return hipGetLastError();
}

// ---- Assembly of the symmetrised DF tensor from the iaP / miP / abP blocks -----------------
// X[n][m][:] of the caller's blocks (quasix/src/evgw.rs:161-186): miP [nmo][nocc] for
// m < nocc, iaP [nocc][nvir] for n < nocc <= m, abP [nvir][nvir] for both virtual; the
// symmetrised tensor is (X[n][m] + X[m][n]) / 2 (evgw.rs:188-197).
struct Blocks {
    const double* ia;
    const double* mi;
    const double* ab;
    int nocc, nvir, naux;
};

__device__ __forceinline__ const double* xrow(const Blocks& b, int n, int m) {
    if (m < b.nocc) return b.mi + ((size_t)n * b.nocc + m) * b.naux;
    if (n < b.nocc) return b.ia + ((size_t)n * b.nvir + (m - b.nocc)) * b.naux;
    return b.ab + ((size_t)(n - b.nocc) * b.nvir + (m - b.nocc)) * b.naux;
}

// flags[t * nmo + m] = 1 when the symmetrised row (targets[t], m) has a non-zero entry.
// One wavefront per row; the two source rows are contiguous over the auxiliary index.
__global__ __launch_bounds__(256) void pair_flags_kernel(Blocks b, const int* __restrict__ targets,
                                                         int nt, int* __restrict__ flags) {
    const int nmo = b.nocc + b.nvir;
    const int lane = threadIdx.x & 63;
    const long long rows = (long long)nt * nmo;
    for (long long r = blockIdx.x * 4ll + (threadIdx.x >> 6); r < rows; r += gridDim.x * 4ll) {
        const int t = (int)(r / nmo), m = (int)(r % nmo), n = targets[t];
        const double* x1 = xrow(b, n, m);
        const double* x2 = xrow(b, m, n);
        bool nz = false;
        for (int p = lane; p < b.naux; p += 64) nz |= 0.5 * (x1[p] + x2[p]) != 0.0;
        nz = __any(nz);
        if (lane == 0) flags[r] = nz ? 1 : 0;
    }
}

// Exclusive prefix sum of flags[0..count) into pos, the total into total[0]; one workgroup.
__global__ __launch_bounds__(1024) void pair_scan_kernel(const int* __restrict__ flags,
                                                         long long count, int* __restrict__ pos,
                                                         int* __restrict__ total) {
    __shared__ int part[1024];
    const int tid = threadIdx.x;
    const long long per = (count + 1023) / 1024;
    const long long lo = tid * per, hi = lo + per < count ? lo + per : count;
    int s = 0;
    for (long long i = lo; i < hi; ++i) s += flags[i];
    part[tid] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {          // inclusive scan of the chunk sums
        const int v = tid >= off ? part[tid - off] : 0;
        __syncthreads();
        part[tid] += v;
        __syncthreads();
    }
    int run = part[tid] - s;
    for (long long i = lo; i < hi; ++i) {
        pos[i] = run;
        run += flags[i];
    }
    if (tid == 1023) total[0] = part[1023];
}

// Writes every flagged row, in (target, m) order, to rows[pos] with its target slot and m.
__global__ __launch_bounds__(256) void pair_scatter_kernel(Blocks b, const int* __restrict__ targets,
                                                           int nt, const int* __restrict__ flags,
                                                           const int* __restrict__ pos,
                                                           double* __restrict__ out_rows,
                                                           int* __restrict__ pt, int* __restrict__ pm) {
    const int nmo = b.nocc + b.nvir;
    const int lane = threadIdx.x & 63;
    const long long rows = (long long)nt * nmo;
    for (long long r = blockIdx.x * 4ll + (threadIdx.x >> 6); r < rows; r += gridDim.x * 4ll) {
        if (!flags[r]) continue;
        const int t = (int)(r / nmo), m = (int)(r % nmo), n = targets[t];
        const double* x1 = xrow(b, n, m);
        const double* x2 = xrow(b, m, n);
        const int j = pos[r];
        double* o = out_rows + (size_t)j * b.naux;
        for (int p = lane; p < b.naux; p += 64) o[p] = 0.5 * (x1[p] + x2[p]);
        if (lane == 0) {
            pt[j] = t;
            pm[j] = m;
        }
    }
}

// ov block of the symmetrised tensor, [nocc * nvir][naux] (the polarizability's operand), and
// its [nmo][nocc][naux] block (the exchange term's df_mi, quasix/src/evgw.rs:305).
__global__ void sym_blocks_kernel(Blocks b, double* __restrict__ ov, double* __restrict__ mi) {
    const int nmo = b.nocc + b.nvir;
    const size_t n_ov = ov ? (size_t)b.nocc * b.nvir * b.naux : 0;
    const size_t n_mi = mi ? (size_t)nmo * b.nocc * b.naux : 0;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n_ov + n_mi;
         i += (size_t)gridDim.x * blockDim.x) {
        if (i < n_ov) {
            const size_t r = i / b.naux;
            const int p = (int)(i % b.naux), occ = (int)(r / b.nvir), a = (int)(r % b.nvir);
            ov[i] = 0.5 * (xrow(b, occ, b.nocc + a)[p] + xrow(b, b.nocc + a, occ)[p]);
        } else {
            const size_t k = i - n_ov, r = k / b.naux;
            const int p = (int)(k % b.naux), m = (int)(r / b.nocc), occ = (int)(r % b.nocc);
            mi[k] = 0.5 * (xrow(b, m, occ)[p] + xrow(b, occ, m)[p]);
        }
    }
}

}  // namespace

hipError_t launch_assemble_pairs(const double* ia, const double* mi, const double* ab, int nocc,
                                 int nvir, int naux, const int* targets, int nt, int* flags,
                                 int* pos, int* total, hipStream_t st) {
    const Blocks b{ia, mi, ab, nocc, nvir, naux};
    const long long rows = (long long)nt * (nocc + nvir);
    if (rows <= 0) return hipErrorInvalidValue;
    const unsigned g = (unsigned)std::min<long long>((rows + 3) / 4, 65535);
    hipLaunchKernelGGL(pair_flags_kernel, dim3(g), dim3(256), 0, st, b, targets, nt, flags);
    hipLaunchKernelGGL(pair_scan_kernel, dim3(1), dim3(1024), 0, st, flags, rows, pos, total);
    return hipGetLastError();
}

hipError_t launch_scatter_pairs(const double* ia, const double* mi, const double* ab, int nocc,
                                int nvir, int naux, const int* targets, int nt, const int* flags,
                                const int* pos, double* rows_out, int* pair_t, int* pair_m,
                                hipStream_t st) {
    const Blocks b{ia, mi, ab, nocc, nvir, naux};
    const long long rows = (long long)nt * (nocc + nvir);
    const unsigned g = (unsigned)std::min<long long>((rows + 3) / 4, 65535);
    hipLaunchKernelGGL(pair_scatter_kernel, dim3(g), dim3(256), 0, st, b, targets, nt, flags, pos,
                       rows_out, pair_t, pair_m);
    return hipGetLastError();
}

hipError_t launch_sym_blocks(const double* ia, const double* mi, const double* ab, int nocc,
                             int nvir, int naux, double* ov_sym, double* mi_sym, hipStream_t st) {
    const Blocks b{ia, mi, ab, nocc, nvir, naux};
    size_t n = 0;
    if (ov_sym) n += (size_t)nocc * nvir * naux;
    if (mi_sym) n += (size_t)(nocc + nvir) * nocc * naux;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(sym_blocks_kernel, grid_for(n), dim3(256), 0, st, b, ov_sym, mi_sym);
    return hipGetLastError();
}

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
    if (mode == 0) return reduce_mode<0>(a, b, count, out, st);
    if (mode == 1) return reduce_mode<1>(a, b, count, out, st);
    if (mode == 2) return reduce_mode<2>(a, b, count, out, st);
    return hipErrorInvalidValue;
}

hipError_t launch_axpby(double alpha, const double* x, double beta, double* y, size_t count,
                        hipStream_t st) {
    hipLaunchKernelGGL(axpby_kernel, grid_for(count), dim3(256), 0, st, alpha, x, beta, y, count);
    return hipGetLastError();
}

}  // namespace qx
