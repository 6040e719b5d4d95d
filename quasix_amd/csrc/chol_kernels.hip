// Batched FP64 factorisation of the dielectric matrices eps_f = I - Pi(i w_f).
//
// Reference: quasix_core/src/selfenergy/correlation_ac.rs:376-401 inverts eps with
// LAPACK DPOTRF+DPOTRI (``invc``), falls back to LU (``inv``), then to the identity.
// Here the same chain runs on the GPU for all frequencies at once:
//   * right-looking blocked Cholesky, block NB = 64: a one-workgroup diagonal factor that
//     also inverts the diagonal block, a panel product with that inverse, and a batched
//     MFMA rank-64 update of the trailing lower triangle;
//   * X = L^-1 by block forward substitution (one launch per block row, MFMA tiles);
//   * Gauss-Jordan with partial pivoting (one workgroup per matrix) for any frequency
//     whose Cholesky hit a non-positive pivot; a singular matrix gives Wc = 0.
// The downstream contraction needs only X (Wc_nm = |X b|^2 - |b|^2), so DPOTRI's
// explicit inverse is formed only for the stage API.
#include "qx_internal.h"

namespace qx {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
constexpr int P = NB + 1;   // LDS pitch (doubles): row-strided MFMA reads stay conflict-free

__device__ __forceinline__ d4 mfma64(double a, double b, d4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ void load_block(double (*dst)[P], const double* __restrict__ src,
                                           int ld, int tid) {
    for (int idx = tid; idx < NB * NB; idx += 256) {
        const int r = idx / NB, c = idx % NB;
        dst[r][c] = src[(size_t)r * ld + c];
    }
}

// acc (this wave: rows 16w..16w+15, all 64 cols) += A[r][q] * Bm[q][c]
//   a_rowmajor: A stored [r][q];  b_kmajor: Bm stored [q][c] (if transB, Bm[q][c] = S[c][q])
template <bool TransB>
__device__ __forceinline__ void mma_64x64(d4 (&acc)[4], const double (*A)[P],
                                          const double (*S)[P], int wave, int lane) {
    const int r = wave * 16 + (lane & 15);
    #pragma unroll
    for (int kk = 0; kk < NB; kk += 4) {
        const int q = kk + (lane >> 4);
        const double a = A[r][q];
        #pragma unroll
        for (int bj = 0; bj < 4; ++bj) {
            const int c = bj * 16 + (lane & 15);
            const double b = TransB ? S[c][q] : S[q][c];
            acc[bj] = mfma64(a, b, acc[bj]);
        }
    }
}

template <typename F>
__device__ __forceinline__ void for_acc(const d4 (&acc)[4], int wave, int lane, F fn) {
    #pragma unroll
    for (int bj = 0; bj < 4; ++bj)
        #pragma unroll
        for (int v = 0; v < 4; ++v)
            fn(wave * 16 + mfma_row_f64(lane, v), bj * 16 + (lane & 15), acc[bj][v]);
}

__global__ __launch_bounds__(256) void potrf_diag_kernel(double* __restrict__ a,
                                                         double* __restrict__ linv_diag,
                                                         int npad, int j, int* __restrict__ fail) {
    __shared__ double L[NB][P];
    __shared__ double Li[NB][P];
    __shared__ int bad;
    const int f = blockIdx.x, tid = threadIdx.x;
    const int nblk = npad / NB;
    double* A = a + (size_t)f * npad * npad + (size_t)j * NB * npad + (size_t)j * NB;
    if (tid == 0) bad = 0;
    for (int idx = tid; idx < NB * NB; idx += 256) {
        const int r = idx / NB, c = idx % NB;
        L[r][c] = c <= r ? A[(size_t)r * npad + c] : 0.0;
    }
    __syncthreads();
    for (int k = 0; k < NB; ++k) {
        if (tid == 0) {
            double d = L[k][k];
            if (!(d > 0.0)) { bad = 1; d = 1.0; }
            L[k][k] = sqrt(d);
        }
        __syncthreads();
        if (tid > k && tid < NB) L[tid][k] /= L[k][k];
        __syncthreads();
        for (int idx = tid; idx < NB * NB; idx += 256) {
            const int r = idx / NB, c = idx % NB;
            if (c > k && c <= r) L[r][c] -= L[r][k] * L[c][k];
        }
        __syncthreads();
    }
    // Inverse of the lower-triangular block: thread c owns column c.
    if (tid < NB) {
        const int c = tid;
        for (int r = 0; r < c; ++r) Li[r][c] = 0.0;
        Li[c][c] = 1.0 / L[c][c];
        for (int r = c + 1; r < NB; ++r) {
            double sum = 0.0;
            for (int k = c; k < r; ++k) sum += L[r][k] * Li[k][c];
            Li[r][c] = -sum / L[r][r];
        }
    }
    __syncthreads();
    double* D = linv_diag + ((size_t)f * nblk + j) * NB * NB;
    for (int idx = tid; idx < NB * NB; idx += 256) {
        const int r = idx / NB, c = idx % NB;
        A[(size_t)r * npad + c] = c <= r ? L[r][c] : 0.0;
        D[idx] = Li[r][c];
    }
    if (tid == 0 && bad) fail[f] = 1;
}

// L[i][j] = A[i][j] * Linv_jj^T for the block rows below the diagonal block j.
__global__ __launch_bounds__(256) void potrf_panel_kernel(double* __restrict__ a,
                                                          const double* __restrict__ linv_diag,
                                                          int npad, int j) {
    __shared__ double As[NB][P];
    __shared__ double Ls[NB][P];
    const int i = j + 1 + blockIdx.x, f = blockIdx.y;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int nblk = npad / NB;
    double* A = a + (size_t)f * npad * npad + (size_t)i * NB * npad + (size_t)j * NB;
    load_block(As, A, npad, tid);
    load_block(Ls, linv_diag + ((size_t)f * nblk + j) * NB * NB, NB, tid);
    __syncthreads();
    d4 acc[4] = {d4{0, 0, 0, 0}, d4{0, 0, 0, 0}, d4{0, 0, 0, 0}, d4{0, 0, 0, 0}};
    mma_64x64<true>(acc, As, Ls, wave, lane);
    for_acc(acc, wave, lane, [&](int r, int c, double v) { A[(size_t)r * npad + c] = v; });
}

// A[i][k] -= L[i][j] L[k][j]^T for all trailing lower-triangle blocks j < k <= i.
__global__ __launch_bounds__(256) void potrf_update_kernel(double* __restrict__ a, int npad,
                                                           int j) {
    __shared__ double Pi_[NB][P];
    __shared__ double Pk[NB][P];
    const int f = blockIdx.y, t = blockIdx.x;
    int ii = (int)((sqrtf(8.0f * (float)t + 1.0f) - 1.0f) * 0.5f);
    while ((ii + 1) * (ii + 2) / 2 <= t) ++ii;
    while (ii * (ii + 1) / 2 > t) --ii;
    const int kk = t - ii * (ii + 1) / 2;
    const int i = j + 1 + ii, k = j + 1 + kk;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    double* base = a + (size_t)f * npad * npad;
    load_block(Pi_, base + (size_t)i * NB * npad + (size_t)j * NB, npad, tid);
    load_block(Pk, base + (size_t)k * NB * npad + (size_t)j * NB, npad, tid);
    __syncthreads();
    d4 acc[4] = {d4{0, 0, 0, 0}, d4{0, 0, 0, 0}, d4{0, 0, 0, 0}, d4{0, 0, 0, 0}};
    mma_64x64<true>(acc, Pi_, Pk, wave, lane);
    double* C = base + (size_t)i * NB * npad + (size_t)k * NB;
    for_acc(acc, wave, lane, [&](int r, int c, double v) { C[(size_t)r * npad + c] -= v; });
}

// Block row i of X = L^-1:  X_ii = Linv_ii,  X_ij = -Linv_ii * sum_{j<=k<i} L_ik X_kj.
__global__ __launch_bounds__(256) void trtri_row_kernel(const double* __restrict__ l,
                                                        const double* __restrict__ linv_diag,
                                                        double* __restrict__ x, int npad, int i) {
    __shared__ double As[NB][P];
    __shared__ double Bs[NB][P];
    const int jb = blockIdx.x, f = blockIdx.y;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int nblk = npad / NB;
    const double* L = l + (size_t)f * npad * npad;
    double* X = x + (size_t)f * npad * npad;
    const double* Dii = linv_diag + ((size_t)f * nblk + i) * NB * NB;
    double* out = X + (size_t)i * NB * npad + (size_t)jb * NB;
    if (jb == i) {
        for (int idx = tid; idx < NB * NB; idx += 256)
            out[(size_t)(idx / NB) * npad + idx % NB] = Dii[idx];
        return;
    }
    d4 acc[4] = {d4{0, 0, 0, 0}, d4{0, 0, 0, 0}, d4{0, 0, 0, 0}, d4{0, 0, 0, 0}};
    for (int k = jb; k < i; ++k) {
        load_block(As, L + (size_t)i * NB * npad + (size_t)k * NB, npad, tid);
        load_block(Bs, X + (size_t)k * NB * npad + (size_t)jb * NB, npad, tid);
        __syncthreads();
        mma_64x64<false>(acc, As, Bs, wave, lane);
        __syncthreads();
    }
    // S -> Bs, Linv_ii -> As, then X_ij = -Linv_ii S
    for_acc(acc, wave, lane, [&](int r, int c, double v) { Bs[r][c] = v; });
    load_block(As, Dii, NB, tid);
    __syncthreads();
    d4 acc2[4] = {d4{0, 0, 0, 0}, d4{0, 0, 0, 0}, d4{0, 0, 0, 0}, d4{0, 0, 0, 0}};
    mma_64x64<false>(acc2, As, Bs, wave, lane);
    for_acc(acc2, wave, lane, [&](int r, int c, double v) { out[(size_t)r * npad + c] = -v; });
}

// Wc = X^T X - I over the first naux rows/cols, full symmetric output [nf][naux][naux].
__global__ __launch_bounds__(256) void wc_from_linv_kernel(const double* __restrict__ x,
                                                           double* __restrict__ wc, int npad,
                                                           int naux) {
    __shared__ double Xa[16][17];
    __shared__ double Xb[16][17];
    const int f = blockIdx.z;
    const int p0 = blockIdx.y * 16, q0 = blockIdx.x * 16;
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    const double* X = x + (size_t)f * npad * npad;
    double sum = 0.0;
    const int rstart = (p0 > q0 ? p0 : q0) / 16 * 16;
    for (int r0 = rstart; r0 < npad; r0 += 16) {
        Xa[ty][tx] = X[(size_t)(r0 + ty) * npad + p0 + tx];
        Xb[ty][tx] = X[(size_t)(r0 + ty) * npad + q0 + tx];
        __syncthreads();
        for (int r = 0; r < 16; ++r) sum += Xa[r][ty] * Xb[r][tx];
        __syncthreads();
    }
    const int p = p0 + ty, q = q0 + tx;
    if (p < naux && q < naux) wc[((size_t)f * naux + p) * naux + q] = sum - (p == q ? 1.0 : 0.0);
}

// Gauss-Jordan inverse with partial pivoting on [A | I] (n x 2n), one workgroup per matrix.
__global__ __launch_bounds__(1024) void gauss_jordan_kernel(double* __restrict__ work, int n,
                                                            const int* __restrict__ active,
                                                            int* __restrict__ singular) {
    const int f = blockIdx.x;
    if (active && !active[f]) return;
    double* W = work + (size_t)f * n * 2 * n;
    const int ld = 2 * n, tid = threadIdx.x;
    __shared__ double red_v[1024];
    __shared__ int red_i[1024];
    __shared__ int is_sing;
    if (tid == 0) is_sing = 0;
    __syncthreads();
    for (int k = 0; k < n; ++k) {
        double best = -1.0;
        int bi = k;
        for (int r = k + tid; r < n; r += blockDim.x) {
            const double v = fabs(W[(size_t)r * ld + k]);
            if (v > best) { best = v; bi = r; }
        }
        red_v[tid] = best;
        red_i[tid] = bi;
        __syncthreads();
        for (int s = blockDim.x / 2; s > 0; s >>= 1) {
            if (tid < s && (red_v[tid + s] > red_v[tid] ||
                            (red_v[tid + s] == red_v[tid] && red_i[tid + s] < red_i[tid]))) {
                red_v[tid] = red_v[tid + s];
                red_i[tid] = red_i[tid + s];
            }
            __syncthreads();
        }
        const int piv = red_i[0];
        const double pv = red_v[0];
        __syncthreads();
        if (!(pv > 1e-300)) {
            if (tid == 0) is_sing = 1;
            break;
        }
        if (piv != k)
            for (int c = tid; c < ld; c += blockDim.x) {
                const double t = W[(size_t)k * ld + c];
                W[(size_t)k * ld + c] = W[(size_t)piv * ld + c];
                W[(size_t)piv * ld + c] = t;
            }
        __syncthreads();
        const double inv = 1.0 / W[(size_t)k * ld + k];
        __syncthreads();
        for (int c = tid; c < ld; c += blockDim.x) W[(size_t)k * ld + c] *= inv;
        __syncthreads();
        for (int idx = tid; idx < n * ld; idx += blockDim.x) {
            const int r = idx / ld, c = idx % ld;
            if (r != k && c != k) W[idx] -= W[(size_t)r * ld + k] * W[(size_t)k * ld + c];
        }
        __syncthreads();
        for (int r = tid; r < n; r += blockDim.x)
            if (r != k) W[(size_t)r * ld + k] = 0.0;
        __syncthreads();
    }
    __syncthreads();
    if (tid == 0 && singular) singular[f] = is_sing;
}

__global__ void gj_setup_kernel(const double* __restrict__ src, int is_eps,
                                double* __restrict__ work, int n, const int* __restrict__ active) {
    const int f = blockIdx.y;
    if (active && !active[f]) return;
    const double* S = src + (size_t)f * n * n;
    double* W = work + (size_t)f * n * 2 * n;
    for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < (size_t)n * n;
         idx += (size_t)gridDim.x * blockDim.x) {
        const int r = idx / n, c = idx % n;
        const double v = S[idx];
        W[(size_t)r * 2 * n + c] = is_eps ? v : (r == c ? 1.0 : 0.0) - v;
        W[(size_t)r * 2 * n + n + c] = r == c ? 1.0 : 0.0;
    }
}

__global__ void gj_finish_kernel(const double* __restrict__ work, double* __restrict__ wc, int n,
                                 const int* __restrict__ active, const int* __restrict__ singular) {
    const int f = blockIdx.y;
    if (active && !active[f]) return;
    const double* W = work + (size_t)f * n * 2 * n;
    double* O = wc + (size_t)f * n * n;
    const bool sing = singular[f] != 0;
    for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < (size_t)n * n;
         idx += (size_t)gridDim.x * blockDim.x) {
        const int r = idx / n, c = idx % n;
        O[idx] = sing ? 0.0 : W[(size_t)r * 2 * n + n + c] - (r == c ? 1.0 : 0.0);
    }
}

}  // namespace

hipError_t launch_potrf_batched(double* a, double* linv_diag, int npad, int nf, int* fail,
                                hipStream_t st) {
    const int nblk = npad / NB;
    for (int j = 0; j < nblk; ++j) {
        hipLaunchKernelGGL(potrf_diag_kernel, dim3(nf), dim3(256), 0, st, a, linv_diag, npad, j,
                           fail);
        const int rem = nblk - j - 1;
        if (rem > 0) {
            hipLaunchKernelGGL(potrf_panel_kernel, dim3(rem, nf), dim3(256), 0, st, a, linv_diag,
                               npad, j);
            hipLaunchKernelGGL(potrf_update_kernel, dim3(rem * (rem + 1) / 2, nf), dim3(256), 0,
                               st, a, npad, j);
        }
    }
    return hipGetLastError();
}

hipError_t launch_trtri_batched(const double* l, const double* linv_diag, double* x, int npad,
                                int nf, hipStream_t st) {
    const int nblk = npad / NB;
    for (int i = 0; i < nblk; ++i)
        hipLaunchKernelGGL(trtri_row_kernel, dim3(i + 1, nf), dim3(256), 0, st, l, linv_diag, x,
                           npad, i);
    return hipGetLastError();
}

hipError_t launch_wc_from_linv(const double* x, double* wc, int npad, int naux, int nf,
                               hipStream_t st) {
    dim3 grid(ceil_div(naux, 16), ceil_div(naux, 16), nf);
    hipLaunchKernelGGL(wc_from_linv_kernel, grid, dim3(256), 0, st, x, wc, npad, naux);
    return hipGetLastError();
}

// src: [nf][n][n] (either Pi or eps), work: [nf][n][2n], wc: [nf][n][n]
hipError_t launch_lu_fallback_wc(const double* src, int is_eps, double* work, double* wc, int n,
                                 int nf, const int* active, hipStream_t st) {
    // singular flags live in the tail of the work buffer
    int* sing = reinterpret_cast<int*>(work + (size_t)nf * n * 2 * n);
    dim3 g(ceil_div(n * n, 256) < 4096 ? ceil_div(n * n, 256) : 4096, nf);
    hipLaunchKernelGGL(gj_setup_kernel, g, dim3(256), 0, st, src, is_eps, work, n, active);
    hipLaunchKernelGGL(gauss_jordan_kernel, dim3(nf), dim3(1024), 0, st, work, n, active, sing);
    hipLaunchKernelGGL(gj_finish_kernel, g, dim3(256), 0, st, work, wc, n, active, sing);
    return hipGetLastError();
}

}  // namespace qx
