// DF tensor construction: AO -> MO transform of Cholesky-normalised 3-centre vectors.
//
// Reference: quasix_core/src/df/mo_transform.rs:136-260 (transform_mo_3center:
// (ia|P) = sum_{mu nu} C_{mu i} (mu nu|P) C_{nu a}, output [nocc * nvir][naux]) and the
// driver's AO -> MO step quasix/python/gw_driver.py:584-624.  Two FP64 MFMA GEMM stages:
//   T[i][nu P]  = sum_mu C_occ[mu][i] J[mu][nu P]            (M = nocc, K = nao, N = nao naux)
//   B[i a][P]   = sum_nu C_vir[nu][a] T[i][nu][P]   for each i (M = nvir, K = nao, N = naux)
// with one batched kernel: C_b = A_b^T B_b, A_b [K][M] (the coefficient matrix as stored),
// B_b [K][N], C_b [M][N], all row-major with leading dimensions and batch strides.
#include "qx_internal.h"

namespace qx {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
constexpr int GT = 64, GKC = 16;   // output tile and K chunk

// One 64x64 tile of C_b = A_b^T B_b; 256 threads, wave w owns rows 16w..16w+15.
__global__ __launch_bounds__(256) void gemm_tn_f64_kernel(
    const double* __restrict__ A, long long lda, long long sa, const double* __restrict__ B,
    long long ldb, long long sb, double* __restrict__ C, long long ldc, long long sc, int M,
    int N, int K) {
    __shared__ double As[GKC][GT + 1];    // [k][m]
    __shared__ double Bs[GKC][GT + 1];    // [k][n]
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int n0 = blockIdx.x * GT, m0 = blockIdx.y * GT, b = blockIdx.z;
    const double* Ab = A + (size_t)b * sa;
    const double* Bb = B + (size_t)b * sb;
    double* Cb = C + (size_t)b * sc;
    // staging: 16 k x 64 columns per operand = 1024 doubles, 4 per thread (coalesced rows)
    double ra[4], rb[4];
    auto fetch = [&](int k0) {
        #pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int idx = tid + 256 * q, kk = idx / GT, cc = idx % GT;
            const int k = k0 + kk;
            ra[q] = (k < K && m0 + cc < M) ? Ab[(size_t)k * lda + m0 + cc] : 0.0;
            rb[q] = (k < K && n0 + cc < N) ? Bb[(size_t)k * ldb + n0 + cc] : 0.0;
        }
    };
    d4 acc[4] = {d4{0, 0, 0, 0}, d4{0, 0, 0, 0}, d4{0, 0, 0, 0}, d4{0, 0, 0, 0}};
    fetch(0);
    for (int k0 = 0; k0 < K; k0 += GKC) {
        #pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int idx = tid + 256 * q;
            As[idx / GT][idx % GT] = ra[q];
            Bs[idx / GT][idx % GT] = rb[q];
        }
        __syncthreads();
        if (k0 + GKC < K) fetch(k0 + GKC);
        const int r = wave * 16 + (lane & 15);
        #pragma unroll
        for (int kk = 0; kk < GKC; kk += 4) {
            const int q = kk + (lane >> 4);
            const double av = As[q][r];
            #pragma unroll
            for (int bj = 0; bj < 4; ++bj)
                acc[bj] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, Bs[q][bj * 16 + (lane & 15)],
                                                               acc[bj], 0, 0, 0);
        }
        __syncthreads();
    }
    #pragma unroll
    for (int bj = 0; bj < 4; ++bj)
        #pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int row = m0 + wave * 16 + mfma_row_f64(lane, v);
            const int col = n0 + bj * 16 + (lane & 15);
            if (row < M && col < N) Cb[(size_t)row * ldc + col] = acc[bj][v];
        }
}

hipError_t gemm_tn(const double* A, long long lda, long long sa, const double* B, long long ldb,
                   long long sb, double* C, long long ldc, long long sc, int M, int N, int K,
                   int batch, hipStream_t st) {
    const long long gx = (N + GT - 1) / GT, gy = (M + GT - 1) / GT;
    if (gx > 0x7fffffff || gy > 65535 || batch > 65535) return hipErrorInvalidValue;
    hipLaunchKernelGGL(gemm_tn_f64_kernel, dim3((unsigned)gx, (unsigned)gy, batch), dim3(256), 0,
                       st, A, lda, sa, B, ldb, sb, C, ldc, sc, M, N, K);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_mo_transform(const double* j3c, const double* c_occ, const double* c_vir,
                               int nao, int nocc, int nvir, int naux, double* tmp, double* out,
                               hipStream_t st) {
    const long long nn = (long long)nao * naux;
    // T[i][nu P] = sum_mu C_occ[mu][i] J[mu][nu P]
    hipError_t e = gemm_tn(c_occ, nocc, 0, j3c, nn, 0, tmp, nn, 0, nocc, (int)nn, nao, 1, st);
    if (e != hipSuccess) return e;
    // B[i][a][P] = sum_nu C_vir[nu][a] T[i][nu][P], batched over i
    return gemm_tn(c_vir, nvir, 0, tmp, naux, nn, out, naux, (long long)nvir * naux, nvir, naux,
                   nao, nocc, st);
}

}  // namespace qx
