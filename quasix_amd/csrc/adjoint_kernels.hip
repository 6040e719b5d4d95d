// Adjoint of the G0W0 solve: dE_QP/dB and dE_QP/deps (no reference counterpart; the
// reference has no derivatives).  Implicit differentiation of the QP equation
//   f(E) = E - eps_n - s - Re Sigma_n(E) = 0,   dE/dtheta = Z (deps_n/dtheta + dReSigma_n/dtheta)
// with Z = 1 / (1 - dReSigma_n/dE), and reverse mode through
//   Sigma_n = -(1/pi) sum_k sum_m w_nm(k) g(E - eps_m, w_k),   w_nm(k) = b_nm^T Wc_k b_nm,
//   dWc_k = eps_k^-1 dPi_k eps_k^-1,   Pi_k = sum_r s_k(r) b_r b_r^T,  s_k = 4 e/(w_k^2 + e^2).
// The per-frequency checkpoint is X_k = L_k^-1 from the forward factorisation
// (eps_k^-1 = X_k^T X_k).  All arithmetic is FP64.
#include "qx_internal.h"

namespace qx {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
constexpr double kPi = 3.14159265358979323846;
constexpr int TB = 64;           // output tile
constexpr int GK = 32;           // K chunk

__device__ __forceinline__ d4 mfma64(double a, double b, d4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ double wave_sum(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// One wavefront per target: Z, abar[k][j] = zc (-1/pi) Re g_jk, and the energy gradients
// of the Green's function (partners) and of the QP-equation anchor (target).
__global__ __launch_bounds__(64) void adj_quad_kernel(
    const int* __restrict__ targets, const int* __restrict__ pair_off,
    const int* __restrict__ pair_m, const double* __restrict__ wnm, int npair, int nf,
    const double* __restrict__ eps, const double* __restrict__ freqs,
    const double* __restrict__ wts, double eta, const double* __restrict__ E,
    const double* __restrict__ gE, int sld, double* __restrict__ abar,
    double* __restrict__ geps) {
    const int t = blockIdx.x, lane = threadIdx.x;
    const int n = targets[t];
    const double Et = E[t];
    const int j0 = pair_off[t], j1 = pair_off[t + 1];
    // pass 1: dReSigma/dE
    double dsum = 0.0;
    for (int j = j0 + lane; j < j1; j += 64) {
        const int m = pair_m[j];
        const double a = Et - eps[m];
        const double c = -eta * (eps[n] > eps[m] ? 1.0 : -1.0);
        for (int k = 0; k < nf; ++k) {
            const double w2 = freqs[k] * freqs[k];
            // emo = a + i c; dg = W (w2 - emo^2) / (emo^2 + w2)^2
            const double e2r = a * a - c * c, e2i = 2.0 * a * c;
            const double nr = w2 - e2r, ni = -e2i;
            const double dr = e2r + w2, di = e2i;
            const double d2r = dr * dr - di * di, d2i = 2.0 * dr * di;
            const double inv = wts[k] / (d2r * d2r + d2i * d2i);
            dsum += wnm[(size_t)k * npair + j] * (nr * d2r + ni * d2i) * inv;
        }
    }
    const double dsig = -wave_sum(dsum) / kPi;
    const double zc = gE[t] / (1.0 - dsig);
    if (lane == 0) atomicAdd(&geps[n], zc);
    // pass 2: abar and partner-energy gradients
    for (int j = j0 + lane; j < j1; j += 64) {
        const int m = pair_m[j];
        const double a = Et - eps[m];
        const double c = -eta * (eps[n] > eps[m] ? 1.0 : -1.0);
        double gm = 0.0;
        for (int k = 0; k < nf; ++k) {
            const double w2 = freqs[k] * freqs[k];
            const double e2r = a * a - c * c, e2i = 2.0 * a * c;
            const double dr = e2r + w2, di = e2i;
            const double gi = wts[k] / (dr * dr + di * di);
            const double gre = (a * dr + c * di) * gi;            // Re[W emo / den]
            const double nr = w2 - e2r, ni = -e2i;
            const double d2r = dr * dr - di * di, d2i = 2.0 * dr * di;
            const double dgre = wts[k] * (nr * d2r + ni * d2i) / (d2r * d2r + d2i * d2i);
            const double w = wnm[(size_t)k * npair + j];
            abar[(size_t)k * sld + j] = zc * (-1.0 / kPi) * gre;
            gm += w * dgre;
        }
        atomicAdd(&geps[m], zc * gm / kPi);
    }
}

// C_b = op(A_b) op(B_b) for square [n][n] matrices, batched over blockIdx.z with element
// strides sa / sb / sc between the matrices of a batch (0 broadcasts one operand); 64x64
// tiles, K streamed in chunks of 32 with register prefetch.
template <bool TA, bool TB_>
__global__ __launch_bounds__(256) void batched_gemm_kernel(const double* __restrict__ A,
                                                           const double* __restrict__ Bm,
                                                           double* __restrict__ C, int n,
                                                           size_t sa, size_t sb, size_t sc) {
    __shared__ double As[TB * (GK + 1)];
    __shared__ double Bs[TB * (GK + 1)];
    const int m0 = blockIdx.y * TB, n0 = blockIdx.x * TB, b = blockIdx.z;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const double* Ab = A + (size_t)b * sa;
    const double* Bb = Bm + (size_t)b * sb;
    double ra[8], rb[8];
    // LDS layouts: As "[m][k]" (pitch GK+1) or "[k][m]" (pitch TB+1); same for Bs with n.
    auto fetch = [&](int k0) {
        #pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int idx = tid + 256 * q;
            if (!TA) { const int mm = idx / GK, kk = idx % GK; ra[q] = Ab[(size_t)(m0 + mm) * n + k0 + kk]; }
            else     { const int kk = idx / TB, mm = idx % TB; ra[q] = Ab[(size_t)(k0 + kk) * n + m0 + mm]; }
            if (!TB_) { const int kk = idx / TB, nn = idx % TB; rb[q] = Bb[(size_t)(k0 + kk) * n + n0 + nn]; }
            else      { const int nn = idx / GK, kk = idx % GK; rb[q] = Bb[(size_t)(n0 + nn) * n + k0 + kk]; }
        }
    };
    d4 acc[4] = {d4{0, 0, 0, 0}, d4{0, 0, 0, 0}, d4{0, 0, 0, 0}, d4{0, 0, 0, 0}};
    fetch(0);
    for (int k0 = 0; k0 < n; k0 += GK) {
        #pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int idx = tid + 256 * q;
            if (!TA) As[(idx / GK) * (GK + 1) + idx % GK] = ra[q];
            else     As[(idx / TB) * (TB + 1) + idx % TB] = ra[q];
            if (!TB_) Bs[(idx / TB) * (TB + 1) + idx % TB] = rb[q];
            else      Bs[(idx / GK) * (GK + 1) + idx % GK] = rb[q];
        }
        __syncthreads();
        if (k0 + GK < n) fetch(k0 + GK);
        const int r = wave * 16 + (lane & 15);
        #pragma unroll
        for (int kk = 0; kk < GK; kk += 4) {
            const int q = kk + (lane >> 4);
            const double av = !TA ? As[r * (GK + 1) + q] : As[q * (TB + 1) + r];
            #pragma unroll
            for (int bj = 0; bj < 4; ++bj) {
                const int c = bj * 16 + (lane & 15);
                const double bv = !TB_ ? Bs[q * (TB + 1) + c] : Bs[c * (GK + 1) + q];
                acc[bj] = mfma64(av, bv, acc[bj]);
            }
        }
        __syncthreads();
    }
    double* Cb = C + (size_t)b * sc;
    #pragma unroll
    for (int bj = 0; bj < 4; ++bj)
        #pragma unroll
        for (int v = 0; v < 4; ++v)
            Cb[(size_t)(m0 + wave * 16 + mfma_row_f64(lane, v)) * n + n0 + bj * 16 + (lane & 15)] =
                acc[bj][v];
}

// out[r][P] += alpha sum_k sum_Q s1_k(r) A[r][Q] M_k[Q][P]      (P, Q < ncol)
// part[pt][r] = sum_{P in tile pt} (sum_k sum_Q s2_k(r) A[r][Q] M_k[Q][P]) A[r][P]  (if s2)
// A: [nrow][ncol]; s1, s2: [nf][sld]; M: [nf][npad][npad] (symmetric use: M_k[Q][P]).
__global__ __launch_bounds__(256) void scaled_gemm_kernel(
    const double* __restrict__ A, int nrow, int ncol, const double* __restrict__ s1,
    const double* __restrict__ s2, int sld, const double* __restrict__ M, int npad, int nf,
    double alpha, double* __restrict__ out, double* __restrict__ part) {
    __shared__ double As[TB][GK + 1];
    __shared__ double Ms[GK][TB + 1];
    const int r0 = blockIdx.x * TB, P0 = blockIdx.y * TB;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int rr = wave * 16 + (lane & 15);
    d4 acc1[4] = {d4{0, 0, 0, 0}, d4{0, 0, 0, 0}, d4{0, 0, 0, 0}, d4{0, 0, 0, 0}};
    d4 acc2[4] = {d4{0, 0, 0, 0}, d4{0, 0, 0, 0}, d4{0, 0, 0, 0}, d4{0, 0, 0, 0}};
    for (int k = 0; k < nf; ++k) {
        const double* Mk = M + (size_t)k * npad * npad;
        const double sc1 = r0 + rr < nrow ? s1[(size_t)k * sld + r0 + rr] : 0.0;
        const double sc2 = (s2 && r0 + rr < nrow) ? s2[(size_t)k * sld + r0 + rr] : 0.0;
        for (int Q0 = 0; Q0 < ncol; Q0 += GK) {
            #pragma unroll
            for (int q = 0; q < 8; ++q) {
                const int idx = tid + 256 * q;
                const int ar = idx / GK, ac = idx % GK;
                As[ar][ac] = (r0 + ar < nrow && Q0 + ac < ncol) ? A[(size_t)(r0 + ar) * ncol + Q0 + ac] : 0.0;
                const int mr = idx / TB, mc = idx % TB;
                Ms[mr][mc] = (Q0 + mr < ncol && P0 + mc < ncol) ? Mk[(size_t)(Q0 + mr) * npad + P0 + mc] : 0.0;
            }
            __syncthreads();
            #pragma unroll
            for (int kk = 0; kk < GK; kk += 4) {
                const int q = kk + (lane >> 4);
                const double av = As[rr][q];
                #pragma unroll
                for (int bj = 0; bj < 4; ++bj) {
                    const double bv = Ms[q][bj * 16 + (lane & 15)];
                    acc1[bj] = mfma64(av * sc1, bv, acc1[bj]);
                    if (s2) acc2[bj] = mfma64(av * sc2, bv, acc2[bj]);
                }
            }
            __syncthreads();
        }
    }
    double rs[4] = {0, 0, 0, 0};
    #pragma unroll
    for (int bj = 0; bj < 4; ++bj)
        #pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int r = r0 + wave * 16 + mfma_row_f64(lane, v);
            const int P = P0 + bj * 16 + (lane & 15);
            if (r < nrow && P < ncol) {
                out[(size_t)r * ncol + P] += alpha * acc1[bj][v];
                if (s2) rs[v] += acc2[bj][v] * A[(size_t)r * ncol + P];
            }
        }
    if (s2) {
        #pragma unroll
        for (int v = 0; v < 4; ++v)
            for (int off = 1; off < 16; off <<= 1) rs[v] += __shfl_xor(rs[v], off, 64);
        if ((lane & 15) == 0)
            #pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int r = r0 + wave * 16 + mfma_row_f64(lane, v);
                if (r < nrow) part[(size_t)blockIdx.y * nrow + r] = rs[v];
            }
    }
}

// ebar[r] = sum over column tiles of part; then geps_i += sum_a ebar(ia), geps_a -= sum_i.
__global__ void adj_eps_from_ebar_kernel(const double* __restrict__ part, int nparts, int nocc,
                                         int nvir, double* __restrict__ geps) {
    const int nov = nocc * nvir;
    for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < nov; r += gridDim.x * blockDim.x) {
        double e = 0.0;
        for (int p = 0; p < nparts; ++p) e += part[(size_t)p * nov + r];
        atomicAdd(&geps[r / nvir], e);
        atomicAdd(&geps[nocc + r % nvir], -e);
    }
}

// s_k(r) and s'_k(r) = ds/de for the rows of B_ov: [nf][sld] each.
__global__ void adj_weights_kernel(const double* __restrict__ eps, int nocc, int nvir, int nf,
                                   const double* __restrict__ freqs, int sld,
                                   double* __restrict__ s, double* __restrict__ ds) {
    const int nov = nocc * nvir, f = blockIdx.y;
    const double w2 = freqs[f] * freqs[f];
    for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < sld; r += gridDim.x * blockDim.x) {
        double a = 0.0, b = 0.0;
        if (r < nov) {
            const double e = eps[r / nvir] - eps[nocc + r % nvir];
            const double d = w2 + e * e;
            a = 4.0 * e / d;
            b = 4.0 * (w2 - e * e) / (d * d);
        }
        s[(size_t)f * sld + r] = a;
        ds[(size_t)f * sld + r] = b;
    }
}

// gBp[j] -= 2 (sum_k abar[k][j]) b_j, then scatter-add pair rows into gB rows (ov pairs).
__global__ void adj_pair_scatter_kernel(const double* __restrict__ gbp,
                                        const double* __restrict__ bp,
                                        const double* __restrict__ abar, int sld, int nf,
                                        const int* __restrict__ targets,
                                        const int* __restrict__ pair_t,
                                        const int* __restrict__ pair_m, int npair, int nocc,
                                        int nvir, int naux, double* __restrict__ gB) {
    for (int j = blockIdx.x; j < npair; j += gridDim.x) {
        double asum = 0.0;
        for (int k = 0; k < nf; ++k) asum += abar[(size_t)k * sld + j];
        const int n = targets[pair_t[j]], m = pair_m[j];
        const int r = n < nocc ? n * nvir + (m - nocc) : m * nvir + (n - nocc);
        for (int c = threadIdx.x; c < naux; c += blockDim.x)
            atomicAdd(&gB[(size_t)r * naux + c],
                      gbp[(size_t)j * naux + c] - 2.0 * asum * bp[(size_t)j * naux + c]);
    }
}

__global__ void zero_upper_blocks_kernel(double* __restrict__ x, int npad, int nf) {
    const size_t total = (size_t)nf * npad * npad;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total;
         i += (size_t)gridDim.x * blockDim.x) {
        const size_t rc = i % ((size_t)npad * npad);
        const int r = rc / npad, c = rc % npad;
        if (c / NB > r / NB) x[i] = 0.0;
    }
}

}  // namespace

hipError_t launch_zero_upper_blocks(double* x, int npad, int nf, hipStream_t st) {
    hipLaunchKernelGGL(zero_upper_blocks_kernel, dim3(8192), dim3(256), 0, st, x, npad, nf);
    return hipGetLastError();
}

hipError_t launch_adjoint_quad(const int* targets, const int* pair_off, const int* pair_m,
                               const double* wnm, int nt, int npair, int nf, const double* eps,
                               const double* freqs, const double* wts, double eta,
                               const double* E, const double* gE, int sld, double* abar,
                               double* geps, hipStream_t st) {
    hipLaunchKernelGGL(adj_quad_kernel, dim3(nt), dim3(64), 0, st, targets, pair_off, pair_m,
                       wnm, npair, nf, eps, freqs, wts, eta, E, gE, sld, abar, geps);
    return hipGetLastError();
}

hipError_t launch_strided_gemm(bool ta, bool tb, const double* A, size_t sa, const double* B,
                               size_t sb, double* C, size_t sc, int n, int batch,
                               hipStream_t st) {
    if (n % TB != 0 || n <= 0 || batch <= 0) return hipErrorInvalidValue;
    dim3 grid(n / TB, n / TB, batch);
    if (!ta && !tb)
        hipLaunchKernelGGL((batched_gemm_kernel<false, false>), grid, dim3(256), 0, st, A, B, C, n, sa, sb, sc);
    else if (ta && !tb)
        hipLaunchKernelGGL((batched_gemm_kernel<true, false>), grid, dim3(256), 0, st, A, B, C, n, sa, sb, sc);
    else if (!ta && tb)
        hipLaunchKernelGGL((batched_gemm_kernel<false, true>), grid, dim3(256), 0, st, A, B, C, n, sa, sb, sc);
    else
        hipLaunchKernelGGL((batched_gemm_kernel<true, true>), grid, dim3(256), 0, st, A, B, C, n, sa, sb, sc);
    return hipGetLastError();
}

hipError_t launch_batched_gemm(bool ta, bool tb, const double* A, const double* B, double* C,
                               int n, int batch, hipStream_t st) {
    const size_t nn = (size_t)n * n;
    return launch_strided_gemm(ta, tb, A, nn, B, nn, C, nn, n, batch, st);
}

hipError_t launch_scaled_gemm(const double* A, int nrow, int ncol, const double* s1,
                              const double* s2, int sld, const double* M, int npad, int nf,
                              double alpha, double* out, double* part, hipStream_t st) {
    dim3 grid(ceil_div(nrow, TB), ceil_div(ncol, TB));
    hipLaunchKernelGGL(scaled_gemm_kernel, grid, dim3(256), 0, st, A, nrow, ncol, s1, s2, sld, M,
                       npad, nf, alpha, out, part);
    return hipGetLastError();
}

hipError_t launch_adjoint_eps_from_ebar(const double* part, int nparts, int nocc, int nvir,
                                        double* geps, hipStream_t st) {
    hipLaunchKernelGGL(adj_eps_from_ebar_kernel, dim3(ceil_div(nocc * nvir, 256)), dim3(256), 0,
                       st, part, nparts, nocc, nvir, geps);
    return hipGetLastError();
}

hipError_t launch_adjoint_weights(const double* eps, int nocc, int nvir, int nf,
                                  const double* freqs, int sld, double* s, double* ds,
                                  hipStream_t st) {
    dim3 grid(ceil_div(sld, 256) < 1024 ? ceil_div(sld, 256) : 1024, nf);
    hipLaunchKernelGGL(adj_weights_kernel, grid, dim3(256), 0, st, eps, nocc, nvir, nf, freqs, sld,
                       s, ds);
    return hipGetLastError();
}

hipError_t launch_adjoint_pair_scatter(const double* gbp, const double* bp, const double* abar,
                                       int sld, int nf, const int* targets, const int* pair_t,
                                       const int* pair_m, int npair, int nocc, int nvir,
                                       int naux, double* gB, hipStream_t st) {
    hipLaunchKernelGGL(adj_pair_scatter_kernel, dim3(npair < 4096 ? (npair > 0 ? npair : 1) : 4096),
                       dim3(256), 0, st, gbp, bp, abar, sld, nf, targets, pair_t, pair_m, npair,
                       nocc, nvir, naux, gB);
    return hipGetLastError();
}

}  // namespace qx
