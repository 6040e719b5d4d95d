// Real-axis screened interaction of the contour-deformation residues on the GPU:
//   w_j = b_j^T [(I - Pi(z_j))^-1 - I] b_j,   Pi(z) = sum_r B_r B_r^T c_r(z),
//   c_r(z) = 2 [1/(z + e_r) - 1/(z - e_r)],   z_j = omega_j + i eta,   e_r = eps_i - eps_a,
// for a batch of jobs (one pair row b_j and one frequency each).
//
// Re Pi and Im Pi of a job are two real symmetric products B^T diag(Re c) B and
// B^T diag(Im c) B: they run as the two frequencies of one workgroup of the FP64
// frequency-pair polarizability kernel, which stages the rows of B once for both.  The
// complex matrix eps = I - Pi is stored as two planes [npad][npad] (real, imaginary) and
// factored P eps = L U by a right-looking blocked LU with partial pivoting (panels of 64):
//   panel       one workgroup per job: pivot search, whole-row swaps (matrix and right-hand
//               side), multipliers and the rank-1 updates inside the 64-column panel;
//   trsm        U12 = L11^-1 A12 for every 64-column block right of the panel;
//   update      A22 -= L21 U12 on FP64 MFMA (complex product as four real ones), all jobs
//               and all 64x64 tiles of the trailing matrix in one launch;
// then one workgroup per job solves L y = P b, U x = y and forms b^T x - b^T b.  The padding
// block of eps is the identity, so it factors trivially and does not couple to the rest.
// Reference: the oracle's dense complex solve (numpy LU with partial pivoting) of the same
// eps(z); the contour-deformation residues follow Golze, Dvorak & Rinke, Front. Chem. 7, 377
// (2019), as cited by quasix_core/src/selfenergy/correlation_pyscf_cd.rs:46.
#include "qx_internal.h"

#include <algorithm>
#include <cstdlib>

namespace qx {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
constexpr int LB = 64;                 // panel width and tile size

__device__ __forceinline__ d4 mfma64(double a, double b, d4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// s[2j][r] = Re c_r(z_j), s[2j + 1][r] = Im c_r(z_j) for r < nov, zero up to ld.
__global__ void res_weights_kernel(const double* __restrict__ eps, int nocc, int nvir,
                                   const double* __restrict__ omegas, double eta, int ld,
                                   double* __restrict__ s) {
    const int j = blockIdx.y;
    const int nov = nocc * nvir;
    const double zr = omegas[j], zi = eta;
    for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < ld; r += gridDim.x * blockDim.x) {
        double cr = 0.0, ci = 0.0;
        if (r < nov) {
            const double e = eps[r / nvir] - eps[nocc + r % nvir];
            const double ar = zr + e, br = zr - e;
            const double i1 = 1.0 / (ar * ar + zi * zi), i2 = 1.0 / (br * br + zi * zi);
            cr = 2.0 * (ar * i1 - br * i2);
            ci = 2.0 * (-zi * i1 + zi * i2);
        }
        s[(size_t)(2 * j) * ld + r] = cr;
        s[(size_t)(2 * j + 1) * ld + r] = ci;
    }
}

// In place: planes (2j, 2j+1) of P hold Re Pi, Im Pi inside n; they become Re eps = I - Re Pi
// and Im eps = -Im Pi, with the identity on the padding.  rhs[j] = (b_j, 0) zero padded.
__global__ void res_setup_kernel(double* __restrict__ P, int n, int npad,
                                 const double* __restrict__ brows, double* __restrict__ rhs) {
    const int j = blockIdx.y;
    const size_t nn = (size_t)npad * npad;
    double* Ar = P + (size_t)(2 * j) * nn;
    double* Ai = Ar + nn;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < nn;
         i += (size_t)gridDim.x * blockDim.x) {
        const int r = (int)(i / npad), c = (int)(i % npad);
        const double d = r == c ? 1.0 : 0.0;
        if (r < n && c < n) {
            Ar[i] = d - Ar[i];
            Ai[i] = -Ai[i];
        } else {
            Ar[i] = d;
            Ai[i] = 0.0;
        }
        if (c == 0) {
            rhs[(size_t)j * 2 * npad + r] = r < n ? brows[(size_t)j * n + r] : 0.0;
            rhs[(size_t)j * 2 * npad + npad + r] = 0.0;
        }
    }
}

// 1 / p for the pivot p = pr + i pi, with |p|^2 formed without fused multiply-adds, so both
// panel kernels get the same bits whatever the surrounding code; ok = false for p = 0.
__device__ __forceinline__ void pivot_inverse(double pr, double pi, double& ir, double& ii,
                                              bool& ok) {
#pragma clang fp contract(off)
    const double pd = pr * pr + pi * pi;
    ok = pd > 0.0;
    ir = ok ? pr / pd : 0.0;
    ii = ok ? -pi / pd : 0.0;
}

// Panel k0 .. k0 + 63 of job blockIdx.x: partial pivoting over rows c .. npad - 1, whole-row
// swaps of both planes and of the right-hand side, multipliers stored below the diagonal,
// rank-1 updates of the remaining panel columns.  A zero pivot sets fail[j].
__global__ __launch_bounds__(256) void lu_panel_kernel(double* __restrict__ P, int npad, int k0,
                                                       double* __restrict__ rhs,
                                                       int* __restrict__ fail) {
    __shared__ double bv[256];
    __shared__ int bi[256];
    __shared__ double piv_inv[2];
    const int j = blockIdx.x, tid = threadIdx.x;
    const size_t nn = (size_t)npad * npad;
    double* Ar = P + (size_t)(2 * j) * nn;
    double* Ai = Ar + nn;
    double* rr = rhs + (size_t)j * 2 * npad;
    double* ri = rr + npad;
    for (int c = k0; c < k0 + LB; ++c) {
        double best = -1.0;
        int br = c;
        for (int r = c + tid; r < npad; r += 256) {
            const double x = Ar[(size_t)r * npad + c], y = Ai[(size_t)r * npad + c];
            const double v = x * x + y * y;
            if (v > best) {
                best = v;
                br = r;
            }
        }
        bv[tid] = best;
        bi[tid] = br;
        __syncthreads();
        for (int s = 128; s > 0; s >>= 1) {
            if (tid < s && (bv[tid + s] > bv[tid] || (bv[tid + s] == bv[tid] && bi[tid + s] < bi[tid]))) {
                bv[tid] = bv[tid + s];
                bi[tid] = bi[tid + s];
            }
            __syncthreads();
        }
        const int piv = bi[0];
        if (piv != c) {
            for (int q = tid; q < npad; q += 256) {
                const size_t a = (size_t)c * npad + q, b = (size_t)piv * npad + q;
                double t = Ar[a];
                Ar[a] = Ar[b];
                Ar[b] = t;
                t = Ai[a];
                Ai[a] = Ai[b];
                Ai[b] = t;
            }
            if (tid == 0) {
                double t = rr[c];
                rr[c] = rr[piv];
                rr[piv] = t;
                t = ri[c];
                ri[c] = ri[piv];
                ri[piv] = t;
            }
        }
        __syncthreads();
        if (tid == 0) {
            bool ok;
            pivot_inverse(Ar[(size_t)c * npad + c], Ai[(size_t)c * npad + c], piv_inv[0],
                          piv_inv[1], ok);
            if (!ok) fail[j] = 1;
        }
        __syncthreads();
        const double ir = piv_inv[0], ii = piv_inv[1];
        for (int r = c + 1 + tid; r < npad; r += 256) {
            const size_t o = (size_t)r * npad + c;
            const double x = Ar[o], y = Ai[o];
            Ar[o] = x * ir - y * ii;
            Ai[o] = x * ii + y * ir;
        }
        __syncthreads();
        const int w = k0 + LB - c - 1;
        if (w > 0) {
            const long long cnt = (long long)(npad - c - 1) * w;
            for (long long idx = tid; idx < cnt; idx += 256) {
                const int r = c + 1 + (int)(idx / w), q = c + 1 + (int)(idx % w);
                const size_t orc = (size_t)r * npad + c, ocq = (size_t)c * npad + q;
                const size_t o = (size_t)r * npad + q;
                const double lr = Ar[orc], li = Ai[orc], ur = Ar[ocq], ui = Ai[ocq];
                Ar[o] -= lr * ur - li * ui;
                Ai[o] -= lr * ui + li * ur;
            }
        }
        __syncthreads();
    }
}

// Argmax over the workgroup of (v, row): the larger v, the smaller row on ties (so the
// pivot is the first row of maximal |a|^2, as in lu_panel_kernel).  Every thread passes
// its candidate; the result is returned to every thread.  Three barriers.
constexpr int PNW = 16;                // waves of lu_panel_rows_kernel
__device__ __forceinline__ int block_argmax(double v, int r, double* sv, int* sr) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    #pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const double ov = __shfl_xor(v, off, 64);
        const int orow = __shfl_xor(r, off, 64);
        if (ov > v || (ov == v && orow < r)) {
            v = ov;
            r = orow;
        }
    }
    if (lane == 0) {
        sv[wv] = v;
        sr[wv] = r;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < PNW; ++w)
            if (sv[w] > v || (sv[w] == v && sr[w] < r)) {
                v = sv[w];
                r = sr[w];
            }
        sr[PNW] = r;
    }
    __syncthreads();
    const int piv = sr[PNW];
    __syncthreads();                      // sv/sr are rewritten by the next call
    return piv;
}

// The panel of lu_panel_kernel with few dependent memory round trips per column: 1024
// threads as 128 row lanes x 8 column groups.  Per column c the pivot row is chosen from
// candidates tracked during the previous column's update (a search pass only for the
// panel's first column); one pass swaps rows c and piv and puts the pivot inverse and the
// U entries of row c right of c into LDS; then each thread takes rows r = c + 1 + (tid / 8)
// + 128 k and,
// in each, the panel columns c + 1 + (tid % 8) + 8 i right of c (at most 8): all of a
// row's loads are issued together, the multiplier of column c is formed (and stored by
// column group 0), the entries are updated and stored, and column group 0 tracks |a|^2 of
// column c + 1 as the next pivot candidate.  So a column costs one round trip per 128 rows
// instead of one per 64 (8 columns per thread is what fits in the registers of 16 waves).  Same pivots (the first row of maximal |a|^2) and the same
// arithmetic as lu_panel_kernel, so the factors are bit-identical.
__global__ __launch_bounds__(64 * PNW) void lu_panel_rows_kernel(double* __restrict__ P,
                                                                 int npad, int k0,
                                                                 double* __restrict__ rhs,
                                                                 int* __restrict__ fail) {
    __shared__ double sv[PNW];
    __shared__ int sr[PNW + 1];
    __shared__ double Ur[LB], Ui[LB], pinv[2];
    constexpr int NT = 64 * PNW, CG = 8, RG = NT / CG, CPT = LB / CG;
    const int j = blockIdx.x, tid = threadIdx.x, rg = tid / CG, cg = tid % CG;
    const size_t nn = (size_t)npad * npad;
    double* Ar = P + (size_t)(2 * j) * nn;
    double* Ai = Ar + nn;
    double* rr = rhs + (size_t)j * 2 * npad;
    double* ri = rr + npad;
    double best = -1.0;
    int brow = k0;
    for (int r = k0 + tid; r < npad; r += NT) {
        const double x = Ar[(size_t)r * npad + k0], y = Ai[(size_t)r * npad + k0];
        const double v = x * x + y * y;
        if (v > best) {
            best = v;
            brow = r;
        }
    }
    for (int c = k0; c < k0 + LB; ++c) {
        const int piv = block_argmax(best, brow, sv, sr);
        const int w = k0 + LB - c - 1;            // panel columns right of c
        // one pass: swap rows c and piv (whole rows) and hand the new row c's panel entries
        // to LDS: the pivot's inverse from column c, the U entries from columns c + 1 ..
        for (int q = tid; q < npad; q += NT) {
            const bool inpanel = q >= c && q <= c + w;
            if (piv == c && !inpanel) continue;
            const size_t ac = (size_t)c * npad + q;
            double cr = Ar[ac], ci = Ai[ac];
            if (piv != c) {
                const size_t ap = (size_t)piv * npad + q;
                const double xr = Ar[ap], xi = Ai[ap];
                Ar[ac] = xr;
                Ai[ac] = xi;
                Ar[ap] = cr;
                Ai[ap] = ci;
                cr = xr;
                ci = xi;
            }
            if (q == c) {
                bool ok;
                pivot_inverse(cr, ci, pinv[0], pinv[1], ok);
                if (!ok) fail[j] = 1;
            } else if (inpanel) {
                Ur[q - c - 1] = cr;
                Ui[q - c - 1] = ci;
            }
        }
        if (piv != c && tid == 0) {
            double t = rr[c];
            rr[c] = rr[piv];
            rr[piv] = t;
            t = ri[c];
            ri[c] = ri[piv];
            ri[piv] = t;
        }
        __syncthreads();
        const double ir = pinv[0], ii = pinv[1];
        best = -1.0;
        brow = c + 1;
        for (int r = c + 1 + rg; r < npad; r += RG) {
            double* rowr = Ar + (size_t)r * npad;
            double* rowi = Ai + (size_t)r * npad;
            const double x = rowr[c], y = rowi[c];
            double a[CPT], b[CPT];
            #pragma unroll
            for (int i = 0; i < CPT; ++i) {
                const int qi = cg + CG * i;            // column c + 1 + qi
                a[i] = qi < w ? rowr[c + 1 + qi] : 0.0;
                b[i] = qi < w ? rowi[c + 1 + qi] : 0.0;
            }
            const double lr = x * ir - y * ii, li = x * ii + y * ir;
            #pragma unroll
            for (int i = 0; i < CPT; ++i) {
                const int qi = cg + CG * i;
                if (qi < w) {
                    const double ur = Ur[qi], ui = Ui[qi];
                    const double ar = a[i] - (lr * ur - li * ui);
                    const double ai = b[i] - (lr * ui + li * ur);
                    rowr[c + 1 + qi] = ar;
                    rowi[c + 1 + qi] = ai;
                    if (qi == 0) {
                        const double v = ar * ar + ai * ai;
                        if (v > best) {
                            best = v;
                            brow = r;
                        }
                    }
                }
            }
            if (cg == 0) {
                rowr[c] = lr;
                rowi[c] = li;
            }
        }
        __syncthreads();
    }
}

// U12 = L11^-1 A12 for the 64-column block blockIdx.x right of the panel, job blockIdx.y.
// One wavefront; lane = column, L11 (unit lower) and the block in LDS.
__global__ __launch_bounds__(64) void lu_trsm_kernel(double* __restrict__ P, int npad, int k0) {
    __shared__ double Lr[LB][LB + 1], Li[LB][LB + 1];
    __shared__ double Xr[LB][LB + 1], Xi[LB][LB + 1];
    const int j = blockIdx.y, lane = threadIdx.x;
    const int c0 = k0 + LB + blockIdx.x * LB;
    const size_t nn = (size_t)npad * npad;
    double* Ar = P + (size_t)(2 * j) * nn;
    double* Ai = Ar + nn;
    for (int r = 0; r < LB; ++r) {
        Lr[r][lane] = Ar[(size_t)(k0 + r) * npad + k0 + lane];
        Li[r][lane] = Ai[(size_t)(k0 + r) * npad + k0 + lane];
        Xr[r][lane] = Ar[(size_t)(k0 + r) * npad + c0 + lane];
        Xi[r][lane] = Ai[(size_t)(k0 + r) * npad + c0 + lane];
    }
    __syncthreads();
    for (int i = 1; i < LB; ++i) {
        double sr = Xr[i][lane], si = Xi[i][lane];
        for (int k = 0; k < i; ++k) {
            const double lr = Lr[i][k], li = Li[i][k], xr = Xr[k][lane], xi = Xi[k][lane];
            sr -= lr * xr - li * xi;
            si -= lr * xi + li * xr;
        }
        Xr[i][lane] = sr;
        Xi[i][lane] = si;
    }
    for (int r = 1; r < LB; ++r) {
        Ar[(size_t)(k0 + r) * npad + c0 + lane] = Xr[r][lane];
        Ai[(size_t)(k0 + r) * npad + c0 + lane] = Xi[r][lane];
    }
}

// A22 -= L21 U12 for the 64x64 tile (blockIdx.y, blockIdx.x) of the trailing matrix of job
// blockIdx.z, K = 64 (the panel), in K chunks of 16 through LDS; four waves of 32x32, each
// 2x2 blocks of 16x16x4 FP64 MFMAs for the real and the imaginary part.
__global__ __launch_bounds__(256) void lu_update_kernel(double* __restrict__ P, int npad,
                                                        int k0) {
    constexpr int KC = 16;
    __shared__ double Lr[LB][KC + 1], Li[LB][KC + 1];
    __shared__ double Ur[KC][LB + 1], Ui[KC][LB + 1];
    const int j = blockIdx.z, tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wr = wave >> 1, wc = wave & 1;
    const int r0 = k0 + LB + blockIdx.y * LB, c0 = k0 + LB + blockIdx.x * LB;
    const size_t nn = (size_t)npad * npad;
    double* Ar = P + (size_t)(2 * j) * nn;
    double* Ai = Ar + nn;
    d4 accr[2][2], acci[2][2];
    #pragma unroll
    for (int bi = 0; bi < 2; ++bi)
        #pragma unroll
        for (int bj = 0; bj < 2; ++bj) accr[bi][bj] = acci[bi][bj] = d4{0, 0, 0, 0};
    for (int kc = 0; kc < LB; kc += KC) {
        for (int idx = tid; idx < LB * KC; idx += 256) {
            const int r = idx / KC, k = idx % KC;
            Lr[r][k] = Ar[(size_t)(r0 + r) * npad + k0 + kc + k];
            Li[r][k] = Ai[(size_t)(r0 + r) * npad + k0 + kc + k];
            const int kk = idx / LB, c = idx % LB;
            Ur[kk][c] = Ar[(size_t)(k0 + kc + kk) * npad + c0 + c];
            Ui[kk][c] = Ai[(size_t)(k0 + kc + kk) * npad + c0 + c];
        }
        __syncthreads();
        #pragma unroll
        for (int ks = 0; ks < KC; ks += 4) {
            const int k = ks + (lane >> 4);
            double lr[2], li[2], ur[2], ui[2];
            #pragma unroll
            for (int q = 0; q < 2; ++q) {
                lr[q] = Lr[wr * 32 + q * 16 + (lane & 15)][k];
                li[q] = Li[wr * 32 + q * 16 + (lane & 15)][k];
                ur[q] = Ur[k][wc * 32 + q * 16 + (lane & 15)];
                ui[q] = Ui[k][wc * 32 + q * 16 + (lane & 15)];
            }
            #pragma unroll
            for (int bi = 0; bi < 2; ++bi)
                #pragma unroll
                for (int bj = 0; bj < 2; ++bj) {
                    accr[bi][bj] = mfma64(lr[bi], ur[bj], accr[bi][bj]);
                    accr[bi][bj] = mfma64(-li[bi], ui[bj], accr[bi][bj]);
                    acci[bi][bj] = mfma64(lr[bi], ui[bj], acci[bi][bj]);
                    acci[bi][bj] = mfma64(li[bi], ur[bj], acci[bi][bj]);
                }
        }
        __syncthreads();
    }
    #pragma unroll
    for (int bi = 0; bi < 2; ++bi)
        #pragma unroll
        for (int bj = 0; bj < 2; ++bj)
            #pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int r = r0 + wr * 32 + bi * 16 + mfma_row_f64(lane, v);
                const int c = c0 + wc * 32 + bj * 16 + (lane & 15);
                const size_t o = (size_t)r * npad + c;
                Ar[o] -= accr[bi][bj][v];
                Ai[o] -= acci[bi][bj][v];
            }
}

// One workgroup per job: L y = P b (unit lower), U x = y, w = b^T x - b^T b.  Rows of a
// 64-row block are shared over the four waves (16 rows each, lanes over columns, so the
// row reads are contiguous); the 64x64 triangle of the block is solved by wave 0.
__global__ __launch_bounds__(256) void lu_solve_kernel(const double* __restrict__ P, int n,
                                                       int npad, double* __restrict__ rhs,
                                                       const double* __restrict__ brows,
                                                       double* __restrict__ w_out) {
    __shared__ double tr[LB], ti[LB];
    __shared__ double red_r[256], red_i[256];
    const int j = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const size_t nn = (size_t)npad * npad;
    const double* Ar = P + (size_t)(2 * j) * nn;
    const double* Ai = Ar + nn;
    double* yr = rhs + (size_t)j * 2 * npad;
    double* yi = yr + npad;
    const int nb = npad / LB;
    // forward: y_b -= L_b,<b y_<b, then the unit-lower triangle of block b
    for (int b = 0; b < nb; ++b) {
        const int r0 = b * LB;
        for (int q = 0; q < 16; ++q) {
            const int r = r0 + wave * 16 + q;
            double sr = 0.0, si = 0.0;
            for (int k = lane; k < r0; k += 64) {
                const double lr = Ar[(size_t)r * npad + k], li = Ai[(size_t)r * npad + k];
                sr += lr * yr[k] - li * yi[k];
                si += lr * yi[k] + li * yr[k];
            }
            for (int off = 32; off > 0; off >>= 1) {
                sr += __shfl_xor(sr, off, 64);
                si += __shfl_xor(si, off, 64);
            }
            if (lane == 0) {
                tr[wave * 16 + q] = yr[r] - sr;
                ti[wave * 16 + q] = yi[r] - si;
            }
        }
        __syncthreads();
        if (wave == 0) {
            double vr = tr[lane], vi = ti[lane];
            for (int k = 0; k < LB; ++k) {
                const double xr = __shfl(vr, k, 64), xi = __shfl(vi, k, 64);
                if (lane > k) {
                    const double lr = Ar[(size_t)(r0 + lane) * npad + r0 + k];
                    const double li = Ai[(size_t)(r0 + lane) * npad + r0 + k];
                    vr -= lr * xr - li * xi;
                    vi -= lr * xi + li * xr;
                }
            }
            yr[r0 + lane] = vr;
            yi[r0 + lane] = vi;
        }
        __syncthreads();
    }
    // backward: x_b = U_bb^-1 (y_b - U_b,>b x_>b)
    for (int b = nb - 1; b >= 0; --b) {
        const int r0 = b * LB;
        for (int q = 0; q < 16; ++q) {
            const int r = r0 + wave * 16 + q;
            double sr = 0.0, si = 0.0;
            for (int k = r0 + LB + lane; k < npad; k += 64) {
                const double ur = Ar[(size_t)r * npad + k], ui = Ai[(size_t)r * npad + k];
                sr += ur * yr[k] - ui * yi[k];
                si += ur * yi[k] + ui * yr[k];
            }
            for (int off = 32; off > 0; off >>= 1) {
                sr += __shfl_xor(sr, off, 64);
                si += __shfl_xor(si, off, 64);
            }
            if (lane == 0) {
                tr[wave * 16 + q] = yr[r] - sr;
                ti[wave * 16 + q] = yi[r] - si;
            }
        }
        __syncthreads();
        if (wave == 0) {
            double vr = tr[lane], vi = ti[lane];
            for (int k = LB - 1; k >= 0; --k) {
                if (lane == k) {
                    const double ur = Ar[(size_t)(r0 + k) * npad + r0 + k];
                    const double ui = Ai[(size_t)(r0 + k) * npad + r0 + k];
                    const double d = ur * ur + ui * ui;
                    const double xr = (vr * ur + vi * ui) / d, xi = (vi * ur - vr * ui) / d;
                    vr = xr;
                    vi = xi;
                }
                const double xr = __shfl(vr, k, 64), xi = __shfl(vi, k, 64);
                if (lane < k) {
                    const double ur = Ar[(size_t)(r0 + lane) * npad + r0 + k];
                    const double ui = Ai[(size_t)(r0 + lane) * npad + r0 + k];
                    vr -= ur * xr - ui * xi;
                    vi -= ur * xi + ui * xr;
                }
            }
            yr[r0 + lane] = vr;
            yi[r0 + lane] = vi;
        }
        __syncthreads();
    }
    // w = b^T x - b^T b
    double sr = 0.0, si = 0.0;
    for (int r = tid; r < n; r += 256) {
        const double bb = brows[(size_t)j * n + r];
        sr += bb * yr[r] - bb * bb;
        si += bb * yi[r];
    }
    red_r[tid] = sr;
    red_i[tid] = si;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (tid < s) {
            red_r[tid] += red_r[tid + s];
            red_i[tid] += red_i[tid + s];
        }
        __syncthreads();
    }
    if (tid == 0) {
        w_out[2 * j] = red_r[0];
        w_out[2 * j + 1] = red_i[0];
    }
}

}  // namespace

hipError_t launch_residue_weights(const double* eps, int nocc, int nvir, int njobs,
                                  const double* omegas, double eta, int ld, double* s,
                                  hipStream_t st) {
    if (njobs <= 0) return hipSuccess;
    dim3 g(ceil_div(ld, 256) < 256 ? ceil_div(ld, 256) : 256, njobs);
    hipLaunchKernelGGL(res_weights_kernel, g, dim3(256), 0, st, eps, nocc, nvir, omegas, eta, ld,
                       s);
    return hipGetLastError();
}

hipError_t launch_residue_lu_solve(double* P, int n, int npad, int njobs, const double* brows,
                                   double* rhs, int* fail, double* w_out, hipStream_t st) {
    if (njobs <= 0) return hipSuccess;
    if (npad % LB) return hipErrorInvalidValue;
    const size_t nn = (size_t)npad * npad;
    dim3 gs((unsigned)std::min<size_t>((nn + 255) / 256, 4096), njobs);
    hipLaunchKernelGGL(res_setup_kernel, gs, dim3(256), 0, st, P, n, npad, brows, rhs);
    const int nb = npad / LB;
    // QXGW_LU_PANEL=0: the four-wave panel kernel with separate search, scale and update
    // passes instead of lu_panel_rows_kernel
    const char* pv = std::getenv("QXGW_LU_PANEL");
    const bool fused = !(pv && std::atoi(pv) == 0);
    for (int k = 0; k < nb; ++k) {
        const int k0 = k * LB, rest = nb - k - 1;
        if (fused)
            hipLaunchKernelGGL(lu_panel_rows_kernel, dim3(njobs), dim3(64 * PNW), 0, st, P,
                               npad, k0, rhs, fail);
        else
            hipLaunchKernelGGL(lu_panel_kernel, dim3(njobs), dim3(256), 0, st, P, npad, k0, rhs,
                               fail);
        if (rest > 0) {
            hipLaunchKernelGGL(lu_trsm_kernel, dim3(rest, njobs), dim3(64), 0, st, P, npad, k0);
            hipLaunchKernelGGL(lu_update_kernel, dim3(rest, rest, njobs), dim3(256), 0, st, P,
                               npad, k0);
        }
    }
    hipLaunchKernelGGL(lu_solve_kernel, dim3(njobs), dim3(256), 0, st, P, n, npad, rhs, brows,
                       w_out);
    return hipGetLastError();
}

}  // namespace qx
