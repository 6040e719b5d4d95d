// Batched FP64 factorisation of the dielectric matrices eps_f = I - Pi(i w_f).
//
// Reference: quasix_core/src/selfenergy/correlation_ac.rs:376-401 inverts eps with
// LAPACK DPOTRF+DPOTRI (``invc``), falls back to LU (``inv``), then to the identity.
// Here the same chain runs on the GPU for all frequencies at once:
//   * left-looking blocked Cholesky, block NB = 64, with X = L^-1 formed in the same step
//     loop.  Step j: one MFMA launch updates block column j of A from all finished columns
//     (A_ij -= L_i,<j L_j,<j^T) and forms the block-row sums S_jk = L_j,k:j X_k:j,k of the
//     inverse (long K, pipelined); a one-wavefront kernel factors and inverts A_jj; one
//     launch forms the panel L_ij = A_ij Linv_jj^T and X_jk = -Linv_jj S_jk.  Only block
//     column j of A and block row j of X are written per step, so the HBM traffic is the
//     left-looking reads, not a rewrite of the trailing matrix every step;
//   * Gauss-Jordan with partial pivoting (one workgroup per matrix) for any frequency
//     whose Cholesky hit a non-positive pivot; a singular matrix gives Wc = 0.
// The downstream contraction needs only X (Wc_nm = |X b|^2 - |b|^2), so DPOTRI's
// explicit inverse is formed only for the stage API.
#include "qx_internal.h"

#include <algorithm>

namespace qx {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
constexpr int P = NB + 1;   // LDS pitch (doubles): row-strided MFMA reads stay conflict-free

__device__ __forceinline__ d4 mfma64(double a, double b, d4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// 64x64 block into LDS by 256 threads: all 16 loads of a thread are issued before the LDS
// writes, so the block costs one memory latency, not sixteen.
__device__ __forceinline__ void load_block(double (*dst)[P], const double* __restrict__ src,
                                           int ld, int tid) {
    constexpr int PER = NB * NB / 256;
    double v[PER];
    #pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int idx = tid + 256 * q;
        v[q] = src[(size_t)(idx / NB) * ld + idx % NB];
    }
    #pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int idx = tid + 256 * q;
        dst[idx / NB][idx % NB] = v[q];
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

// Diagonal block j: L_jj = chol(A_jj) and Linv_jj, one wavefront per frequency.
// Both loops over the 64 pivots stay rolled (fully unrolled, the body is ~100 KB of code and
// runs instruction-fetch bound).  Registers keep static indices by rotation: at pivot k,
// r[c] holds row `lane`, column k + c of the partially eliminated block; each step applies
// the rank-1 update and shifts the row down by one slot.  The LDS vectors are stored
// pre-shifted the same way (entry c = row k + c), so every read has a static, 16-byte
// aligned offset; slots past row 63 are written as zero.  Factorisation: column k is
// broadcast through `col` (all lanes read the same addresses).  Inverse: lane c solves
// L x = e_c by column-oriented forward substitution with the same rotation, reading
// column k of L from Ls[k][.].
// 1/sqrt(d) to full FP64 precision: the hardware estimate (v_rsq_f64) refined by two
// Newton steps y <- y + y (1/2 - d y^2 / 2), which keeps the pivot's dependency chain to a
// few FMAs instead of an IEEE sqrt followed by a division.
__device__ __forceinline__ double rsqrt_nr(double d) {
    double y = __builtin_amdgcn_rsq(d);
    const double h = 0.5 * d;
    #pragma unroll
    for (int i = 0; i < 2; ++i) y = fma(y, fma(-h * y, y, 0.5), y);
    return y;
}

// col: NB doubles of LDS, Ls: NB * NB doubles of LDS (both 16-byte aligned), idg: NB doubles
// of LDS (1 / L_kk).
__device__ __forceinline__ void diag_body(double* __restrict__ a, double* __restrict__ x,
                                          double* __restrict__ linv_diag, int npad, int j,
                                          int* __restrict__ fail, int f, double* col,
                                          double* Ls, double* idg) {
    const int lane = threadIdx.x;
    const int nblk = npad / NB;
    const size_t off = (size_t)f * npad * npad + (size_t)j * NB * npad + (size_t)j * NB;
    double* A = a + off;
    double* X = x + off;
    double* D = linv_diag + ((size_t)f * nblk + j) * NB * NB;
    // This wave is on the critical path of the step loop and shares its CU with the column
    // updates of other frequency groups: give it instruction-issue priority.
    __builtin_amdgcn_s_setprio(3);
    bool bad = false;
    {
        double r[NB];
        #pragma unroll
        for (int c = 0; c < NB; ++c) r[c] = A[(size_t)lane * npad + c];
        for (int k = 0; k < NB; ++k) {
            const int slot = lane >= k ? lane - k : lane - k + NB;   // every slot written once
            col[slot] = lane >= k ? r[0] : 0.0;
            __syncthreads();
            double d = col[0];
            if (!(d > 0.0)) { bad = true; d = 1.0; }
            const double isd = rsqrt_nr(d), sd = d * isd;
            if (lane == k) idg[k] = isd;
            const double l = lane > k ? r[0] * isd : sd;             // L[lane][k] for lane >= k
            Ls[k * NB + slot] = lane >= k ? l : 0.0;
            const double t = lane > k ? l * isd : 0.0;
            #pragma unroll
            for (int c = 1; c < NB; ++c) r[c - 1] = r[c] - t * col[c];
            __syncthreads();
        }
    }
    for (int q = 0; q < NB; ++q) A[(size_t)q * npad + lane] = q >= lane ? Ls[lane * NB + q - lane] : 0.0;
    {
        double y[NB];
        #pragma unroll
        for (int c = 0; c < NB; ++c) y[c] = c == lane ? 1.0 : 0.0;
        for (int k = 0; k < NB; ++k) {
            const double* Lk = Ls + k * NB;      // Lk[c] = L[k + c][k]
            const double yk = y[0] * idg[k];     // Linv[k][lane]
            X[(size_t)k * npad + lane] = yk;
            D[k * NB + lane] = yk;
            #pragma unroll
            for (int c = 1; c < NB; ++c) y[c - 1] = y[c] - Lk[c] * yk;
        }
    }
    if (__any(bad) && lane == 0) fail[f] = 1;
}

__global__ __launch_bounds__(64) void potrf_diag_kernel(double* __restrict__ a,
                                                        double* __restrict__ x,
                                                        double* __restrict__ linv_diag,
                                                        int npad, int j, int* __restrict__ fail) {
    __shared__ __attribute__((aligned(16))) double col[NB];
    __shared__ __attribute__((aligned(16))) double Ls[NB * NB];
    __shared__ double idg[NB];
    diag_body(a, x, linv_diag, npad, j, fail, blockIdx.x, col, Ls, idg);
}

// After the diagonal factor of step j:
//   blockIdx.x <  j : finish block (j, bx) of X = L^-1:  X_j,bx = -Linv_jj S_j,bx
//   blockIdx.x >= j : panel block i = j+1+(bx-j):         L_ij  =  A_ij Linv_jj^T
__device__ __forceinline__ void rowpanel_body(double* __restrict__ a, double* __restrict__ x,
                                              const double* __restrict__ linv_diag, int npad,
                                              int j, int bx, int f, double (*As)[P],
                                              double (*Ls)[P]) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int nblk = npad / NB;
    const double* Djj = linv_diag + ((size_t)f * nblk + j) * NB * NB;
    __builtin_amdgcn_s_setprio(2);           // critical path of the step loop, as above
    load_block(Ls, Djj, NB, tid);
    d4 acc[4] = {d4{0, 0, 0, 0}, d4{0, 0, 0, 0}, d4{0, 0, 0, 0}, d4{0, 0, 0, 0}};
    if (bx < j) {
        double* S = x + (size_t)f * npad * npad + (size_t)j * NB * npad + (size_t)bx * NB;
        load_block(As, S, npad, tid);
        __syncthreads();
        mma_64x64<false>(acc, Ls, As, wave, lane);          // Linv_jj * S
        for_acc(acc, wave, lane, [&](int r, int c, double v) { S[(size_t)r * npad + c] = -v; });
    } else {
        const int i = j + 1 + (bx - j);
        double* A = a + (size_t)f * npad * npad + (size_t)i * NB * npad + (size_t)j * NB;
        load_block(As, A, npad, tid);
        __syncthreads();
        mma_64x64<true>(acc, As, Ls, wave, lane);            // A_ij * Linv_jj^T
        for_acc(acc, wave, lane, [&](int r, int c, double v) { A[(size_t)r * npad + c] = v; });
    }
}

__global__ __launch_bounds__(256) void potrf_rowpanel_kernel(double* __restrict__ a,
                                                             double* __restrict__ x,
                                                             const double* __restrict__ linv_diag,
                                                             int npad, int j) {
    __shared__ double As[NB][P];
    __shared__ double Ls[NB][P];
    rowpanel_body(a, x, linv_diag, npad, j, blockIdx.x, blockIdx.y, As, Ls);
}

// Left-looking update of step j, one 64x64 output tile per workgroup, K streamed through
// LDS in chunks of 32 with the next chunk's global loads in flight during the MFMAs:
//   bx <  nblk - j : Cholesky tile i = j + bx:   A_ij -= sum_{m<j} L_im L_jm^T   (K = 64 j)
//   bx >= nblk - j : inverse sums k = bx - (nblk - j) < j:
//                    X_jk := S_jk = sum_{k<=m<j} L_jm X_mk                    (K = 64 (j - k))
// LDS tiles are unpadded with XOR-swizzled columns: [row][k] tiles store k at
// k ^ 4 (row % 8), the [k][col] tile of the NN case stores col at col ^ 16 (k % 2).  The
// MFMA operand reads (16 rows x 4 k per wave) then touch every double-bank exactly twice,
// the minimum for 64 lanes, and the staging writes stay conflict-free per row.
constexpr int GK = 32;
__device__ __forceinline__ int swz_rk(int row, int k) { return row * GK + (k ^ (4 * (row & 7))); }
__device__ __forceinline__ int swz_kc(int k, int c) { return k * NB + (c ^ (16 * (k & 1))); }

// As, Bs: NB * GK doubles of LDS each.  Chol: Cholesky tile (NT product) or inverse sum (NN);
// the choice is uniform per workgroup, so each instance carries only its own addressing.
// Staging uses 16-byte loads: a thread moves 4 pairs of consecutive doubles per operand,
// and the swizzles keep each pair adjacent and 16-byte aligned in LDS.
template <bool Chol>
__device__ __forceinline__ void colupdate_body(double* __restrict__ a, double* __restrict__ x,
                                               int npad, int j, int bx, int f, double* As,
                                               double* Bs) {
    const int nblk = npad / NB;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    double* Af = a + (size_t)f * npad * npad;
    double* Xf = x + (size_t)f * npad * npad;
    const int i = Chol ? j + bx : j;                 // output block row
    const int k = Chol ? j : bx - (nblk - j);        // output block column
    const int kbeg = Chol ? 0 : k * NB, kend = j * NB;
    // A operand rows: L[i-block][kbeg:kend]; B operand: NT -> L[j-block][kbeg:kend] (rows),
    // NN -> X[kbeg:kend][k-block].
    const double* arow = Af + (size_t)i * NB * npad;
    const double* brow = Chol ? Af + (size_t)j * NB * npad : Xf + (size_t)k * NB;
    // [row][k] staging: pair p = tid + 256 q -> row p / 16, k 2 (p % 16)
    // [k][col] staging (NN B): pair p -> k p / 32, col 2 (p % 32)
    const int rr0 = tid >> 4, kk2 = (tid & 15) * 2;
    const int kr0 = tid >> 5, cc2 = (tid & 31) * 2;
    double ax[4], ay[4], bxv[4], byv[4];     // pairs held as scalars (kept in registers)
    auto fetch = [&](int k0) {
        #pragma unroll
        for (int q = 0; q < 4; ++q) {
            const double2 va =
                *reinterpret_cast<const double2*>(arow + (size_t)(rr0 + 16 * q) * npad + k0 + kk2);
            const double2 vb =
                Chol ? *reinterpret_cast<const double2*>(brow + (size_t)(rr0 + 16 * q) * npad + k0 + kk2)
                     : *reinterpret_cast<const double2*>(brow + (size_t)(k0 + kr0 + 8 * q) * npad + cc2);
            ax[q] = va.x;
            ay[q] = va.y;
            bxv[q] = vb.x;
            byv[q] = vb.y;
        }
    };
    d4 acc[4] = {d4{0, 0, 0, 0}, d4{0, 0, 0, 0}, d4{0, 0, 0, 0}, d4{0, 0, 0, 0}};
    if (kbeg < kend) fetch(kbeg);
    for (int k0 = kbeg; k0 < kend; k0 += GK) {
        #pragma unroll
        for (int q = 0; q < 4; ++q) {
            *reinterpret_cast<double2*>(As + swz_rk(rr0 + 16 * q, kk2)) = double2{ax[q], ay[q]};
            *reinterpret_cast<double2*>(Bs + (Chol ? swz_rk(rr0 + 16 * q, kk2)
                                                   : swz_kc(kr0 + 8 * q, cc2))) =
                double2{bxv[q], byv[q]};
        }
        __syncthreads();
        if (k0 + GK < kend) fetch(k0 + GK);
        const int rr = wave * 16 + (lane & 15);
        #pragma unroll
        for (int kk = 0; kk < GK; kk += 4) {
            const int q = kk + (lane >> 4);
            const double av = As[swz_rk(rr, q)];
            #pragma unroll
            for (int bj = 0; bj < 4; ++bj) {
                const int c = bj * 16 + (lane & 15);
                const double bv = Bs[Chol ? swz_rk(c, q) : swz_kc(q, c)];
                acc[bj] = mfma64(av, bv, acc[bj]);
            }
        }
        __syncthreads();
    }
    if (Chol) {
        if (j == 0) return;
        double* C = Af + (size_t)i * NB * npad + (size_t)j * NB;
        for_acc(acc, wave, lane, [&](int r, int c, double v) { C[(size_t)r * npad + c] -= v; });
    } else {
        double* C = Xf + (size_t)j * NB * npad + (size_t)k * NB;
        for_acc(acc, wave, lane, [&](int r, int c, double v) { C[(size_t)r * npad + c] = v; });
    }
}

__device__ __forceinline__ void colupdate_tile(double* a, double* x, int npad, int j, int bx,
                                               int f, double* As, double* Bs) {
    if (bx < npad / NB - j) colupdate_body<true>(a, x, npad, j, bx, f, As, Bs);
    else colupdate_body<false>(a, x, npad, j, bx, f, As, Bs);
}

__global__ __launch_bounds__(256, 3) void potrf_colupdate_kernel(double* __restrict__ a,
                                                              double* __restrict__ x, int npad,
                                                              int j) {
    __shared__ __attribute__((aligned(16))) double As[NB * GK];
    __shared__ __attribute__((aligned(16))) double Bs[NB * GK];
    colupdate_tile(a, x, npad, j, blockIdx.x, blockIdx.y, As, Bs);
}

// ---- Grouped forms: one launch per step covers every active instance of a batch --------
// Instances are sorted by block count, descending, so at step j the active ones (nblk > j)
// are a prefix of the table.  off[q] is the first work item of instance q in a launch
// whose per-instance grid is (items per frequency) x nf.
__global__ __launch_bounds__(256, 3) void potrf_colupdate_group_kernel(
    const PotrfInst* __restrict__ inst, const int* __restrict__ off, int nact, int j) {
    __shared__ __attribute__((aligned(16))) double As[NB * GK];
    __shared__ __attribute__((aligned(16))) double Bs[NB * GK];
    const int w = blockIdx.x;
    const int q = find_inst(off, nact, w);
    const PotrfInst in = inst[q];
    const int nblk = in.npad / NB, local = w - off[q];
    colupdate_tile(in.a, in.x, in.npad, j, local % nblk, local / nblk, As, Bs);
}

__global__ __launch_bounds__(64) void potrf_diag_group_kernel(const PotrfInst* __restrict__ inst,
                                                              int nf, int j) {
    __shared__ __attribute__((aligned(16))) double col[NB];
    __shared__ __attribute__((aligned(16))) double Ls[NB * NB];
    __shared__ double idg[NB];
    const PotrfInst in = inst[blockIdx.x / nf];
    diag_body(in.a, in.x, in.linv, in.npad, j, in.fail, blockIdx.x % nf, col, Ls, idg);
}

__global__ __launch_bounds__(256) void potrf_rowpanel_group_kernel(
    const PotrfInst* __restrict__ inst, const int* __restrict__ off, int nact, int j) {
    __shared__ double As[NB][P];
    __shared__ double Ls[NB][P];
    const int w = blockIdx.x;
    const int q = find_inst(off, nact, w);
    const PotrfInst in = inst[q];
    const int nb1 = in.npad / NB - 1, local = w - off[q];
    rowpanel_body(in.a, in.x, in.linv, in.npad, j, local % nb1, local / nb1, As, Ls);
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

hipError_t launch_potrf_batched(double* a, double* x, double* linv_diag, int npad, int nf,
                                int* fail, hipStream_t st, const StreamFan* fan, int groups) {
    const int nblk = npad / NB;
    int ng = fan ? std::min(groups, fan->n + 1) : 1;
    ng = std::max(1, std::min(ng, nf));
    hipStream_t sg[StreamFan::kMax + 1];
    sg[0] = st;
    if (ng > 1) {
        hipError_t e = hipEventRecord(fan->fork, st);
        for (int g = 1; g < ng && e == hipSuccess; ++g) {
            sg[g] = fan->side[g - 1];
            e = hipStreamWaitEvent(sg[g], fan->fork, 0);
        }
        if (e != hipSuccess) return e;
    }
    const size_t mat = (size_t)npad * npad, dblk = (size_t)nblk * NB * NB;
    for (int j = 0; j < nblk; ++j) {
        for (int g = 0; g < ng; ++g) {
            const int f0 = g * nf / ng, nfg = (g + 1) * nf / ng - f0;
            double* ag = a + f0 * mat;
            double* xg = x + f0 * mat;
            double* dg = linv_diag + f0 * dblk;
            if (j > 0)
                hipLaunchKernelGGL(potrf_colupdate_kernel, dim3(nblk, nfg), dim3(256), 0, sg[g],
                                   ag, xg, npad, j);
            hipLaunchKernelGGL(potrf_diag_kernel, dim3(nfg), dim3(64), 0, sg[g], ag, xg, dg,
                               npad, j, fail + f0);
            if (nblk > 1)
                hipLaunchKernelGGL(potrf_rowpanel_kernel, dim3(nblk - 1, nfg), dim3(256), 0,
                                   sg[g], ag, xg, dg, npad, j);
        }
    }
    for (int g = 1; g < ng; ++g) {
        hipError_t e = hipEventRecord(fan->join[g - 1], sg[g]);
        if (e == hipSuccess) e = hipStreamWaitEvent(st, fan->join[g - 1], 0);
        if (e != hipSuccess) return e;
    }
    return hipGetLastError();
}

hipError_t launch_potrf_group(const PotrfInst* inst_dev, const int* nblk_sorted, int ninst,
                              int nf, const int* cu_off_dev, const int* rp_off_dev,
                              const int* cu_off_host, const int* rp_off_host, hipStream_t st) {
    if (ninst <= 0) return hipSuccess;
    const int maxblk = nblk_sorted[0];
    int nact = ninst;
    for (int j = 0; j < maxblk; ++j) {
        while (nact > 0 && nblk_sorted[nact - 1] <= j) --nact;
        if (j > 0 && cu_off_host[nact] > 0)
            hipLaunchKernelGGL(potrf_colupdate_group_kernel, dim3(cu_off_host[nact]), dim3(256), 0,
                               st, inst_dev, cu_off_dev, nact, j);
        hipLaunchKernelGGL(potrf_diag_group_kernel, dim3(nact * nf), dim3(64), 0, st, inst_dev,
                           nf, j);
        if (rp_off_host[nact] > 0)
            hipLaunchKernelGGL(potrf_rowpanel_group_kernel, dim3(rp_off_host[nact]), dim3(256), 0,
                               st, inst_dev, rp_off_dev, nact, j);
    }
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
