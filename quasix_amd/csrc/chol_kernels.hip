// Batched FP64 factorisation of the dielectric matrices eps_f = I - Pi(i w_f).
//
// Reference: quasix_core/src/selfenergy/correlation_ac.rs:376-401 inverts eps with
// LAPACK DPOTRF+DPOTRI (``invc``), falls back to LU (``inv``), then to the identity.
// Here the same chain runs on the GPU for all frequencies at once:
//   * left-looking blocked Cholesky, block NB = 64, with X = L^-1 formed in the same step
//     loop.  Step j: one MFMA launch updates block column j of A from all finished columns
//     (A_ij -= L_i,<j L_j,<j^T) and forms the block-row sums S_jk = L_j,k:j X_k:j,k of the
//     inverse (long K, pipelined); a one-workgroup kernel factors and inverts A_jj; one
//     launch forms the panel L_ij = A_ij Linv_jj^T and X_jk = -Linv_jj S_jk.  Only block
//     column j of A and block row j of X are written per step, so the HBM traffic is the
//     left-looking reads, not a rewrite of the trailing matrix every step.  The single-
//     instance path runs the steps in pairs (j, j + 1): one wide update of both block
//     columns (and both block rows of X) over m < j reads the operand rows once for two
//     steps, and a K = 64 update adds block column j to column j + 1 after panel j; the
//     diagonal factor runs inside the update that finishes its block (first-dispatched
//     workgroups), so a step needs no separate diagonal launch;
//   * Gauss-Jordan with partial pivoting (one workgroup per matrix) for any frequency
//     whose Cholesky hit a non-positive pivot; a singular matrix gives Wc = 0.
// The downstream contraction needs only X (Wc_nm = |X b|^2 - |b|^2), so DPOTRI's
// explicit inverse is formed only for the stage API.
#include "qx_internal.h"

#include <algorithm>
#include <cstdlib>

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

// Diagonal block j: L_jj = chol(A_jj) and Linv_jj, one 256-thread workgroup per frequency.
// Thread t owns row i = t / 4 of the block in the factorisation and column t / 4 of the
// inverse, and within it the 16 entries g + 4 m (g = t % 4, m = 0..15).  Both loops run over
// 16 groups of four pivots.  Registers keep static indices by rotation: in group kk, slot m
// holds entry g + 4 (m + kk), so slot 0 of the four threads of a row (a quad) is that row's
// entries in the group's four columns; it is finished after the group and shifted out.
// Only the first 16 - kk slots are live, which the four unrolled phases (16, 12, 8, 4
// slots updated) exploit.
// Factorisation, blocked right-looking on the lower triangle only, two barriers per group:
// every thread publishes its slot 0 (the panel of the four columns); every thread factors
// the 4x4 diagonal block redundantly, solves its own row of the panel against it and
// records the row in Lp; after the second barrier each thread applies the rank-4 update
// A_ic -= sum_s L_i,s L_c,s to its live slots with the rows L_c read from Lp.
// Inverse: the quad of column jc solves L x = e_jc by forward substitution, four pivots per
// group: the holder of row 4 kk + s scales it by 1 / L_kk and a DPP broadcast hands x to
// the quad, then all apply the rank-4 update with rows of L from Lp.  No barrier is needed.
// Lp holds L grouped by four columns, [group][row][4], so a row's four entries are one
// 32-byte run; the rows past 63 are padding that only dead slots read.
constexpr int LPR = 80;         // rows per Lp group (64 + padding for dead slots)

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

// Broadcast lane S of every quad (DPP quad_perm [S, S, S, S]).
template <int S>
__device__ __forceinline__ double quad_bcast(double v) {
    constexpr int ctrl = S | (S << 2) | (S << 4) | (S << 6);
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)(unsigned)b, ctrl, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_mov_dpp((int)(unsigned)(b >> 32), ctrl, 0xf, 0xf, false);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) |
                                            (unsigned long long)(unsigned)lo));
}

__device__ __forceinline__ void rotate16(double (&r)[16]) {
    #pragma unroll
    for (int m = 0; m < 15; ++m) r[m] = r[m + 1];
    r[15] = 0.0;
}

__device__ __forceinline__ double4 lds4(const double* p) {
    const double2 a = *reinterpret_cast<const double2*>(p);
    const double2 b = *reinterpret_cast<const double2*>(p + 2);
    return double4{a.x, a.y, b.x, b.y};
}

// Factorisation groups kk in [KK0, KK0 + 4).  P: 64 x 4 panel, Lp: 16 x LPR x 4.
template <int KK0>
__device__ __forceinline__ void chol_phase(double (&r)[16], int i, int g, double* P, double* Lp,
                                           double* idg, bool& bad) {
    constexpr int NL = 16 - KK0;
    #pragma unroll 1
    for (int kk = KK0; kk < KK0 + 4; ++kk) {
        P[i * 4 + g] = r[0];
        __syncthreads();
        double q[4][4];
        #pragma unroll
        for (int a = 0; a < 4; ++a) {
            const double4 v = lds4(P + (4 * kk + a) * 4);
            q[a][0] = v.x; q[a][1] = v.y; q[a][2] = v.z; q[a][3] = v.w;
        }
        const double4 own = lds4(P + i * 4);
        const double ai[4] = {own.x, own.y, own.z, own.w};
        // 4x4 diagonal block, in every thread
        double lb[4][4], isd[4];
        #pragma unroll
        for (int s = 0; s < 4; ++s) {
            double dd = q[s][s];
            #pragma unroll
            for (int t = 0; t < s; ++t) dd = fma(-lb[s][t], lb[s][t], dd);
            if (!(dd > 0.0)) { bad = true; dd = 1.0; }
            isd[s] = rsqrt_nr(dd);
            #pragma unroll
            for (int a = s + 1; a < 4; ++a) {
                double v = q[a][s];
                #pragma unroll
                for (int t = 0; t < s; ++t) v = fma(-lb[a][t], lb[s][t], v);
                lb[a][s] = v * isd[s];
            }
        }
        // own row of the panel: L_i,4kk+s (zero above the diagonal)
        double li[4];
        #pragma unroll
        for (int s = 0; s < 4; ++s) {
            double v = ai[s];
            #pragma unroll
            for (int t = 0; t < s; ++t) v = fma(-li[t], lb[s][t], v);
            li[s] = i >= 4 * kk + s ? v * isd[s] : 0.0;
        }
        double* L = Lp + kk * LPR * 4;
        if (g == 0) {
            *reinterpret_cast<double2*>(L + i * 4) = double2{li[0], li[1]};
            *reinterpret_cast<double2*>(L + i * 4 + 2) = double2{li[2], li[3]};
        }
        if (i == 0) idg[4 * kk + g] = g == 0 ? isd[0] : g == 1 ? isd[1] : g == 2 ? isd[2] : isd[3];
        __syncthreads();
        #pragma unroll
        for (int m = 0; m < NL; ++m) {
            const double4 lc = lds4(L + (g + 4 * (m + kk)) * 4);
            double v = r[m];
            v = fma(-li[0], lc.x, v);
            v = fma(-li[1], lc.y, v);
            v = fma(-li[2], lc.z, v);
            r[m] = fma(-li[3], lc.w, v);
        }
        rotate16(r);
    }
}

// Inverse groups kk in [KK0, KK0 + 4): column jc, rows g + 4 (m + kk).
template <int KK0>
__device__ __forceinline__ void inv_phase(double (&y)[16], int jc, int g, const double* Lp,
                                          const double* idg, double* X, double* D, int npad) {
    constexpr int NL = 16 - KK0;
    #pragma unroll 1
    for (int kk = KK0; kk < KK0 + 4; ++kk) {
        const double* L = Lp + kk * LPR * 4;
        const double4 lo = lds4(L + (g + 4 * kk) * 4);     // own holder row 4 kk + g
        const double lrow[4] = {lo.x, lo.y, lo.z, lo.w};
        double y0 = y[0], xs[4];
        #pragma unroll
        for (int s = 0; s < 4; ++s) {
            const double v = y0 * idg[4 * kk + s];
            xs[s] = s == 0 ? quad_bcast<0>(v) : s == 1 ? quad_bcast<1>(v)
                  : s == 2 ? quad_bcast<2>(v) : quad_bcast<3>(v);
            if (g > s) y0 = fma(-lrow[s], xs[s], y0);
        }
        const double xg = g == 0 ? xs[0] : g == 1 ? xs[1] : g == 2 ? xs[2] : xs[3];
        X[(size_t)(4 * kk + g) * npad + jc] = xg;
        D[(4 * kk + g) * NB + jc] = xg;
        #pragma unroll
        for (int m = 1; m < NL; ++m) {
            const double4 lr = lds4(L + (g + 4 * (m + kk)) * 4);
            double v = y[m];
            v = fma(-lr.x, xs[0], v);
            v = fma(-lr.y, xs[1], v);
            v = fma(-lr.z, xs[2], v);
            y[m] = fma(-lr.w, xs[3], v);
        }
        rotate16(y);
    }
}

// P: 4 * NB doubles, Lp: 16 * LPR * 4 doubles, idg: NB doubles of LDS (16-byte aligned).
__device__ __forceinline__ void diag_body(double* __restrict__ a, double* __restrict__ x,
                                          double* __restrict__ linv_diag, int npad, int j,
                                          int* __restrict__ fail, int f, double* P, double* Lp,
                                          double* idg) {
    const int t = threadIdx.x, i = t >> 2, g = t & 3;
    const int nblk = npad / NB;
    const size_t off = (size_t)f * npad * npad + (size_t)j * NB * npad + (size_t)j * NB;
    double* A = a + off;
    double* X = x + off;
    double* D = linv_diag + ((size_t)f * nblk + j) * NB * NB;
    // These waves are on the critical path of the step loop and share their CU with the
    // column updates of other frequency groups: give them instruction-issue priority.
    __builtin_amdgcn_s_setprio(3);
    bool bad = false;
    double r[16];
    #pragma unroll
    for (int m = 0; m < 16; ++m) r[m] = A[(size_t)i * npad + g + 4 * m];
    chol_phase<0>(r, i, g, P, Lp, idg, bad);
    chol_phase<4>(r, i, g, P, Lp, idg, bad);
    chol_phase<8>(r, i, g, P, Lp, idg, bad);
    chol_phase<12>(r, i, g, P, Lp, idg, bad);
    __syncthreads();
    #pragma unroll
    for (int m = 0; m < 16; ++m) {
        const int c = g + 4 * m;
        A[(size_t)i * npad + c] = c <= i ? Lp[(m * LPR + i) * 4 + g] : 0.0;
    }
    double y[16];
    #pragma unroll
    for (int m = 0; m < 16; ++m) y[m] = g + 4 * m == i ? 1.0 : 0.0;
    inv_phase<0>(y, i, g, Lp, idg, X, D, npad);
    inv_phase<4>(y, i, g, Lp, idg, X, D, npad);
    inv_phase<8>(y, i, g, Lp, idg, X, D, npad);
    inv_phase<12>(y, i, g, Lp, idg, X, D, npad);
    if (__any(bad) && (t & 63) == 0) fail[f] = 1;
}

__global__ __launch_bounds__(256) void potrf_diag_kernel(double* __restrict__ a,
                                                         double* __restrict__ x,
                                                         double* __restrict__ linv_diag,
                                                         int npad, int j, int* __restrict__ fail) {
    __shared__ __attribute__((aligned(16))) double P[4 * NB];
    __shared__ __attribute__((aligned(16))) double Lp[16 * LPR * 4];
    __shared__ double idg[NB];
    diag_body(a, x, linv_diag, npad, j, fail, blockIdx.x, P, Lp, idg);
}

// Work item -> (tile bx, frequency f) for nb tiles per frequency and nfg frequencies, with
// the tiles of frequency f on XCD f % 8.  Workgroups go to the 8 XCDs round-robin, so the
// item at launch position p = o + l (l local to this instance, o = the first position of
// the instance mod 8) runs on XCD p % 8.  The items are listed by home XCD (frequencies
// h, h + 8, ... for home h; with diag, the diagonal tiles bx = 0 of a home block first,
// so they are dispatched first on their XCD), and XCD x takes, in order, as many of them
// as it has positions in [o, o + W), W = nb nfg: a bijection of [0, W) for any nb, nfg, o.
// Every tile of a frequency then reads the operand it shares with the others (the rows of
// block columns j, j + 1 for the update, Linv_jj for the panel) through one L2, not eight.
__device__ __forceinline__ int xcd_count(int m, int y) {   // #{p < m : p % 8 == y}
    return m > y ? (m - y + 7) >> 3 : 0;
}
__device__ __forceinline__ void xcd_item(int l, int o, int nb, int nfg, bool diag, int& bx,
                                         int& f) {
    const int W = nb * nfg, p = o + l, x = p & 7;
    int item = (p >> 3) - xcd_count(o, x);
    for (int y = 0; y < x; ++y) item += xcd_count(o + W, y) - xcd_count(o, y);
    int hs = 0;
    for (int h = 0; h < 8; ++h) {
        const int c = h < nfg ? (nfg - h + 7) >> 3 : 0;
        if (item < hs + c * nb) {
            int r = item - hs;
            if (diag && r < c) {
                f = h + 8 * r;
                bx = 0;
            } else {
                if (diag) r -= c;
                const int per = diag ? nb - 1 : nb;
                f = h + 8 * (r / per);
                bx = (diag ? 1 : 0) + r % per;
            }
            return;
        }
        hs += c * nb;
    }
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
                                                             int npad, int j, int xmap) {
    __shared__ double As[NB][P];
    __shared__ double Ls[NB][P];
    int bx = blockIdx.x, f = blockIdx.y;
    if (xmap) xcd_item(bx + gridDim.x * f, 0, gridDim.x, gridDim.y, false, bx, f);
    rowpanel_body(a, x, linv_diag, npad, j, bx, f, As, Ls);
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

// Chol: Cholesky tile (NT product) or inverse sum (NN); the choice is uniform per
// workgroup, so each instance carries only its own addressing.  Wide: the tile spans two
// block columns j, j + 1 (Chol: 64 x 128 output, B operand L rows of both blocks) or two
// block rows j, j + 1 of X (NN: 128 x 64 output, A operand L rows of both blocks), so the
// operand that is unique to the tile is read once for two steps.  kmin: the contraction
// starts at max(kbeg, kmin); an NN tile whose K range was cut this way adds to the partial
// sum already stored (the rest of its range came from the wide update before it).
// LDS: As = (Wide NN ? 2 : 1) * NB * GK doubles, then Bs = (Wide Chol ? 2 : 1) * NB * GK.
// Staging uses 16-byte loads: a thread moves 4 pairs of consecutive doubles per 64 operand
// rows, and the swizzles keep each pair adjacent and 16-byte aligned in LDS.
template <bool Chol, bool Wide>
__device__ __forceinline__ void colupdate_body(double* __restrict__ a, double* __restrict__ x,
                                               int npad, int j, int bx, int f, int kmin,
                                               double* S) {
    constexpr int RA = (Wide && !Chol) ? 2 : 1;    // 64-row groups of the output / A operand
    constexpr int RB = (Wide && Chol) ? 2 : 1;     // 64-column groups of the output / B operand
    double* As = S;
    double* Bs = S + RA * NB * GK;
    const int nblk = npad / NB;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    double* Af = a + (size_t)f * npad * npad;
    double* Xf = x + (size_t)f * npad * npad;
    const int i = Chol ? j + bx : j;                 // output block row
    const int k = Chol ? j : bx - (nblk - j);        // output block column
    const int kbeg = Chol ? kmin : max(k * NB, kmin), kend = j * NB;
    const bool accum = !Chol && k * NB < kmin;
    if (Chol && kbeg >= kend) return;
    // A operand rows: L[i-block(s)][kbeg:kend]; B operand: NT -> L[j-block(s)][kbeg:kend]
    // (rows), NN -> X[kbeg:kend][k-block].
    const double* arow = Af + (size_t)i * NB * npad;
    const double* brow = Chol ? Af + (size_t)j * NB * npad : Xf + (size_t)k * NB;
    // [row][k] staging: pair p = tid + 256 q -> row p / 16, k 2 (p % 16)
    // [k][col] staging (NN B): pair p -> k p / 32, col 2 (p % 32)
    const int rr0 = tid >> 4, kk2 = (tid & 15) * 2;
    const int kr0 = tid >> 5, cc2 = (tid & 31) * 2;
    constexpr int NPA = 4 * RA, NPB = 4 * RB;
    double ax[NPA], ay[NPA], bxv[NPB], byv[NPB];   // pairs held as scalars (kept in registers)
    auto fetch = [&](int k0) {
        #pragma unroll
        for (int q = 0; q < NPA; ++q) {
            const double2 va =
                *reinterpret_cast<const double2*>(arow + (size_t)(rr0 + 16 * q) * npad + k0 + kk2);
            ax[q] = va.x;
            ay[q] = va.y;
        }
        #pragma unroll
        for (int q = 0; q < NPB; ++q) {
            const double2 vb =
                Chol ? *reinterpret_cast<const double2*>(brow + (size_t)(rr0 + 16 * q) * npad + k0 + kk2)
                     : *reinterpret_cast<const double2*>(brow + (size_t)(k0 + kr0 + 8 * q) * npad + cc2);
            bxv[q] = vb.x;
            byv[q] = vb.y;
        }
    };
    d4 acc[RA][4 * RB];
    #pragma unroll
    for (int ra = 0; ra < RA; ++ra)
        #pragma unroll
        for (int c = 0; c < 4 * RB; ++c) acc[ra][c] = d4{0, 0, 0, 0};
    if (kbeg < kend) fetch(kbeg);
    for (int k0 = kbeg; k0 < kend; k0 += GK) {
        #pragma unroll
        for (int q = 0; q < NPA; ++q)
            *reinterpret_cast<double2*>(As + swz_rk(rr0 + 16 * q, kk2)) = double2{ax[q], ay[q]};
        #pragma unroll
        for (int q = 0; q < NPB; ++q)
            *reinterpret_cast<double2*>(Bs + (Chol ? swz_rk(rr0 + 16 * q, kk2)
                                                   : swz_kc(kr0 + 8 * q, cc2))) =
                double2{bxv[q], byv[q]};
        __syncthreads();
        if (k0 + GK < kend) fetch(k0 + GK);
        #pragma unroll
        for (int kk = 0; kk < GK; kk += 4) {
            const int q = kk + (lane >> 4);
            double av[RA];
            #pragma unroll
            for (int ra = 0; ra < RA; ++ra) av[ra] = As[swz_rk(ra * NB + wave * 16 + (lane & 15), q)];
            #pragma unroll
            for (int cb = 0; cb < 4 * RB; ++cb) {
                const int c = cb * 16 + (lane & 15);
                const double bv = Bs[Chol ? swz_rk(c, q) : swz_kc(q, c)];
                #pragma unroll
                for (int ra = 0; ra < RA; ++ra) acc[ra][cb] = mfma64(av[ra], bv, acc[ra][cb]);
            }
        }
        __syncthreads();
    }
    double* C = Chol ? Af + (size_t)i * NB * npad + (size_t)j * NB
                     : Xf + (size_t)j * NB * npad + (size_t)k * NB;
    #pragma unroll
    for (int ra = 0; ra < RA; ++ra)
        #pragma unroll
        for (int rb = 0; rb < RB; ++rb) {
            if (Chol && rb == 1 && i == j) continue;     // block (j, j + 1) is above the diagonal
            const d4 t[4] = {acc[ra][4 * rb], acc[ra][4 * rb + 1], acc[ra][4 * rb + 2],
                             acc[ra][4 * rb + 3]};
            double* Cs = C + (size_t)ra * NB * npad + rb * NB;
            for_acc(t, wave, lane, [&](int r, int c, double v) {
                double& o = Cs[(size_t)r * npad + c];
                if (Chol) o -= v;
                else if (accum) o += v;
                else o = v;
            });
        }
}

__device__ __forceinline__ void colupdate_tile(double* a, double* x, int npad, int j, int bx,
                                               int f, int kmin, double* S) {
    if (bx < npad / NB - j) colupdate_body<true, false>(a, x, npad, j, bx, f, kmin, S);
    else colupdate_body<false, false>(a, x, npad, j, bx, f, kmin, S);
}

// Diagonal factor fused into the update that finishes the diagonal block: with linv_diag
// set, the workgroups of the diagonal tiles (bx = 0, one per frequency) are dispatched
// first and, once their tile is written, factor and invert it (diag_body) while the other
// tiles are still being updated, so the step needs no separate diagonal launch.
constexpr int DIAG_LDS = 4 * NB + 16 * LPR * 4 + NB;     // P, Lp, idg of diag_body (doubles)

__device__ __forceinline__ void diag_first(int nblk, int& bx, int& f) {
    const int w = blockIdx.x + gridDim.x * blockIdx.y, nfg = gridDim.y;
    if (w < nfg) {
        bx = 0;
        f = w;
    } else {
        const int r = w - nfg;
        f = r / (nblk - 1);
        bx = 1 + r % (nblk - 1);
    }
}

__device__ __forceinline__ void fused_diag(double* a, double* x, double* linv_diag, int npad,
                                           int j, int* fail, int f, double* S) {
    __threadfence();                          // the tile's stores, before any thread reads it
    __syncthreads();
    diag_body(a, x, linv_diag, npad, j, fail, f, S, S + 4 * NB, S + 4 * NB + 16 * LPR * 4);
}

__global__ __launch_bounds__(256, 3) void potrf_colupdate_kernel(double* __restrict__ a,
                                                              double* __restrict__ x, int npad,
                                                              int j, int kmin,
                                                              double* __restrict__ linv_diag,
                                                              int* __restrict__ fail, int xmap) {
    __shared__ __attribute__((aligned(16))) double S[2 * NB * GK > DIAG_LDS ? 2 * NB * GK : DIAG_LDS];
    int bx = blockIdx.x, f = blockIdx.y;
    if (xmap) xcd_item(bx + gridDim.x * f, 0, gridDim.x, gridDim.y, true, bx, f);
    else if (linv_diag) diag_first(npad / NB, bx, f);
    colupdate_tile(a, x, npad, j, bx, f, kmin, S);
    if (linv_diag && bx == 0) fused_diag(a, x, linv_diag, npad, j, fail, f, S);
}

// Wide update for the step pair (j, j + 1): block columns j and j + 1 of A (tiles i >= j)
// and block rows j and j + 1 of the inverse sums (k < j), all over m < j.
__global__ __launch_bounds__(256, 2) void potrf_colupdate2_kernel(double* __restrict__ a,
                                                               double* __restrict__ x, int npad,
                                                               int j,
                                                               double* __restrict__ linv_diag,
                                                               int* __restrict__ fail, int xmap) {
    __shared__ __attribute__((aligned(16))) double S[3 * NB * GK > DIAG_LDS ? 3 * NB * GK : DIAG_LDS];
    int bx = blockIdx.x, f = blockIdx.y;
    if (xmap) xcd_item(bx + gridDim.x * f, 0, gridDim.x, gridDim.y, true, bx, f);
    else if (linv_diag) diag_first(npad / NB, bx, f);
    if (bx < npad / NB - j) colupdate_body<true, true>(a, x, npad, j, bx, f, 0, S);
    else colupdate_body<false, true>(a, x, npad, j, bx, f, 0, S);
    if (linv_diag && bx == 0) fused_diag(a, x, linv_diag, npad, j, fail, f, S);
}

// QXGW_POTRF_FUSE=0: separate diagonal launches.
bool potrf_fuse_diag() {
    const char* v = std::getenv("QXGW_POTRF_FUSE");
    return !(v && std::atoi(v) == 0);
}

// QXGW_POTRF_XCD=0: workgroups in launch order (tiles of a frequency consecutive) instead
// of the per-XCD frequency mapping of xcd_item.
bool potrf_xcd_map() {
    const char* v = std::getenv("QXGW_POTRF_XCD");
    return !(v && std::atoi(v) == 0);
}

// QXGW_POTRF_PAIR=0: one full column update per step instead of step pairs.
bool potrf_pair_steps() {
    const char* v = std::getenv("QXGW_POTRF_PAIR");
    return !(v && std::atoi(v) == 0);
}

// ---- Grouped forms: one launch per step covers every active instance of a batch --------
// Instances are sorted by block count, descending, so at step j the active ones (nblk > j)
// are a prefix of the table.  off[q] is the first work item of instance q in a launch
// whose per-instance grid is (items per frequency) x nf.
__global__ __launch_bounds__(256, 3) void potrf_colupdate_group_kernel(
    const PotrfInst* __restrict__ inst, const int* __restrict__ off, int nact, int j, int nf,
    int xmap) {
    __shared__ __attribute__((aligned(16))) double S[2 * NB * GK];
    const int w = blockIdx.x;
    const int q = find_inst(off, nact, w);
    const PotrfInst in = inst[q];
    const int nblk = in.npad / NB, local = w - off[q];
    int bx = local % nblk, f = local / nblk;
    if (xmap) xcd_item(local, off[q] & 7, nblk, nf, false, bx, f);
    colupdate_tile(in.a, in.x, in.npad, j, bx, f, 0, S);
}

__global__ __launch_bounds__(256) void potrf_diag_group_kernel(const PotrfInst* __restrict__ inst,
                                                               int nf, int j) {
    __shared__ __attribute__((aligned(16))) double P[4 * NB];
    __shared__ __attribute__((aligned(16))) double Lp[16 * LPR * 4];
    __shared__ double idg[NB];
    const PotrfInst in = inst[blockIdx.x / nf];
    diag_body(in.a, in.x, in.linv, in.npad, j, in.fail, blockIdx.x % nf, P, Lp, idg);
}

__global__ __launch_bounds__(256) void potrf_rowpanel_group_kernel(
    const PotrfInst* __restrict__ inst, const int* __restrict__ off, int nact, int j, int nf,
    int xmap) {
    __shared__ double As[NB][P];
    __shared__ double Ls[NB][P];
    const int w = blockIdx.x;
    const int q = find_inst(off, nact, w);
    const PotrfInst in = inst[q];
    const int nb1 = in.npad / NB - 1, local = w - off[q];
    int bx = local % nb1, f = local / nb1;
    if (xmap) xcd_item(local, off[q] & 7, nb1, nf, false, bx, f);
    rowpanel_body(in.a, in.x, in.linv, in.npad, j, bx, f, As, Ls);
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
                                int* fail, hipStream_t st, const StreamFan* fan, int groups,
                                const PotrfAfter* after, const PotrfAfter* lead) {
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
    if (lead) {
        const hipError_t e = (*lead)(st, 0, nf / ng);
        if (e != hipSuccess) return e;
    }
    const size_t mat = (size_t)npad * npad, dblk = (size_t)nblk * NB * NB;
    // Step pairs (j, j + 1): one wide update of both block columns over m < j, the diagonal
    // and panel kernels of j, a K = 64 update of column j + 1 by block column j (kmin = 64 j),
    // then the diagonal and panel kernels of j + 1.  A last odd step runs alone.
    // With fusion the diagonal factor of every step but the first runs inside its update.
    const bool pairs = potrf_pair_steps(), fuse = potrf_fuse_diag();
    const int xmap = potrf_xcd_map() ? 1 : 0;
    for (int j = 0; j < nblk;) {
        const bool pair = pairs && j + 1 < nblk;
        for (int g = 0; g < ng; ++g) {
            const int f0 = g * nf / ng, nfg = (g + 1) * nf / ng - f0;
            double* ag = a + f0 * mat;
            double* xg = x + f0 * mat;
            double* dg = linv_diag + f0 * dblk;
            for (int s = 0; s < (pair ? 2 : 1); ++s) {
                const int js = j + s;
                double* dfuse = fuse && js > 0 ? dg : nullptr;
                if (pair && s == 0 && js > 0)
                    hipLaunchKernelGGL(potrf_colupdate2_kernel, dim3(nblk, nfg), dim3(256), 0,
                                       sg[g], ag, xg, npad, js, dfuse, fail + f0, xmap);
                else if (js > 0)
                    hipLaunchKernelGGL(potrf_colupdate_kernel, dim3(nblk, nfg), dim3(256), 0,
                                       sg[g], ag, xg, npad, js, pair ? j * NB : 0, dfuse,
                                       fail + f0, xmap);
                if (!dfuse)
                    hipLaunchKernelGGL(potrf_diag_kernel, dim3(nfg), dim3(256), 0, sg[g], ag,
                                       xg, dg, npad, js, fail + f0);
                if (nblk > 1)
                    hipLaunchKernelGGL(potrf_rowpanel_kernel, dim3(nblk - 1, nfg), dim3(256), 0,
                                       sg[g], ag, xg, dg, npad, js, xmap);
            }
        }
        j += pair ? 2 : 1;
    }
    if (after)
        for (int g = 0; g < ng; ++g) {
            const int f0 = g * nf / ng, nfg = (g + 1) * nf / ng - f0;
            const hipError_t e = (*after)(sg[g], f0, nfg);
            if (e != hipSuccess) return e;
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
    const int maxblk = nblk_sorted[0], xmap = potrf_xcd_map() ? 1 : 0;
    int nact = ninst;
    for (int j = 0; j < maxblk; ++j) {
        while (nact > 0 && nblk_sorted[nact - 1] <= j) --nact;
        if (j > 0 && cu_off_host[nact] > 0)
            hipLaunchKernelGGL(potrf_colupdate_group_kernel, dim3(cu_off_host[nact]), dim3(256), 0,
                               st, inst_dev, cu_off_dev, nact, j, nf, xmap);
        hipLaunchKernelGGL(potrf_diag_group_kernel, dim3(nact * nf), dim3(256), 0, st, inst_dev,
                           nf, j);
        if (rp_off_host[nact] > 0)
            hipLaunchKernelGGL(potrf_rowpanel_group_kernel, dim3(rp_off_host[nact]), dim3(256), 0,
                               st, inst_dev, rp_off_dev, nact, j, nf, xmap);
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
