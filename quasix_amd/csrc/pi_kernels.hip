// RI polarizability Pi(i w) = 4 * sum_ia B_ia B_ia^T * e_ia / (w^2 + e_ia^2) on MFMA.
//
// Reference: quasix_core/src/selfenergy/correlation_ac.rs:184-211 (get_rho_response:
// scale the ia rows by e/(w^2+e^2), then one DGEMM) and
// quasix_core/src/dielectric/polarizability.rs:169-279 (stage API, same weights).
//
// One launch covers every frequency.  A workgroup owns one 128x128 tile of the lower
// triangle of Pi for one frequency and streams the N_ov rows of B through LDS; the
// frequency weight s_f(r) is folded into the second operand while it is staged, so the
// product is B^T (S_f B) and the operand B is never rescaled in HBM.  Tiles are assigned
// XCD-aware: consecutive logical work items (same tile, neighbouring frequencies, so the
// same B columns) land on the same XCD and share its L2.
//
// Precisions: FP64 (v_mfma_f64_16x16x4f64), FP32 (v_mfma_f32_16x16x4f32, FP32 accumulate)
// and BF16x3 (B = hi + lo, S B = hi + lo, products hi*hi + hi*lo + lo*hi on
// v_mfma_f32_16x16x32_bf16 with FP32 accumulation).  Each K-chunk's global loads are issued
// into registers before the MFMA phase of the previous chunk, and 16x16 blocks outside naux
// or above the diagonal are skipped.
#include "qx_internal.h"

#include <cstdlib>

#include <algorithm>

namespace qx {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));

constexpr int TILE = 128;

__device__ __forceinline__ int xcd_remap(int bid, int n) {
    // Blocks are dispatched round-robin over the 8 XCDs; give each XCD a contiguous
    // range of logical work so that neighbouring items share its L2.
    const int per = n / 8, rem = n % 8;
    const int xcd = bid % 8, slot = bid / 8;
    return xcd < rem ? xcd * (per + 1) + slot : rem * (per + 1) + (xcd - rem) * per + slot;
}

// Lower-triangle tiles (I >= J) of a T x T tile grid enumerated column by column (J outer,
// I = J .. T - 1 inner): neighbours share the column operand rows J in L2, and the cheaper
// diagonal (first of a column) and edge (last, I = T - 1) tiles are spread over the order,
// so the contiguous range each XCD gets holds its share of them.
__device__ __forceinline__ void tri_index_cols(int t, int T, int& I, int& J) {
    int j = 0, start = 0;
    while (j < T - 1 && t >= start + (T - j)) {
        start += T - j;
        ++j;
    }
    J = j;
    I = j + (t - start);
}

// The same for a kernel that knows only the tile count T (T + 1) / 2.
__device__ __forceinline__ void tri_index_cols_n(int t, int ntiles, int& I, int& J) {
    const int T = (int)((sqrtf(8.0f * (float)ntiles + 1.0f) - 1.0f) * 0.5f + 0.5f);
    tri_index_cols(t, T, I, J);
}

__device__ __forceinline__ void tri_index(int t, int& I, int& J) {
    // t enumerates lower-triangle tiles (I >= J) row by row.
    int i = (int)((sqrtf(8.0f * (float)t + 1.0f) - 1.0f) * 0.5f);
    while ((i + 1) * (i + 2) / 2 <= t) ++i;
    while (i * (i + 1) / 2 > t) --i;
    I = i;
    J = t - i * (i + 1) / 2;
}

__device__ __forceinline__ unsigned short f32_to_bf16_rne(float x) {
    unsigned int u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf16_to_f32(unsigned short h) {
    return __uint_as_float(((unsigned int)h) << 16);
}

// Weights s[f][r] = 4 e_r / (w_f^2 + e_r^2); rows r >= nov (padding) get 0.
__global__ void pi_weights_kernel(const double* __restrict__ eps, int nocc, int nvir, int nf,
                                  const double* __restrict__ freqs, int ld,
                                  double* __restrict__ s, float* __restrict__ s32) {
    const int nov = nocc * nvir;
    const int f = blockIdx.y;
    const double w = freqs[f];
    for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < ld; r += gridDim.x * blockDim.x) {
        double v = 0.0;
        if (r < nov) {
            const int i = r / nvir, a = r % nvir;
            const double e = eps[i] - eps[nocc + a];
            v = 4.0 * e / (w * w + e * e);
        }
        s[(size_t)f * ld + r] = v;
        if (s32) s32[(size_t)f * ld + r] = (float)v;
    }
}

// ---- FP64 / FP32 kernel ------------------------------------------------------------------
template <typename T> struct Mfma;
template <> struct Mfma<double> {
    typedef d4 acc_t;
    static __device__ __forceinline__ acc_t run(double a, double b, acc_t c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
};
template <> struct Mfma<float> {
    typedef f4 acc_t;
    static __device__ __forceinline__ acc_t run(float a, float b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
};

// F64L: the accumulators come from v_mfma_f64_16x16x4f64 (row = 4 v + lane/16); otherwise
// from an FP32-output MFMA (row = 4 (lane/16) + v).
template <typename T, bool F64L>
__device__ __forceinline__ void store_tile(const typename Mfma<T>::acc_t (&acc)[4][4],
                                           int I0, int J0, int wr, int wc, int lane,
                                           int npad, int naux, int f, PiOut mode,
                                           double* __restrict__ out, double sign_scale) {
    #pragma unroll
    for (int bi = 0; bi < 4; ++bi)
        #pragma unroll
        for (int bj = 0; bj < 4; ++bj)
            #pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = I0 + wr * 64 + bi * 16 +
                                (F64L ? mfma_row_f64(lane, v) : mfma_row_f32(lane, v));
                const int col = J0 + wc * 64 + bj * 16 + (lane & 15);
                const double p = (double)acc[bi][bj][v] * sign_scale;
                if (mode == PiOut::EpsLower) {
                    if (row < npad && col < npad)
                        out[((size_t)f * npad + row) * npad + col] = (row == col ? 1.0 : 0.0) - p;
                } else {
                    // full symmetric Pi with leading dimension npad; on diagonal tiles only
                    // the lower triangle is mirrored so the result is exactly symmetric
                    // (the reference symmetrises explicitly, polarizability.rs:264-272)
                    if (row < naux && col < naux && (I0 != J0 || row >= col)) {
                        out[((size_t)f * npad + row) * npad + col] = p;
                        out[((size_t)f * npad + col) * npad + row] = p;
                    }
                }
            }
}

// Which 16x16 MFMA blocks of this wave hold entries of the lower triangle inside naux.
__device__ __forceinline__ bool block_live(int I0, int J0, int wr, int wc, int bi, int bj,
                                           int naux) {
    const int r = I0 + wr * 64 + bi * 16, c = J0 + wc * 64 + bj * 16;
    return r < naux && c < naux && (I0 != J0 || r >= c);
}

// One 128x128 lower-triangle tile (I0, J0) of frequency f.  As/Bs: KC x (TILE + 2) in LDS.
// Every 16x16 block of every wave of the tile lies in the lower triangle and inside naux.
__device__ __forceinline__ bool tile_full(int I0, int J0, int naux) {
    return I0 != J0 && I0 + TILE <= naux && J0 + TILE <= naux;
}

template <typename T, int KC, bool Guarded>
__device__ __forceinline__ void pi_tile_body(const double* __restrict__ B,
                                             const double* __restrict__ s, int nov, int naux,
                                             int npad, int sld, double* __restrict__ out,
                                             PiOut mode, int I0, int J0, int f,
                                             T (*As)[TILE + 2], T (*Bs)[TILE + 2]) {
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);   // wave-uniform (scalar)
    const int wr = wave >> 1, wc = wave & 1;
    const double* sf = s + (size_t)f * sld;

    // Staging: thread owns column c of the tile and rows kr0 + 2q of each chunk.
    constexpr int PER = KC * TILE / 256;
    const int c = tid % TILE, kr0 = tid / TILE;
    const bool ca = !Guarded || I0 + c < naux, cb = !Guarded || J0 + c < naux;
    double ra[PER], rb[PER];
    auto fetch = [&](int r0) {
        #pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int r = r0 + kr0 + 2 * q;
            double a = 0.0, b = 0.0;
            if (r < nov) {
                const double* row = B + (size_t)r * naux;
                if (ca) a = row[I0 + c];
                if (cb) b = row[J0 + c] * sf[r];
            }
            ra[q] = a;
            rb[q] = b;
        }
    };

    typename Mfma<T>::acc_t acc[4][4];
    #pragma unroll
    for (int bi = 0; bi < 4; ++bi)
        #pragma unroll
        for (int bj = 0; bj < 4; ++bj) acc[bi][bj] = typename Mfma<T>::acc_t{0, 0, 0, 0};

    fetch(0);
    for (int r0 = 0; r0 < nov; r0 += KC) {
        #pragma unroll
        for (int q = 0; q < PER; ++q) {
            As[kr0 + 2 * q][c] = (T)ra[q];
            Bs[kr0 + 2 * q][c] = (T)rb[q];
        }
        __syncthreads();
        if (r0 + KC < nov) fetch(r0 + KC);      // in flight during the MFMA phase
        #pragma unroll
        for (int kk = 0; kk < KC; kk += 4) {
            const int k = kk + (lane >> 4);
            T av[4], bv[4];
            #pragma unroll
            for (int q = 0; q < 4; ++q) {
                av[q] = As[k][wr * 64 + q * 16 + (lane & 15)];
                bv[q] = Bs[k][wc * 64 + q * 16 + (lane & 15)];
            }
            #pragma unroll
            for (int bi = 0; bi < 4; ++bi)
                #pragma unroll
                for (int bj = 0; bj < 4; ++bj)
                    if (!Guarded || block_live(I0, J0, wr, wc, bi, bj, naux))
                        acc[bi][bj] = Mfma<T>::run(av[bi], bv[bj], acc[bi][bj]);
        }
        __syncthreads();
    }
    store_tile<T, sizeof(T) == 8>(acc, I0, J0, wr, wc, lane, npad, naux, f, mode, out, 1.0);
}

template <typename T, int KC>
__global__ __launch_bounds__(256, 2) void pi_syrk_kernel(
    const double* __restrict__ B, const double* __restrict__ s, int nov, int naux, int npad,
    int nf, int ntiles, int sld, double* __restrict__ out, PiOut mode) {
    __shared__ T As[KC][TILE + 2];
    __shared__ T Bs[KC][TILE + 2];
    const int work = xcd_remap(blockIdx.x, ntiles * nf);
    const int tile = work / nf, f = work % nf;
    int I, J;
    tri_index(tile, I, J);
    if (tile_full(I * TILE, J * TILE, naux))
        pi_tile_body<T, KC, false>(B, s, nov, naux, npad, sld, out, mode, I * TILE, J * TILE, f,
                                   As, Bs);
    else
        pi_tile_body<T, KC, true>(B, s, nov, naux, npad, sld, out, mode, I * TILE, J * TILE, f,
                                  As, Bs);
}

// ---- FP64, two frequencies per tile ---------------------------------------------------------
// 512 threads own one 128x128 lower tile for the frequencies f0 and f0 + 1: the N_ov rows of
// B are staged once, unscaled, and serve both; the weight s_f(r) multiplies the row-side
// MFMA operand in registers (Pi_IJ = sum_r (s_f(r) B_rI) B_rJ).  Eight waves of 64 x 32
// (4 x 2 blocks of 16x16x4 MFMAs, 2 x 32 accumulators); per k step a wave reads 6 operand
// values from LDS for 16 MFMAs, and the global operand traffic per flop is half that of the
// one-frequency kernel.
// DB: two LDS buffers, so the next chunk is stored while other waves still multiply the
// current one and one barrier per chunk suffices (K32: 134 KB of LDS, one workgroup per CU).
// NFR = 1: the first frequency only (tail items): the second one's MFMAs and store are skipped.
template <int P2_KC, bool DB, bool Guarded, int NFR = 2>
__device__ __forceinline__ void pi_f64_pair_body(const double* __restrict__ B,
                                                 const double* __restrict__ s, int nov,
                                                 int naux, int npad, int sld, int nf,
                                                 double* __restrict__ out, PiOut mode, int I0,
                                                 int J0, int f0, double (*As)[TILE + 2],
                                                 double (*Bs)[TILE + 2],
                                                 double (*Ss)[P2_KC]) {
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    // 64-row x 32-column wave tile (wr, wc).  On a diagonal tile the live 16x16 blocks per
    // role are 7, 3, 0, 0 (wr = 0) and 8, 8, 7, 3 (wr = 1); waves w and w + 4 share a SIMD, so
    // there the roles are permuted to pair heavy with light ones (8+0, 8+0, 7+3, 7+3 instead
    // of 7+8, 3+8, 0+7, 0+3).
    const int role = (Guarded && I0 == J0) ? (0x17326054 >> (4 * wave)) & 15 : wave;
    const int wr = role >> 2, wc = role & 3;
    const bool two = f0 + 1 < nf;                       // wave-uniform
    constexpr int PER = P2_KC * TILE / 512;
    const int c = tid % TILE, kr0 = tid / TILE;         // column, first row of the thread
    const bool ca = !Guarded || I0 + c < naux, cb = !Guarded || J0 + c < naux;
    double ra[PER], rb[PER], rs = 0.0;
    const int sf_i = tid / P2_KC, sk = tid % P2_KC;     // weight staging: threads 0 .. 31
    auto fetch = [&](int r0) {
        #pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int r = r0 + kr0 + 4 * q;
            double a = 0.0, b = 0.0;
            if (r < nov) {
                const double* row = B + (size_t)r * naux;
                if (ca) a = row[I0 + c];
                if (cb) b = row[J0 + c];
            }
            ra[q] = a;
            rb[q] = b;
        }
        if (tid < 2 * P2_KC) {
            const int f = f0 + sf_i, r = r0 + sk;
            rs = (f < nf && r < nov) ? s[(size_t)f * sld + r] : 0.0;
        }
    };
    d4 acc[2][4][2];
    #pragma unroll
    for (int f = 0; f < 2; ++f)
        #pragma unroll
        for (int bi = 0; bi < 4; ++bi)
            #pragma unroll
            for (int bj = 0; bj < 2; ++bj) acc[f][bi][bj] = d4{0, 0, 0, 0};
    auto live = [&](int bi, int bj) {
        const int r = I0 + wr * 64 + bi * 16, cc = J0 + wc * 32 + bj * 16;
        return !Guarded || (r < naux && cc < naux && (I0 != J0 || r >= cc));
    };
    auto stage = [&](int buf) {                         // registers -> LDS buffer buf
        #pragma unroll
        for (int q = 0; q < PER; ++q) {
            As[buf * P2_KC + kr0 + 4 * q][c] = ra[q];
            Bs[buf * P2_KC + kr0 + 4 * q][c] = rb[q];
        }
        if (tid < 2 * P2_KC) Ss[2 * buf + sf_i][sk] = rs;
    };
    auto compute = [&](int buf) {
        #pragma unroll
        for (int kk = 0; kk < P2_KC; kk += 4) {
            const int k = buf * P2_KC + kk + (lane >> 4);
            double av[4], bv[2];
            #pragma unroll
            for (int q = 0; q < 4; ++q) av[q] = As[k][wr * 64 + q * 16 + (lane & 15)];
            #pragma unroll
            for (int q = 0; q < 2; ++q) bv[q] = Bs[k][wc * 32 + q * 16 + (lane & 15)];
            const double s0 = Ss[2 * buf][k - buf * P2_KC], s1 = Ss[2 * buf + 1][k - buf * P2_KC];
            #pragma unroll
            for (int bi = 0; bi < 4; ++bi) {
                const double a0 = av[bi] * s0, a1 = av[bi] * s1;
                // no test of `two` here: a branch per MFMA pair splits the cluster into basic
                // blocks and the next k step's LDS reads can no longer overlap it; past nf the
                // weights are zero and the second accumulator is not stored
                #pragma unroll
                for (int bj = 0; bj < 2; ++bj)
                    if (live(bi, bj)) {
                        acc[0][bi][bj] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, bv[bj],
                                                                             acc[0][bi][bj], 0, 0, 0);
                        if (NFR == 2)
                            acc[1][bi][bj] = __builtin_amdgcn_mfma_f64_16x16x4f64(
                                a1, bv[bj], acc[1][bi][bj], 0, 0, 0);
                    }
            }
        }
    };
    fetch(0);
    if constexpr (DB) {
        stage(0);
        __syncthreads();
        int cur = 0;
        for (int r0 = 0; r0 < nov; r0 += P2_KC) {
            const bool more = r0 + P2_KC < nov;         // uniform over the workgroup
            if (more) fetch(r0 + P2_KC);                // in flight during the MFMA phase
            compute(cur);
            // buffer cur ^ 1 was last read before the previous barrier
            if (more) stage(cur ^ 1);
            __syncthreads();
            cur ^= 1;
        }
    } else {
        for (int r0 = 0; r0 < nov; r0 += P2_KC) {
            stage(0);
            __syncthreads();
            if (r0 + P2_KC < nov) fetch(r0 + P2_KC);    // in flight during the MFMA phase
            compute(0);
            __syncthreads();
        }
    }
    auto put = [&](int f, int row, int col, double p) {
        const size_t fo = (size_t)(f0 + f) * npad;
        if (mode == PiOut::EpsLower) {
            if (row < npad && col < npad)
                out[(fo + row) * npad + col] = (row == col ? 1.0 : 0.0) - p;
        } else if (row < naux && col < naux && (I0 != J0 || row >= col)) {
            out[(fo + row) * npad + col] = p;
            out[(fo + col) * npad + row] = p;
        }
    };
    #pragma unroll
    for (int f = 0; f < 2; ++f) {
        if (f == 1 && (NFR == 1 || !two)) break;
        #pragma unroll
        for (int bi = 0; bi < 4; ++bi)
            #pragma unroll
            for (int bj = 0; bj < 2; ++bj)
                #pragma unroll
                for (int v = 0; v < 4; ++v)
                    put(f, I0 + wr * 64 + bi * 16 + mfma_row_f64(lane, v),
                        J0 + wc * 32 + bj * 16 + (lane & 15), acc[f][bi][bj][v]);
    }
}

// Items are (tile, frequency pair), tile-major, tiles column by column (so that no XCD's
// contiguous range holds only the cheap edge tiles of the last tile row).  The first nfull
// items run as pairs; the rest (the part of the last round of workgroups that would leave
// CUs idle) run as two single-frequency workgroups each, dispatched last.  Grid:
// nfull + 2 (ntiles * nfp - nfull).
template <int P2_KC, bool DB>
__global__ __launch_bounds__(512, 1) void pi_f64_pair_kernel(
    const double* __restrict__ B, const double* __restrict__ s, int nov, int naux, int npad,
    int nf, int ntiles, int nfull, int sld, double* __restrict__ out, PiOut mode) {
    constexpr int NBUF = DB ? 2 : 1;
    __shared__ double As[NBUF * P2_KC][TILE + 2];
    __shared__ double Bs[NBUF * P2_KC][TILE + 2];
    __shared__ double Ss[2 * NBUF][P2_KC];
    const int nfp = (nf + 1) / 2;
    const int bid = blockIdx.x;
    const bool single = bid >= nfull;
    int work, f0;
    if (!single) {
        work = xcd_remap(bid, nfull);
        f0 = 2 * (work % nfp);
    } else {
        work = nfull + (bid - nfull) / 2;
        f0 = 2 * (work % nfp) + (bid - nfull) % 2;
        if (f0 >= nf) return;                     // odd nf: the pair's second slot is empty
    }
    int I, J;
    tri_index_cols_n(work / nfp, ntiles, I, J);
    const bool full = tile_full(I * TILE, J * TILE, naux);
    if (single) {
        if (full)
            pi_f64_pair_body<P2_KC, DB, false, 1>(B, s, nov, naux, npad, sld, nf, out, mode,
                                                  I * TILE, J * TILE, f0, As, Bs, Ss);
        else
            pi_f64_pair_body<P2_KC, DB, true, 1>(B, s, nov, naux, npad, sld, nf, out, mode,
                                                 I * TILE, J * TILE, f0, As, Bs, Ss);
    } else if (full) {
        pi_f64_pair_body<P2_KC, DB, false>(B, s, nov, naux, npad, sld, nf, out, mode, I * TILE,
                                           J * TILE, f0, As, Bs, Ss);
    } else {
        pi_f64_pair_body<P2_KC, DB, true>(B, s, nov, naux, npad, sld, nf, out, mode, I * TILE,
                                          J * TILE, f0, As, Bs, Ss);
    }
}

// ---- BF16x3 kernel ----------------------------------------------------------------------
// Operands are staged K-contiguous: LDS rows are auxiliary indices, 32 N_ov rows per
// chunk, so each lane fetches its 8 consecutive k values with one 16-byte read.
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));

// hi + lo BF16 split of an FP32 quad (round to nearest even): hi by the hardware
// conversion, its FP32 value by shifting the packed halves back, lo from the remainder.
__device__ __forceinline__ void split_bf16x4(const f4 v, bf16x4& h, bf16x4& l) {
    typedef unsigned int u2 __attribute__((ext_vector_type(2)));
    h = __builtin_convertvector(v, bf16x4);
    const u2 hb = __builtin_bit_cast(u2, h);
    const f4 hf = f4{__uint_as_float(hb.x << 16), __uint_as_float(hb.x & 0xffff0000u),
                     __uint_as_float(hb.y << 16), __uint_as_float(hb.y & 0xffff0000u)};
    l = __builtin_convertvector(v - hf, bf16x4);
}

// The three split products over a wave's 4x4 blocks, product-outer so that 16 independent
// MFMAs separate dependent accumulations.  Guarded: blocks outside the triangle/naux are
// skipped (scalar branches; the wave index is scalar).  Tiles whose blocks are all live take
// the unguarded instance, decided once per tile outside the K loop, so their MFMAs form one
// straight-line block.
template <bool Guarded>
__device__ __forceinline__ void mfma_bf16x3(f4 (&acc)[4][4], const bf16x8 (&ah)[4],
                                            const bf16x8 (&al)[4], const bf16x8 (&ch)[4],
                                            const bf16x8 (&cl)[4], const bool (&live)[4][4]) {
    #pragma unroll
    for (int bi = 0; bi < 4; ++bi)
        #pragma unroll
        for (int bj = 0; bj < 4; ++bj)
            if (!Guarded || live[bi][bj])
                acc[bi][bj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[bi], ch[bj], acc[bi][bj], 0, 0, 0);
    #pragma unroll
    for (int bi = 0; bi < 4; ++bi)
        #pragma unroll
        for (int bj = 0; bj < 4; ++bj)
            if (!Guarded || live[bi][bj])
                acc[bi][bj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[bi], cl[bj], acc[bi][bj], 0, 0, 0);
    #pragma unroll
    for (int bi = 0; bi < 4; ++bi)
        #pragma unroll
        for (int bj = 0; bj < 4; ++bj)
            if (!Guarded || live[bi][bj])
                acc[bi][bj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[bi], ch[bj], acc[bi][bj], 0, 0, 0);
}

__global__ void pi_prepare_mixed_kernel(const double* __restrict__ B, int nov, int naux,
                                        int npad, int kpad, uint16_t* __restrict__ hi,
                                        uint16_t* __restrict__ lo, float* __restrict__ f32) {
    // Transpose [nov][naux] -> [npad][kpad] through a 32x32 LDS tile.
    __shared__ double t[32][33];
    const int k0 = blockIdx.x * 32, c0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;   // 32 x 8
    for (int j = ty; j < 32; j += 8) {
        const int r = k0 + j, c = c0 + tx;
        t[j][tx] = (r < nov && c < naux) ? B[(size_t)r * naux + c] : 0.0;
    }
    __syncthreads();
    for (int j = ty; j < 32; j += 8) {
        const int c = c0 + j, r = k0 + tx;
        if (c < npad && r < kpad) {
            const double x = t[tx][j];
            const float xf = (float)x;
            const unsigned short h = f32_to_bf16_rne(xf);
            const unsigned short l = f32_to_bf16_rne((float)(x - (double)bf16_to_f32(h)));
            const size_t o = (size_t)c * kpad + r;
            hi[o] = h;
            lo[o] = l;
            f32[o] = xf;
        }
    }
}

// LDS row layout of a KB-wide chunk.  KB = 64: unpadded 128-byte rows with the 16-byte
// chunk index XOR-swizzled by (row / 2) % 8; measured on gfx950 this streams MFMA fragment
// reads (16 rows x 4 chunks) at the full LDS rate, where a padded 144-byte pitch reaches
// about half of it.  KB = 32: rows padded to 40 bf16.
template <int KB> struct PiLds {
    static constexpr int pitch = KB == 64 ? 64 : KB + 8;
    static __device__ __forceinline__ int col(int row, int chunk) {
        return KB == 64 ? (chunk ^ ((row >> 1) & 7)) << 3 : chunk << 3;
    }
};

// KB: N_ov rows per LDS chunk (32 or 64).  Staging map: each thread owns R = KB/16 tile
// rows and 8 consecutive k of each; a row's KB values are read by KB/8 consecutive threads
// (coalesced), and a thread needs only its 8 frequency weights per chunk.
template <int KB, bool Guarded>
__device__ __forceinline__ void pi_tile_bf16x3_body(
    const uint16_t* __restrict__ bt_hi, const uint16_t* __restrict__ bt_lo,
    const float* __restrict__ bt_f32, const float* __restrict__ s32, int kpad, int naux,
    int npad, double* __restrict__ out, PiOut mode, int I0, int J0, int f,
    __bf16 (*Ah)[PiLds<KB>::pitch], __bf16 (*Al)[PiLds<KB>::pitch],
    __bf16 (*Ch)[PiLds<KB>::pitch], __bf16 (*Cl)[PiLds<KB>::pitch]) {
    typedef PiLds<KB> Lay;
    constexpr int TPR = KB / 8;                  // threads per row
    constexpr int RSTEP = 256 / TPR;             // row stride between a thread's rows
    constexpr int R = TILE / RSTEP;              // rows per thread
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);   // wave-uniform (scalar)
    const int wr = wave >> 1, wc = wave & 1;
    const float* sf = s32 + (size_t)f * kpad;

    const int row0 = tid / TPR, kq0 = (tid % TPR) * 8;
    u16x8 rh[R], rl[R];
    f4 rc[2 * R], rs[2];
    auto fetch = [&](int k0) {
        #pragma unroll
        for (int q = 0; q < R; ++q) {
            const int row = row0 + q * RSTEP;
            // operand rows are allocated and zero-filled up to a multiple of 256 >= the
            // tile grid, so tile rows past npad load zeros without a bounds select
            const size_t oa = (size_t)(I0 + row) * kpad + k0 + kq0;
            const size_t oc = (size_t)(J0 + row) * kpad + k0 + kq0;
            rh[q] = *reinterpret_cast<const u16x8*>(bt_hi + oa);
            rl[q] = *reinterpret_cast<const u16x8*>(bt_lo + oa);
            rc[2 * q] = *reinterpret_cast<const f4*>(bt_f32 + oc);
            rc[2 * q + 1] = *reinterpret_cast<const f4*>(bt_f32 + oc + 4);
        }
        rs[0] = *reinterpret_cast<const f4*>(sf + k0 + kq0);
        rs[1] = *reinterpret_cast<const f4*>(sf + k0 + kq0 + 4);
    };

    f4 acc[4][4];
    #pragma unroll
    for (int bi = 0; bi < 4; ++bi)
        #pragma unroll
        for (int bj = 0; bj < 4; ++bj) acc[bi][bj] = f4{0.f, 0.f, 0.f, 0.f};
    bool live[4][4];
    #pragma unroll
    for (int bi = 0; bi < 4; ++bi)
        #pragma unroll
        for (int bj = 0; bj < 4; ++bj) live[bi][bj] = block_live(I0, J0, wr, wc, bi, bj, naux);

    fetch(0);
    for (int k0 = 0; k0 < kpad; k0 += KB) {
        #pragma unroll
        for (int q = 0; q < R; ++q) {
            const int row = row0 + q * RSTEP;
            const int kc = Lay::col(row, kq0 >> 3);
            *reinterpret_cast<u16x8*>(&Ah[row][kc]) = rh[q];
            *reinterpret_cast<u16x8*>(&Al[row][kc]) = rl[q];
            // C = s * B split into hi + lo on the fly (packed hardware conversions)
            #pragma unroll
            for (int hq = 0; hq < 2; ++hq) {
                const f4 v = rc[2 * q + hq] * rs[hq];
                const bf16x4 h = __builtin_convertvector(v, bf16x4);
                const bf16x4 l = __builtin_convertvector(v - __builtin_convertvector(h, f4), bf16x4);
                *reinterpret_cast<bf16x4*>(&Ch[row][kc + 4 * hq]) = h;
                *reinterpret_cast<bf16x4*>(&Cl[row][kc + 4 * hq]) = l;
            }
        }
        __syncthreads();
        if (k0 + KB < kpad) fetch(k0 + KB);      // in flight during the MFMA phase
        #pragma unroll
        for (int ks = 0; ks < KB; ks += 32) {
            const int cq = ks / 8 + (lane >> 4), r16 = lane & 15;
            // the swizzle of rows wr*64 + q*16 + r16 depends on r16 only
            const int kq = Lay::col(r16, cq);
            bf16x8 ah[4], al[4], ch[4], cl[4];
            #pragma unroll
            for (int q = 0; q < 4; ++q) {
                ah[q] = *reinterpret_cast<const bf16x8*>(&Ah[wr * 64 + q * 16 + r16][kq]);
                al[q] = *reinterpret_cast<const bf16x8*>(&Al[wr * 64 + q * 16 + r16][kq]);
                ch[q] = *reinterpret_cast<const bf16x8*>(&Ch[wc * 64 + q * 16 + r16][kq]);
                cl[q] = *reinterpret_cast<const bf16x8*>(&Cl[wc * 64 + q * 16 + r16][kq]);
            }
            // product-outer order: 16 independent MFMAs between dependent accumulations
            mfma_bf16x3<Guarded>(acc, ah, al, ch, cl, live);
        }
        __syncthreads();
    }
    store_tile<float, false>(acc, I0, J0, wr, wc, lane, npad, naux, f, mode, out, 1.0);
}

// ---- 32x32x16 MFMA form ------------------------------------------------------------------
// A wave's 64x64 block as 2x2 blocks of 32x32: per 16-wide k-step one
// v_mfma_f32_32x32x16_bf16 per block and product, half the instructions of the 16x16x32
// form for the same work.  Measured on MI355X, one wave issuing four independent 32x32x16
// chains runs the matrix pipe near its peak, where 16x16x32 chains need two waves.
// Fragments: lane l holds row l % 32 and the 8 consecutive k at (l / 32) * 8 of the step;
// outputs as mfma_row_32.
typedef float f16v __attribute__((ext_vector_type(16)));

__device__ __forceinline__ bool block32_live(int I0, int J0, int wr, int wc, int bi, int bj,
                                             int naux) {
    const int r = I0 + wr * 64 + bi * 32, c = J0 + wc * 64 + bj * 32;
    return r < naux && c < naux && (I0 != J0 || r + 31 >= c);
}

template <bool Guarded>
__device__ __forceinline__ void mfma32_bf16x3(f16v (&acc)[2][2], const bf16x8 (&ah)[2],
                                              const bf16x8 (&al)[2], const bf16x8 (&ch)[2],
                                              const bf16x8 (&cl)[2], const bool (&live)[2][2]) {
    #pragma unroll
    for (int bi = 0; bi < 2; ++bi)
        #pragma unroll
        for (int bj = 0; bj < 2; ++bj)
            if (!Guarded || live[bi][bj])
                acc[bi][bj] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[bi], ch[bj], acc[bi][bj], 0, 0, 0);
    #pragma unroll
    for (int bi = 0; bi < 2; ++bi)
        #pragma unroll
        for (int bj = 0; bj < 2; ++bj)
            if (!Guarded || live[bi][bj])
                acc[bi][bj] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[bi], cl[bj], acc[bi][bj], 0, 0, 0);
    #pragma unroll
    for (int bi = 0; bi < 2; ++bi)
        #pragma unroll
        for (int bj = 0; bj < 2; ++bj)
            if (!Guarded || live[bi][bj])
                acc[bi][bj] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[bi], ch[bj], acc[bi][bj], 0, 0, 0);
}

// Same output conventions as store_tile (EpsLower: I - Pi over the tile; full: mirrored).
__device__ __forceinline__ void store_tile32(const f16v (&acc)[2][2], int I0, int J0, int wr,
                                             int wc, int lane, int npad, int naux, int f,
                                             PiOut mode, double* __restrict__ out) {
    #pragma unroll
    for (int bi = 0; bi < 2; ++bi)
        #pragma unroll
        for (int bj = 0; bj < 2; ++bj)
            #pragma unroll
            for (int v = 0; v < 16; ++v) {
                const int row = I0 + wr * 64 + bi * 32 + mfma_row_32(lane, v);
                const int col = J0 + wc * 64 + bj * 32 + (lane & 31);
                const double p = (double)acc[bi][bj][v];
                if (mode == PiOut::EpsLower) {
                    if (row < npad && col < npad)
                        out[((size_t)f * npad + row) * npad + col] = (row == col ? 1.0 : 0.0) - p;
                } else if (row < naux && col < naux && (I0 != J0 || row >= col)) {
                    out[((size_t)f * npad + row) * npad + col] = p;
                    out[((size_t)f * npad + col) * npad + row] = p;
                }
            }
}

// The tile body of pi_tile_bf16x3_body (same staging and LDS layout) on 32x32x16 MFMAs.
template <int KB, bool Guarded>
__device__ __forceinline__ void pi_tile_bf16x3_m32_body(
    const uint16_t* __restrict__ bt_hi, const uint16_t* __restrict__ bt_lo,
    const float* __restrict__ bt_f32, const float* __restrict__ s32, int kpad, int naux,
    int npad, double* __restrict__ out, PiOut mode, int I0, int J0, int f,
    __bf16 (*Ah)[PiLds<KB>::pitch], __bf16 (*Al)[PiLds<KB>::pitch],
    __bf16 (*Ch)[PiLds<KB>::pitch], __bf16 (*Cl)[PiLds<KB>::pitch]) {
    typedef PiLds<KB> Lay;
    constexpr int TPR = KB / 8, RSTEP = 256 / TPR, R = TILE / RSTEP;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wr = wave >> 1, wc = wave & 1;
    const float* sf = s32 + (size_t)f * kpad;
    const int row0 = tid / TPR, kq0 = (tid % TPR) * 8;
    u16x8 rh[R], rl[R];
    f4 rc[2 * R], rs[2];
    auto fetch = [&](int k0) {
        #pragma unroll
        for (int q = 0; q < R; ++q) {
            const int row = row0 + q * RSTEP;
            // operand rows are allocated and zero-filled up to a multiple of 256 >= the
            // tile grid, so tile rows past npad load zeros without a bounds select
            const size_t oa = (size_t)(I0 + row) * kpad + k0 + kq0;
            const size_t oc = (size_t)(J0 + row) * kpad + k0 + kq0;
            rh[q] = *reinterpret_cast<const u16x8*>(bt_hi + oa);
            rl[q] = *reinterpret_cast<const u16x8*>(bt_lo + oa);
            rc[2 * q] = *reinterpret_cast<const f4*>(bt_f32 + oc);
            rc[2 * q + 1] = *reinterpret_cast<const f4*>(bt_f32 + oc + 4);
        }
        rs[0] = *reinterpret_cast<const f4*>(sf + k0 + kq0);
        rs[1] = *reinterpret_cast<const f4*>(sf + k0 + kq0 + 4);
    };
    f16v acc[2][2];
    bool live[2][2];
    #pragma unroll
    for (int bi = 0; bi < 2; ++bi)
        #pragma unroll
        for (int bj = 0; bj < 2; ++bj) {
            #pragma unroll
            for (int v = 0; v < 16; ++v) acc[bi][bj][v] = 0.f;
            live[bi][bj] = block32_live(I0, J0, wr, wc, bi, bj, naux);
        }
    fetch(0);
    for (int k0 = 0; k0 < kpad; k0 += KB) {
        #pragma unroll
        for (int q = 0; q < R; ++q) {
            const int row = row0 + q * RSTEP;
            const int kc = Lay::col(row, kq0 >> 3);
            *reinterpret_cast<u16x8*>(&Ah[row][kc]) = rh[q];
            *reinterpret_cast<u16x8*>(&Al[row][kc]) = rl[q];
            #pragma unroll
            for (int hq = 0; hq < 2; ++hq) {
                bf16x4 h, l;
                split_bf16x4(rc[2 * q + hq] * rs[hq], h, l);
                *reinterpret_cast<bf16x4*>(&Ch[row][kc + 4 * hq]) = h;
                *reinterpret_cast<bf16x4*>(&Cl[row][kc + 4 * hq]) = l;
            }
        }
        __syncthreads();
        if (k0 + KB < kpad) fetch(k0 + KB);      // in flight during the MFMA phase
        #pragma unroll
        for (int ks = 0; ks < KB; ks += 16) {
            const int r32 = lane & 31;
            // rows w*64 + q*32 + r32: the swizzle depends on r32 only
            const int kq = Lay::col(r32, ks / 8 + (lane >> 5));
            bf16x8 ah[2], al[2], ch[2], cl[2];
            #pragma unroll
            for (int q = 0; q < 2; ++q) {
                ah[q] = *reinterpret_cast<const bf16x8*>(&Ah[wr * 64 + q * 32 + r32][kq]);
                al[q] = *reinterpret_cast<const bf16x8*>(&Al[wr * 64 + q * 32 + r32][kq]);
                ch[q] = *reinterpret_cast<const bf16x8*>(&Ch[wc * 64 + q * 32 + r32][kq]);
                cl[q] = *reinterpret_cast<const bf16x8*>(&Cl[wc * 64 + q * 32 + r32][kq]);
            }
            mfma32_bf16x3<Guarded>(acc, ah, al, ch, cl, live);
        }
        __syncthreads();
    }
    store_tile32(acc, I0, J0, wr, wc, lane, npad, naux, f, mode, out);
}

// KB = 32 fits three workgroups per CU (40 KB of LDS, at most 168 VGPRs).
template <int KB>
__global__ __launch_bounds__(256, KB == 32 ? 3 : 2) void pi_syrk_bf16x3_m32_kernel(
    const uint16_t* __restrict__ bt_hi, const uint16_t* __restrict__ bt_lo,
    const float* __restrict__ bt_f32, const float* __restrict__ s32, int kpad, int naux,
    int npad, int nf, int ntiles, double* __restrict__ out, PiOut mode) {
    constexpr int KBP = PiLds<KB>::pitch;
    __shared__ __attribute__((aligned(16))) __bf16 Ah[TILE][KBP];
    __shared__ __attribute__((aligned(16))) __bf16 Al[TILE][KBP];
    __shared__ __attribute__((aligned(16))) __bf16 Ch[TILE][KBP];
    __shared__ __attribute__((aligned(16))) __bf16 Cl[TILE][KBP];
    const int work = xcd_remap(blockIdx.x, ntiles * nf);
    const int tile = work / nf, f = work % nf;
    int I, J;
    tri_index(tile, I, J);
    if (tile_full(I * TILE, J * TILE, naux))
        pi_tile_bf16x3_m32_body<KB, false>(bt_hi, bt_lo, bt_f32, s32, kpad, naux, npad, out,
                                           mode, I * TILE, J * TILE, f, Ah, Al, Ch, Cl);
    else
        pi_tile_bf16x3_m32_body<KB, true>(bt_hi, bt_lo, bt_f32, s32, kpad, naux, npad, out,
                                          mode, I * TILE, J * TILE, f, Ah, Al, Ch, Cl);
}

// ---- Producer/consumer form ---------------------------------------------------------------
// 512 threads per 128x128 tile and frequency: waves 0-3 only compute (each a 64x64 block on
// 32x32x16 MFMAs, reading the fragments of the next k-step ahead of the current MFMAs),
// waves 4-7 only stage (global loads, s_f scaling and hi/lo split of the C rows, LDS
// writes).  The LDS holds two K64 chunks (128 KB, one workgroup per CU, two waves per
// SIMD: one of each kind): while the consumers multiply chunk i, the producers stage chunk
// i + 1 into the other buffer and load chunk i + 2 into registers; one barrier per chunk
// hands the buffers over.  The matrix pipe of each SIMD is fed by one wave that does
// nothing else.
// WSK = 64: 128 KB of LDS, one workgroup per CU, consumers read the next k-step ahead.
// WSK = 32: 80 KB, two workgroups per CU (four waves per SIMD, at most 128 VGPRs), no
// read-ahead; a stalled workgroup leaves the CU to the other one.
template <int WSK> struct WsCfg {
    static constexpr int pitch = PiLds<WSK>::pitch;
    static constexpr int buf = 4 * TILE * pitch;  // bf16 per buffer (Ah, Al, Ch, Cl)
};

template <int WSK, int NP, bool Guarded>
__device__ __forceinline__ void pi_tile_bf16x3_ws_body(
    const uint16_t* __restrict__ bt_hi, const uint16_t* __restrict__ bt_lo,
    const float* __restrict__ bt_f32, const float* __restrict__ s32, int kpad, int naux,
    int npad, double* __restrict__ out, PiOut mode, int I0, int J0, int f, __bf16* L) {
    typedef PiLds<WSK> Lay;
    constexpr int P = WsCfg<WSK>::pitch, WSBUF = WsCfg<WSK>::buf;
    constexpr bool Ahead = WSK == 64;
    typedef __bf16 Row[P];
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int nch = kpad / WSK;
    auto arr = [&](int buf, int a) { return reinterpret_cast<Row*>(L + buf * WSBUF + a * TILE * P); };
    if (wave < 4) {
        // ---- consumer ----
        const int wr = wave >> 1, wc = wave & 1, r32 = lane & 31;
        f16v acc[2][2];
        bool live[2][2];
        #pragma unroll
        for (int bi = 0; bi < 2; ++bi)
            #pragma unroll
            for (int bj = 0; bj < 2; ++bj) {
                #pragma unroll
                for (int v = 0; v < 16; ++v) acc[bi][bj][v] = 0.f;
                live[bi][bj] = block32_live(I0, J0, wr, wc, bi, bj, naux);
            }
        __syncthreads();                          // chunk 0 staged
        for (int i = 0; i < nch; ++i) {
            Row* Ah = arr(i & 1, 0);
            Row* Al = arr(i & 1, 1);
            Row* Ch = arr(i & 1, 2);
            Row* Cl = arr(i & 1, 3);
            bf16x8 fa[2][4][2];
            auto read_step = [&](int set, int ks) {
                // rows w*64 + q*32 + r32: the swizzle depends on r32 only
                const int kq = Lay::col(r32, ks / 8 + (lane >> 5));
                #pragma unroll
                for (int q = 0; q < 2; ++q) {
                    fa[set][0][q] = *reinterpret_cast<const bf16x8*>(&Ah[wr * 64 + q * 32 + r32][kq]);
                    fa[set][1][q] = *reinterpret_cast<const bf16x8*>(&Al[wr * 64 + q * 32 + r32][kq]);
                    fa[set][2][q] = *reinterpret_cast<const bf16x8*>(&Ch[wc * 64 + q * 32 + r32][kq]);
                    fa[set][3][q] = *reinterpret_cast<const bf16x8*>(&Cl[wc * 64 + q * 32 + r32][kq]);
                }
            };
            #pragma unroll
            for (int ks = 0; ks < WSK; ks += 16) {
                const int cur = Ahead ? (ks / 16) & 1 : 0;
                if (!Ahead || ks == 0) read_step(cur, ks);
                if (Ahead && ks + 16 < WSK) read_step(cur ^ 1, ks + 16);
                if (Ahead) __builtin_amdgcn_sched_barrier(0);
                mfma32_bf16x3<Guarded>(acc, fa[cur][0], fa[cur][1], fa[cur][2], fa[cur][3], live);
            }
            __syncthreads();                      // buffer i & 1 free, chunk i + 1 staged
        }
        store_tile32(acc, I0, J0, wr, wc, lane, npad, naux, f, mode, out);
    } else {
        // ---- producer (NP waves) ----
        // Two register sets: chunk c is loaded into set c % 2 two iterations before it is
        // staged, which covers the latency of loads that miss the L2.
        constexpr int TPR = WSK / 8, RSTEP = 64 * NP / TPR, R = TILE / RSTEP;
        const int ptid = tid - 256;
        const float* sf = s32 + (size_t)f * kpad;
        const int row0 = ptid / TPR, kq0 = (ptid % TPR) * 8;
        u16x8 rh[2][R], rl[2][R];
        f4 rc[2][2 * R], rs[2][2];
        auto fetch = [&](int set, int k0) {
            #pragma unroll
            for (int q = 0; q < R; ++q) {
                const int row = row0 + q * RSTEP;
                // operand rows are allocated and zero-filled up to a multiple of 256 >= the
                // tile grid, so tile rows past npad load zeros without a bounds select
                const size_t oa = (size_t)(I0 + row) * kpad + k0 + kq0;
                const size_t oc = (size_t)(J0 + row) * kpad + k0 + kq0;
                rh[set][q] = *reinterpret_cast<const u16x8*>(bt_hi + oa);
                rl[set][q] = *reinterpret_cast<const u16x8*>(bt_lo + oa);
                rc[set][2 * q] = *reinterpret_cast<const f4*>(bt_f32 + oc);
                rc[set][2 * q + 1] = *reinterpret_cast<const f4*>(bt_f32 + oc + 4);
            }
            rs[set][0] = *reinterpret_cast<const f4*>(sf + k0 + kq0);
            rs[set][1] = *reinterpret_cast<const f4*>(sf + k0 + kq0 + 4);
        };
        auto stage = [&](int set, int buf) {
            Row* Ah = arr(buf, 0);
            Row* Al = arr(buf, 1);
            Row* Ch = arr(buf, 2);
            Row* Cl = arr(buf, 3);
            #pragma unroll
            for (int q = 0; q < R; ++q) {
                const int row = row0 + q * RSTEP;
                const int kc = Lay::col(row, kq0 >> 3);
                *reinterpret_cast<u16x8*>(&Ah[row][kc]) = rh[set][q];
                *reinterpret_cast<u16x8*>(&Al[row][kc]) = rl[set][q];
                #pragma unroll
                for (int hq = 0; hq < 2; ++hq) {
                    bf16x4 h, l;
                    split_bf16x4(rc[set][2 * q + hq] * rs[set][hq], h, l);
                    *reinterpret_cast<bf16x4*>(&Ch[row][kc + 4 * hq]) = h;
                    *reinterpret_cast<bf16x4*>(&Cl[row][kc + 4 * hq]) = l;
                }
            }
        };
        fetch(0, 0);
        stage(0, 0);
        if (nch > 1) fetch(1, WSK);
        if (nch > 2) fetch(0, 2 * WSK);
        __syncthreads();                          // chunk 0 staged
        // one barrier per chunk, as the consumers: iterations i (even) and i + 1
        for (int i = 0; i < nch; i += 2) {
            if (i + 1 < nch) stage(1, 1);         // chunk i + 1 from set 1 into buffer 1
            if (i + 3 < nch) fetch(1, (i + 3) * WSK);
            __syncthreads();
            if (i + 1 >= nch) break;
            if (i + 2 < nch) stage(0, 0);         // chunk i + 2 from set 0 into buffer 0
            if (i + 4 < nch) fetch(0, (i + 4) * WSK);
            __syncthreads();
        }
    }
}

// NP: producer waves (4, or 8 for twice the loads in flight per CU).
template <int WSK, int NP>
__global__ __launch_bounds__(256 + 64 * NP, WSK == 32 ? 2 : 1) void pi_syrk_bf16x3_ws_kernel(
    const uint16_t* __restrict__ bt_hi, const uint16_t* __restrict__ bt_lo,
    const float* __restrict__ bt_f32, const float* __restrict__ s32, int kpad, int naux,
    int npad, int nf, int ntiles, double* __restrict__ out, PiOut mode) {
    __shared__ __attribute__((aligned(16))) __bf16 L[2 * WsCfg<WSK>::buf];
    const int work = xcd_remap(blockIdx.x, ntiles * nf);
    const int tile = work / nf, f = work % nf;
    int I, J;
    tri_index(tile, I, J);
    if (tile_full(I * TILE, J * TILE, naux))
        pi_tile_bf16x3_ws_body<WSK, NP, false>(bt_hi, bt_lo, bt_f32, s32, kpad, naux, npad,
                                               out, mode, I * TILE, J * TILE, f, L);
    else
        pi_tile_bf16x3_ws_body<WSK, NP, true>(bt_hi, bt_lo, bt_f32, s32, kpad, naux, npad,
                                              out, mode, I * TILE, J * TILE, f, L);
}

// Producer/consumer form with two frequencies per tile (K32 chunks): the A operand and the
// FP32 C rows of a chunk are loaded and staged once and serve both frequencies, which halves
// the global and L2 traffic per MFMA; the producers split C for each frequency, the
// consumers keep two sets of accumulators.  Buffers: Ah, Al, Ch[2], Cl[2] of 128 x 40 bf16
// (60 KB), two of them.  f1 == f0 (odd frequency count) computes one twice, stores it once.
constexpr int WS2K = 32;
constexpr int WS2P = PiLds<WS2K>::pitch;
constexpr int WS2A = TILE * WS2P;                 // bf16 per array
constexpr int WS2BUF = 6 * WS2A;                  // bf16 per buffer

// NFR = 1: one frequency only (f1 is ignored): the second frequency's split, staging and
// MFMAs are skipped (tail items, see pi_syrk_bf16x3_ws2_kernel).
template <bool Guarded, int NFR = 2>
__device__ __forceinline__ void pi_tile_bf16x3_ws2_body(
    const uint16_t* __restrict__ bt_hi, const uint16_t* __restrict__ bt_lo,
    const float* __restrict__ bt_f32, const float* __restrict__ s32, int kpad, int naux,
    int npad, double* __restrict__ out, PiOut mode, int I0, int J0, int f0, int f1,
    __bf16* L) {
    typedef __bf16 Row[WS2P];
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int nch = kpad / WS2K;
    // arrays: 0 Ah, 1 Al, 2 + fi Ch[fi], 4 + fi Cl[fi]
    auto arr = [&](int buf, int a) { return reinterpret_cast<Row*>(L + buf * WS2BUF + a * WS2A); };
    if (wave < 4) {
        // ---- consumer ----
        const int wr = wave >> 1, wc = wave & 1, r32 = lane & 31;
        f16v acc[2][2][2];
        bool live[2][2];
        #pragma unroll
        for (int bi = 0; bi < 2; ++bi)
            #pragma unroll
            for (int bj = 0; bj < 2; ++bj) {
                #pragma unroll
                for (int v = 0; v < 16; ++v) acc[0][bi][bj][v] = acc[1][bi][bj][v] = 0.f;
                live[bi][bj] = block32_live(I0, J0, wr, wc, bi, bj, naux);
            }
        __syncthreads();                          // chunk 0 staged
        for (int i = 0; i < nch; ++i) {
            const int b = i & 1;
            #pragma unroll
            for (int ks = 0; ks < WS2K; ks += 16) {
                const int kq = (ks / 8 + (lane >> 5)) * 8;
                bf16x8 ah[2], al[2];
                #pragma unroll
                for (int q = 0; q < 2; ++q) {
                    ah[q] = *reinterpret_cast<const bf16x8*>(&arr(b, 0)[wr * 64 + q * 32 + r32][kq]);
                    al[q] = *reinterpret_cast<const bf16x8*>(&arr(b, 1)[wr * 64 + q * 32 + r32][kq]);
                }
                #pragma unroll
                for (int fi = 0; fi < NFR; ++fi) {
                    bf16x8 ch[2], cl[2];
                    #pragma unroll
                    for (int q = 0; q < 2; ++q) {
                        ch[q] = *reinterpret_cast<const bf16x8*>(&arr(b, 2 + fi)[wc * 64 + q * 32 + r32][kq]);
                        cl[q] = *reinterpret_cast<const bf16x8*>(&arr(b, 4 + fi)[wc * 64 + q * 32 + r32][kq]);
                    }
                    mfma32_bf16x3<Guarded>(acc[fi], ah, al, ch, cl, live);
                }
            }
            __syncthreads();                      // buffer i & 1 free, chunk i + 1 staged
        }
        store_tile32(acc[0], I0, J0, wr, wc, lane, npad, naux, f0, mode, out);
        if (NFR == 2 && f1 != f0)
            store_tile32(acc[1], I0, J0, wr, wc, lane, npad, naux, f1, mode, out);
    } else {
        // ---- producer: two register sets, chunk c loaded into set c % 2 two iterations
        // before it is staged ----
        constexpr int TPR = WS2K / 8, RSTEP = 256 / TPR, R = TILE / RSTEP;
        const int ptid = tid - 256;
        const float* sf0 = s32 + (size_t)f0 * kpad;
        const float* sf1 = s32 + (size_t)f1 * kpad;
        const int row0 = ptid / TPR, kq0 = (ptid % TPR) * 8;
        u16x8 rh[2][R], rl[2][R];
        f4 rc[2][2 * R], rs[2][2][2];             // rs[set][frequency][half]
        auto fetch = [&](int set, int k0) {
            #pragma unroll
            for (int q = 0; q < R; ++q) {
                const int row = row0 + q * RSTEP;
                // operand rows are allocated and zero-filled up to a multiple of 256 >= the
                // tile grid, so tile rows past npad load zeros without a bounds select
                const size_t oa = (size_t)(I0 + row) * kpad + k0 + kq0;
                const size_t oc = (size_t)(J0 + row) * kpad + k0 + kq0;
                rh[set][q] = *reinterpret_cast<const u16x8*>(bt_hi + oa);
                rl[set][q] = *reinterpret_cast<const u16x8*>(bt_lo + oa);
                rc[set][2 * q] = *reinterpret_cast<const f4*>(bt_f32 + oc);
                rc[set][2 * q + 1] = *reinterpret_cast<const f4*>(bt_f32 + oc + 4);
            }
            rs[set][0][0] = *reinterpret_cast<const f4*>(sf0 + k0 + kq0);
            rs[set][0][1] = *reinterpret_cast<const f4*>(sf0 + k0 + kq0 + 4);
            rs[set][1][0] = *reinterpret_cast<const f4*>(sf1 + k0 + kq0);
            rs[set][1][1] = *reinterpret_cast<const f4*>(sf1 + k0 + kq0 + 4);
        };
        auto stage = [&](int set, int buf) {
            #pragma unroll
            for (int q = 0; q < R; ++q) {
                const int row = row0 + q * RSTEP;
                *reinterpret_cast<u16x8*>(&arr(buf, 0)[row][kq0]) = rh[set][q];
                *reinterpret_cast<u16x8*>(&arr(buf, 1)[row][kq0]) = rl[set][q];
                #pragma unroll
                for (int fi = 0; fi < NFR; ++fi)
                    #pragma unroll
                    for (int hq = 0; hq < 2; ++hq) {
                        bf16x4 h, l;
                        split_bf16x4(rc[set][2 * q + hq] * rs[set][fi][hq], h, l);
                        *reinterpret_cast<bf16x4*>(&arr(buf, 2 + fi)[row][kq0 + 4 * hq]) = h;
                        *reinterpret_cast<bf16x4*>(&arr(buf, 4 + fi)[row][kq0 + 4 * hq]) = l;
                    }
            }
        };
        fetch(0, 0);
        stage(0, 0);
        if (nch > 1) fetch(1, WS2K);
        if (nch > 2) fetch(0, 2 * WS2K);
        __syncthreads();                          // chunk 0 staged
        for (int i = 0; i < nch; i += 2) {
            if (i + 1 < nch) stage(1, 1);
            if (i + 3 < nch) fetch(1, (i + 3) * WS2K);
            __syncthreads();
            if (i + 1 >= nch) break;
            if (i + 2 < nch) stage(0, 0);
            if (i + 4 < nch) fetch(0, (i + 4) * WS2K);
            __syncthreads();
        }
    }
}

// Items are (tile, frequency pair), tile-major.  The first nfull items run as pairs; the
// last ntiles * nfp - nfull items (the part of the last round of workgroups that would leave
// CUs idle) run as two single-frequency workgroups each, dispatched last, so the final round
// is shorter.  Grid: nfull + 2 (ntiles * nfp - nfull).
__global__ __launch_bounds__(512, 1) void pi_syrk_bf16x3_ws2_kernel(
    const uint16_t* __restrict__ bt_hi, const uint16_t* __restrict__ bt_lo,
    const float* __restrict__ bt_f32, const float* __restrict__ s32, int kpad, int naux,
    int npad, int nf, int ntiles, int nfull, double* __restrict__ out, PiOut mode) {
    __shared__ __attribute__((aligned(16))) __bf16 L[2 * WS2BUF];
    const int nfp = (nf + 1) / 2;
    const int bid = blockIdx.x;
    const bool single = bid >= nfull;
    int work, f0;
    if (!single) {
        work = xcd_remap(bid, nfull);
        f0 = 2 * (work % nfp);
    } else {
        work = nfull + (bid - nfull) / 2;
        f0 = 2 * (work % nfp) + (bid - nfull) % 2;
        if (f0 >= nf) return;                     // odd nf: the pair's second slot is empty
    }
    const int T = (int)((sqrtf(8.0f * (float)ntiles + 1.0f) - 1.0f) * 0.5f + 0.5f);
    const int f1 = min(f0 + 1, nf - 1);
    int I, J;
    tri_index_cols(work / nfp, T, I, J);
    const bool full = tile_full(I * TILE, J * TILE, naux);
    if (single) {
        if (full)
            pi_tile_bf16x3_ws2_body<false, 1>(bt_hi, bt_lo, bt_f32, s32, kpad, naux, npad, out,
                                              mode, I * TILE, J * TILE, f0, f0, L);
        else
            pi_tile_bf16x3_ws2_body<true, 1>(bt_hi, bt_lo, bt_f32, s32, kpad, naux, npad, out,
                                             mode, I * TILE, J * TILE, f0, f0, L);
    } else if (full) {
        pi_tile_bf16x3_ws2_body<false>(bt_hi, bt_lo, bt_f32, s32, kpad, naux, npad, out, mode,
                                       I * TILE, J * TILE, f0, f1, L);
    } else {
        pi_tile_bf16x3_ws2_body<true>(bt_hi, bt_lo, bt_f32, s32, kpad, naux, npad, out, mode,
                                      I * TILE, J * TILE, f0, f1, L);
    }
}

// Double-buffered variants: two LDS buffers, one barrier per chunk.  While the MFMAs of
// chunk i read buffer i % 2, the same waves split chunk i + 1 (already in registers) into the
// other buffer and issue the global loads of chunk i + 2, so the split work and the load
// latency overlap the matrix work.  KB = 32: 64 KB of LDS (two workgroups per CU); KB = 64:
// 128 KB (one workgroup per CU, two MFMA k-steps per chunk).  Rows of KB bf16 are unpadded
// with the 16-byte chunk index XOR-swizzled ((row / 4) % 4 for 4 chunks per row, (row / 2) %
// 8 for 8): the 16 rows x 4 chunks of a fragment read or a 64-thread staging write cover all
// 16 bank groups once.
template <int KB>
__device__ __forceinline__ int db_off(int buf, int arr, int row, int chunk) {
    constexpr int CPR = KB / 8;
    const int sw = CPR == 4 ? (row >> 2) & 3 : (row >> 1) & 7;
    return ((buf * 4 + arr) * TILE + row) * KB + ((chunk ^ sw) << 3);
}

// L: 2 * 4 * TILE * KB bf16 of LDS (16-byte aligned).
template <int KB, bool Guarded>
__device__ __forceinline__ void pi_tile_bf16x3_db_body(
    const uint16_t* __restrict__ bt_hi, const uint16_t* __restrict__ bt_lo,
    const float* __restrict__ bt_f32, const float* __restrict__ s32, int kpad, int naux,
    int npad, double* __restrict__ out, PiOut mode, int I0, int J0, int f, __bf16* L) {
    constexpr int TPR = KB / 8;                  // threads per row (one 16-byte chunk each)
    constexpr int RSTEP = 256 / TPR;             // row stride between a thread's rows
    constexpr int R = TILE / RSTEP;              // rows per thread
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);   // wave-uniform (scalar)
    const int wr = wave >> 1, wc = wave & 1;
    const float* sf = s32 + (size_t)f * kpad;
    const int row0 = tid / TPR, ch0 = tid % TPR, kq0 = ch0 * 8;
    u16x8 rh[R], rl[R];
    f4 rc[2 * R], rs[2];
    auto fetch = [&](int k0) {
        #pragma unroll
        for (int q = 0; q < R; ++q) {
            const int row = row0 + q * RSTEP;
            // operand rows are allocated and zero-filled up to a multiple of 256 >= the
            // tile grid, so tile rows past npad load zeros without a bounds select
            const size_t oa = (size_t)(I0 + row) * kpad + k0 + kq0;
            const size_t oc = (size_t)(J0 + row) * kpad + k0 + kq0;
            rh[q] = *reinterpret_cast<const u16x8*>(bt_hi + oa);
            rl[q] = *reinterpret_cast<const u16x8*>(bt_lo + oa);
            rc[2 * q] = *reinterpret_cast<const f4*>(bt_f32 + oc);
            rc[2 * q + 1] = *reinterpret_cast<const f4*>(bt_f32 + oc + 4);
        }
        rs[0] = *reinterpret_cast<const f4*>(sf + k0 + kq0);
        rs[1] = *reinterpret_cast<const f4*>(sf + k0 + kq0 + 4);
    };
    auto stage = [&](int buf) {
        #pragma unroll
        for (int q = 0; q < R; ++q) {
            const int row = row0 + q * RSTEP;
            *reinterpret_cast<u16x8*>(L + db_off<KB>(buf, 0, row, ch0)) = rh[q];
            *reinterpret_cast<u16x8*>(L + db_off<KB>(buf, 1, row, ch0)) = rl[q];
            #pragma unroll
            for (int hq = 0; hq < 2; ++hq) {
                const f4 v = rc[2 * q + hq] * rs[hq];
                const bf16x4 h = __builtin_convertvector(v, bf16x4);
                const bf16x4 l = __builtin_convertvector(v - __builtin_convertvector(h, f4), bf16x4);
                *reinterpret_cast<bf16x4*>(L + db_off<KB>(buf, 2, row, ch0) + 4 * hq) = h;
                *reinterpret_cast<bf16x4*>(L + db_off<KB>(buf, 3, row, ch0) + 4 * hq) = l;
            }
        }
    };

    f4 acc[4][4];
    #pragma unroll
    for (int bi = 0; bi < 4; ++bi)
        #pragma unroll
        for (int bj = 0; bj < 4; ++bj) acc[bi][bj] = f4{0.f, 0.f, 0.f, 0.f};
    bool live[4][4];
    #pragma unroll
    for (int bi = 0; bi < 4; ++bi)
        #pragma unroll
        for (int bj = 0; bj < 4; ++bj) live[bi][bj] = block_live(I0, J0, wr, wc, bi, bj, naux);

    const int nch = kpad / KB;
    fetch(0);
    stage(0);
    if (nch > 1) fetch(KB);
    __syncthreads();
    const int r16 = lane & 15, kc = lane >> 4;
    for (int i = 0; i < nch; ++i) {
        const int b = i & 1;
        #pragma unroll
        for (int ks = 0; ks < KB / 32; ++ks) {
            bf16x8 ah[4], al[4], ch[4], cl[4];
            #pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int ra = wr * 64 + q * 16 + r16, rcn = wc * 64 + q * 16 + r16;
                ah[q] = *reinterpret_cast<const bf16x8*>(L + db_off<KB>(b, 0, ra, 4 * ks + kc));
                al[q] = *reinterpret_cast<const bf16x8*>(L + db_off<KB>(b, 1, ra, 4 * ks + kc));
                ch[q] = *reinterpret_cast<const bf16x8*>(L + db_off<KB>(b, 2, rcn, 4 * ks + kc));
                cl[q] = *reinterpret_cast<const bf16x8*>(L + db_off<KB>(b, 3, rcn, 4 * ks + kc));
            }
            mfma_bf16x3<Guarded>(acc, ah, al, ch, cl, live);
        }
        if (i + 1 < nch) stage(b ^ 1);
        if (i + 2 < nch) fetch((i + 2) * KB);
        __syncthreads();
    }
    store_tile<float, false>(acc, I0, J0, wr, wc, lane, npad, naux, f, mode, out, 1.0);
}

template <int KB>
__global__ __launch_bounds__(256, KB == 32 ? 2 : 1) void pi_syrk_bf16x3_db_kernel(
    const uint16_t* __restrict__ bt_hi, const uint16_t* __restrict__ bt_lo,
    const float* __restrict__ bt_f32, const float* __restrict__ s32, int kpad, int naux,
    int npad, int nf, int ntiles, double* __restrict__ out, PiOut mode) {
    __shared__ __attribute__((aligned(16))) __bf16 L[2 * 4 * TILE * KB];
    const int work = xcd_remap(blockIdx.x, ntiles * nf);
    const int tile = work / nf, f = work % nf;
    int I, J;
    tri_index(tile, I, J);
    if (tile_full(I * TILE, J * TILE, naux))
        pi_tile_bf16x3_db_body<KB, false>(bt_hi, bt_lo, bt_f32, s32, kpad, naux, npad, out, mode,
                                          I * TILE, J * TILE, f, L);
    else
        pi_tile_bf16x3_db_body<KB, true>(bt_hi, bt_lo, bt_f32, s32, kpad, naux, npad, out, mode,
                                         I * TILE, J * TILE, f, L);
}

template <int KB>
__global__ __launch_bounds__(256, 2) void pi_syrk_bf16x3_kernel(
    const uint16_t* __restrict__ bt_hi, const uint16_t* __restrict__ bt_lo,
    const float* __restrict__ bt_f32, const float* __restrict__ s32, int kpad, int naux,
    int npad, int nf, int ntiles, double* __restrict__ out, PiOut mode) {
    constexpr int KBP = PiLds<KB>::pitch;
    __shared__ __attribute__((aligned(16))) __bf16 Ah[TILE][KBP];
    __shared__ __attribute__((aligned(16))) __bf16 Al[TILE][KBP];
    __shared__ __attribute__((aligned(16))) __bf16 Ch[TILE][KBP];
    __shared__ __attribute__((aligned(16))) __bf16 Cl[TILE][KBP];
    const int work = xcd_remap(blockIdx.x, ntiles * nf);
    const int tile = work / nf, f = work % nf;
    int I, J;
    tri_index(tile, I, J);
    if (tile_full(I * TILE, J * TILE, naux))
        pi_tile_bf16x3_body<KB, false>(bt_hi, bt_lo, bt_f32, s32, kpad, naux, npad, out, mode,
                                       I * TILE, J * TILE, f, Ah, Al, Ch, Cl);
    else
        pi_tile_bf16x3_body<KB, true>(bt_hi, bt_lo, bt_f32, s32, kpad, naux, npad, out, mode,
                                      I * TILE, J * TILE, f, Ah, Al, Ch, Cl);
}

// ---- Grouped (batched) kernels ------------------------------------------------------------
// Persistent: gridDim.x (a multiple of 8) workgroups walk the concatenated work items of
// all instances (instance-major, then tile, then frequency).  In every sweep the G/8
// workgroups of one XCD take G/8 consecutive items, so the frequencies of one tile, which
// stream the same operand rows, share that XCD's L2.  Every workgroup leaves the loop once
// its next item is past the end.
__device__ __forceinline__ bool group_item(const PiInst* __restrict__ inst,
                                           const int* __restrict__ off, int ninst, int nf,
                                           int total, int it, PiInst& in, int& I0, int& J0,
                                           int& f) {
    const int G = gridDim.x;
    const int w = it * G + (blockIdx.x & 7) * (G >> 3) + (blockIdx.x >> 3);
    if (w >= total) return false;
    const int q = find_inst(off, ninst, w);
    in = inst[q];
    const int local = w - off[q];
    f = local % nf;
    int I, J;
    tri_index(local / nf, I, J);
    I0 = I * TILE;
    J0 = J * TILE;
    return true;
}

template <int KB>
__global__ __launch_bounds__(256, 2) void pi_group_bf16x3_kernel(
    const PiInst* __restrict__ inst, const int* __restrict__ off, int ninst, int nf, int total) {
    constexpr int KBP = PiLds<KB>::pitch;
    __shared__ __attribute__((aligned(16))) __bf16 Ah[TILE][KBP];
    __shared__ __attribute__((aligned(16))) __bf16 Al[TILE][KBP];
    __shared__ __attribute__((aligned(16))) __bf16 Ch[TILE][KBP];
    __shared__ __attribute__((aligned(16))) __bf16 Cl[TILE][KBP];
    PiInst in;
    int I0, J0, f;
    for (int it = 0; group_item(inst, off, ninst, nf, total, it, in, I0, J0, f); ++it)
        if (tile_full(I0, J0, in.naux))
            pi_tile_bf16x3_body<KB, false>(in.bt_hi, in.bt_lo, in.bt_f32, in.s32, in.kpad,
                                           in.naux, in.npad, in.out, PiOut::EpsLower, I0, J0, f,
                                           Ah, Al, Ch, Cl);
        else
            pi_tile_bf16x3_body<KB, true>(in.bt_hi, in.bt_lo, in.bt_f32, in.s32, in.kpad,
                                          in.naux, in.npad, in.out, PiOut::EpsLower, I0, J0, f,
                                          Ah, Al, Ch, Cl);
}

__global__ __launch_bounds__(256, 2) void pi_group_bf16x3_m32_kernel(
    const PiInst* __restrict__ inst, const int* __restrict__ off, int ninst, int nf, int total) {
    constexpr int KBP = PiLds<64>::pitch;
    __shared__ __attribute__((aligned(16))) __bf16 Ah[TILE][KBP];
    __shared__ __attribute__((aligned(16))) __bf16 Al[TILE][KBP];
    __shared__ __attribute__((aligned(16))) __bf16 Ch[TILE][KBP];
    __shared__ __attribute__((aligned(16))) __bf16 Cl[TILE][KBP];
    PiInst in;
    int I0, J0, f;
    for (int it = 0; group_item(inst, off, ninst, nf, total, it, in, I0, J0, f); ++it)
        if (tile_full(I0, J0, in.naux))
            pi_tile_bf16x3_m32_body<64, false>(in.bt_hi, in.bt_lo, in.bt_f32, in.s32, in.kpad,
                                               in.naux, in.npad, in.out, PiOut::EpsLower, I0, J0,
                                               f, Ah, Al, Ch, Cl);
        else
            pi_tile_bf16x3_m32_body<64, true>(in.bt_hi, in.bt_lo, in.bt_f32, in.s32, in.kpad,
                                              in.naux, in.npad, in.out, PiOut::EpsLower, I0, J0,
                                              f, Ah, Al, Ch, Cl);
}

// Pairs of frequencies (f, f + 1) of one tile and instance, for the two-frequency
// producer/consumer body; valid when nf is even (every instance's item count is even, so
// items 2 w and 2 w + 1 are such a pair).
__device__ __forceinline__ bool group_pair_item(const PiInst* __restrict__ inst,
                                                const int* __restrict__ off, int ninst, int nf,
                                                int total, int it, PiInst& in, int& I0, int& J0,
                                                int& f) {
    const int G = gridDim.x;
    const int w2 = it * G + (blockIdx.x & 7) * (G >> 3) + (blockIdx.x >> 3);
    if (2 * w2 >= total) return false;
    const int w = 2 * w2;
    const int q = find_inst(off, ninst, w);
    in = inst[q];
    const int local = w - off[q];
    f = local % nf;
    int I, J;
    tri_index_cols(local / nf, (in.npad + TILE - 1) / TILE, I, J);
    I0 = I * TILE;
    J0 = J * TILE;
    return true;
}

__global__ __launch_bounds__(512, 1) void pi_group_bf16x3_ws2_kernel(
    const PiInst* __restrict__ inst, const int* __restrict__ off, int ninst, int nf, int total) {
    __shared__ __attribute__((aligned(16))) __bf16 L[2 * WS2BUF];
    PiInst in;
    int I0, J0, f;
    for (int it = 0; group_pair_item(inst, off, ninst, nf, total, it, in, I0, J0, f); ++it)
        if (tile_full(I0, J0, in.naux))
            pi_tile_bf16x3_ws2_body<false>(in.bt_hi, in.bt_lo, in.bt_f32, in.s32, in.kpad,
                                           in.naux, in.npad, in.out, PiOut::EpsLower, I0, J0, f,
                                           f + 1, L);
        else
            pi_tile_bf16x3_ws2_body<true>(in.bt_hi, in.bt_lo, in.bt_f32, in.s32, in.kpad,
                                          in.naux, in.npad, in.out, PiOut::EpsLower, I0, J0, f,
                                          f + 1, L);
}

__global__ __launch_bounds__(256, 2) void pi_group_bf16x3_db_kernel(
    const PiInst* __restrict__ inst, const int* __restrict__ off, int ninst, int nf, int total) {
    __shared__ __attribute__((aligned(16))) __bf16 L[2 * 4 * TILE * 32];
    PiInst in;
    int I0, J0, f;
    for (int it = 0; group_item(inst, off, ninst, nf, total, it, in, I0, J0, f); ++it)
        if (tile_full(I0, J0, in.naux))
            pi_tile_bf16x3_db_body<32, false>(in.bt_hi, in.bt_lo, in.bt_f32, in.s32, in.kpad,
                                          in.naux, in.npad, in.out, PiOut::EpsLower, I0, J0, f,
                                          L);
        else
            pi_tile_bf16x3_db_body<32, true>(in.bt_hi, in.bt_lo, in.bt_f32, in.s32, in.kpad,
                                         in.naux, in.npad, in.out, PiOut::EpsLower, I0, J0, f,
                                         L);
}

__global__ __launch_bounds__(256, 2) void pi_group_f64_kernel(
    const PiInst* __restrict__ inst, const int* __restrict__ off, int ninst, int nf, int total) {
    __shared__ double As[16][TILE + 2];
    __shared__ double Bs[16][TILE + 2];
    PiInst in;
    int I0, J0, f;
    for (int it = 0; group_item(inst, off, ninst, nf, total, it, in, I0, J0, f); ++it)
        if (tile_full(I0, J0, in.naux))
            pi_tile_body<double, 16, false>(in.B, in.s, in.nov, in.naux, in.npad, in.kpad, in.out,
                                            PiOut::EpsLower, I0, J0, f, As, Bs);
        else
            pi_tile_body<double, 16, true>(in.B, in.s, in.nov, in.naux, in.npad, in.kpad, in.out,
                                           PiOut::EpsLower, I0, J0, f, As, Bs);
}

// FP64 with two frequencies per tile (even nf): one workgroup per item, an item being a
// (tile, frequency pair) of one instance (items 2 w, 2 w + 1 of the concatenated list, tiles
// column by column), run by the two-frequency body of the single-instance kernel (K chunks
// of 32 rows, one LDS buffer).  Grid total / 2; xcd_remap gives each XCD a contiguous range,
// so the frequency pairs and neighbouring tiles that stream the same rows share its L2.  Not
// persistent: a loop would keep the instance descriptor live across the body, which then
// spills.
__global__ __launch_bounds__(512, 1) void pi_group_f64_pair_kernel(
    const PiInst* __restrict__ inst, const int* __restrict__ off, int ninst, int nf, int total) {
    __shared__ double As[32][TILE + 2];
    __shared__ double Bs[32][TILE + 2];
    __shared__ double Ss[2][32];
    const int w = 2 * xcd_remap(blockIdx.x, total / 2);
    const int q = find_inst(off, ninst, w);
    const PiInst in = inst[q];
    const int local = w - off[q], f = local % nf;
    int I, J;
    tri_index_cols(local / nf, (in.npad + TILE - 1) / TILE, I, J);
    if (tile_full(I * TILE, J * TILE, in.naux))
        pi_f64_pair_body<32, false, false>(in.B, in.s, in.nov, in.naux, in.npad, in.kpad, nf,
                                           in.out, PiOut::EpsLower, I * TILE, J * TILE, f, As,
                                           Bs, Ss);
    else
        pi_f64_pair_body<32, false, true>(in.B, in.s, in.nov, in.naux, in.npad, in.kpad, nf,
                                          in.out, PiOut::EpsLower, I * TILE, J * TILE, f, As, Bs,
                                          Ss);
}

// Wide variant: 256 (A, unscaled) x 128 (C = s B, split on the fly) tile per 512-thread
// workgroup, 8 waves of 64x64.  Converting only the narrow side halves the split work per
// MFMA.  LDS rows are 64 bf16 (8 chunks of 16 B) with the chunk index XOR-swizzled by
// (row / 2) % 8, which keeps both the staging writes and the MFMA fragment reads free of
// bank conflicts.  Operand rows are allocated to a multiple of 256 (zero padding), so the
// loads need no bounds selects.
constexpr int WTI = 256, WTJ = 128, WKB = 64;

__device__ __forceinline__ int swz(int row, int chunk) { return chunk ^ ((row >> 1) & 7); }

__global__ __launch_bounds__(512, 1) void pi_syrk_bf16x3_wide_kernel(
    const uint16_t* __restrict__ bt_hi, const uint16_t* __restrict__ bt_lo,
    const float* __restrict__ bt_f32, const float* __restrict__ s32, int kpad, int naux,
    int npad, int nf, int ntiles, int TJ, double* __restrict__ out, PiOut mode) {
    __shared__ __attribute__((aligned(16))) __bf16 Ah[WTI * WKB];
    __shared__ __attribute__((aligned(16))) __bf16 Al[WTI * WKB];
    __shared__ __attribute__((aligned(16))) __bf16 Ch[WTJ * WKB];
    __shared__ __attribute__((aligned(16))) __bf16 Cl[WTJ * WKB];
    const int work = xcd_remap(blockIdx.x, ntiles * nf);
    int tile = work / nf;
    const int f = work % nf;
    int a = 0;                                   // tiles of block row a: b = 0..min(2a+1, TJ-1)
    for (;;) {
        const int cnt = 2 * a + 2 < TJ ? 2 * a + 2 : TJ;
        if (tile < cnt) break;
        tile -= cnt;
        ++a;
    }
    const int I0 = a * WTI, J0 = tile * WTJ;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);   // wave-uniform (scalar)
    const int wr = wave >> 1, wc = wave & 1;
    const float* sf = s32 + (size_t)f * kpad;

    const int row0 = tid >> 3, ch0 = tid & 7;    // staging: 8 threads per row, 16 B each
    const uint16_t* ph = bt_hi + (size_t)(I0 + row0) * kpad + ch0 * 8;
    const uint16_t* pl = bt_lo + (size_t)(I0 + row0) * kpad + ch0 * 8;
    const float* pc = bt_f32 + (size_t)(J0 + row0) * kpad + ch0 * 8;
    const size_t rstep = (size_t)64 * kpad;
    u16x8 rh[4], rl[4];
    f4 rc[4], rs[2];
    auto fetch = [&](int k0) {
        #pragma unroll
        for (int q = 0; q < 4; ++q) {
            rh[q] = *reinterpret_cast<const u16x8*>(ph + q * rstep + k0);
            rl[q] = *reinterpret_cast<const u16x8*>(pl + q * rstep + k0);
        }
        #pragma unroll
        for (int q = 0; q < 2; ++q) {
            rc[2 * q] = *reinterpret_cast<const f4*>(pc + q * rstep + k0);
            rc[2 * q + 1] = *reinterpret_cast<const f4*>(pc + q * rstep + k0 + 4);
        }
        rs[0] = *reinterpret_cast<const f4*>(sf + k0 + ch0 * 8);
        rs[1] = *reinterpret_cast<const f4*>(sf + k0 + ch0 * 8 + 4);
    };

    f4 acc[4][4];
    bool live[4][4];
    #pragma unroll
    for (int bi = 0; bi < 4; ++bi)
        #pragma unroll
        for (int bj = 0; bj < 4; ++bj) {
            acc[bi][bj] = f4{0.f, 0.f, 0.f, 0.f};
            const int r = I0 + wr * 64 + bi * 16, c = J0 + wc * 64 + bj * 16;
            live[bi][bj] = r < naux && c < naux && r >= c;
        }
    bool any = false;
    #pragma unroll
    for (int bi = 0; bi < 4; ++bi)
        #pragma unroll
        for (int bj = 0; bj < 4; ++bj) any |= live[bi][bj];

    fetch(0);
    for (int k0 = 0; k0 < kpad; k0 += WKB) {
        #pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int R = row0 + 64 * q;
            const int o = R * WKB + 8 * swz(R, ch0);
            *reinterpret_cast<u16x8*>(&Ah[o]) = rh[q];
            *reinterpret_cast<u16x8*>(&Al[o]) = rl[q];
        }
        #pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int R = row0 + 64 * q;
            const int o = R * WKB + 8 * swz(R, ch0);
            #pragma unroll
            for (int hq = 0; hq < 2; ++hq) {
                const f4 v = rc[2 * q + hq] * rs[hq];
                const bf16x4 h = __builtin_convertvector(v, bf16x4);
                const bf16x4 l = __builtin_convertvector(v - __builtin_convertvector(h, f4), bf16x4);
                *reinterpret_cast<bf16x4*>(&Ch[o + 4 * hq]) = h;
                *reinterpret_cast<bf16x4*>(&Cl[o + 4 * hq]) = l;
            }
        }
        __syncthreads();
        if (k0 + WKB < kpad) fetch(k0 + WKB);
        if (any) {
            #pragma unroll
            for (int ks = 0; ks < WKB; ks += 32) {
                const int cq = ks / 8 + (lane >> 4), r16 = lane & 15;
                bf16x8 ah[4], al[4], chv[4], clv[4];
                #pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int Ra = wr * 64 + q * 16 + r16, Rc = wc * 64 + q * 16 + r16;
                    const int oa = Ra * WKB + 8 * swz(Ra, cq), oc = Rc * WKB + 8 * swz(Rc, cq);
                    ah[q] = *reinterpret_cast<const bf16x8*>(&Ah[oa]);
                    al[q] = *reinterpret_cast<const bf16x8*>(&Al[oa]);
                    chv[q] = *reinterpret_cast<const bf16x8*>(&Ch[oc]);
                    clv[q] = *reinterpret_cast<const bf16x8*>(&Cl[oc]);
                }
                mfma_bf16x3<true>(acc, ah, al, chv, clv, live);
            }
        }
        __syncthreads();
    }
    #pragma unroll
    for (int bi = 0; bi < 4; ++bi)
        #pragma unroll
        for (int bj = 0; bj < 4; ++bj)
            #pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = I0 + wr * 64 + bi * 16 + mfma_row_f32(lane, v);
                const int col = J0 + wc * 64 + bj * 16 + (lane & 15);
                const double p = (double)acc[bi][bj][v];
                if (mode == PiOut::EpsLower) {
                    if (row < npad && col < npad)
                        out[((size_t)f * npad + row) * npad + col] = (row == col ? 1.0 : 0.0) - p;
                } else if (row < naux && col < naux && row >= col) {
                    out[((size_t)f * npad + row) * npad + col] = p;
                    out[((size_t)f * npad + col) * npad + row] = p;
                }
            }
}

}  // namespace

hipError_t launch_pi_weights(const double* eps, int nocc, int nvir, int nf,
                             const double* freqs, double* s, float* s32, hipStream_t st) {
    const int ld = round_up(nocc * nvir, KPAD);
    dim3 grid(ceil_div(ld, 256) < 1024 ? ceil_div(ld, 256) : 1024, nf);
    hipLaunchKernelGGL(pi_weights_kernel, grid, dim3(256), 0, st, eps, nocc, nvir, nf, freqs,
                       ld, s, s32);
    return hipGetLastError();
}

hipError_t launch_pi_prepare_mixed(const double* B, int nov, int naux, int npad, int kpad,
                                   uint16_t* bt_hi, uint16_t* bt_lo, float* bt_f32,
                                   hipStream_t st) {
    dim3 grid(ceil_div(kpad, 32), ceil_div(npad, 32));
    hipLaunchKernelGGL(pi_prepare_mixed_kernel, grid, dim3(256), 0, st, B, nov, naux, npad,
                       kpad, bt_hi, bt_lo, bt_f32);
    return hipGetLastError();
}

// Compute units of the current device (cached per device).
static int device_cu_count() {
    static int cached[64] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 0;
    if (cached[dev] == 0) {
        int n = 0;
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
            return 0;
        cached[dev] = n;
    }
    return cached[dev];
}

hipError_t launch_pi_syrk(int precision, const double* B, const PiOperands& ops,
                          const double* s, int nov, int naux, int npad, int nf, double* out,
                          PiOut mode, hipStream_t st, int variant) {
    const int T = ceil_div(mode == PiOut::EpsLower ? npad : naux, TILE);
    const int ntiles = T * (T + 1) / 2;
    const int sld = round_up(nov, KPAD);
    dim3 grid(ntiles * nf), block(256);
    if (precision == 0) {
        // one 512-thread workgroup per CU: the items past the last full round of CUs run as
        // single-frequency halves when those fit in one round (QXGW_PI_TAIL=0: never)
        const int nitems = ntiles * ((nf + 1) / 2), cus = device_cu_count();
        const int ntail = cus > 0 ? nitems % cus : 0;
        const char* tv = std::getenv("QXGW_PI_TAIL");
        const bool tail = !(tv && std::atoi(tv) == 0) && ntail > 0 && 2 * ntail <= cus;
        const int nfull = tail ? nitems - ntail : nitems;
        const dim3 gp(nfull + 2 * (nitems - nfull));
        if (variant == 1)
            hipLaunchKernelGGL((pi_f64_pair_kernel<16, false>), gp, dim3(512), 0, st, B, s, nov,
                               naux, npad, nf, ntiles, nfull, sld, out, mode);
        else if (variant == 2)
            hipLaunchKernelGGL((pi_f64_pair_kernel<32, false>), gp, dim3(512), 0, st, B, s, nov,
                               naux, npad, nf, ntiles, nfull, sld, out, mode);
        else if (variant == 3)
            hipLaunchKernelGGL((pi_f64_pair_kernel<32, true>), gp, dim3(512), 0, st, B, s, nov,
                               naux, npad, nf, ntiles, nfull, sld, out, mode);
        else if (variant == 4)
            hipLaunchKernelGGL((pi_f64_pair_kernel<16, true>), gp, dim3(512), 0, st, B, s, nov,
                               naux, npad, nf, ntiles, nfull, sld, out, mode);
        else
            hipLaunchKernelGGL((pi_syrk_kernel<double, 16>), grid, block, 0, st, B, s, nov, naux,
                               npad, nf, ntiles, sld, out, mode);
    } else if (precision == 1) {
        hipLaunchKernelGGL((pi_syrk_kernel<float, 32>), grid, block, 0, st, B, s, nov, naux,
                           npad, nf, ntiles, sld, out, mode);
    } else {
        if (!ops.bt_hi || !ops.s32 || ops.kpad != sld || ops.rows < T * TILE)
            return hipErrorInvalidValue;
        if (variant == 2) {
            if (ops.rows < round_up(npad > naux ? npad : naux, WTI)) return hipErrorInvalidValue;
            const int TJ = ceil_div(mode == PiOut::EpsLower ? npad : naux, WTJ);
            const int TIw = ceil_div(mode == PiOut::EpsLower ? npad : naux, WTI);
            int nt = 0;
            for (int a = 0; a < TIw; ++a) nt += 2 * a + 2 < TJ ? 2 * a + 2 : TJ;
            hipLaunchKernelGGL(pi_syrk_bf16x3_wide_kernel, dim3(nt * nf), dim3(512), 0, st,
                               ops.bt_hi, ops.bt_lo, ops.bt_f32, ops.s32, ops.kpad, naux, npad, nf,
                               nt, TJ, out, mode);
        } else if (variant == 3)
            hipLaunchKernelGGL(pi_syrk_bf16x3_db_kernel<32>, grid, block, 0, st, ops.bt_hi,
                               ops.bt_lo, ops.bt_f32, ops.s32, ops.kpad, naux, npad, nf, ntiles,
                               out, mode);
        else if (variant == 4)
            hipLaunchKernelGGL(pi_syrk_bf16x3_db_kernel<64>, grid, block, 0, st, ops.bt_hi,
                               ops.bt_lo, ops.bt_f32, ops.s32, ops.kpad, naux, npad, nf, ntiles,
                               out, mode);
        else if (variant == 5)
            hipLaunchKernelGGL(pi_syrk_bf16x3_m32_kernel<64>, grid, block, 0, st, ops.bt_hi,
                               ops.bt_lo, ops.bt_f32, ops.s32, ops.kpad, naux, npad, nf, ntiles,
                               out, mode);
        else if (variant == 6)
            hipLaunchKernelGGL(pi_syrk_bf16x3_m32_kernel<32>, grid, block, 0, st, ops.bt_hi,
                               ops.bt_lo, ops.bt_f32, ops.s32, ops.kpad, naux, npad, nf, ntiles,
                               out, mode);
        else if (variant == 7 || variant == 9 || variant == 10) {
            if (ops.kpad % 64) return hipErrorInvalidValue;
            if (variant == 10)
                hipLaunchKernelGGL((pi_syrk_bf16x3_ws_kernel<64, 8>), grid, dim3(768), 0, st,
                                   ops.bt_hi, ops.bt_lo, ops.bt_f32, ops.s32, ops.kpad, naux,
                                   npad, nf, ntiles, out, mode);
            else
                hipLaunchKernelGGL(variant == 7 ? (pi_syrk_bf16x3_ws_kernel<64, 4>)
                                                : (pi_syrk_bf16x3_ws_kernel<32, 4>),
                                   grid, dim3(512), 0, st, ops.bt_hi, ops.bt_lo, ops.bt_f32,
                                   ops.s32, ops.kpad, naux, npad, nf, ntiles, out, mode);
        } else if (variant == 8) {
            if (ops.kpad % WS2K) return hipErrorInvalidValue;
            // one 120 KB workgroup per CU: the items past the last full round of CUs run as
            // single-frequency halves when those fit in one round (QXGW_PI_TAIL=0: never)
            const int nitems = ntiles * ((nf + 1) / 2), cus = device_cu_count();
            const int ntail = cus > 0 ? nitems % cus : 0;
            const char* tv = std::getenv("QXGW_PI_TAIL");
            const bool tail = !(tv && std::atoi(tv) == 0) && ntail > 0 && 2 * ntail <= cus;
            const int nfull = tail ? nitems - ntail : nitems;
            hipLaunchKernelGGL(pi_syrk_bf16x3_ws2_kernel, dim3(nfull + 2 * (nitems - nfull)),
                               dim3(512), 0, st, ops.bt_hi, ops.bt_lo, ops.bt_f32, ops.s32,
                               ops.kpad, naux, npad, nf, ntiles, nfull, out, mode);
        }
        else if (variant == 1)
            hipLaunchKernelGGL(pi_syrk_bf16x3_kernel<64>, grid, block, 0, st, ops.bt_hi, ops.bt_lo,
                               ops.bt_f32, ops.s32, ops.kpad, naux, npad, nf, ntiles, out, mode);
        else
            hipLaunchKernelGGL(pi_syrk_bf16x3_kernel<32>, grid, block, 0, st, ops.bt_hi, ops.bt_lo,
                               ops.bt_f32, ops.s32, ops.kpad, naux, npad, nf, ntiles, out, mode);
    }
    return hipGetLastError();
}

int pi_variant_count(int precision) { return precision == 2 ? 11 : precision == 0 ? 5 : 1; }

hipError_t launch_pi_group(int precision, const PiInst* inst_dev, const int* off_dev, int ninst,
                           int nf, int total, int grid, hipStream_t st, int variant) {
    if (ninst <= 0 || total <= 0) return hipSuccess;
    grid = std::max(8, grid / 8 * 8);
    if (precision == 0 && variant == 1) {                  // one workgroup per item
        if (nf % 2) return hipErrorInvalidValue;
        hipLaunchKernelGGL(pi_group_f64_pair_kernel, dim3(total / 2), dim3(512), 0, st, inst_dev,
                           off_dev, ninst, nf, total);
    } else if (precision == 0)
        hipLaunchKernelGGL(pi_group_f64_kernel, dim3(grid), dim3(256), 0, st, inst_dev, off_dev,
                           ninst, nf, total);
    else if (precision == 2 && variant == 4) {
        if (nf % 2) return hipErrorInvalidValue;
        hipLaunchKernelGGL(pi_group_bf16x3_ws2_kernel, dim3(grid), dim3(512), 0, st, inst_dev,
                           off_dev, ninst, nf, total);
    } else if (precision == 2 && variant == 3)
        hipLaunchKernelGGL(pi_group_bf16x3_m32_kernel, dim3(grid), dim3(256), 0, st, inst_dev,
                           off_dev, ninst, nf, total);
    else if (precision == 2 && variant == 2)
        hipLaunchKernelGGL(pi_group_bf16x3_db_kernel, dim3(grid), dim3(256), 0, st, inst_dev,
                           off_dev, ninst, nf, total);
    else if (precision == 2 && variant == 1)
        hipLaunchKernelGGL(pi_group_bf16x3_kernel<64>, dim3(grid), dim3(256), 0, st, inst_dev,
                           off_dev, ninst, nf, total);
    else if (precision == 2)
        hipLaunchKernelGGL(pi_group_bf16x3_kernel<32>, dim3(grid), dim3(256), 0, st, inst_dev,
                           off_dev, ninst, nf, total);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

int pi_group_variant_count(int precision) { return precision == 2 ? 5 : precision == 0 ? 2 : 1; }

int pi_group_occupancy(int precision, int variant) {
    int n = 0;
    hipError_t e;
    if (precision == 0 && variant == 1)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, pi_group_f64_pair_kernel, 512, 0);
    else if (precision == 0)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, pi_group_f64_kernel, 256, 0);
    else if (variant == 4)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, pi_group_bf16x3_ws2_kernel, 512, 0);
    else if (variant == 3)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, pi_group_bf16x3_m32_kernel, 256, 0);
    else if (variant == 2)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, pi_group_bf16x3_db_kernel, 256, 0);
    else if (variant == 1)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, pi_group_bf16x3_kernel<64>, 256, 0);
    else
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, pi_group_bf16x3_kernel<32>, 256, 0);
    return e == hipSuccess && n > 0 ? n : 1;
}

}  // namespace qx
