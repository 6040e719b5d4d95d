// Screened-interaction contraction, CD self-energy quadrature, residues and the Newton
// quasiparticle solve.
//
// Reference:
//   * Wc_nm(i w) = b_nm^T Wc(i w) b_nm: quasix_core/src/selfenergy/correlation_ac.rs:404-421
//     (per-target DGEMM q_nm = Wc^T L[:,n,:], column dot with L[:,:,n]).  Here the pairs of
//     all targets form one matrix Bp [npair][naux] and, with eps = L L^T and M = L^-1 - I,
//         Wc_nm = |L^-1 b|^2 - |b|^2 = 2 b.(M b) + |M b|^2,
//     one triangular GEMM with the row dots fused into its epilogue (no cancellation: the
//     product M b is small exactly when Wc is small).
//   * Imaginary-axis quadrature with the eta-shifted Green's function:
//     correlation_pyscf_cd.rs:209-226 (emo = E - eps_m - i eta sign(eps_n - eps_m),
//     g = w_k emo / (emo^2 + w_k^2), Sigma = -(1/pi) sum W g).
//   * Residues Wc_nm(|E - eps_m| + i eta) for the poles between the Fermi level and E
//     (contour deformation; the reference's CDIntegrator is unwired).
//   * Newton: quasix_core/src/qp/newton_solver.rs:174-391 (Richardson 5-point derivative,
//     adaptive damping, step clip, bisection fallback); iteration-0 linearised guess
//     quasix_core/src/qp/evgw.rs:730-741.
#include "../../include/qxgw.h"
#include "qx_internal.h"
#include "qx_newton.h"

#include <algorithm>
#include <cstdlib>

namespace qx {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
constexpr int P = NB + 1;
constexpr double kPi = 3.14159265358979323846;

__device__ __forceinline__ d4 mfma64(double a, double b, d4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// ---- Wc_nm contraction -----------------------------------------------------------------
// Tile of 64 pairs (from j0) x 64 auxiliary columns (block pb) of frequency f; Bs, Ms:
// NB x P doubles of LDS.  The K loop runs over the blocks qb <= pb of the triangular M; the
// next block's 2 x 16 loads per thread are in flight during the current block's MFMAs, and
// the last staged pair block (qb = pb) is the one the row-dot epilogue needs.
__device__ __forceinline__ void wnm_linv_body(const double* __restrict__ bp,
                                              const double* __restrict__ x, int npair, int npad,
                                              int nblk, double* __restrict__ wpart, int j0,
                                              int pb, int f, double (*Bs)[P], double (*Ms)[P]) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const double* X = x + (size_t)f * npad * npad;
    // staging: thread owns columns 2 c2, 2 c2 + 1 of rows r0 + 8 q (16-byte loads)
    constexpr int PER = NB * NB / 512;
    const int c2 = (tid & 31) * 2, r0 = tid >> 5;
    double2 rb[PER], rm[PER];
    auto fetch = [&](int qb) {
        #pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int r = r0 + 8 * q, j = j0 + r;
            rb[q] = j < npair ? *reinterpret_cast<const double2*>(bp + (size_t)j * npad + qb * NB + c2)
                              : double2{0.0, 0.0};
            double2 m = *reinterpret_cast<const double2*>(X + (size_t)(pb * NB + r) * npad +
                                                          qb * NB + c2);
            if (qb == pb && r == c2) m.x -= 1.0;
            if (qb == pb && r == c2 + 1) m.y -= 1.0;
            rm[q] = m;
        }
    };
    d4 acc[4] = {d4{0, 0, 0, 0}, d4{0, 0, 0, 0}, d4{0, 0, 0, 0}, d4{0, 0, 0, 0}};
    fetch(0);
    for (int qb = 0; qb <= pb; ++qb) {
        #pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int r = r0 + 8 * q;
            Bs[r][c2] = rb[q].x;
            Bs[r][c2 + 1] = rb[q].y;
            Ms[r][c2] = rm[q].x;
            Ms[r][c2 + 1] = rm[q].y;
        }
        __syncthreads();
        if (qb < pb) fetch(qb + 1);
        const int r = wave * 16 + (lane & 15);
        #pragma unroll
        for (int kk = 0; kk < NB; kk += 4) {
            const int q = kk + (lane >> 4);
            const double a = Bs[r][q];
            #pragma unroll
            for (int bj = 0; bj < 4; ++bj) acc[bj] = mfma64(a, Ms[bj * 16 + (lane & 15)][q], acc[bj]);
        }
        if (qb < pb) __syncthreads();
    }
    // epilogue: Bs holds b[j][P] for P in block pb (the last staged block)
    double rs[4] = {0, 0, 0, 0};
    #pragma unroll
    for (int bj = 0; bj < 4; ++bj)
        #pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int r = wave * 16 + mfma_row_f64(lane, v), c = bj * 16 + (lane & 15);
            const double y = acc[bj][v];
            rs[v] += 2.0 * Bs[r][c] * y + y * y;
        }
    #pragma unroll
    for (int v = 0; v < 4; ++v)
        for (int off = 1; off < 16; off <<= 1) rs[v] += __shfl_xor(rs[v], off, 64);
    if ((lane & 15) == 0)
        #pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int j = j0 + wave * 16 + mfma_row_f64(lane, v);
            if (j < npair) wpart[((size_t)f * nblk + pb) * npair + j] = rs[v];
        }
}

// grid (ceil(npair/64), nblk, nf)
__global__ __launch_bounds__(256) void wnm_linv_kernel(const double* __restrict__ bp,
                                                       const double* __restrict__ x, int npair,
                                                       int npad, double* __restrict__ wpart) {
    __shared__ double Bs[NB][P];
    __shared__ double Ms[NB][P];
    wnm_linv_body(bp, x, npair, npad, gridDim.y, wpart, blockIdx.x * NB, blockIdx.y, blockIdx.z,
                  Bs, Ms);
}

// Grouped form: instance q owns ceil(npair_q/64) * nblk_q * nf consecutive workgroups.
__global__ __launch_bounds__(256) void wnm_linv_group_kernel(const WnmInst* __restrict__ inst,
                                                             const int* __restrict__ off,
                                                             int ninst, int nf) {
    __shared__ double Bs[NB][P];
    __shared__ double Ms[NB][P];
    const int w = blockIdx.x;
    const int q = find_inst(off, ninst, w);
    const WnmInst in = inst[q];
    const int nblk = in.npad / NB, njb = (in.npair + NB - 1) / NB;
    const int local = w - off[q];
    const int jb = local % njb, pb = (local / njb) % nblk, f = local / (njb * nblk);
    wnm_linv_body(in.bp, in.x, in.npair, in.npad, nblk, in.wpart, jb * NB, pb, f, Bs, Ms);
}

// Grouped reduction: one thread per (instance, frequency, pair); off: prefix of nf * npair.
__global__ void wnm_reduce_group_kernel(const WnmInst* __restrict__ inst,
                                        const int* __restrict__ off, int ninst, int total) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += gridDim.x * blockDim.x) {
        const int q = find_inst(off, ninst, i);
        const WnmInst in = inst[q];
        const int local = i - off[q];
        const int f = local / in.npair, j = local % in.npair, nblk = in.npad / NB;
        double s = 0.0;
        for (int b = 0; b < nblk; ++b) s += in.wpart[((size_t)f * nblk + b) * in.npair + j];
        in.wnm[(size_t)f * in.npair + j] = s;
    }
}

__global__ void wnm_reduce_kernel(const double* __restrict__ wpart, int npair, int nparts,
                                  double* __restrict__ wnm) {
    const int f = blockIdx.y;
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < npair; j += gridDim.x * blockDim.x) {
        double s = 0.0;
        for (int b = 0; b < nparts; ++b) s += wpart[((size_t)f * nparts + b) * npair + j];
        wnm[(size_t)f * npair + j] = s;
    }
}

// wnm from an explicit Wc [nf][npad][npad]: b^T Wc b, tile 64 pairs x 64 cols.
__global__ __launch_bounds__(256) void wnm_wc_kernel(const double* __restrict__ bp,
                                                     const double* __restrict__ wc, int npair,
                                                     int npad, double* __restrict__ wpart) {
    __shared__ double Bs[NB][P];
    __shared__ double Ws[NB][P];
    const int j0 = blockIdx.x * NB, pb = blockIdx.y, f = blockIdx.z;
    const int nblk = gridDim.y;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const double* W = wc + (size_t)f * npad * npad;
    d4 acc[4] = {d4{0, 0, 0, 0}, d4{0, 0, 0, 0}, d4{0, 0, 0, 0}, d4{0, 0, 0, 0}};
    for (int qb = 0; qb < nblk; ++qb) {
        for (int idx = tid; idx < NB * NB; idx += 256) {
            const int r = idx / NB, c = idx % NB;
            const int j = j0 + r;
            Bs[r][c] = j < npair ? bp[(size_t)j * npad + qb * NB + c] : 0.0;
            Ws[r][c] = W[(size_t)(pb * NB + r) * npad + qb * NB + c];
        }
        __syncthreads();
        const int r = wave * 16 + (lane & 15);
        #pragma unroll
        for (int kk = 0; kk < NB; kk += 4) {
            const int q = kk + (lane >> 4);
            const double a = Bs[r][q];
            #pragma unroll
            for (int bj = 0; bj < 4; ++bj) acc[bj] = mfma64(a, Ws[bj * 16 + (lane & 15)][q], acc[bj]);
        }
        __syncthreads();
    }
    for (int idx = tid; idx < NB * NB; idx += 256) {
        const int r = idx / NB, c = idx % NB;
        const int j = j0 + r;
        Bs[r][c] = j < npair ? bp[(size_t)j * npad + pb * NB + c] : 0.0;
    }
    __syncthreads();
    double rs[4] = {0, 0, 0, 0};
    #pragma unroll
    for (int bj = 0; bj < 4; ++bj)
        #pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int r = wave * 16 + mfma_row_f64(lane, v), c = bj * 16 + (lane & 15);
            rs[v] += Bs[r][c] * acc[bj][v];
        }
    #pragma unroll
    for (int v = 0; v < 4; ++v)
        for (int off = 1; off < 16; off <<= 1) rs[v] += __shfl_xor(rs[v], off, 64);
    if ((lane & 15) == 0)
        #pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int j = j0 + wave * 16 + mfma_row_f64(lane, v);
            if (j < npair) wpart[((size_t)f * nblk + pb) * npair + j] = rs[v];
        }
}

// ---- Wc_nm contraction, BF16x3 --------------------------------------------------------------
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef double dv4 __attribute__((ext_vector_type(4)));
constexpr int WT = 128;          // tile (pairs x auxiliary columns)
constexpr int WK = 32;           // K chunk
constexpr int WKP = 40;          // padded LDS pitch (bf16)

__device__ __forceinline__ void split4(const dv4 x, bf16x4& h, bf16x4& l) {
    const f4 xf = __builtin_convertvector(x, f4);
    h = __builtin_convertvector(xf, bf16x4);
    const dv4 r = x - __builtin_convertvector(__builtin_convertvector(h, f4), dv4);
    l = __builtin_convertvector(__builtin_convertvector(r, f4), bf16x4);
}

// hi/lo split of row P of M_f = X_f - I (row r = f npad + P of the [nf npad][npad] arrays).
// The Wc_nm kernels read row P only at columns Q < min(128 (P / 128) + 128, npad), their K
// range for the 128-row tile of P, so only those are written: X over the lower 64-blocks,
// zeros in the 64-block above the diagonal of an even block row.  The rest of the arrays is
// never read (and not written).
__device__ __forceinline__ void split_m_row(const double* __restrict__ x, int npad, size_t r,
                                            __bf16* __restrict__ mh, __bf16* __restrict__ ml) {
    const int P = r % npad;
    const int qlim = min((P / WT + 1) * WT, npad);
    const size_t base = r * npad;
    for (int Q = threadIdx.x * 4; Q < qlim; Q += blockDim.x * 4) {
        dv4 v = dv4{0.0, 0.0, 0.0, 0.0};
        if (Q / NB <= P / NB) {
            v = *reinterpret_cast<const dv4*>(x + base + Q);
            if (P >= Q && P < Q + 4) v[P - Q] -= 1.0;
        }
        bf16x4 h, l;
        split4(v, h, l);
        *reinterpret_cast<bf16x4*>(mh + base + Q) = h;
        *reinterpret_cast<bf16x4*>(ml + base + Q) = l;
    }
}

__global__ void split_m_kernel(const double* __restrict__ x, int npad, int nf,
                               __bf16* __restrict__ mh, __bf16* __restrict__ ml) {
    const size_t rows = (size_t)nf * npad;
    for (size_t r = blockIdx.x; r < rows; r += gridDim.x) split_m_row(x, npad, r, mh, ml);
}

// hi/lo split of the pair rows, rows padded with zeros to npair_pad.
__global__ void split_rows_kernel(const double* __restrict__ bp, int npair, int npair_pad,
                                  int npad, __bf16* __restrict__ bh, __bf16* __restrict__ bl) {
    const size_t total = (size_t)npair_pad * npad / 4;
    for (size_t i4 = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i4 < total;
         i4 += (size_t)gridDim.x * blockDim.x) {
        const size_t i = i4 * 4;
        const int r = i / npad;
        const dv4 v = r < npair ? *reinterpret_cast<const dv4*>(bp + i) : dv4{0.0, 0.0, 0.0, 0.0};
        bf16x4 h, l;
        split4(v, h, l);
        *reinterpret_cast<bf16x4*>(bh + i) = h;
        *reinterpret_cast<bf16x4*>(bl + i) = l;
    }
}

// Y = Bp M^T on BF16 MFMA (hi*hi + hi*lo + lo*hi, FP32 accumulation) over Q < P0 + 128
// (M is lower block-triangular); epilogue 2 b.y + y.y reduced per 64-column block.
__global__ __launch_bounds__(256, 2) void wnm_bf16x3_kernel(
    const __bf16* __restrict__ bh, const __bf16* __restrict__ bl, const __bf16* __restrict__ mh,
    const __bf16* __restrict__ ml, const double* __restrict__ bp, int npair, int npad,
    double* __restrict__ wpart) {
    __shared__ __attribute__((aligned(16))) __bf16 Ah[WT][WKP];
    __shared__ __attribute__((aligned(16))) __bf16 Al[WT][WKP];
    __shared__ __attribute__((aligned(16))) __bf16 Mh[WT][WKP];
    __shared__ __attribute__((aligned(16))) __bf16 Ml[WT][WKP];
    const int j0 = blockIdx.x * WT, P0 = blockIdx.y * WT, f = blockIdx.z;
    const int nblk = npad / NB;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wr = wave >> 1, wc = wave & 1;
    const int row = tid >> 1, kq0 = (tid & 1) * 16;
    const bool vm = P0 + row < npad;
    const __bf16* pah = bh + (size_t)(j0 + row) * npad + kq0;
    const __bf16* pal = bl + (size_t)(j0 + row) * npad + kq0;
    const size_t moff = (size_t)f * npad * npad + (size_t)(P0 + row) * npad + kq0;
    const __bf16* pmh = mh + moff;
    const __bf16* pml = ml + moff;
    bf16x8 rah[2], ral[2], rmh[2], rml[2];
    const bf16x8 z8 = bf16x8{0, 0, 0, 0, 0, 0, 0, 0};
    auto fetch = [&](int k0) {
        #pragma unroll
        for (int q = 0; q < 2; ++q) {
            rah[q] = *reinterpret_cast<const bf16x8*>(pah + k0 + 8 * q);
            ral[q] = *reinterpret_cast<const bf16x8*>(pal + k0 + 8 * q);
            rmh[q] = vm ? *reinterpret_cast<const bf16x8*>(pmh + k0 + 8 * q) : z8;
            rml[q] = vm ? *reinterpret_cast<const bf16x8*>(pml + k0 + 8 * q) : z8;
        }
    };
    f4 acc[4][4];
    #pragma unroll
    for (int bi = 0; bi < 4; ++bi)
        #pragma unroll
        for (int bj = 0; bj < 4; ++bj) acc[bi][bj] = f4{0.f, 0.f, 0.f, 0.f};
    const int kend = P0 + WT < npad ? P0 + WT : npad;
    fetch(0);
    for (int k0 = 0; k0 < kend; k0 += WK) {
        #pragma unroll
        for (int q = 0; q < 2; ++q) {
            *reinterpret_cast<bf16x8*>(&Ah[row][kq0 + 8 * q]) = rah[q];
            *reinterpret_cast<bf16x8*>(&Al[row][kq0 + 8 * q]) = ral[q];
            *reinterpret_cast<bf16x8*>(&Mh[row][kq0 + 8 * q]) = rmh[q];
            *reinterpret_cast<bf16x8*>(&Ml[row][kq0 + 8 * q]) = rml[q];
        }
        __syncthreads();
        if (k0 + WK < kend) fetch(k0 + WK);
        const int kq = (lane >> 4) * 8, r16 = lane & 15;
        bf16x8 ah[4], al[4], xh[4], xl[4];
        #pragma unroll
        for (int q = 0; q < 4; ++q) {
            ah[q] = *reinterpret_cast<const bf16x8*>(&Ah[wr * 64 + q * 16 + r16][kq]);
            al[q] = *reinterpret_cast<const bf16x8*>(&Al[wr * 64 + q * 16 + r16][kq]);
            xh[q] = *reinterpret_cast<const bf16x8*>(&Mh[wc * 64 + q * 16 + r16][kq]);
            xl[q] = *reinterpret_cast<const bf16x8*>(&Ml[wc * 64 + q * 16 + r16][kq]);
        }
        #pragma unroll
        for (int bi = 0; bi < 4; ++bi)
            #pragma unroll
            for (int bj = 0; bj < 4; ++bj) {
                f4 c = acc[bi][bj];
                c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[bi], xh[bj], c, 0, 0, 0);
                c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[bi], xl[bj], c, 0, 0, 0);
                c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[bi], xh[bj], c, 0, 0, 0);
                acc[bi][bj] = c;
            }
        __syncthreads();
    }
    #pragma unroll
    for (int bi = 0; bi < 4; ++bi) {
        double rs[4] = {0, 0, 0, 0};
        #pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int j = j0 + wr * 64 + bi * 16 + mfma_row_f32(lane, v);
            #pragma unroll
            for (int bj = 0; bj < 4; ++bj) {
                const int Pc = P0 + wc * 64 + bj * 16 + (lane & 15);
                const double y = (double)acc[bi][bj][v];
                const double b = (j < npair && Pc < npad) ? bp[(size_t)j * npad + Pc] : 0.0;
                rs[v] += 2.0 * b * y + y * y;
            }
        }
        #pragma unroll
        for (int v = 0; v < 4; ++v)
            for (int off = 1; off < 16; off <<= 1) rs[v] += __shfl_xor(rs[v], off, 64);
        const int pb = P0 / NB + wc;
        if ((lane & 15) == 0 && pb < nblk)
            #pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int j = j0 + wr * 64 + bi * 16 + mfma_row_f32(lane, v);
                if (j < npair) wpart[((size_t)f * nblk + pb) * npair + j] = rs[v];
            }
    }
}

// Work items of the two-frequency kernel: (pair tile, frequency pair, auxiliary row tile).
// A tile's K range (0 .. P0 + 128) grows with P0, so the row tiles are taken from the last
// one down: the longest workgroups are dispatched first and the short ones fill the tail.
//   blocked = 0: grid (pair tiles, frequency pairs, row tiles) in that order.
//   blocked = 1: the items of one row tile are cut into blocks of bp pair tiles x bf
//   frequency pairs (bp * bf <= the 64 workgroups one XCD holds at a time); block b runs on
//   XCD b % 8 (workgroups are dispatched round-robin over the 8 XCDs), as the slots
//   (b / 8) * bsz .. + bsz of that XCD.  A block's workgroups start together and share K
//   ranges, so each chunk of its bp pair-row tiles and bf M blocks is fetched into that
//   XCD's L2 once and read there by the others, instead of every workgroup streaming both
//   operands from beyond L2.  Slots past the items of a partial block return at once.
struct WnmMap {
    int blocked, npt, nfp, nP, bp, bf, nbp, nbf;
};

__device__ __forceinline__ bool wnm_item(const WnmMap& m, int& pt, int& fp, int& pz) {
    if (!m.blocked) {
        pt = blockIdx.x;
        fp = blockIdx.y;
        pz = blockIdx.z;
        return true;
    }
    const int bsz = m.bp * m.bf, xcd = blockIdx.x % 8, slot = blockIdx.x / 8;
    const int blk = (slot / bsz) * 8 + xcd, it = slot % bsz;
    const int per = m.nbp * m.nbf;
    if (blk >= m.nP * per) return false;
    pz = blk / per;
    const int b = blk % per;
    pt = (b / m.nbf) * m.bp + it / m.bf;
    fp = (b % m.nbf) * m.bf + it % m.bf;
    return pt < m.npt && fp < m.nfp;
}

// Two frequencies (f0, f1) per workgroup on 32x32x16 MFMAs: the pair-row operand Bp of a
// chunk is loaded and staged once for both, which cuts the operand traffic per MFMA by a
// quarter.  f1 == f0 (odd frequency count) computes one twice and writes it once.  One item:
// pair tile pt, frequency pair fp, row tile P0.
struct WnmLds {
    __bf16 Ah[WT][WKP], Al[WT][WKP], Mh[2][WT][WKP], Ml[2][WT][WKP];
};

__device__ __forceinline__ void wnm_f2_item(const __bf16* __restrict__ bh,
                                            const __bf16* __restrict__ bl,
                                            const __bf16* __restrict__ mh,
                                            const __bf16* __restrict__ ml,
                                            const double* __restrict__ bp, int npair, int npad,
                                            int nf, int pt, int fp, int P0,
                                            double* __restrict__ wpart, WnmLds& L) {
    typedef float f16v __attribute__((ext_vector_type(16)));
    auto& Ah = L.Ah;
    auto& Al = L.Al;
    auto& Mh = L.Mh;
    auto& Ml = L.Ml;
    const int j0 = pt * WT;
    const int f0 = 2 * fp, f1 = min(f0 + 1, nf - 1);
    const int nblk = npad / NB;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wr = wave >> 1, wc = wave & 1, r32 = lane & 31;
    const int row = tid >> 1, kq0 = (tid & 1) * 16;
    const bool vm = P0 + row < npad;
    const __bf16* pah = bh + (size_t)(j0 + row) * npad + kq0;
    const __bf16* pal = bl + (size_t)(j0 + row) * npad + kq0;
    const size_t moff = (size_t)(P0 + row) * npad + kq0;
    const size_t fstride[2] = {(size_t)f0 * npad * npad, (size_t)f1 * npad * npad};
    bf16x8 rah[2], ral[2], rmh[2][2], rml[2][2];
    const bf16x8 z8 = bf16x8{0, 0, 0, 0, 0, 0, 0, 0};
    auto fetch = [&](int k0) {
        #pragma unroll
        for (int q = 0; q < 2; ++q) {
            rah[q] = *reinterpret_cast<const bf16x8*>(pah + k0 + 8 * q);
            ral[q] = *reinterpret_cast<const bf16x8*>(pal + k0 + 8 * q);
            #pragma unroll
            for (int fi = 0; fi < 2; ++fi) {
                const size_t o = fstride[fi] + moff + k0 + 8 * q;
                rmh[fi][q] = vm ? *reinterpret_cast<const bf16x8*>(mh + o) : z8;
                rml[fi][q] = vm ? *reinterpret_cast<const bf16x8*>(ml + o) : z8;
            }
        }
    };
    f16v acc[2][2][2];
    #pragma unroll
    for (int fi = 0; fi < 2; ++fi)
        #pragma unroll
        for (int bi = 0; bi < 2; ++bi)
            #pragma unroll
            for (int bj = 0; bj < 2; ++bj)
                #pragma unroll
                for (int v = 0; v < 16; ++v) acc[fi][bi][bj][v] = 0.f;
    const int kend = P0 + WT < npad ? P0 + WT : npad;
    fetch(0);
    for (int k0 = 0; k0 < kend; k0 += WK) {
        #pragma unroll
        for (int q = 0; q < 2; ++q) {
            *reinterpret_cast<bf16x8*>(&Ah[row][kq0 + 8 * q]) = rah[q];
            *reinterpret_cast<bf16x8*>(&Al[row][kq0 + 8 * q]) = ral[q];
            #pragma unroll
            for (int fi = 0; fi < 2; ++fi) {
                *reinterpret_cast<bf16x8*>(&Mh[fi][row][kq0 + 8 * q]) = rmh[fi][q];
                *reinterpret_cast<bf16x8*>(&Ml[fi][row][kq0 + 8 * q]) = rml[fi][q];
            }
        }
        __syncthreads();
        if (k0 + WK < kend) fetch(k0 + WK);
        #pragma unroll
        for (int ks = 0; ks < WK; ks += 16) {
            const int kq = ks + (lane >> 5) * 8;
            bf16x8 ah[2], al[2];
            #pragma unroll
            for (int q = 0; q < 2; ++q) {
                ah[q] = *reinterpret_cast<const bf16x8*>(&Ah[wr * 64 + q * 32 + r32][kq]);
                al[q] = *reinterpret_cast<const bf16x8*>(&Al[wr * 64 + q * 32 + r32][kq]);
            }
            #pragma unroll
            for (int fi = 0; fi < 2; ++fi) {
                bf16x8 xh[2], xl[2];
                #pragma unroll
                for (int q = 0; q < 2; ++q) {
                    xh[q] = *reinterpret_cast<const bf16x8*>(&Mh[fi][wc * 64 + q * 32 + r32][kq]);
                    xl[q] = *reinterpret_cast<const bf16x8*>(&Ml[fi][wc * 64 + q * 32 + r32][kq]);
                }
                #pragma unroll
                for (int bi = 0; bi < 2; ++bi)
                    #pragma unroll
                    for (int bj = 0; bj < 2; ++bj)
                        acc[fi][bi][bj] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[bi], xh[bj], acc[fi][bi][bj], 0, 0, 0);
                #pragma unroll
                for (int bi = 0; bi < 2; ++bi)
                    #pragma unroll
                    for (int bj = 0; bj < 2; ++bj)
                        acc[fi][bi][bj] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[bi], xl[bj], acc[fi][bi][bj], 0, 0, 0);
                #pragma unroll
                for (int bi = 0; bi < 2; ++bi)
                    #pragma unroll
                    for (int bj = 0; bj < 2; ++bj)
                        acc[fi][bi][bj] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[bi], xh[bj], acc[fi][bi][bj], 0, 0, 0);
            }
        }
        __syncthreads();
    }
    const int pb = P0 / NB + wc;
    #pragma unroll
    for (int fi = 0; fi < 2; ++fi) {
        if (fi == 1 && f1 == f0) break;
        const int f = fi ? f1 : f0;
        #pragma unroll
        for (int bi = 0; bi < 2; ++bi) {
            double rs[16];
            #pragma unroll
            for (int v = 0; v < 16; ++v) {
                const int j = j0 + wr * 64 + bi * 32 + mfma_row_32(lane, v);
                rs[v] = 0.0;
                #pragma unroll
                for (int bj = 0; bj < 2; ++bj) {
                    const int Pc = P0 + wc * 64 + bj * 32 + r32;
                    const double y = (double)acc[fi][bi][bj][v];
                    const double b = (j < npair && Pc < npad) ? bp[(size_t)j * npad + Pc] : 0.0;
                    rs[v] += 2.0 * b * y + y * y;
                }
                for (int off = 1; off < 32; off <<= 1) rs[v] += __shfl_xor(rs[v], off, 64);
            }
            if (r32 == 0 && pb < nblk)
                #pragma unroll
                for (int v = 0; v < 16; ++v) {
                    const int j = j0 + wr * 64 + bi * 32 + mfma_row_32(lane, v);
                    if (j < npair) wpart[((size_t)f * nblk + pb) * npair + j] = rs[v];
                }
        }
    }
}

__global__ __launch_bounds__(256, 2) void wnm_bf16x3_f2_kernel(
    const __bf16* __restrict__ bh, const __bf16* __restrict__ bl, const __bf16* __restrict__ mh,
    const __bf16* __restrict__ ml, const double* __restrict__ bp, int npair, int npad, int nf,
    WnmMap map, double* __restrict__ wpart) {
    __shared__ __attribute__((aligned(16))) WnmLds L;
    int pt, fp, pz;
    if (!wnm_item(map, pt, fp, pz)) return;             // uniform over the workgroup
    wnm_f2_item(bh, bl, mh, ml, bp, npair, npad, nf, pt, fp, (map.nP - 1 - pz) * WT, wpart, L);
}

// Grouped form for a batch: instance q owns npt * nfp * nP consecutive workgroups (pair tile
// fastest, row tiles from the last one down), off is their prefix.
__global__ __launch_bounds__(256, 2) void wnm_bf16x3_f2_group_kernel(
    const WnmInst* __restrict__ inst, const int* __restrict__ off, int ninst, int nf) {
    __shared__ __attribute__((aligned(16))) WnmLds L;
    const int w = blockIdx.x;
    const int q = find_inst(off, ninst, w);
    const WnmInst in = inst[q];
    const int npt = in.npp / WT, nfp = (nf + 1) / 2, nP = (in.npad + WT - 1) / WT;
    const int local = w - off[q];
    const int pt = local % npt, fp = (local / npt) % nfp, pz = local / (npt * nfp);
    wnm_f2_item(in.bh, in.bl, in.mh, in.ml, in.bp, in.npair, in.npad, nf, pt, fp,
                (nP - 1 - pz) * WT, in.wpart, L);
}

// Grouped hi/lo split for a batch, blockIdx.y = instance: the pair rows (zero padded to npp
// rows) and the rows of M_f = X_f - I that the Wc_nm kernels read, as split_rows_kernel and
// split_m_kernel do for one instance.
__global__ void split_group_kernel(const WnmInst* __restrict__ inst, int nf) {
    const WnmInst in = inst[blockIdx.y];
    const size_t rtot = (size_t)in.npp * in.npad / 4;
    for (size_t i4 = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i4 < rtot;
         i4 += (size_t)gridDim.x * blockDim.x) {
        const size_t i = i4 * 4;
        const int r = i / in.npad;
        const dv4 v = r < in.npair ? *reinterpret_cast<const dv4*>(in.bp + i)
                                   : dv4{0.0, 0.0, 0.0, 0.0};
        bf16x4 h, l;
        split4(v, h, l);
        *reinterpret_cast<bf16x4*>(in.bh + i) = h;
        *reinterpret_cast<bf16x4*>(in.bl + i) = l;
    }
    const size_t rows = (size_t)nf * in.npad;
    for (size_t r = blockIdx.x; r < rows; r += gridDim.x) split_m_row(in.x, in.npad, r, in.mh, in.ml);
}

__device__ __forceinline__ double wave_sum(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// ---- Exchange self-energy (diagonal) -----------------------------------------------------
// sigma_x[p] = -sum_i sum_R mi[p][i][R]^2 over the [nmo][nocc][naux] block (the driver's
// compute_exchange, quasix_core/src/qp/evgw.rs:1167-1193, diagonal).  One workgroup per p
// streams its contiguous nocc * naux row once (memory-bound).
__global__ __launch_bounds__(256) void exchange_diag_kernel(const double* __restrict__ mi,
                                                            long long row, double* __restrict__ sx) {
    __shared__ double red[4];
    const int p = blockIdx.x, tid = threadIdx.x;
    const double* r = mi + (size_t)p * row;
    double acc = 0.0;
    for (long long k = tid; k < row; k += 256) acc += r[k] * r[k];
    acc = wave_sum(acc);
    if ((tid & 63) == 0) red[tid >> 6] = acc;
    __syncthreads();
    if (tid == 0) sx[p] = -((red[0] + red[1]) + (red[2] + red[3]));
}

// ---- CD quadrature ------------------------------------------------------------------------


// I_n(E_q) for nq energies of target t (whole workgroup of NW wavefronts: the pairs are
// spread over all its lanes); returns 1 if any E_q has a residue pole among the target's
// partners.  Every thread returns the same values (the NW partial sums are added in one
// fixed order), so callers may branch on them uniformly.
template <int NQ, int NW = 1>
__device__ int eval_imag(const SigmaProblem& p, int t, const double (&E)[NQ], int nq,
                         C2 (&out)[NQ]) {
    const int lane = threadIdx.x & 63;
    const int n = p.targets[t];
    const double en_ref = p.eps_ref[n];
    double sre[NQ], sim[NQ];
    #pragma unroll
    for (int q = 0; q < NQ; ++q) { sre[q] = 0.0; sim[q] = 0.0; }
    int pole = 0;
    for (int j = p.pair_off[t] + (int)threadIdx.x; j < p.pair_off[t + 1]; j += 64 * NW) {
        const int m = p.pair_m[j];
        const double em = p.eps_g[m];
        const double c = -p.eta * (en_ref > p.eps_ref[m] ? 1.0 : -1.0);
        #pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const double Eq = E[q];
            if (q < nq &&
                ((Eq > p.ef && p.ef < em && em < Eq) || (Eq <= p.ef && Eq < em && em < p.ef)))
                pole = 1;
        }
        // all NQ slots are evaluated (slots past nq hold a placeholder energy and are
        // never returned), so the accumulators stay in registers
        for (int k = 0; k < p.nf; ++k) {
            const double w = p.wnm[(size_t)k * p.npair + j];
            const double wk = p.wts[k], w2 = p.freqs[k] * p.freqs[k];
            #pragma unroll
            for (int q = 0; q < NQ; ++q) {
                const double a = E[q] - em;
                // emo = a + i c ; den = emo^2 + w2 ; g = wk emo / den
                const double dre = a * a - c * c + w2, dim = 2.0 * a * c;
                const double inv = wk / (dre * dre + dim * dim);
                const double gre = (a * dre + c * dim) * inv;
                const double gim = (c * dre - a * dim) * inv;
                sre[q] += w * gre;
                sim[q] += w * gim;
            }
        }
    }
    for (int off = 32; off > 0; off >>= 1) pole |= __shfl_xor(pole, off, 64);
    if (NW == 1) {
        #pragma unroll
        for (int q = 0; q < NQ; ++q)
            if (q < nq) {
                out[q].re = -wave_sum(sre[q]) / kPi;
                out[q].im = -wave_sum(sim[q]) / kPi;
            }
        return pole;
    }
    __shared__ double red[NW][2 * NQ + 1];
    const int w = threadIdx.x >> 6;
    #pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const double r = wave_sum(sre[q]), i = wave_sum(sim[q]);
        if (lane == 0) {
            red[w][2 * q] = r;
            red[w][2 * q + 1] = i;
        }
    }
    if (lane == 0) red[w][2 * NQ] = pole;
    __syncthreads();
    #pragma unroll
    for (int q = 0; q < NQ; ++q) {
        double r = 0.0, i = 0.0;
        #pragma unroll
        for (int v = 0; v < NW; ++v) {
            r += red[v][2 * q];
            i += red[v][2 * q + 1];
        }
        if (q < nq) {
            out[q].re = -r / kPi;
            out[q].im = -i / kPi;
        }
    }
    int any = 0;
    #pragma unroll
    for (int v = 0; v < NW; ++v) any |= red[v][2 * NQ] != 0.0;
    __syncthreads();                          // red is reused by the next call
    return any;
}

__global__ __launch_bounds__(64) void sigma_imag_eval_kernel(SigmaProblem p,
                                                             const double* __restrict__ energies,
                                                             int nq, double* __restrict__ out) {
    const int t = blockIdx.x;
    for (int q0 = 0; q0 < nq; q0 += 5) {
        double E[5];
        const int cnt = nq - q0 < 5 ? nq - q0 : 5;
        #pragma unroll
        for (int q = 0; q < 5; ++q) E[q] = q < cnt ? energies[(size_t)t * nq + q0 + q] : 0.0;
        C2 s[5];
        eval_imag<5>(p, t, E, cnt, s);
        if ((threadIdx.x & 63) == 0)
            #pragma unroll
            for (int q = 0; q < 5; ++q)
                if (q < cnt) {
                    out[((size_t)t * nq + q0 + q) * 2] = s[q].re;
                    out[((size_t)t * nq + q0 + q) * 2 + 1] = s[q].im;
                }
    }
}

// AC self-energy on the imaginary grid (quasix_core/src/selfenergy/correlation_ac.rs:
// 283-311, 404-438): for target t (orbital n) and grid point q, omega_q = -i nu_q for an
// occupied n and +i nu_q for a virtual one (nu_0 = 0),
//   Sigma_t(omega_q) = -(1/pi) sum_k sum_j W_k wnm[k][j] z / (z^2 + w_k^2),
//   z = omega_q + ef - eps_m(j).
// One wavefront per target, lanes over its pairs, QB grid points per pass.
__global__ __launch_bounds__(64) void sigma_ac_kernel(SigmaProblem p,
                                                      const double* __restrict__ nu, int nws,
                                                      double* __restrict__ out) {
    constexpr int QB = 8;
    const int t = blockIdx.x, lane = threadIdx.x & 63;
    const double sgn = p.targets[t] < p.nocc ? -1.0 : 1.0;
    for (int q0 = 0; q0 < nws; q0 += QB) {
        double nv[QB], sre[QB], sim[QB];
        #pragma unroll
        for (int q = 0; q < QB; ++q) {
            nv[q] = q0 + q < nws ? sgn * nu[q0 + q] : 0.0;
            sre[q] = 0.0;
            sim[q] = 0.0;
        }
        for (int j = p.pair_off[t] + lane; j < p.pair_off[t + 1]; j += 64) {
            const double a = p.ef - p.eps_g[p.pair_m[j]];
            for (int k = 0; k < p.nf; ++k) {
                const double w = p.wnm[(size_t)k * p.npair + j] * p.wts[k];
                const double w2 = p.freqs[k] * p.freqs[k];
                #pragma unroll
                for (int q = 0; q < QB; ++q) {
                    // z = a + i v;  z^2 + w2 = (a^2 - v^2 + w2) + i 2 a v
                    const double v = nv[q];
                    const double dre = a * a - v * v + w2, dim = 2.0 * a * v;
                    const double inv = w / (dre * dre + dim * dim);
                    sre[q] += (a * dre + v * dim) * inv;
                    sim[q] += (v * dre - a * dim) * inv;
                }
            }
        }
        #pragma unroll
        for (int q = 0; q < QB; ++q) {
            const double r = -wave_sum(sre[q]) / kPi, i = -wave_sum(sim[q]) / kPi;
            if (lane == 0 && q0 + q < nws) {
                out[((size_t)t * nws + q0 + q) * 2] = r;
                out[((size_t)t * nws + q0 + q) * 2 + 1] = i;
            }
        }
    }
}

template <int NW = 1>
struct DeviceEval {
    const SigmaProblem* p;
    int t;
    __device__ bool operator()(const double* E, int nq, C2* out) {
        double Ev[5];
        C2 s[5];
        #pragma unroll
        for (int q = 0; q < 5; ++q) Ev[q] = q < nq ? E[q] : 0.0;
        const int pole = eval_imag<5, NW>(*p, t, Ev, nq, s);
        #pragma unroll
        for (int q = 0; q < 5; ++q)
            if (q < nq) out[q] = s[q];
        return pole != 0;
    }
};

// One workgroup of NW wavefronts per target: every thread runs the same Newton iteration on
// the same reduced values, the pair sums are split over all lanes.

// The Fermi level is formed here from the device energies (the same expression as on the
// host), so the host needs no copy of them to launch.  qp_state (optional): the C API's code
// for targets that finished on the device (QX_QP_NEWTON, QX_QP_BISECTION, QX_QP_LINEARIZED);
// targets that need the residue path are left for the host to fill.
template <int NW>
__global__ __launch_bounds__(64 * NW) void newton_cd_kernel(
    SigmaProblem p0, NewtonParams np, const double* __restrict__ guess, double* e_out,
    double* z_out, double* sig_re, double* sig_im, int* state, int* iters, int* qp_state) {
    SigmaProblem p = p0;
    p.ef = 0.5 * (p.eps_g[p.nocc - 1] + p.eps_g[p.nocc]);
    const int t = blockIdx.x;
    const int n = p.targets[t];
    DeviceEval<NW> ev{&p, t};
    double E = 0.0, Z = 0.0;
    C2 s = {0.0, 0.0};
    int it = 0;
    const int stt = newton_core(ev, np, p.eps_ref[n], p.static_terms[t], p.ef, guess != nullptr,
                                guess ? guess[t] : 0.0, E, Z, s, it);
    if (threadIdx.x == 0) {
        e_out[t] = E;
        z_out[t] = Z;
        sig_re[t] = s.re;
        sig_im[t] = s.im;
        state[t] = stt;
        iters[t] = it;
        if (qp_state) {
            if (stt == kNewtonConverged) qp_state[t] = QX_QP_NEWTON;
            else if (stt == kNewtonBisection) qp_state[t] = QX_QP_BISECTION;
            else if (stt == kNewtonLinearized) qp_state[t] = QX_QP_LINEARIZED;
        }
    }
}

// Grouped Newton: one wavefront per target of every instance; t_off: prefix of ntarget.
// The Fermi level is recomputed from the instance's energies on the device, so a replayed
// launch sequence follows the current contents of the energy arrays.
__global__ __launch_bounds__(64) void newton_group_kernel(const NewtonInst* __restrict__ inst,
                                                          const int* __restrict__ t_off,
                                                          int ninst, NewtonParams np,
                                                          double* e_out, double* z_out,
                                                          double* sig_re, double* sig_im,
                                                          int* state, int* iters) {
    const int g = blockIdx.x;
    const int q = find_inst(t_off, ninst, g);
    SigmaProblem p = inst[q].p;
    p.ef = 0.5 * (p.eps_g[p.nocc - 1] + p.eps_g[p.nocc]);
    const int t = g - t_off[q], o = inst[q].out_off + t;
    const int n = p.targets[t];
    DeviceEval<> ev{&p, t};
    double E = 0.0, Z = 0.0;
    C2 s = {0.0, 0.0};
    int it = 0;
    const int stt = newton_core(ev, np, p.eps_ref[n], p.static_terms[t], p.ef, false, 0.0, E, Z,
                                s, it);
    if ((threadIdx.x & 63) == 0) {
        e_out[o] = E;
        z_out[o] = Z;
        sig_re[o] = s.re;
        sig_im[o] = s.im;
        state[o] = stt;
        iters[o] = it;
    }
}

}  // namespace

hipError_t launch_wnm_split_rows(const double* bpairs, int npair, int npad,
                                 const WnmMixed* mixed, hipStream_t st) {
    if (!mixed) return hipErrorInvalidValue;
    hipLaunchKernelGGL(split_rows_kernel, dim3(2048), dim3(256), 0, st, bpairs, npair,
                       round_up(npair, WT), npad, mixed->bh, mixed->bl);
    return hipGetLastError();
}

hipError_t launch_split_m(const double* x, int npad, int nf, __bf16* mh, __bf16* ml,
                          hipStream_t st) {
    if (nf <= 0) return hipSuccess;
    hipLaunchKernelGGL(split_m_kernel, dim3(8192), dim3(256), 0, st, x, npad, nf, mh, ml);
    return hipGetLastError();
}

hipError_t launch_wnm(int precision, const double* bpairs, const double* x, int npair, int npad,
                      int nf, double* wpart, double* wnm, const WnmMixed* mixed, hipStream_t st,
                      bool split_rows, bool split_m) {
    const int nblk = npad / NB;
    if (precision == 2) {
        if (!mixed) return hipErrorInvalidValue;
        const int npp = round_up(npair, WT);
        if (split_rows)
            hipLaunchKernelGGL(split_rows_kernel, dim3(2048), dim3(256), 0, st, bpairs, npair, npp,
                               npad, mixed->bh, mixed->bl);
        if (split_m)
            hipLaunchKernelGGL(split_m_kernel, dim3(8192), dim3(256), 0, st, x, npad, nf,
                               mixed->mh, mixed->ml);
        // QXGW_WNM_F2=0 selects the one-frequency kernel
        const char* fv = std::getenv("QXGW_WNM_F2");
        const bool f2 = !(fv && std::atoi(fv) == 0);
        if (f2) {
            // XCD blocks when there are at least 16 (pair tile, frequency pair) items per row
            // tile; QXGW_WNM_BLOCKED=0 / 1 forces the plain grid order / the blocks
            const char* bv = std::getenv("QXGW_WNM_BLOCKED");
            WnmMap m{};
            m.npt = npp / WT;
            m.nfp = (nf + 1) / 2;
            m.nP = ceil_div(npad, WT);
            m.blocked = bv ? std::atoi(bv) != 0 : m.nfp * m.npt >= 16;

            if (m.blocked) {
                // up to 8 frequency pairs per block, then as many pair tiles as fit in 64
                m.bf = std::min(m.nfp, 8);
                m.bp = std::min(m.npt, std::max(1, 64 / m.bf));
                // even out the pair-tile blocks: bp = ceil(npt / ceil(npt / bp))
                m.bp = ceil_div(m.npt, ceil_div(m.npt, m.bp));
                m.nbp = ceil_div(m.npt, m.bp);
                m.nbf = ceil_div(m.nfp, m.bf);
                const long long nblk = (long long)m.nP * m.nbp * m.nbf;
                const long long slots = ceil_div((int)nblk, 8) * 8LL * m.bp * m.bf;
                hipLaunchKernelGGL(wnm_bf16x3_f2_kernel, dim3((unsigned)slots), dim3(256), 0, st,
                                   mixed->bh, mixed->bl, mixed->mh, mixed->ml, bpairs, npair,
                                   npad, nf, m, wpart);
            } else {
                dim3 grid(m.npt, m.nfp, m.nP);
                hipLaunchKernelGGL(wnm_bf16x3_f2_kernel, grid, dim3(256), 0, st, mixed->bh,
                                   mixed->bl, mixed->mh, mixed->ml, bpairs, npair, npad, nf, m,
                                   wpart);
            }
        } else {
            dim3 grid(npp / WT, ceil_div(npad, WT), nf);
            hipLaunchKernelGGL(wnm_bf16x3_kernel, grid, dim3(256), 0, st, mixed->bh, mixed->bl,
                               mixed->mh, mixed->ml, bpairs, npair, npad, wpart);
        }
    } else {
        dim3 grid(ceil_div(npair, NB), nblk, nf);
        hipLaunchKernelGGL(wnm_linv_kernel, grid, dim3(256), 0, st, bpairs, x, npair, npad, wpart);
    }
    dim3 g2(ceil_div(npair, 256), nf);
    hipLaunchKernelGGL(wnm_reduce_kernel, g2, dim3(256), 0, st, wpart, npair, nblk, wnm);
    return hipGetLastError();
}

hipError_t launch_wnm_group(const WnmInst* inst_dev, const int* off_dev, const int* red_off_dev,
                            int ninst, int nf, int total_tiles, int total_red, hipStream_t st,
                            bool mixed) {
    if (ninst <= 0) return hipSuccess;
    if (mixed) {
        hipLaunchKernelGGL(split_group_kernel, dim3(256, ninst), dim3(256), 0, st, inst_dev, nf);
        if (total_tiles > 0)
            hipLaunchKernelGGL(wnm_bf16x3_f2_group_kernel, dim3(total_tiles), dim3(256), 0, st,
                               inst_dev, off_dev, ninst, nf);
    } else if (total_tiles > 0) {
        hipLaunchKernelGGL(wnm_linv_group_kernel, dim3(total_tiles), dim3(256), 0, st, inst_dev,
                           off_dev, ninst, nf);
    }
    if (total_red > 0)
        hipLaunchKernelGGL(wnm_reduce_group_kernel, dim3(std::min(ceil_div(total_red, 256), 4096)),
                           dim3(256), 0, st, inst_dev, red_off_dev, ninst, total_red);
    return hipGetLastError();
}

hipError_t launch_newton_group(const NewtonInst* inst_dev, const int* t_off_dev, int ninst,
                               int total_targets, const NewtonParams& np, double* e_out,
                               double* z_out, double* sig_re, double* sig_im, int* state,
                               int* iters, hipStream_t st) {
    if (ninst <= 0 || total_targets <= 0) return hipSuccess;
    hipLaunchKernelGGL(newton_group_kernel, dim3(total_targets), dim3(64), 0, st, inst_dev,
                       t_off_dev, ninst, np, e_out, z_out, sig_re, sig_im, state, iters);
    return hipGetLastError();
}

hipError_t launch_wnm_from_wc(const double* bpairs, const double* wc, int npair, int npad, int nf,
                              double* wpart, double* wnm, hipStream_t st) {
    const int nblk = npad / NB;
    dim3 grid(ceil_div(npair, NB), nblk, nf);
    hipLaunchKernelGGL(wnm_wc_kernel, grid, dim3(256), 0, st, bpairs, wc, npair, npad, wpart);
    dim3 g2(ceil_div(npair, 256), nf);
    hipLaunchKernelGGL(wnm_reduce_kernel, g2, dim3(256), 0, st, wpart, npair, nblk, wnm);
    return hipGetLastError();
}

hipError_t launch_newton_cd(const SigmaProblem& p, const NewtonParams& np, const double* guess,
                            double* e_out, double* z_out, double* sig_re, double* sig_im,
                            int* state, int* iters, hipStream_t st, int* qp_state) {
    if (p.nt <= 0) return hipSuccess;
    // QXGW_NEWTON_NW = 8 (default) or 4 wavefronts per target (16 spills)
    const char* v = std::getenv("QXGW_NEWTON_NW");
    const int nw = v ? std::atoi(v) : 8;
    if (nw == 8)
        hipLaunchKernelGGL(newton_cd_kernel<8>, dim3(p.nt), dim3(512), 0, st, p, np, guess,
                           e_out, z_out, sig_re, sig_im, state, iters, qp_state);
    else
        hipLaunchKernelGGL(newton_cd_kernel<4>, dim3(p.nt), dim3(256), 0, st, p, np, guess,
                           e_out, z_out, sig_re, sig_im, state, iters, qp_state);
    return hipGetLastError();
}

hipError_t launch_exchange_diag(const double* mi, int nmo, int nocc, int naux, double* sx,
                                hipStream_t st) {
    if (nmo <= 0) return hipSuccess;
    hipLaunchKernelGGL(exchange_diag_kernel, dim3(nmo), dim3(256), 0, st, mi,
                       (long long)nocc * naux, sx);
    return hipGetLastError();
}

hipError_t launch_sigma_ac(const SigmaProblem& p, const double* nu, int nws, double* out,
                           hipStream_t st) {
    if (p.nt <= 0 || nws <= 0) return hipSuccess;
    hipLaunchKernelGGL(sigma_ac_kernel, dim3(p.nt), dim3(64), 0, st, p, nu, nws, out);
    return hipGetLastError();
}

hipError_t launch_sigma_imag_eval(const SigmaProblem& p, const double* energies, int nq,
                                  double* out, hipStream_t st) {
    if (p.nt <= 0 || nq <= 0) return hipSuccess;
    hipLaunchKernelGGL(sigma_imag_eval_kernel, dim3(p.nt), dim3(64), 0, st, p, energies, nq, out);
    return hipGetLastError();
}

}  // namespace qx
