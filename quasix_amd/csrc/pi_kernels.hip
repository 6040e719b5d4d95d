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
// v_mfma_f32_16x16x32_bf16 with FP32 accumulation, flushed to FP64 every kFlush rows).
#include "qx_internal.h"

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
                                  double* __restrict__ s) {
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

template <typename T, int KC>
__global__ __launch_bounds__(256) void pi_syrk_kernel(
    const double* __restrict__ B, const double* __restrict__ s, int nov, int naux, int npad,
    int nf, int ntiles, int sld, double* __restrict__ out, PiOut mode) {
    __shared__ T As[KC][TILE + 2];
    __shared__ T Bs[KC][TILE + 2];
    const int work = xcd_remap(blockIdx.x, ntiles * nf);
    const int tile = work / nf, f = work % nf;
    int I, J;
    tri_index(tile, I, J);
    const int I0 = I * TILE, J0 = J * TILE;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wr = wave >> 1, wc = wave & 1;
    const double* sf = s + (size_t)f * sld;

    typename Mfma<T>::acc_t acc[4][4];
    #pragma unroll
    for (int bi = 0; bi < 4; ++bi)
        #pragma unroll
        for (int bj = 0; bj < 4; ++bj) acc[bi][bj] = typename Mfma<T>::acc_t{0, 0, 0, 0};

    for (int r0 = 0; r0 < nov; r0 += KC) {
        for (int idx = tid; idx < KC * TILE; idx += 256) {
            const int kr = idx / TILE, c = idx % TILE;
            const int r = r0 + kr;
            double a = 0.0, b = 0.0;
            if (r < nov) {
                const double* row = B + (size_t)r * naux;
                if (I0 + c < naux) a = row[I0 + c];
                if (J0 + c < naux) b = row[J0 + c] * sf[r];
            }
            As[kr][c] = (T)a;
            Bs[kr][c] = (T)b;
        }
        __syncthreads();
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
                for (int bj = 0; bj < 4; ++bj) acc[bi][bj] = Mfma<T>::run(av[bi], bv[bj], acc[bi][bj]);
        }
        __syncthreads();
    }
    store_tile<T, sizeof(T) == 8>(acc, I0, J0, wr, wc, lane, npad, naux, f, mode, out, 1.0);
}

// ---- BF16x3 kernel ----------------------------------------------------------------------
// Operands are staged K-contiguous: LDS rows are auxiliary indices, 32 N_ov rows per
// chunk, so each lane fetches its 8 consecutive k values with one 16-byte read.
constexpr int KB = 32;
constexpr int KBP = 40;          // padded LDS pitch (bf16 elements): conflict-free rows
constexpr int kFlush = 2048;     // FP32 partial sums are folded into FP64 every kFlush rows

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

__global__ __launch_bounds__(256) void pi_syrk_bf16x3_kernel(
    const uint16_t* __restrict__ bt_hi, const uint16_t* __restrict__ bt_lo,
    const float* __restrict__ bt_f32, const double* __restrict__ s, int kpad, int naux,
    int npad, int nf, int ntiles, int sld, double* __restrict__ out, PiOut mode) {
    __shared__ __attribute__((aligned(16))) unsigned short Ah[TILE][KBP];
    __shared__ __attribute__((aligned(16))) unsigned short Al[TILE][KBP];
    __shared__ __attribute__((aligned(16))) unsigned short Ch[TILE][KBP];
    __shared__ __attribute__((aligned(16))) unsigned short Cl[TILE][KBP];
    __shared__ float sk[KB];
    const int work = xcd_remap(blockIdx.x, ntiles * nf);
    const int tile = work / nf, f = work % nf;
    int I, J;
    tri_index(tile, I, J);
    const int I0 = I * TILE, J0 = J * TILE;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wr = wave >> 1, wc = wave & 1;
    const double* sf = s + (size_t)f * sld;

    f4 acc[4][4];
    d4 acc64[4][4];
    #pragma unroll
    for (int bi = 0; bi < 4; ++bi)
        #pragma unroll
        for (int bj = 0; bj < 4; ++bj) {
            acc[bi][bj] = f4{0.f, 0.f, 0.f, 0.f};
            acc64[bi][bj] = d4{0.0, 0.0, 0.0, 0.0};
        }

    for (int k0 = 0; k0 < kpad; k0 += KB) {
        if (tid < KB) sk[tid] = (float)sf[k0 + tid];
        // A operand (unscaled, pre-split): 128 rows x 32 k, 8 bf16 per thread per part.
        for (int idx = tid; idx < TILE * (KB / 8); idx += 256) {
            const int row = idx / (KB / 8), kq = (idx % (KB / 8)) * 8;
            const int c = I0 + row;
            u16x8 h = u16x8{0, 0, 0, 0, 0, 0, 0, 0}, l = h;
            if (c < npad) {
                h = *reinterpret_cast<const u16x8*>(bt_hi + (size_t)c * kpad + k0 + kq);
                l = *reinterpret_cast<const u16x8*>(bt_lo + (size_t)c * kpad + k0 + kq);
            }
            *reinterpret_cast<u16x8*>(&Ah[row][kq]) = h;
            *reinterpret_cast<u16x8*>(&Al[row][kq]) = l;
        }
        __syncthreads();   // sk visible
        // C operand = s * B, split on the fly.
        for (int idx = tid; idx < TILE * (KB / 4); idx += 256) {
            const int row = idx / (KB / 4), kq = (idx % (KB / 4)) * 4;
            const int c = J0 + row;
            float4 x = make_float4(0.f, 0.f, 0.f, 0.f);
            if (c < npad) x = *reinterpret_cast<const float4*>(bt_f32 + (size_t)c * kpad + k0 + kq);
            const float v[4] = {x.x * sk[kq], x.y * sk[kq + 1], x.z * sk[kq + 2], x.w * sk[kq + 3]};
            #pragma unroll
            for (int q = 0; q < 4; ++q) {
                const unsigned short h = f32_to_bf16_rne(v[q]);
                Ch[row][kq + q] = h;
                Cl[row][kq + q] = f32_to_bf16_rne(v[q] - bf16_to_f32(h));
            }
        }
        __syncthreads();
        {
            const int kq = (lane >> 4) * 8, r16 = lane & 15;
            bf16x8 ah[4], al[4], ch[4], cl[4];
            #pragma unroll
            for (int q = 0; q < 4; ++q) {
                ah[q] = *reinterpret_cast<const bf16x8*>(&Ah[wr * 64 + q * 16 + r16][kq]);
                al[q] = *reinterpret_cast<const bf16x8*>(&Al[wr * 64 + q * 16 + r16][kq]);
                ch[q] = *reinterpret_cast<const bf16x8*>(&Ch[wc * 64 + q * 16 + r16][kq]);
                cl[q] = *reinterpret_cast<const bf16x8*>(&Cl[wc * 64 + q * 16 + r16][kq]);
            }
            #pragma unroll
            for (int bi = 0; bi < 4; ++bi)
                #pragma unroll
                for (int bj = 0; bj < 4; ++bj) {
                    f4 c = acc[bi][bj];
                    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[bi], ch[bj], c, 0, 0, 0);
                    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[bi], cl[bj], c, 0, 0, 0);
                    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[bi], ch[bj], c, 0, 0, 0);
                    acc[bi][bj] = c;
                }
        }
        if (((k0 + KB) % kFlush) == 0 || k0 + KB >= kpad) {
            #pragma unroll
            for (int bi = 0; bi < 4; ++bi)
                #pragma unroll
                for (int bj = 0; bj < 4; ++bj) {
                    #pragma unroll
                    for (int v = 0; v < 4; ++v) acc64[bi][bj][v] += (double)acc[bi][bj][v];
                    acc[bi][bj] = f4{0.f, 0.f, 0.f, 0.f};
                }
        }
        __syncthreads();
    }
    store_tile<double, false>(acc64, I0, J0, wr, wc, lane, npad, naux, f, mode, out, 1.0);
}

}  // namespace

hipError_t launch_pi_weights(const double* eps, int nocc, int nvir, int nf,
                             const double* freqs, double* s, hipStream_t st) {
    const int ld = round_up(nocc * nvir, 32);
    dim3 grid(ceil_div(ld, 256) < 1024 ? ceil_div(ld, 256) : 1024, nf);
    hipLaunchKernelGGL(pi_weights_kernel, grid, dim3(256), 0, st, eps, nocc, nvir, nf, freqs,
                       ld, s);
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

hipError_t launch_pi_syrk(int precision, const double* B, const PiOperands& ops,
                          const double* s, int nov, int naux, int npad, int nf, double* out,
                          PiOut mode, hipStream_t st) {
    const int T = ceil_div(mode == PiOut::EpsLower ? npad : naux, TILE);
    const int ntiles = T * (T + 1) / 2;
    const int sld = round_up(nov, 32);
    dim3 grid(ntiles * nf), block(256);
    if (precision == 0) {
        hipLaunchKernelGGL((pi_syrk_kernel<double, 16>), grid, block, 0, st, B, s, nov, naux,
                           npad, nf, ntiles, sld, out, mode);
    } else if (precision == 1) {
        hipLaunchKernelGGL((pi_syrk_kernel<float, 32>), grid, block, 0, st, B, s, nov, naux,
                           npad, nf, ntiles, sld, out, mode);
    } else {
        if (!ops.bt_hi || ops.kpad != sld) return hipErrorInvalidValue;
        hipLaunchKernelGGL(pi_syrk_bf16x3_kernel, grid, block, 0, st, ops.bt_hi, ops.bt_lo,
                           ops.bt_f32, s, ops.kpad, naux, npad, nf, ntiles, sld, out, mode);
    }
    return hipGetLastError();
}

}  // namespace qx
