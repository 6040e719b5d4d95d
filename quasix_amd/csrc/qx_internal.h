// Internal declarations shared by the qxgw HIP sources (not part of the C ABI).
//
// Device data layout (all row-major, FP64 unless stated):
//   eps        [nmo]                      orbital energies
//   B_ov       [N_ov][naux]               DF tensor, ia rows (reference iaP layout)
//   eps_mat    [nf][npad][npad]           dielectric matrices eps = I - Pi, then L (lower)
//   xinv       [nf][npad][npad]           L^-1 (lower)
//   linv_diag  [nf][nblk][NB][NB]         inverses of the diagonal NB x NB blocks of L
//   B_pairs    [npair][npad]              target-partner rows b_nm (zero padded)
//   wnm        [nf][npair]                Wc_nm(i w_f) = b_nm^T Wc(i w_f) b_nm
// npad = naux rounded up to a multiple of NB; the padding block of eps is the identity.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstddef>
#include <functional>

namespace qx {

constexpr int NB = 64;          // Cholesky block size
constexpr int KPAD = 64;        // N_ov is padded to a multiple of KPAD for the Pi operands
constexpr int kMaxTargetsPerLaunch = 1 << 20;

inline int round_up(int x, int m) { return (x + m - 1) / m * m; }
inline int ceil_div(int x, int m) { return (x + m - 1) / m; }

// Row of accumulator element v held by `lane` in a 16x16 MFMA output block (column is
// lane % 16 for all three).  Measured on gfx950 with an in-bounds probe:
//   v_mfma_f32_16x16x4f32, v_mfma_f32_16x16x32_bf16:  row = 4 * (lane / 16) + v
//   v_mfma_f64_16x16x4f64:                            row = 4 * v + lane / 16
__device__ __forceinline__ int mfma_row_f32(int lane, int v) { return 4 * (lane >> 4) + v; }
__device__ __forceinline__ int mfma_row_f64(int lane, int v) { return 4 * v + (lane >> 4); }
// 32x32 output block of v_mfma_f32_32x32x16_bf16 (16 registers, column lane % 32), measured
// the same way: row = 8 * (v / 4) + 4 * (lane / 32) + v % 4.
__device__ __forceinline__ int mfma_row_32(int lane, int v) {
    return 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
}

// Work tables of grouped (batched) launches: off[0..n] are the first work items of the n
// instances; returns the instance that owns item w (off[0] <= w < off[n]).  The search
// runs on wave-uniform values.
__device__ __forceinline__ int find_inst(const int* __restrict__ off, int n, int w) {
    int lo = 0, hi = n;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (off[mid] <= w) lo = mid;
        else hi = mid;
    }
    return lo;
}

// ---- Polarizability (pi_kernels.hip) ---------------------------------------------------
// Pi is built from the frequency weights s_f(r) = 4 e_r / (w_f^2 + e_r^2),
// e_r = eps_i - eps_a for row r = i*nvir + a.
enum class PiOut { EpsLower = 0, PiFull = 1 };

struct PiOperands {              // per-solve operand copies for the mixed kernels
    const uint16_t* bt_hi = nullptr;   // [npad][kpad] bf16 high part of B^T
    const uint16_t* bt_lo = nullptr;   // [npad][kpad] bf16 low part of B^T
    const float*    bt_f32 = nullptr;  // [npad][kpad] B^T in FP32
    const float*    s32 = nullptr;     // [nf][kpad] frequency weights in FP32
    int kpad = 0;                      // N_ov rounded up to KPAD
    int rows = 0;                      // allocated rows of bt_* (zero padded, multiple of 256)
};

// s[f][r] table (FP64, and FP32 into s32 if non-null); eia from eps.  Layout [nf][kpad].
hipError_t launch_pi_weights(const double* eps, int nocc, int nvir, int nf,
                             const double* freqs, double* s, float* s32, hipStream_t st);
// B^T split copies for the mixed kernels; rows [naux, nrows) are zero.
hipError_t launch_pi_prepare_mixed(const double* B, int nov, int naux, int nrows, int kpad,
                                   uint16_t* bt_hi, uint16_t* bt_lo, float* bt_f32,
                                   hipStream_t st);
// precision: 0 = FP64 MFMA, 1 = FP32 MFMA, 2 = BF16x3 MFMA (FP32 accumulation).
// variant selects the kernel instance: BF16x3 has 128x128 tiles with K chunks 32 (0) and
// 64 (1), 256x128 tiles with a swizzled LDS layout (2), double-buffered 128x128 tiles with
// K chunks of 32 (3) and 64 (4), 128x128 tiles on 32x32x16 MFMAs with K chunks of 64 (5)
// and 32 (6), and the producer/consumer form with 512 threads: one frequency and K64 chunks
// (7), two frequencies and K32 chunks (8), one frequency and K32 chunks with two workgroups
// per CU (9), eight producer waves with K64 (10).  FP64 has one frequency per 128x128 tile
// (0) and two frequencies per tile with K chunks of 16 (1) and 32 (2), and the same with
// double-buffered LDS, K32 (3) and K16 (4); FP32 has one instance.
// pi_variant_count() tells how many exist per precision.
hipError_t launch_pi_syrk(int precision, const double* B, const PiOperands& ops,
                          const double* s, int nov, int naux, int npad, int nf,
                          double* out, PiOut mode, hipStream_t st, int variant = 0);
int pi_variant_count(int precision);

// One instance of a grouped Pi launch.  Mixed (BF16x3) kernels read bt_hi/bt_lo/bt_f32/s32,
// FP64 kernels read B/s; both write eps = I - Pi (lower tiles) into out [nf][npad][npad].
struct PiInst {
    const uint16_t* bt_hi;
    const uint16_t* bt_lo;
    const float* bt_f32;
    const float* s32;
    const double* B;
    const double* s;
    double* out;
    int nov, kpad, naux, npad;
};
// Persistent grouped Pi over all instances; off_dev: [ninst + 1] prefix sums of
// ntiles_q * nf with ntiles_q = T (T + 1) / 2, T = ceil(npad_q / 128).  grid is rounded
// down to a multiple of 8.  precision 0 (FP64, variants 0 = one frequency per tile, 1 = two
// frequencies per tile, even nf only) or 2 (BF16x3, variants 0 = K32, 1 = K64,
// 2 = double-buffered K32, 3 = K64 on 32x32x16 MFMAs, 4 = producer/consumer with two
// frequencies per tile, even nf only).
hipError_t launch_pi_group(int precision, const PiInst* inst_dev, const int* off_dev, int ninst,
                           int nf, int total, int grid, hipStream_t st, int variant = 0);
int pi_group_variant_count(int precision);
int pi_group_occupancy(int precision, int variant);   // resident workgroups per CU

// Side streams for running independent groups of a batch concurrently: work forks from
// the caller's stream through `fork` and joins back through `join[g]`.  All streams have
// the default priority.
struct StreamFan {
    static constexpr int kMax = 3;
    int n = 0;                              // side streams available
    hipStream_t side[kMax] = {};
    hipEvent_t fork = nullptr, join[kMax] = {};
};

// ---- Factorisation (chol_kernels.hip) --------------------------------------------------
// Cholesky eps = L L^T (L into a, lower blocks) and X = L^-1 (lower blocks of x) together.
// With a fan, the frequencies are split into `groups` contiguous groups whose step loops run
// on separate streams, so one group's latency-bound diagonal and panel kernels overlap the
// other groups' column updates.
// after (optional): called once per group, on that group's stream after its last kernel and
// before the join, with the group's first frequency and frequency count.
// lead (optional): called once on st after the fork and before the first group's kernels
// (same arguments for that group): work queued there runs beside the other groups' steps.
typedef std::function<hipError_t(hipStream_t, int, int)> PotrfAfter;
hipError_t launch_potrf_batched(double* a, double* x, double* linv_diag, int npad, int nf,
                                int* fail, hipStream_t st, const StreamFan* fan = nullptr,
                                int groups = 1, const PotrfAfter* after = nullptr,
                                const PotrfAfter* lead = nullptr);
// Grouped factorisation of a batch of instances with different sizes and a common nf.
// inst_dev: [ninst] sorted by npad descending (nblk_sorted[q] = npad_q / NB on the host);
// cu_off / rp_off: [ninst + 1] prefix sums of nblk_q * nf and (nblk_q - 1) * nf, on the
// device and on the host.  Each step is one launch per kernel kind over all instances
// that still have that block column.
struct PotrfInst {
    double* a;          // [nf][npad][npad]  eps -> L
    double* x;          // [nf][npad][npad]  X = L^-1
    double* linv;       // [nf][nblk][NB][NB]
    int* fail;          // [nf]
    int npad;
};
hipError_t launch_potrf_group(const PotrfInst* inst_dev, const int* nblk_sorted, int ninst,
                              int nf, const int* cu_off_dev, const int* rp_off_dev,
                              const int* cu_off_host, const int* rp_off_host, hipStream_t st);
// eps^-1 - I from X = L^-1 :  W = X^T X - I  (full symmetric)
hipError_t launch_wc_from_linv(const double* x, double* wc, int npad, int naux, int nf,
                               hipStream_t st);
// Dense fallback: Gauss-Jordan with partial pivoting, one workgroup per matrix, for the
// frequencies whose Cholesky failed (fail[f] != 0).  Writes (I - Pi)^-1 - I into wc or,
// if singular, zero (the identity fallback of correlation_ac.rs:396).
hipError_t launch_lu_fallback_wc(const double* pi_full_or_eps, int is_eps, double* work,
                                 double* wc, int npad, int nf, const int* fail,
                                 hipStream_t st);

// ---- Contractions and self-energy (sigma_kernels.hip) ----------------------------------
// Workspace of the BF16x3 contraction: hi/lo splits of the pair rows ([round_up(npair,128)]
// [npad]) and of M_f = X_f - I ([nf][npad][npad]).
struct WnmMixed {
    __bf16 *bh, *bl, *mh, *ml;
};
// wnm[f][j] = 2 b_j.(M b_j) + |M b_j|^2,  M = X - I.  precision 2 runs the product on BF16
// MFMA (hi/lo split, FP32 accumulation), otherwise FP64 MFMA.
// split_rows = false reuses bh/bl from an earlier launch_wnm_split_rows.
hipError_t launch_wnm(int precision, const double* bpairs, const double* x, int npair,
                      int npad, int nf, double* wpart, double* wnm, const WnmMixed* mixed,
                      hipStream_t st, bool split_rows = true, bool split_m = true);
// The BF16 split of M_f = X_f - I (what the mixed Wc_nm kernels read) for nf frequencies
// starting at x, into mh / ml at the same offsets; launch_wnm with split_m = false expects it.
hipError_t launch_split_m(const double* x, int npad, int nf, __bf16* mh, __bf16* ml,
                          hipStream_t st);
hipError_t launch_wnm_split_rows(const double* bpairs, int npair, int npad,
                                 const WnmMixed* mixed, hipStream_t st);
// wnm from an explicit Wc: wnm[f][j] = b_j^T Wc_f b_j  (FP64)
hipError_t launch_wnm_from_wc(const double* bpairs, const double* wc, int npair, int npad,
                              int nf, double* wpart, double* wnm, hipStream_t st);

// Grouped Wc_nm (batched instances).  off: prefix of the work items per instance (see
// launch_wnm_group), red_off: prefix of nf * npair_q (reduction items).
struct WnmInst {
    const double* bp;   // [npair][npad] pair rows, zero padded
    const double* x;    // [nf][npad][npad] X = L^-1
    double* wpart;      // [nf][nblk][npair]
    double* wnm;        // [nf][npair]
    int npair, npad;
    // mixed only: hi/lo BF16 copies of the pair rows ([npp][npad], npp = npair rounded up to
    // 128) and of M = X - I ([nf][npad][npad])
    __bf16 *bh = nullptr, *bl = nullptr, *mh = nullptr, *ml = nullptr;
    int npp = 0;
};
// mixed: split the pair rows and M per instance, then the two-frequency BF16x3 kernel over
// total_tiles = sum of npp / 128 * ceil(nf / 2) * ceil(npad / 128) items; otherwise FP64
// over sum of ceil(npair / 64) * nblk * nf tiles.  Then the per-pair reduction.
hipError_t launch_wnm_group(const WnmInst* inst_dev, const int* off_dev, const int* red_off_dev,
                            int ninst, int nf, int total_tiles, int total_red, hipStream_t st,
                            bool mixed = false);

struct NewtonParams {
    int max_iterations;
    double energy_tolerance;
    double initial_damping, min_damping, max_damping, max_step, derivative_delta;
    double z_min, z_max;
    int use_richardson, use_bisection, max_bisection_iterations;
    int linearized;              // 1: linearised QP step instead of the Newton solve
};

struct SigmaProblem {            // everything the CD quadrature needs (device pointers)
    int nt, npair, nf, nocc, nmo;
    const int* targets;          // [nt] orbital index
    const int* pair_off;         // [nt+1] CSR offsets into pairs (pairs sorted by target)
    const int* pair_m;           // [npair]
    const double* wnm;           // [nf][npair]
    const double* eps_g;         // [nmo] energies of the Green's function, poles and ef
    const double* eps_ref;       // [nmo] mean-field energies (QP equation, eta sign)
    const double* freqs;         // [nf]
    const double* wts;           // [nf]
    const double* static_terms;  // [nt] Sigma_x - V_xc
    double eta;
    double ef;
};

// Status codes written per target by the Newton kernel.
enum NewtonState : int {
    kNewtonConverged = 0, kNewtonBisection = 1, kNewtonFailed = 2, kNewtonNeedsResidue = 3,
    kNewtonLinearized = 4
};

// Device Newton loop (one wavefront per target), imaginary-axis part only; targets whose
// iterates reach a residue pole stop with kNewtonNeedsResidue.  The Fermi level is formed on
// the device from p.eps_g (p.ef is not read).  qp_state (nullable): the C API's QX_QP_* code
// of every target that finished on the device.
hipError_t launch_newton_cd(const SigmaProblem& p, const NewtonParams& np,
                            const double* guess,        // nullable: linearised guess
                            double* e_out, double* z_out, double* sig_re, double* sig_im,
                            int* state, int* iters, hipStream_t st, int* qp_state = nullptr);
// Grouped Newton over the targets of a batch (no guesses); target t of instance q writes
// output index out_off_q + t.
struct NewtonInst {
    SigmaProblem p;
    int out_off;
};
hipError_t launch_newton_group(const NewtonInst* inst_dev, const int* t_off_dev, int ninst,
                               int total_targets, const NewtonParams& np, double* e_out,
                               double* z_out, double* sig_re, double* sig_im, int* state,
                               int* iters, hipStream_t st);
// sx[p] = -sum_{i,R} mi[p][i][R]^2 for the [nmo][nocc][naux] block (exchange diagonal).
hipError_t launch_exchange_diag(const double* mi, int nmo, int nocc, int naux, double* sx,
                                hipStream_t st);
// AC self-energy Sigma_t(omega_q) on the imaginary grid nu[0..nws) (omega_q = -i nu_q for
// occupied targets, +i nu_q for virtual ones); out [nt][nws] complex (re, im).
hipError_t launch_sigma_ac(const SigmaProblem& p, const double* nu, int nws, double* out,
                           hipStream_t st);
// I_n(E) at nE energies per target: E[t][q] -> out[t][q] (complex, interleaved re/im)
hipError_t launch_sigma_imag_eval(const SigmaProblem& p, const double* energies, int nq,
                                  double* out, hipStream_t st);

// ---- Residues (residue_kernels.hip) ------------------------------------------------------
// For job j, w_j = b_j^T [(I - Pi(z_j))^-1 - I] b_j with z_j = omegas[j] + i eta.
// launch_residue_weights: s[2j][r] = Re c_r(z_j), s[2j + 1][r] = Im c_r(z_j),
// c_r(z) = 2 [1/(z + e_r) - 1/(z - e_r)], rows r < nocc * nvir, zero up to ld; the two
// frequencies (2j, 2j + 1) of the FP64 polarizability with these weights (PiOut::PiFull,
// leading dimension npad) are Re Pi and Im Pi of job j.
hipError_t launch_residue_weights(const double* eps, int nocc, int nvir, int njobs,
                                  const double* omegas, double eta, int ld, double* s,
                                  hipStream_t st);
// P: [2 njobs][npad][npad] holding Re Pi, Im Pi of each job inside n (npad % 64 == 0);
// overwritten by eps = I - Pi and its blocked LU with partial pivoting.  brows [njobs][n],
// rhs: [njobs][2][npad] workspace, fail [njobs] (zeroed by the caller; 1 on a zero
// pivot), w_out [njobs][2] (re, im).
hipError_t launch_residue_lu_solve(double* P, int n, int npad, int njobs, const double* brows,
                                   double* rhs, int* fail, double* w_out, hipStream_t st);

// ---- DF construction (df_kernels.hip) -----------------------------------------------------
// out[i * nvir + a][P] = sum_{mu,nu} c_occ[mu][i] j3c[mu][nu][P] c_vir[nu][a];
// tmp: nocc * nao * naux doubles.
hipError_t launch_mo_transform(const double* j3c, const double* c_occ, const double* c_vir,
                               int nao, int nocc, int nvir, int naux, double* tmp, double* out,
                               hipStream_t st);

// ---- Adjoint (adjoint_kernels.hip) -----------------------------------------------------
hipError_t launch_adjoint_quad(const int* targets, const int* pair_off, const int* pair_m,
                               const double* wnm, int nt, int npair, int nf, const double* eps,
                               const double* freqs, const double* wts, double eta,
                               const double* E, const double* gE, int sld, double* abar,
                               double* geps, hipStream_t st);
// C_b = op(A_b) op(B_b), square n x n (n % 64 == 0), batched.
hipError_t launch_batched_gemm(bool ta, bool tb, const double* A, const double* B, double* C,
                               int n, int batch, hipStream_t st);
// The same with element strides between the matrices of a batch (0 broadcasts one operand).
hipError_t launch_strided_gemm(bool ta, bool tb, const double* A, size_t sa, const double* B,
                               size_t sb, double* C, size_t sc, int n, int batch,
                               hipStream_t st);

// ---- Stage entry points on the reference's intermediates (stage_kernels.hip) ------------
// Pair rows of a full DF tensor lpq [naux][nmo][nmo] for every (n, m), j = n * nmo + m:
// u_j = (lpq[:, n, m] + lpq[:, m, n]) / 2 and v_j = (lpq[:, n, m] - lpq[:, m, n]) / 2, both
// [nmo * nmo][npad] zero padded; *any_v (device, cleared by the launcher) becomes 1 when
// some v entry is non-zero.  b_n^T W b_m = u^T W u - v^T W v for symmetric W.
hipError_t launch_lpq_pair_rows(const double* lpq, int naux, int nmo, int npad, double* u,
                                double* v, int* any_v, hipStream_t st);
// Target-partner rows of the symmetrised tensor assembled from the iaP [nocc][nvir][naux],
// miP [nmo][nocc][naux] and abP [nvir][nvir][naux] blocks (quasix/src/evgw.rs:161-205), for
// the targets [nt], rows that are identically zero dropped.  launch_assemble_pairs writes
// flags[t * nmo + m] (row non-zero), their exclusive prefix pos[] and the row count total[0]
// (flags, pos: nt * nmo ints); launch_scatter_pairs then writes row pos[r] of rows_out
// ([total][naux]) with its target slot and partner, in (target, partner) order.
hipError_t launch_assemble_pairs(const double* ia, const double* mi, const double* ab, int nocc,
                                 int nvir, int naux, const int* targets, int nt, int* flags,
                                 int* pos, int* total, hipStream_t st);
hipError_t launch_scatter_pairs(const double* ia, const double* mi, const double* ab, int nocc,
                                int nvir, int naux, const int* targets, int nt, const int* flags,
                                const int* pos, double* rows_out, int* pair_t, int* pair_m,
                                hipStream_t st);
// The ov block [nocc * nvir][naux] and the [nmo][nocc][naux] block of the same symmetrised
// tensor (either output nullable).
hipError_t launch_sym_blocks(const double* ia, const double* mi, const double* ab, int nocc,
                             int nvir, int naux, double* ov_sym, double* mi_sym, hipStream_t st);
// Static CD self-energy at E = 0 for every orbital n (sigma: [nmo] complex, re/im):
// -1/pi sum_f w_f sum_m W_nm(f) emo / (emo^2 + w_f^2), emo = -eps_m - i eta s,
// s = +1 if eps_n > eps_m else -1; W_nm(f) = wu[f][n nmo + m] - wv[f][n nmo + m] (wv
// nullable).
hipError_t launch_sigma_static_cd(const double* wu, const double* wv, int nmo, int nf,
                                  const double* eps, const double* freqs, const double* wts,
                                  double eta, double* sigma, hipStream_t st);
// dst[b][npad][npad] = src[b] ([n][n], batch stride src_stride elements, 0 broadcasts),
// zero padded (all zero when src is NULL); diag_pad puts 1 on the padding diagonal.
hipError_t launch_pad_square(const double* src, size_t src_stride, int n, int npad, int batch,
                             bool diag_pad, double* dst, hipStream_t st);
// eps[b] = I - (M[b] + M[b]^T) / 2 over [npad][npad].
hipError_t launch_eps_sym(const double* M, int npad, int batch, double* eps, hipStream_t st);
// W[b] ([n][n] compact) = (T[b] + T[b]^T) / 2 from [npad][npad].
hipError_t launch_sym_compact(const double* T, int npad, int n, int batch, double* W,
                              hipStream_t st);
// a[b] += alpha I over [npad][npad] (the first n diagonal entries).
hipError_t launch_add_identity(double* a, int npad, int n, int batch, double alpha,
                               hipStream_t st);
// out[0] = sum a^2 (mode 0), max |a - b| (mode 1) or max |a| (mode 2) over count entries;
// out holds 1 + 1024 doubles (the result, then the partials of the first pass).
hipError_t launch_reduce(const double* a, const double* b, size_t count, int mode, double* out,
                         hipStream_t st);
// y = alpha x + beta y elementwise over count entries; x == nullptr means y = beta y.
hipError_t launch_axpby(double alpha, const double* x, double beta, double* y, size_t count,
                        hipStream_t st);
// out[r][P] += alpha sum_k sum_Q s1_k(r) A[r][Q] M_k[Q][P]; part[pt][r] = row dots with s2.
hipError_t launch_scaled_gemm(const double* A, int nrow, int ncol, const double* s1,
                              const double* s2, int sld, const double* M, int npad, int nf,
                              double alpha, double* out, double* part, hipStream_t st);
hipError_t launch_adjoint_eps_from_ebar(const double* part, int nparts, int nocc, int nvir,
                                        double* geps, hipStream_t st);
hipError_t launch_adjoint_weights(const double* eps, int nocc, int nvir, int nf,
                                  const double* freqs, int sld, double* s, double* ds,
                                  hipStream_t st);
hipError_t launch_adjoint_pair_scatter(const double* gbp, const double* bp, const double* abar,
                                       int sld, int nf, const int* targets, const int* pair_t,
                                       const int* pair_m, int npair, int nocc, int nvir,
                                       int naux, double* gB, hipStream_t st);
// Zero the strictly-upper NB-blocks of every [npad][npad] matrix of the batch.
hipError_t launch_zero_upper_blocks(double* x, int npad, int nf, hipStream_t st);

}  // namespace qx
