/*
 * qxgw.h -- C ABI of the MI355X (gfx950) G0W0/evGW correlation path.
 *
 * Drop-in boundary for QuasiX's correlation hot path: the polarizability Pi(i w), the
 * screened interaction Wc = (I - Pi)^-1 - I, the contour-deformation self-energy
 * (imaginary-axis quadrature + residues) and the Newton quasiparticle solve.  Every entry
 * point names the reference interface it replaces (paths relative to the reference root).
 *
 * Conventions
 *   - Plain pointers and sizes; no framework types.  Functions ending in _dev take DEVICE
 *     pointers (HIP device memory of the handle's device) for every array argument; the
 *     descriptor and stats structs are host memory.  Functions without the suffix take
 *     host pointers and copy internally.
 *   - Row-major ("C order") arrays, FP64 unless stated.  B_ov is the reference's iaP block,
 *     [nocc*nvir][naux] (quasix/src/evgw.rs:105-114).
 *   - Status codes 1-5 carry the meaning of the QuasixError variants the path can raise
 *     (InvalidInput, DimensionMismatch, NumericalError, ConvergenceError, NotImplemented;
 *     quasix_core/src/common/mod.rs:15-70); the reference's other variants (I/O, HDF5,
 *     basis-set, SCF, ...) belong to subsystems outside this path and have no code here.
 *     qx_last_error() returns the message of the last failure on the handle.
 *   - One handle per host thread; calls on one handle are serialised on its stream.
 */
#ifndef QXGW_H
#define QXGW_H

#ifdef __cplusplus
extern "C" {
#endif

#define QXGW_API __attribute__((visibility("default")))
#define QXGW_VERSION_MAJOR 0
#define QXGW_VERSION_MINOR 2

typedef struct qx_context* qx_handle;

typedef enum {
    QX_OK = 0,
    QX_ERR_INVALID_INPUT = 1,       /* QuasixError::InvalidInput */
    QX_ERR_DIMENSION_MISMATCH = 2,  /* QuasixError::DimensionMismatch */
    QX_ERR_NUMERICAL = 3,           /* QuasixError::NumericalError */
    QX_ERR_CONVERGENCE = 4,         /* QuasixError::ConvergenceError */
    QX_ERR_NOT_IMPLEMENTED = 5,     /* QuasixError::NotImplemented */
    QX_ERR_DEVICE = 6,              /* HIP runtime failure (no QuasixError counterpart) */
    QX_ERR_ALLOC = 7                /* device allocation failure */
} qx_status;

typedef enum {
    QX_PREC_FP64 = 0,    /* FP64 MFMA everywhere (the GPU oracle path) */
    QX_PREC_FP32 = 1,    /* Pi on FP32 MFMA, FP32 accumulation */
    QX_PREC_BF16X3 = 2   /* Pi on BF16 MFMA with hi+lo split operands, FP32 accumulation */
} qx_precision;

/* NewtonSolverConfig, quasix_core/src/qp/newton_solver.rs:56-103 (same fields/defaults). */
typedef struct {
    int max_iterations;             /* 50 */
    double energy_tolerance;        /* 1e-8 Ha */
    double initial_damping;         /* 1.0 */
    double min_damping;             /* 0.1 */
    double max_damping;             /* 1.0 */
    double max_step;                /* 1.0 Ha */
    double derivative_delta;        /* 1e-6 Ha */
    int use_richardson;             /* 1 */
    double z_min;                   /* 0.1 */
    double z_max;                   /* 0.999 */
    int use_bisection_fallback;     /* 1 */
    int max_bisection_iterations;   /* 30 */
} qx_newton_config;

/* Shape and options of one G0W0 solve. */
typedef struct {
    int nocc, nvir, naux;      /* orbital / auxiliary dimensions */
    int nfreq;                 /* Gauss-Legendre imaginary frequencies */
    double x0;                 /* grid scale, w = x0 (1+t)/(1-t)  (correlation_ac.rs:94-114) */
    double eta;                /* broadening of the CD Green's function and residues */
    int precision;             /* qx_precision of the Pi build (BF16X3 also runs the Wc_nm
                                  product on BF16 MFMA; factorisation stays FP64) */
    int ntarget;               /* number of target orbitals */
    int npair;                 /* number of (target, partner) rows of B_pair */
    qx_newton_config newton;
    int qp_solver;             /* QX_SOLVER_NEWTON (default) or QX_SOLVER_LINEARIZED */
} qx_g0w0_desc;

/* QP solvers of qx_g0w0_desc.qp_solver (quasix_core/src/qp/evgw.rs:705-820). */
#define QX_SOLVER_NEWTON 0      /* NewtonQPSolver::solve with the linearised fallback */
#define QX_SOLVER_LINEARIZED 1  /* E = eps + Z (Sigma_x - V_xc + Re Sigma_c(E0)) at E0 = the
                                   guess (else eps), forward-difference Z clamped */

typedef struct {
    int chol_fallbacks;        /* frequencies whose Cholesky failed (Gauss-Jordan used) */
    int singular_fallbacks;    /* frequencies replaced by Wc = 0 (identity fallback) */
    int residue_targets;       /* targets solved on the residue (host-driven) path */
    int residue_jobs;          /* real-axis Wc evaluations performed */
    int newton_failures;       /* targets that fell back to the linearised solution */
    double seconds;            /* wall time of the call */
    int pi_variant;            /* Pi kernel instance chosen by the autotuner */
    int waves;                 /* batched solves: workspace waves */
    int graph_replay;          /* batched solves: 1 if the call replayed a captured graph */
    int resolved_instances;    /* batched solves: instances re-solved one by one */
} qx_stats;

/* Per-target state reported in `state` arrays. */
#define QX_QP_NEWTON 0
#define QX_QP_BISECTION 1
#define QX_QP_LINEARIZED 2

/* ---- handle ------------------------------------------------------------------------ */
QXGW_API qx_status qx_create(qx_handle* out, int device);
QXGW_API void qx_destroy(qx_handle h);
QXGW_API const char* qx_last_error(qx_handle h);
/* Use a caller-owned hipStream_t (NULL = the handle's own stream). */
QXGW_API qx_status qx_set_stream(qx_handle h, void* hip_stream);
QXGW_API qx_status qx_synchronize(qx_handle h);
/* Release the cached device workspace. */
QXGW_API qx_status qx_release_workspace(qx_handle h);
QXGW_API void qx_newton_config_default(qx_newton_config* c);
QXGW_API void qx_g0w0_desc_default(qx_g0w0_desc* d);

/* ---- profiling (no reference counterpart: per-phase device time for roofline reports) -- */
/* Phases timed with HIP events on the handle's stream while profiling is enabled. */
#define QX_PHASE_PREP 0      /* frequency weights, mixed-precision operand split */
#define QX_PHASE_PI 1        /* Pi build (one launch per solve, all frequencies) */
#define QX_PHASE_POTRF 2     /* blocked Cholesky of all eps_f, with L^-1 in the same loop */
#define QX_PHASE_TRTRI 3     /* reserved (L^-1 is formed inside the Cholesky step loop) */
#define QX_PHASE_WNM 4       /* Wc_nm contraction */
#define QX_PHASE_NEWTON 5    /* device Newton kernel */
#define QX_PHASE_RESIDUE 6   /* residue (real-axis Wc) jobs */
#define QX_NPHASE 8
/* Enable (1) or disable (0) profiling; both reset the accumulated times. */
QXGW_API qx_status qx_set_profiling(qx_handle h, int enable);
/* ms[QX_NPHASE]: accumulated milliseconds; counts[QX_NPHASE]: number of timed intervals. */
QXGW_API qx_status qx_get_profile(qx_handle h, double* ms, long long* counts);

/* ---- A1: frequency grid (host only) -------------------------------------------------- */
/* Gauss-Legendre nodes/weights on [-1,1], ascending.
 * Replaces quasix_core::freq::imaginary_grid::gauss_legendre_grid (imaginary_grid.rs:42;
 * PyO3 quasix_core.gauss_legendre_grid, quasix_core/src/python/mod.rs:29). */
QXGW_API qx_status qx_gauss_legendre_grid(int n, double* nodes, double* weights);
/* w = x0 (1+t)/(1-t), W = w_GL 2 x0 / (1-t)^2.
 * Replaces correlation_ac::get_scaled_legendre_roots (correlation_ac.rs:94-114). */
QXGW_API qx_status qx_scaled_legendre_roots(int nw, double x0, double* freqs, double* wts);

/* ---- A2: polarizability ------------------------------------------------------------- */
/* pi_out[f][P][Q] = 4 sum_ia B[ia][P] B[ia][Q] e_ia / (freqs[f]^2 + e_ia^2), full symmetric.
 * Replaces correlation_ac::get_rho_response (correlation_ac.rs:184-211) for all
 * frequencies at once and dielectric::compute_polarizability_p0 (polarizability.rs:169). */
QXGW_API qx_status qx_rho_response_dev(qx_handle h, int precision, int nocc, int nvir,
                                       int naux, const double* mo_energy, const double* B_ov,
                                       int nfreq, const double* freqs, double* pi_out);

/* ---- A3: screened interaction -------------------------------------------------------- */
/* wc_out[f] = (I - pi[f])^-1 - I with the Cholesky -> LU -> identity chain of
 * correlation_ac.rs:376-401 (and dielectric::compute_W_single_frequency, screening.rs:334).
 * fallbacks (nullable, host): number of frequencies that needed the LU fallback. */
QXGW_API qx_status qx_screened_wc_dev(qx_handle h, int naux, int nfreq, const double* pi,
                                      double* wc_out, int* fallbacks);

/* w_out[f] ([naux][naux], device) = vs (I - vs p0[f] vs^T)^-1 vs^T, symmetrised, with vs =
 * v_sqrt when given, else the principal square root of the SPD Coulomb metric `metric`
 * (Newton-Schulz on FP64 MFMA, accepted when |vs vs^T - v|_max <= 1e-10), else vs = I.
 * eps = I - vs p0 vs^T is symmetrised and inverted by Cholesky, LU when that fails; a singular
 * eps or a non-finite W is QX_ERR_NUMERICAL.  Replaces
 * dielectric::screening::compute_screened_interaction{,_with_vsqrt} and
 * compute_W_single_frequency (quasix_core/src/dielectric/screening.rs:164-420) behind the
 * quasix_core.compute_screening_w stage function (quasix_core/src/python/g0w0.rs:102-130). */
QXGW_API qx_status qx_screening_w_dev(qx_handle h, int naux, int nfreq, const double* p0,
                                      const double* metric, const double* v_sqrt,
                                      double* w_out);

/* ---- A4: Wc_nm on the imaginary axis --------------------------------------------------- */
/* wnm[f][j] = B_pair[j]^T Wc(i freqs[f]) B_pair[j] for every pair row; Pi from B_ov and
 * mo_energy.  Replaces the per-target contraction of correlation_ac::get_sigma_diag
 * (correlation_ac.rs:404-421) and selfenergy::frequency_cache (frequency_cache.rs:220). */
QXGW_API qx_status qx_wmn_imag_dev(qx_handle h, int precision, int nocc, int nvir, int naux,
                                   const double* mo_energy, const double* B_ov, int nfreq,
                                   const double* freqs, int npair, const double* B_pair,
                                   double* wnm);

/* ---- A5/A6: CD self-energy at given energies ---------------------------------------------- */
/* sigma[t][q] (complex, re/im interleaved) = I_n(E) + R_n(E) at E = energies[t][q] for target
 * n = targets[t], given wnm from qx_wmn_imag_dev.  Pairs must be sorted by target.
 * Replaces selfenergy::correlation_pyscf_cd::compute_sigma_c_pyscf_cd
 * (correlation_pyscf_cd.rs:60-298, there at E = 0 without residues) and the unwired
 * freq::contour_deformation::CDIntegrator (contour_deformation.rs:142-422). */
QXGW_API qx_status qx_sigma_cd_dev(qx_handle h, const qx_g0w0_desc* d, const double* mo_energy,
                                   const double* mo_energy_ref, const double* B_ov,
                                   const double* B_pair, const int* pair_target,
                                   const int* pair_m, const int* targets, const double* freqs,
                                   const double* wts, const double* wnm, int nq,
                                   const double* energies, double* sigma);

/* Static CD self-energy of every orbital at E = 0, without residues, from a given
 * polarizability p0_iw [nfreq][naux][naux] and the full DF tensor lpq [naux][nmo][nmo]
 * (device pointers): Wc_f = (I - p0_f)^-1 - I, W_nm = lpq[:, n, m]^T Wc_f lpq[:, m, n],
 * sigma[n] = -1/pi sum_f w_f sum_m W_nm emo / (emo^2 + freqs_f^2), emo = -eps_m - i eta s,
 * s = +1 if eps_n > eps_m else -1.  sigma: [nmo] complex (re, im), device.  A singular
 * I - p0_f is QX_ERR_NUMERICAL.  Replaces selfenergy::correlation_pyscf_cd::
 * compute_sigma_c_pyscf_cd (quasix_core/src/selfenergy/correlation_pyscf_cd.rs:60-298) and
 * its stage function (quasix_core/src/python/g0w0.rs:597-631). */
QXGW_API qx_status qx_sigma_c_pyscf_cd_dev(qx_handle h, int nmo, int naux, int nfreq, double eta,
                                           const double* mo_energy, const double* p0_iw,
                                           const double* lpq, const double* freqs,
                                           const double* weights, double* sigma);

/* Host-only linearised QP equation with Z = 1: qp = mo_energy + sigma_x + sigma_c - vxc,
 * QX_ERR_NUMERICAL if a result is not finite.  Replaces qp::solver::
 * solve_quasiparticle_linearized (quasix_core/src/qp/solver.rs:137-221) behind the
 * quasix_core.solve_qp_equation stage function (quasix_core/src/python/g0w0.rs:423-450). */
QXGW_API qx_status qx_solve_qp_linearized(int nmo, const double* mo_energy,
                                          const double* sigma_x, const double* sigma_c,
                                          const double* vxc, double* qp_energy,
                                          double* z_factor);

/* ---- A9/A10: fused G0W0 solve ---------------------------------------------------------- */
/* One G0W0 solve: Pi at all frequencies -> FP64 Cholesky -> Wc_nm -> Newton on the CD
 * self-energy for every target.  Pairs (pair_target, pair_m, B_pair[npair][naux]) list the
 * non-zero rows b_nm of each target, sorted by target.  static_terms[t] = Sigma_x - V_xc.
 * mo_energy_ref (nullable): mean-field energies anchoring the QP equation and the eta sign
 * (defaults to mo_energy).  guess (nullable): Newton start; NULL = linearised guess.
 * Outputs per target: qp_energy, z_factor, sigma_re/sigma_im = Sigma_c(E_QP), state.
 * Replaces the iteration-0 path of qp::evgw::EvGWDriver::run_evgw_blocks
 * (quasix_core/src/qp/evgw.rs:603-783) with NewtonQPSolver::solve (newton_solver.rs:174). */
QXGW_API qx_status qx_g0w0_cd_dev(qx_handle h, const qx_g0w0_desc* d, const double* mo_energy,
                                  const double* mo_energy_ref, const double* B_ov,
                                  const double* B_pair, const int* pair_target,
                                  const int* pair_m, const int* targets,
                                  const double* static_terms, const double* guess,
                                  double* qp_energy, double* z_factor, double* sigma_re,
                                  double* sigma_im, int* state, qx_stats* stats);
/* qx_g0w0_cd_dev for an ov-only tensor (oo and vv blocks zero): the pairs are every
 * virtual partner of an occupied target and every occupied partner of a virtual target,
 * gathered on the device from B_ov; desc->npair is ignored.  Same reference counterpart. */
QXGW_API qx_status qx_g0w0_cd_ov_dev(qx_handle h, const qx_g0w0_desc* d,
                                     const double* mo_energy, const double* mo_energy_ref,
                                     const double* B_ov, const int* targets,
                                     const double* static_terms, const double* guess,
                                     double* qp_energy, double* z_factor, double* sigma_re,
                                     double* sigma_im, int* state, qx_stats* stats);
/* qx_g0w0_cd_ov_dev that allocates nothing.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory during the solve (tight serving loops,
 * harnesses that forbid hipMalloc).  Every workspace buffer is carved from a static device
 * array of QXGW_NOALLOC_SCRATCH_BYTES per GPU (held until the process ends; its address is
 * resolved per device in qx_create); the caller still passes inputs and outputs.  A problem
 * whose workspace exceeds the scratch is rejected with QX_ERR_INVALID_INPUT before anything
 * runs (in FP64 about naux <= 950 at nfreq 16, <= 640 at nfreq 32); profiling must be off;
 * the BF16X3 Pi
 * instance is the tuned one if this handle tuned that shape before, else the first
 * candidate (no timing); the factorisation runs on the handle's stream only.  All calls
 * on one GPU share its scratch, which keeps the last call's contents: calls that overlap
 * on one GPU (other handles, other threads) race. */
#define QXGW_NOALLOC_SCRATCH_BYTES (256ull << 20)
QXGW_API qx_status qx_g0w0_cd_ov_noalloc_dev(qx_handle h, const qx_g0w0_desc* d,
                                             const double* mo_energy,
                                             const double* mo_energy_ref, const double* B_ov,
                                             const int* targets, const double* static_terms,
                                             const double* guess, double* qp_energy,
                                             double* z_factor, double* sigma_re,
                                             double* sigma_im, int* state, qx_stats* stats);
/* ---- DF tensor construction ----------------------------------------------------------- */
/* Assembly of the symmetrised DF tensor from the caller's blocks, on the device: with
 * X[:, :nocc] = miP [nmo][nocc][naux], X[:nocc, nocc:] = iaP [nocc][nvir][naux] and
 * X[nocc:, nocc:] = abP [nvir][nvir][naux], the symmetrised rows (X[n][m] + X[m][n]) / 2.
 * Outputs (each nullable): B_ov_sym [nocc*nvir][naux] (the polarizability's operand),
 * mi_sym [nmo][nocc][naux] (the exchange block), and for the ntarget targets the
 * target-partner rows that are not identically zero, in (target, partner) order:
 * B_pair [npair][naux], pair_target (target slot), pair_m (partner), *npair <= pair_capacity
 * (ntarget * nmo always suffices; a smaller capacity that cannot hold them is
 * QX_ERR_DIMENSION_MISMATCH, as is a target outside the orbitals).  abP may be NULL when
 * ntarget is 0.  The [nmo][nmo][naux] tensor itself is never formed.  Replaces the block
 * assembly of quasix/src/evgw.rs:161-205 and the df_mi slice of :305. */
QXGW_API qx_status qx_assemble_blocks_dev(qx_handle h, int nocc, int nvir, int naux,
                                          const double* iaP, const double* miP,
                                          const double* abP, int ntarget, const int* targets,
                                          double* B_ov_sym, double* mi_sym, int pair_capacity,
                                          double* B_pair, int* pair_target, int* pair_m,
                                          int* npair);
/* AO -> MO transform of Cholesky-normalised 3-centre vectors (device pointers, row-major):
 * b_ov[i * nvir + a][P] = sum_{mu,nu} c_occ[mu][i] j3c_ao[mu][nu][P] c_vir[nu][a], with
 * j3c_ao [nao][nao][naux], c_occ [nao][nocc], c_vir [nao][nvir]; FP64 MFMA GEMMs through a
 * [nocc][nao][naux] workspace.  Replaces df::mo_transform::transform_mo_3center
 * (quasix_core/src/df/mo_transform.rs:136-260). */
QXGW_API qx_status qx_df_mo_transform_dev(qx_handle h, int nao, int nocc, int nvir, int naux,
                                          const double* j3c_ao, const double* c_occ,
                                          const double* c_vir, double* b_ov);

/* ---- Exchange ------------------------------------------------------------------------- */
/* Diagonal of the exchange self-energy from the [nmo][nocc][naux] block mi (the reference's
 * ijP / miP tensor): sigma_x[p] = -sum_i sum_R mi[p][i][R]^2 (device pointers).  Replaces
 * EvGWDriver::compute_exchange (quasix_core/src/qp/evgw.rs:1167-1193), whose diagonal the
 * driver uses (:501-502) and quasix/src/evgw.rs:330 returns as "sigma_x". */
QXGW_API qx_status qx_exchange_diag_dev(qx_handle h, int nmo, int nocc, int naux,
                                        const double* mi, double* sigma_x);

/* ---- Analytic continuation (the reference's live self-energy method) ------------------- */
/* Sigma_c(i nu) of the targets on the AC grid: Wc_nm at desc->nfreq scaled Gauss-Legendre
 * frequencies (desc->x0; the reference default is 100), grid nu = {0} + {w_k < iw_cutoff},
 * omega_q = -i nu_q for occupied and +i nu_q for virtual targets.  Pairs as in
 * qx_g0w0_cd_dev (desc->npair rows).  sigma, omega: device [ntarget][nws] complex (re, im
 * interleaved) with nws <= nfreq + 1 returned in *nws.  Energies in mo_energy serve the
 * polarizability, the Green's function and the Fermi level, as in the reference.
 * Replaces correlation_ac::get_sigma_diag (quasix_core/src/selfenergy/correlation_ac.rs:
 * 241-449) and the quasix_core.compute_sigma_c_ac stage function (python/g0w0.rs:721-751). */
QXGW_API qx_status qx_sigma_ac_dev(qx_handle h, const qx_g0w0_desc* d, const double* mo_energy,
                                   const double* B_ov, const double* B_pair,
                                   const int* pair_target, const int* pair_m,
                                   const int* targets, double iw_cutoff, double* sigma,
                                   double* omega, int* nws);
/* Host-only Pade continuation (complex values as (re, im) pairs).
 * qx_thiele: reciprocal-difference coefficients of n points (correlation_ac.rs:470-509).
 * qx_pade_thiele: the continued fraction at a real omega (correlation_ac.rs:530-566).
 * qx_ac_pade_thiele_diag: per-orbital fit on points 1, 7, ..., 37 then every 4th (even
 * count) of sigma/omega [norbs][nw]; coeff [nfit][norbs], ofit [norbs][nfit] (allocate nw
 * entries per orbital); *nfit returned (correlation_ac.rs:587-654). */
QXGW_API qx_status qx_thiele(int n, const double* fn, const double* zn, double* coeff);
QXGW_API qx_status qx_pade_thiele(int n, double omega, const double* zn, const double* coeff,
                                  double* out);
QXGW_API qx_status qx_ac_pade_thiele_diag(int norbs, int nw, const double* sigma,
                                          const double* omega, double* coeff, double* ofit,
                                          int* nfit);
/* Host-only QP step of the evGW driver on the Pade model (quasix_core/src/qp/evgw.rs:
 * 705-820): solver 0 = linearised at prev_qp[p] (or mo_energy[p] when prev_qp is NULL),
 * solver 1 = Newton (newton_solver.rs) from prev_qp[p] or, when NULL, from the linearised
 * value at mo_energy[p], with the linearised fallback on failure.  Z clamped to
 * [cfg->z_min, cfg->z_max]; ef from mo_energy.  Outputs per listed orbital (host). */
QXGW_API qx_status qx_qp_solve_ac(int norbs, int nfit, const double* coeff, const double* ofit,
                                  const int* orbs, int nmo, int nocc, const double* mo_energy,
                                  const double* sigma_x, const double* vxc, int solver,
                                  const double* prev_qp, const qx_newton_config* cfg,
                                  double* qp_energy, double* z_factor, double* sigma_re,
                                  double* sigma_im, int* state);

/* ---- Batched solves (no reference counterpart: the reference solves one system per call;
 * this is for callers with many small systems, e.g. a molecule set or a screening loop) -- */
/* One instance of a batch.  Device pointers except `targets`. */
typedef struct {
    int nocc, nvir, naux;
    int ntarget;
    const double* mo_energy;      /* [nocc + nvir] */
    const double* B_ov;           /* [nocc * nvir][naux], ov-only tensor */
    const int* targets;           /* [ntarget] orbital indices, HOST memory */
    const double* static_terms;   /* [ntarget] Sigma_x - V_xc, or NULL for zero */
} qx_batch_instance;

/* qx_g0w0_cd_ov_dev for many instances of different sizes in grouped launches: one
 * persistent Pi launch over every instance's tiles and frequencies, one launch per
 * factorisation kernel and step over every instance that still has that block column, one
 * grouped Wc_nm contraction (BF16X3: a split launch and the two-frequency BF16 kernel, as
 * the single-instance path; FP64: the FP64 kernel) with its reduction, and one Newton
 * launch.  d supplies nfreq, x0, eta, precision (FP64 or BF16X3) and newton; its sizes are
 * ignored.  Outputs are device arrays of sum(ntarget) entries, instance after instance in
 * the given order.  The
 * first call with a batch (sizes, pointers, targets, options, outputs) runs the launches
 * directly and tunes them; later calls with the same batch replay them as a captured HIP
 * graph, so they follow the current contents of mo_energy, B_ov and static_terms.
 * Instances whose Cholesky fails at a frequency, or with a target whose Newton iterates
 * reach a residue pole, are re-solved by qx_g0w0_cd_ov_dev.  Instances run in waves that fit
 * the workspace budget (environment QXGW_BATCH_GB, default 40% of device memory). */
QXGW_API qx_status qx_g0w0_batch_ov_dev(qx_handle h, const qx_g0w0_desc* d, int ninst,
                                        const qx_batch_instance* inst, double* qp_energy,
                                        double* z_factor, double* sigma_re, double* sigma_im,
                                        int* state, qx_stats* stats);

/* ---- Adjoint (no reference counterpart: the reference has no derivatives) -------------- */
/* grad_B[N_ov][naux] = dL/dB_ov and grad_eps[nmo] = dL/d mo_energy for
 * L = sum_t grad_qp[t] * qp_energy[t], where qp_energy comes from qx_g0w0_cd_ov_dev (FP64) with
 * the same inputs and Sigma_x - V_xc held fixed.  Implicit differentiation of the Newton QP
 * equation, reverse mode through the quadrature, Wc = eps^-1 - I and Pi; re-forms the FP64
 * factorisation and keeps X_k = L_k^-1 of every frequency as the checkpoint.  Requires that
 * no residue pole is active at the solutions (ov-only tensors, frontier targets). */
QXGW_API qx_status qx_g0w0_adjoint_ov_dev(qx_handle h, const qx_g0w0_desc* d,
                                          const double* mo_energy, const double* B_ov,
                                          const int* targets, const double* qp_energy,
                                          const double* grad_qp, double* grad_B,
                                          double* grad_eps);
/* Host-pointer variant of qx_g0w0_cd_dev. */
QXGW_API qx_status qx_g0w0_cd(qx_handle h, const qx_g0w0_desc* d, const double* mo_energy,
                              const double* mo_energy_ref, const double* B_ov,
                              const double* B_pair, const int* pair_target, const int* pair_m,
                              const int* targets, const double* static_terms,
                              const double* guess, double* qp_energy, double* z_factor,
                              double* sigma_re, double* sigma_im, int* state, qx_stats* stats);

#ifdef __cplusplus
}
#endif
#endif /* QXGW_H */
