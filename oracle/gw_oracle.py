"""CPU FP64 restatement of QuasiX's G0W0/evGW correlation path -- TEST INFRASTRUCTURE ONLY.

This module is the oracle: the checker that the MI355X HIP path is compared against.
Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py``
may import it, and only as the checker or as the timed CPU baseline -- never as the
product path.  The product path (``quasix_amd``) fails loudly if its HIP library is
missing; it never routes through this file.

Every function cites the reference file:line it restates (paths relative to the
reference root).  numpy/scipy supply DGEMM / DPOTRF / DPOTRI / DGETRF through their
bundled OpenBLAS, mirroring the reference's ndarray-linalg + OpenBLAS backend.

Conventions (fixed once, identical in the oracle and on the GPU):
  * e_ia = eps_i - eps_a (< 0);  Pi(iw) = 4 * sum_ia B_ia B_ia^T * e_ia / (w^2 + e_ia^2)
    (quasix_core/src/selfenergy/correlation_ac.rs:184-211).  Pi is negative semidefinite,
    eps = I - Pi is SPD,  Wc = eps^-1 - I  (correlation_ac.rs:376-401).
  * B_nm for the self-energy is the symmetrised full tensor assembled from the
    iaP / miP / abP blocks (quasix/src/evgw.rs:161-205).
  * Contour deformation (CD): Sigma_c(E) = I(E) + R(E), with the imaginary-axis
    Green's function shifted by -i*eta*sign(eps_n - eps_m)
    (quasix_core/src/selfenergy/correlation_pyscf_cd.rs:209-226) and residues of
    Wc at z = |E - eps_m| + i*eta for the poles between the Fermi level and E
    (Golze, Dvorak & Rinke, Front. Chem. 7, 377 (2019); PySCF gw_cd, which
    correlation_pyscf_cd.rs:46 cites).  The reference has no live residue code
    (freq/contour_deformation.rs:158-221 is unwired), so this part is pinned by the
    brute-force Casida result in ``oracle/casida.py`` rather than by reference values.

Parity status per function (checks live in tests/test_oracle_*.py):
  gauss_legendre_nodes_weights, get_scaled_legendre_roots  pinned (closed form n=2,
      numpy leggauss, reference tests correlation_ac.rs:722-752, imaginary_grid.rs:276+)
  get_rho_response / compute_polarizability_p0              pinned (brute-force outer
      products, reference known answer polarizability.rs:432+, symmetry :813)
  screened_wc                                               pinned (Woodbury / eigen
      identities, zero-Pi test correlation_fixed.rs)
  wmn_imag, sigma_cd, residues                              pinned by Casida brute force
  get_sigma_diag, thiele, pade_thiele, ac_pade_thiele_diag  pinned (rational-function
      exactness, reference tests correlation_ac.rs:755-810)
  qp_step_ac                                                pinned (Pade model of a known
      rational Sigma: the Newton root solves the QP equation with the mean-field Fermi
      level and not with the current one; linearised step in closed form,
      tests/test_oracle_evgw_pins.py)
  evgw_ac, evgw_cd, convergence_window                      pinned (closed-form evGW fixed
      point of the one-occupied / one-virtual / one-auxiliary model, with and without the
      mean-field Fermi-level shift of the AC QP step; fixed-point residuals of converged
      runs evaluated independently of the driver loop; hand-counted frontier windows and a
      case where the frontier and all-orbital stop rules part, tests/test_oracle_evgw_pins.py)
  metric_sqrt, compute_screening_w                          pinned (scipy.linalg.sqrtm;
      the Dyson identity W = v + v P0 W; identity metric gives (I - P0)^-1;
      tests/test_oracle_stage.py)
  sigma_c_pyscf_cd                                          pinned (equals the imaginary-axis
      part of sigma_cd_eval at E = 0, itself pinned by Casida, when no residue pole lies
      between the Fermi level and 0; zero P0 gives zero; tests/test_oracle_stage.py)
  solve_qp_linearized                                       pinned (closed form, tests/
      test_oracle_stage.py)
  NewtonQPSolver                                            pinned (reference tests
      newton_solver.rs:609-732 restated as known answers)
  transform_mo_3center                                      pinned (identity coefficients
      select (i, nocc + a) AO pairs; orthogonal rotation preserves the Frobenius norm)
  exchange_diag                                             pinned (rank-one closed form
      -a_p^2 |b|^2 |c|^2, tests/test_oracle_core.py)
  diis_solve, DIISMonitor, evgw_step_mix                    pinned (numpy.linalg.solve,
      constrained least-squares closed form of the DIIS combination, the singular
      zero-error case that the driver's call order produces, tests/test_convergence.py)
  g0w0_adjoint                                              pinned (central finite
      differences of g0w0_cd)
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np
import scipy.linalg as sla

__all__ = [
    "gauss_legendre_nodes_weights", "get_scaled_legendre_roots",
    "transform_to_imaginary_axis_pyscf", "transform_weights_pyscf",
    "make_inputs", "assemble_mnP", "pair_list",
    "get_rho_response", "compute_polarizability_p0", "rho_response_complex",
    "screened_wc", "wmn_imag", "sigma_cd_eval", "NewtonQPSolver", "g0w0_cd",
    "evgw_cd", "get_sigma_diag", "thiele", "pade_thiele", "ac_pade_thiele_diag",
    "linearized_qp_ac", "qp_step_ac", "evgw_ac", "convergence_window", "metric_sqrt",
    "compute_screening_w", "sigma_c_pyscf_cd", "solve_qp_linearized", "CONFIGS",
]

# --------------------------------------------------------------------------------------
# A1. Frequency grid
# --------------------------------------------------------------------------------------

def gauss_legendre_nodes_weights(n: int):
    """Golub-Welsch GL nodes/weights on [-1, 1].

    Restates quasix_core/src/selfenergy/correlation_ac.rs:122-160: tridiagonal Jacobi
    matrix with beta_i = i / sqrt((2i-1)(2i+1)), nodes = eigenvalues (DSYEV, ascending),
    weights = 2 * v_0^2.
    """
    if n == 0:
        return np.zeros(0), np.zeros(0)
    if n == 1:
        return np.array([0.0]), np.array([2.0])
    i = np.arange(1, n, dtype=np.float64)
    beta = i / np.sqrt((2.0 * i - 1.0) * (2.0 * i + 1.0))
    J = np.diag(beta, 1) + np.diag(beta, -1)
    w, v = np.linalg.eigh(J)
    return w, 2.0 * v[0, :] ** 2


def get_scaled_legendre_roots(nw: int, x0: float = 0.5):
    """Map GL nodes to [0, inf): w = x0 (1+t)/(1-t), W = w_GL * 2 x0 / (1-t)^2.

    Restates correlation_ac.rs:94-114.
    """
    t, w = gauss_legendre_nodes_weights(nw)
    omt = 1.0 - t
    return x0 * (1.0 + t) / omt, w * 2.0 * x0 / (omt * omt)


def transform_to_imaginary_axis_pyscf(nodes, x0=0.5):
    """Restates quasix_core/src/freq/imaginary_grid.rs:130-145 (1e12 guard at t -> 1)."""
    nodes = np.asarray(nodes, dtype=np.float64)
    out = np.empty_like(nodes)
    for k, x in enumerate(nodes):
        out[k] = 1e12 if abs(1.0 - x) < 1e-14 else x0 * (1.0 + x) / (1.0 - x)
    return out


def transform_weights_pyscf(weights, nodes, x0=0.5):
    """Restates quasix_core/src/freq/imaginary_grid.rs:196-213 (0 weight at t -> 1)."""
    out = np.empty(len(weights))
    for k, (w, x) in enumerate(zip(weights, nodes)):
        out[k] = 0.0 if abs(1.0 - x) < 1e-14 else w * 2.0 * x0 / ((1.0 - x) * (1.0 - x))
    return out


# --------------------------------------------------------------------------------------
# Synthetic inputs (the five workloads the path is measured on)
# --------------------------------------------------------------------------------------

CONFIGS = {
    0: dict(nocc=5, nvir=19, naux=84, nfreq=32),
    1: dict(nocc=40, nvir=400, naux=1800, nfreq=48),
    3: dict(nocc=20, nvir=200, naux=900, nfreq=32),
    4: dict(nocc=100, nvir=1200, naux=5000, nfreq=64),
}


def make_inputs(nocc: int, nvir: int, naux: int, seed: int = 0):
    """Seeded synthetic G0W0 inputs.

    eps_occ ~ U[-1.5, -0.3], eps_vir ~ U[0.05, 3.0] (each sorted);
    B = U diag(s) V^T with Gaussian QR-orthonormalised U (N_ov x r), V (naux x r),
    r = min(N_ov, naux), s_k = 0.5 exp(-k/20).  B is returned as [nocc*nvir, naux]
    (the reference's iaP layout, quasix/src/evgw.rs:105-114).  The draws happen in the
    order eps_occ, eps_vir, U, V.
    """
    rng = np.random.default_rng(seed)
    e_occ = np.sort(rng.uniform(-1.5, -0.3, nocc))
    e_vir = np.sort(rng.uniform(0.05, 3.0, nvir))
    nov = nocc * nvir
    r = min(nov, naux)
    U = np.linalg.qr(rng.standard_normal((nov, r)))[0]
    V = np.linalg.qr(rng.standard_normal((naux, r)))[0]
    s = 0.5 * np.exp(-np.arange(r) / 20.0)
    B = (U * s) @ V.T
    return np.concatenate([e_occ, e_vir]), np.ascontiguousarray(B)


def assemble_mnP(iaP, miP, abP, nocc, nvir, naux):
    """Full symmetric (n_mo, n_mo, n_aux) tensor from the three DF blocks.

    Restates quasix/src/evgw.rs:161-205: fill mn[:, :nocc] from miP, mn[:nocc, nocc:]
    from iaP, mn[nocc:, nocc:] from abP, then symmetrise mn[m,n] = mn[n,m] = average.
    """
    nmo = nocc + nvir
    mn = np.zeros((nmo, nmo, naux))
    mn[:, :nocc, :] = np.asarray(miP).reshape(nmo, nocc, naux)
    mn[:nocc, nocc:, :] = np.asarray(iaP).reshape(nocc, nvir, naux)
    mn[nocc:, nocc:, :] = np.asarray(abP).reshape(nvir, nvir, naux)
    return 0.5 * (mn + mn.transpose(1, 0, 2))


def ov_only_blocks(B, nocc, nvir):
    """miP/abP blocks consistent with an ov-only tensor (oo and vv blocks are zero)."""
    naux = B.shape[1]
    nmo = nocc + nvir
    miP = np.zeros((nmo, nocc, naux))
    miP[nocc:, :, :] = B.reshape(nocc, nvir, naux).transpose(1, 0, 2)
    abP = np.zeros((nvir * nvir, naux))
    return miP.reshape(nmo * nocc, naux), abP


def pair_list(mnP, targets):
    """(target index, partner m) pairs whose B_nm row is not identically zero.

    Rows that are exactly zero contribute exactly zero to Wc_nm, so dropping them
    changes no result; the GPU path receives the same list.
    """
    pt, pm, rows = [], [], []
    for t, n in enumerate(targets):
        for m in range(mnP.shape[1]):
            row = mnP[n, m]
            if np.any(row != 0.0):
                pt.append(t)
                pm.append(m)
                rows.append(row)
    naux = mnP.shape[2]
    return (np.asarray(pt, dtype=np.int32), np.asarray(pm, dtype=np.int32),
            np.asarray(rows, dtype=np.float64).reshape(len(rows), naux))


# --------------------------------------------------------------------------------------
# A2. Polarizability
# --------------------------------------------------------------------------------------

def get_rho_response(omega: float, mo_energy, B_ov, nocc: int):
    """Pi(i omega) = 4 * B^T diag(e/(w^2+e^2)) B,  B_ov laid out [nocc*nvir, naux].

    Restates correlation_ac.rs:184-211 (there lpq_ov is [naux, nocc, nvir]; the product
    is the same).
    """
    e = np.asarray(mo_energy)
    eia = (e[:nocc, None] - e[None, nocc:]).reshape(-1)
    chi = eia / (omega * omega + eia * eia)
    return 4.0 * (B_ov * chi[:, None]).T @ B_ov


def compute_polarizability_p0(ia_P, mo_energy, mo_occ, freqs):
    """Stage API: ia_P [nocc, nvir, naux] -> P0 [nf, naux, naux].

    Restates quasix_core/src/dielectric/polarizability.rs:169-279 (same weights
    4 e_ia/(w^2+e_ia^2), occupied = mo_occ > 0.5, explicit symmetrisation).
    """
    ia_P = np.asarray(ia_P)
    nocc, nvir, naux = ia_P.shape
    occ = [i for i, o in enumerate(mo_occ) if o > 0.5]
    vir = [i for i, o in enumerate(mo_occ) if o < 0.5]
    if len(occ) != nocc or len(vir) != nvir:
        raise ValueError("Occupation mismatch")
    e = np.asarray(mo_energy)
    eia = (e[occ][:, None] - e[vir][None, :]).reshape(-1)
    B = ia_P.reshape(nocc * nvir, naux)
    out = np.empty((len(freqs), naux, naux))
    for k, w in enumerate(freqs):
        wgt = 4.0 * eia / (w * w + eia * eia)
        p = (B * wgt[:, None]).T @ B
        out[k] = 0.5 * (p + p.T)
    return out


def rho_response_complex(z: complex, mo_energy, B_ov, nocc: int):
    """Pi(z) = 2 * B^T diag(1/(z + e) - 1/(z - e)) B at complex z.

    At z = i*omega this equals get_rho_response exactly.  Used for the CD residues at
    z = omega_r + i*eta (PySCF gw_cd real-axis response, cited by
    correlation_pyscf_cd.rs:46; the reference's own residue code is unwired).
    """
    e = np.asarray(mo_energy)
    eia = (e[:nocc, None] - e[None, nocc:]).reshape(-1)
    d = 2.0 * (1.0 / (z + eia) - 1.0 / (z - eia))
    return (B_ov * d[:, None]).T @ B_ov.astype(np.complex128)


# --------------------------------------------------------------------------------------
# A3. Screened interaction
# --------------------------------------------------------------------------------------

def screened_wc(pi):
    """Wc = (I - Pi)^-1 - I with the reference's fallback chain.

    Restates correlation_ac.rs:376-401: Cholesky inverse (DPOTRF+DPOTRI, ``invc``) ->
    LU inverse -> identity (i.e. Wc = 0) if both fail.
    """
    n = pi.shape[0]
    eps = np.eye(n) - pi
    try:
        c, low = sla.cho_factor(eps, lower=True)
        inv = sla.cho_solve((c, low), np.eye(n))
    except (np.linalg.LinAlgError, sla.LinAlgError):
        try:
            inv = np.linalg.inv(eps)
        except np.linalg.LinAlgError:
            inv = np.eye(n)
    return inv - np.eye(n)


def metric_sqrt(v):
    """Symmetric square root of the DF metric: eigendecomposition, eigenvalues floored at
    1e-14, Q diag(sqrt(l)) Q^T, error if |vs vs^T - v|_max > 1e-10.
    Restates quasix_core/src/dielectric/screening.rs:580-639."""
    lam, Q = np.linalg.eigh(np.asarray(v, dtype=np.float64))
    vs = (Q * np.sqrt(np.maximum(lam, 1e-14))) @ Q.T
    err = np.max(np.abs(vs @ vs.T - v))
    if err > 1e-10:
        raise ArithmeticError(f"Eigendecomposition sqrt reconstruction error: {err:.3e}")
    return vs


def compute_screening_w(p0, metric, v_sqrt=None):
    """W_f = vs eps_f^-1 vs^T with eps_f = sym(I - vs P0_f vs^T), Cholesky inverse (LU if it
    fails), W symmetrised.  Restates quasix_core/src/dielectric/screening.rs:164-420."""
    vs = metric_sqrt(metric) if v_sqrt is None else np.asarray(v_sqrt, dtype=np.float64)
    p0 = np.asarray(p0, dtype=np.float64)
    n = p0.shape[1]
    out = np.empty_like(p0)
    for f in range(p0.shape[0]):
        m = (vs @ p0[f]) @ vs.T
        eps = np.eye(n) - m
        eps = 0.5 * (eps + eps.T)
        try:
            c = sla.cho_factor(eps, lower=True)
            inv = sla.cho_solve(c, np.eye(n))
        except (np.linalg.LinAlgError, sla.LinAlgError):
            inv = np.linalg.inv(eps)
        w = (vs @ inv) @ vs.T
        out[f] = 0.5 * (w + w.T)
    return out


def sigma_c_pyscf_cd(mo_energy, eta, p0_iw, lpq, freqs, weights):
    """Static CD self-energy of every orbital (no residues, E = 0).

    Restates quasix_core/src/selfenergy/correlation_pyscf_cd.rs:60-298: Wc_f = inv(I - P0_f)
    - I (general inverse); W_mn = sum_PQ lpq[P,n,m] Wc[P,Q] lpq[Q,m,n];
    Sigma_n = -1/pi sum_f w_f sum_m W_mn emo / (emo^2 + w_f^2),
    emo = -eps_m - i eta sign, sign = +1 if eps_n > eps_m else -1.
    """
    e = np.asarray(mo_energy, dtype=np.float64)
    nmo = e.size
    naux = p0_iw.shape[1]
    sig = np.zeros(nmo, dtype=np.complex128)
    for f, (w, wt) in enumerate(zip(freqs, weights)):
        wc = np.linalg.inv(np.eye(naux) - p0_iw[f]) - np.eye(naux)
        for n in range(nmo):
            W = np.einsum("pm,pq,qm->m", lpq[:, n, :], wc, lpq[:, :, n])
            sign = np.where(e[n] > e, 1.0, -1.0)
            emo = -e - 1j * eta * sign
            sig[n] += np.sum(W * wt * emo / (emo * emo + w * w))
    return sig * (-1.0 / math.pi)


def solve_qp_linearized(mo_energy, sigma_x, sigma_c, vxc):
    """E = eps + Sigma_x + Sigma_c - V_xc with Z = 1 (quasix_core/src/qp/solver.rs:137-221)."""
    e = np.asarray(mo_energy) + np.asarray(sigma_x) + np.asarray(sigma_c) - np.asarray(vxc)
    return e, np.ones_like(e)


def wmn_imag(mo_energy, B_ov, nocc, B_pair, freqs):
    """Wc_nm(i w_k) = b_nm^T Wc(i w_k) b_nm for every pair row, every frequency.

    Restates the per-target contraction of correlation_ac.rs:404-421
    (q_nm = Wc^T L[:, n, :], w_m = sum_Q q_nm * L[:, :, n]) for an explicit pair list.
    Returns [nf, npair].
    """
    out = np.empty((len(freqs), B_pair.shape[0]))
    for k, w in enumerate(freqs):
        wc = screened_wc(get_rho_response(w, mo_energy, B_ov, nocc))
        out[k] = np.einsum("jp,jp->j", B_pair @ wc, B_pair)
    return out


# --------------------------------------------------------------------------------------
# A4-A6. CD self-energy: imaginary-axis quadrature + residues
# --------------------------------------------------------------------------------------

@dataclass
class CDContext:
    """Everything the CD self-energy of a set of targets needs, once Wc_nm(iw) is known."""
    mo_energy: np.ndarray          # energies in the Green's function / poles / Pi
    mo_energy_ref: np.ndarray      # energies deciding the eta sign (mean-field eps)
    nocc: int
    B_ov: np.ndarray               # [N_ov, naux] for the real-axis Pi of residues
    targets: np.ndarray            # orbital index per target
    pair_t: np.ndarray             # target slot per pair
    pair_m: np.ndarray             # partner orbital per pair
    B_pair: np.ndarray             # [npair, naux]
    freqs: np.ndarray
    wts: np.ndarray
    wnm: np.ndarray                # [nf, npair]
    eta: float = 0.01
    res_cache: dict = field(default_factory=dict)

    @property
    def ef(self):
        e = self.mo_energy
        return 0.5 * (e[self.nocc - 1] + e[self.nocc])


def wc_real_axis(ctx: CDContext, z: complex, b: np.ndarray) -> complex:
    """b^T [(I - Pi(z))^-1 - I] b for one pair row at complex frequency z (dense LU)."""
    key = (z, id(b))
    if key in ctx.res_cache:
        return ctx.res_cache[key]
    pi = rho_response_complex(z, ctx.mo_energy, ctx.B_ov, ctx.nocc)
    eps = np.eye(pi.shape[0], dtype=np.complex128) - pi
    x = np.linalg.solve(eps, b.astype(np.complex128))
    val = complex(b @ x - b @ b)
    ctx.res_cache[key] = val
    return val


def sigma_cd_eval(ctx: CDContext, t: int, E: float) -> complex:
    """Sigma_c_n(E) = I_n(E) + R_n(E) for target slot t.

    I_n(E) = -(1/pi) sum_k W_k sum_m Wc_nm(i w_k) * g,
      g = emo / (emo^2 + w_k^2),  emo = E - eps_m - i eta sign(eps_n - eps_m)
      (sign = +1 if eps_n > eps_m else -1; correlation_pyscf_cd.rs:209-226, evaluated at
      the QP energy E instead of 0).
    R_n(E) = +sum_{ef < eps_m < E} Wc_nm(E - eps_m + i eta)
             -sum_{E < eps_m < ef} Wc_nm(eps_m - E + i eta)    (CD residues).
    """
    n = ctx.targets[t]
    sel = ctx.pair_t == t
    m = ctx.pair_m[sel]
    w = ctx.wnm[:, sel]                              # [nf, np]
    em = ctx.mo_energy[m]
    sign = np.where(ctx.mo_energy_ref[n] > ctx.mo_energy_ref[m], 1.0, -1.0)
    emo = (E - em) - 1j * ctx.eta * sign             # [np]
    g = ctx.wts[:, None] * emo[None, :] / (emo[None, :] ** 2 + (ctx.freqs ** 2)[:, None])
    sig = -np.sum(w * g) / math.pi
    ef = ctx.ef
    rows = np.nonzero(sel)[0]
    for j, mm in zip(rows, m):
        e_m = ctx.mo_energy[mm]
        if E > ef and ef < e_m < E:
            sig += wc_real_axis(ctx, complex(E - e_m, ctx.eta), ctx.B_pair[j])
        elif E <= ef and E < e_m < ef:
            sig -= wc_real_axis(ctx, complex(e_m - E, ctx.eta), ctx.B_pair[j])
    return complex(sig)


# --------------------------------------------------------------------------------------
# A9. Newton QP solver
# --------------------------------------------------------------------------------------

@dataclass
class NewtonConfig:
    """Defaults restate quasix_core/src/qp/newton_solver.rs:84-103."""
    max_iterations: int = 50
    energy_tolerance: float = 1e-8
    initial_damping: float = 1.0
    min_damping: float = 0.1
    max_damping: float = 1.0
    max_step: float = 1.0
    derivative_delta: float = 1e-6
    use_richardson: bool = True
    z_min: float = 0.1
    z_max: float = 0.999
    use_bisection_fallback: bool = True
    max_bisection_iterations: int = 30


@dataclass
class NewtonResult:
    qp_energy: float
    residual: float
    iterations: int
    converged: bool
    sigma_c: complex
    z_factor: float
    method: str


class ConvergenceError(RuntimeError):
    pass


class NewtonQPSolver:
    """Restates quasix_core/src/qp/newton_solver.rs:128-402.

    sigma_c_eval takes omega = E - fermi_level (as the reference's closure does).
    """

    def __init__(self, epsilon_hf, sigma_x, vxc, fermi_level, config=None):
        self.c = config or NewtonConfig()
        self.eps = epsilon_hf
        self.sx = sigma_x
        self.vxc = vxc
        self.ef = fermi_level

    def _deriv(self, f, omega):          # newton_solver.rs:282-303
        de = self.c.derivative_delta
        if self.c.use_richardson:
            p1 = f(omega + de).real
            m1 = f(omega - de).real
            p2 = f(omega + 2 * de).real
            m2 = f(omega - 2 * de).real
            return (8.0 * (p1 - m1) - (p2 - m2)) / (12.0 * de)
        return (f(omega + de).real - f(omega - de).real) / (2 * de)

    def _z(self, f, energy):             # newton_solver.rs:306-315
        d = self._deriv(f, energy - self.ef)
        return min(max(1.0 / (1.0 - d), self.c.z_min), self.c.z_max)

    def _residual(self, f, energy):      # newton_solver.rs:394-401
        return energy - self.eps - self.sx - f(energy - self.ef).real + self.vxc

    def solve(self, f, initial_guess):   # newton_solver.rs:174-279
        c = self.c
        energy = initial_guess
        damping = c.initial_damping
        prev = math.inf
        for it in range(c.max_iterations):
            omega = energy - self.ef
            sc = f(omega)
            res = energy - self.eps - self.sx - sc.real + self.vxc
            ares = abs(res)
            if ares < c.energy_tolerance:
                return NewtonResult(energy, ares, it + 1, True, sc, self._z(f, energy), "newton")
            fp = 1.0 - self._deriv(f, omega)
            if abs(fp) < 1e-10:
                step = -math.copysign(1.0, res) * min(c.max_step, ares)
            else:
                step = -res / fp
            step = min(max(step, -c.max_step), c.max_step)
            if ares > prev:
                damping = max(damping * 0.5, c.min_damping)
            elif ares < 0.5 * prev and damping < c.max_damping:
                damping = min(damping * 1.2, c.max_damping)
            energy += damping * step
            prev = ares
        if c.use_bisection_fallback:
            return self._bisect(f, initial_guess)
        raise ConvergenceError("Newton-Raphson failed to converge")

    def _bisect(self, f, guess):         # newton_solver.rs:317-391
        c = self.c
        width = 2.0
        lo, hi = guess - width, guess + width
        if self._residual(f, lo) * self._residual(f, hi) > 0.0:
            for k in range(1, 5):
                w = width * 2.0 ** k
                lo, hi = guess - w, guess + w
                if self._residual(f, lo) * self._residual(f, hi) <= 0.0:
                    break
        for it in range(c.max_bisection_iterations):
            mid = 0.5 * (lo + hi)
            fm = self._residual(f, mid)
            if abs(fm) < c.energy_tolerance or 0.5 * (hi - lo) < c.energy_tolerance:
                return NewtonResult(mid, abs(fm), it + 1, True, f(mid - self.ef),
                                    self._z(f, mid), "bisection")
            if fm * self._residual(f, lo) < 0.0:
                hi = mid
            else:
                lo = mid
        raise ConvergenceError("Bisection fallback failed to converge")


def linearized_guess(f, eps_p, static, ef, de=1e-6, z_min=0.1, z_max=0.999):
    """Iteration-0 Newton start: E = eps + Z (static + ReSigma(eps)), forward-difference Z.

    Restates quasix_core/src/qp/evgw.rs:730-741.
    """
    om = eps_p - ef
    s0 = f(om).real
    s1 = f(om + de).real
    z = min(max(1.0 / (1.0 - (s1 - s0) / de), z_min), z_max)
    return eps_p + z * (static + s0)


# --------------------------------------------------------------------------------------
# Full G0W0 solve (CD + Newton) and evGW
# --------------------------------------------------------------------------------------

def frontier_targets(nocc, nmo, below=5, above=5):
    """HOMO-below .. LUMO+above (clipped to the orbital range)."""
    lo = max(0, nocc - 1 - below)
    hi = min(nmo, nocc + above + 1)
    return np.arange(lo, hi, dtype=np.int64)


def g0w0_cd(mo_energy, B_ov, nocc, targets, freqs, wts, *, pairs=None, static=None,
            eta=0.01, mo_energy_ref=None, newton=None, guesses=None):
    """One G0W0 solve: Pi -> Wc -> Wc_nm(iw) -> Newton on Sigma_c = I + R.

    Follows the iteration-0 branch of quasix_core/src/qp/evgw.rs:603-783 with the AC
    Sigma replaced by the CD Sigma.  ``pairs`` = (pair_t, pair_m, B_pair); default is the
    ov-only pair list.  ``static`` = Sigma_x - V_xc per target (default 0).
    Returns dict(E, Z, sigma, wnm, ...).
    """
    e = np.asarray(mo_energy, dtype=np.float64)
    targets = np.asarray(targets)
    nt = len(targets)
    if pairs is None:
        miP, abP = ov_only_blocks(B_ov, nocc, len(e) - nocc)
        mn = assemble_mnP(B_ov, miP, abP, nocc, len(e) - nocc, B_ov.shape[1])
        pairs = pair_list(mn, targets)
    pt, pm, Bp = pairs
    static = np.zeros(nt) if static is None else np.asarray(static, dtype=np.float64)
    wnm = wmn_imag(e, B_ov, nocc, Bp, freqs)
    ctx = CDContext(e, e if mo_energy_ref is None else np.asarray(mo_energy_ref), nocc, B_ov,
                    targets, pt, pm, Bp, np.asarray(freqs), np.asarray(wts), wnm, eta)
    ef = ctx.ef
    E = np.empty(nt)
    Z = np.empty(nt)
    S = np.empty(nt, dtype=np.complex128)
    conv = np.empty(nt, dtype=bool)
    for t, n in enumerate(targets):
        f = lambda om, t=t: sigma_cd_eval(ctx, t, om + ef)
        g = linearized_guess(f, e[n], static[t], ef) if guesses is None else guesses[t]
        solver = NewtonQPSolver(e[n], static[t], 0.0, ef, newton)
        try:
            r = solver.solve(f, g)
            E[t], Z[t], S[t], conv[t] = r.qp_energy, r.z_factor, r.sigma_c, r.converged
        except ConvergenceError:        # evgw.rs:766-781: fall back to linearized
            om = g - ef
            s0 = f(om)
            s1 = f(om + 1e-6)
            z = min(max(1.0 / (1.0 - (s1.real - s0.real) / 1e-6), 0.1), 0.999)
            E[t], Z[t], S[t], conv[t] = e[n] + z * (static[t] + s0.real), z, s0, False
    return dict(E=E, Z=Z, sigma=S, converged=conv, wnm=wnm, pairs=pairs, ef=ef)


def transform_mo_3center(j3c_ao, c_occ, c_vir):
    """(ia|P) = sum_{mu nu} C_occ[mu, i] (mu nu|P) C_vir[nu, a] as [nocc * nvir, naux].
    Restates quasix_core/src/df/mo_transform.rs:136-260 (the AO -> MO transform)."""
    t = np.einsum("mi,mnP->inP", c_occ, j3c_ao)
    return np.einsum("inP,na->iaP", t, c_vir).reshape(c_occ.shape[1] * c_vir.shape[1], -1)


def exchange_diag(mi, nmo, nocc):
    """Diagonal of EvGWDriver::compute_exchange (quasix_core/src/qp/evgw.rs:1167-1193):
    Sigma_x[p, p] = -sum_{i < nocc} sum_R (p i | R)^2 over the [nmo, nocc, naux] block."""
    m = np.asarray(mi, dtype=np.float64).reshape(nmo, nocc, -1)
    return -np.einsum("piR,piR->p", m, m)


def diis_solve(a, b):
    """Gaussian elimination with partial pivoting and the reference's checks.

    Restates ConvergenceMonitor::solve_linear_system (quasix_core/src/qp/convergence.rs:
    440-540): pivot tolerance 1e-12 * max(max|a|, 1), rows eliminated only where
    |a_ki| > 1e-10 |pivot|, coefficient sum (all but the multiplier) within 0.5 of 1.
    Raises ArithmeticError where the reference returns NumericalError.
    """
    a = np.array(a, dtype=np.float64)
    n = len(a)
    m = np.hstack([a, np.asarray(b, dtype=np.float64).reshape(n, 1)])
    tol = 1e-12 * max(np.abs(a).max(), 1.0)
    for i in range(n - 1):
        p = i + int(np.argmax(np.abs(m[i:, i])))
        if abs(m[p, i]) < tol:
            raise ArithmeticError("singular DIIS matrix")
        m[[i, p]] = m[[p, i]]
        for k in range(i + 1, n):
            if abs(m[k, i]) > 1e-10 * abs(m[i, i]):
                m[k] = m[k] - m[k, i] / m[i, i] * m[i]
    if abs(m[n - 1, n - 1]) < tol:
        raise ArithmeticError("singular DIIS matrix")
    x = np.zeros(n)
    for i in reversed(range(n)):
        x[i] = (m[i, n] - np.dot(m[i, i + 1:n], x[i + 1:])) / m[i, i]
    if abs(np.sum(x[:-1]) - 1.0) > 0.5 or not np.all(np.isfinite(x)):
        raise ArithmeticError("DIIS coefficients rejected")
    return x


class DIISMonitor:
    """The (E, Z) history and DIIS extrapolation of quasix_core/src/qp/convergence.rs:
    146-189 (check_convergence: append, keep the last ``space`` entries) and 264-328
    (apply_diis: error = E - E_last, Pulay matrix with a -1 border, keep ``space``
    vectors)."""

    def __init__(self, space=6):
        self.space = space
        self.hist_e, self.hist_z, self.vecs, self.errs = [], [], [], []

    def record(self, e, z):
        self.hist_e.append(np.array(e, dtype=float))
        self.hist_z.append(np.array(z, dtype=float))
        self.hist_e, self.hist_z = self.hist_e[-self.space:], self.hist_z[-self.space:]

    def extrapolate(self, e, z):
        e = np.asarray(e, dtype=float)
        z = np.asarray(z, dtype=float)
        if not self.hist_e:
            return e, z
        self.vecs = (self.vecs + [np.concatenate([e, z])])[-self.space:]
        self.errs = (self.errs + [e - self.hist_e[-1]])[-self.space:]
        n = len(self.vecs)
        if n < 2:
            return e, z
        bm = -np.ones((n + 1, n + 1))
        bm[n, n] = 0.0
        bm[:n, :n] = np.array([[ei @ ej for ej in self.errs] for ei in self.errs])
        c = diis_solve(bm, np.r_[np.zeros(n), -1.0])
        x = sum(ci * v for ci, v in zip(c[:n], self.vecs))
        return x[:e.size], x[e.size:]


def evgw_step_mix(it, mon, e_old, z_old, e_new, z_new, damping, diis, z_min=0.1, z_max=0.999):
    """Mixing step of quasix_core/src/qp/evgw.rs:832-927: new values at iteration 0, damping
    at iteration 1 and whenever DIIS is off, fails or moves E by >= 0.5 Ha or Z by >= 0.2;
    the new values are recorded in the history before the DIIS call (:848)."""
    damped = ((1 - damping) * e_old + damping * e_new, (1 - damping) * z_old + damping * z_new)
    if it == 0:
        mon.record(e_new, z_new)
        return np.array(e_new, dtype=float), np.array(z_new, dtype=float)
    if it == 1:
        mon.record(*damped)
        return damped
    if not diis:
        return damped
    mon.record(e_new, z_new)
    try:
        de, dz = mon.extrapolate(e_new, z_new)
        ok = np.max(np.abs(de - e_new)) < 0.5 and np.max(np.abs(dz - z_new)) < 0.2
        e, z = (de, dz) if ok else damped
    except ArithmeticError:
        e, z = damped
    return e, np.clip(z, z_min, z_max)


def convergence_window(mo_energy, nocc, frontier_convergence=True, frontier_window=1.0):
    """[start, end) of the orbitals whose change decides convergence.

    Restates EvGWDriver::determine_convergence_window (quasix_core/src/qp/evgw.rs:1121-1160),
    evaluated on the mean-field energies: every orbital when frontier_convergence is off;
    otherwise the occupied orbitals found scanning down from the HOMO while within
    frontier_window of it, and the virtual ones scanning up from the LUMO while within
    frontier_window of it (HOMO and LUMO always included).
    """
    e = np.asarray(mo_energy, dtype=np.float64)
    nmo = e.size
    if not frontier_convergence:
        return 0, nmo
    start = max(nocc - 1, 0)
    for i in range(nocc - 1, -1, -1):
        if e[i] >= e[nocc - 1] - frontier_window:
            start = i
        else:
            break
    end = nocc + 1
    for i in range(nocc, nmo):
        if e[i] <= e[nocc] + frontier_window:
            end = i + 1
        else:
            break
    return start, end


def _window_changes(new, old, window):
    """(max, rms) of |new - old| over the window (evgw.rs:1002-1020)."""
    lo, hi = window
    d = np.abs(np.asarray(new) - np.asarray(old))[lo:hi]
    return float(np.max(d)), float(np.sqrt(np.sum(d * d) / (hi - lo)))


def evgw_cd(mo_energy, B_ov, nocc, targets, freqs, wts, *, n_iter=8, damping=0.5,
            eta=0.01, tol=1e-4, static=None, diis=False, diis_space=6,
            frontier_convergence=True, frontier_window=1.0, newton=None, pairs=None,
            solver="newton"):
    """Eigenvalue-self-consistent GW on the CD self-energy.

    Follows quasix_core/src/qp/evgw.rs:603-1027 with the AC self-energy replaced by the CD
    one: iteration 0 is G0W0 on the mean-field energies; iteration k > 0 rebuilds Pi, the
    Green's function poles and the residues from the previous QP energies (:626-633); the
    (E, Z) update is ``evgw_step_mix`` (:832-927, damping and DIIS); Newton restarts from
    the previous QP energy (:742).  Orbitals outside ``targets`` keep their mean-field
    energies.  Stop rule (:929-1027): after every iteration, including the first, the max
    |dE| over ``convergence_window`` of the mean-field energies is compared with ``tol``.
    ``pairs`` = (pair_t, pair_m, B_pair) of the targets (default: the ov-only pair list);
    ``solver`` "newton" or "linearized" (the linearised step of :786-820 at the mean-field
    energy in iteration 0 and at the previous QP energy after it).
    """
    e0 = np.asarray(mo_energy, dtype=np.float64)
    targets = np.asarray(targets)
    win = convergence_window(e0, nocc, frontier_convergence, frontier_window)
    cur = e0.copy()
    z = np.ones(len(targets))
    mon = DIISMonitor(diis_space)
    hist = []
    res = None
    converged = False
    for it in range(n_iter):
        guesses = None if it == 0 else cur[targets]
        res = _g0w0_iter(cur, e0, B_ov, nocc, targets, freqs, wts, static, eta, guesses,
                         newton, pairs, solver)
        et, z = evgw_step_mix(it, mon, cur[targets], z, res["E"], res["Z"], damping, diis)
        new = cur.copy()
        new[targets] = et
        dmax, drms = _window_changes(new, cur, win)
        zf = np.ones(e0.size)
        zf[targets] = z
        hist.append(dict(cycle=it, qp_energies=new.copy(), z_factors=zf, energy_change=dmax,
                         rms_change=drms))
        cur = new
        if dmax < tol:
            converged = True
            break
    return dict(E=cur, Z=z, history=hist, last=res, converged=converged, n_cycles=len(hist),
                window=win)


def _g0w0_iter(cur, e0, B_ov, nocc, targets, freqs, wts, static, eta, guesses, newton=None,
               pairs=None, solver="newton"):
    """One evGW step: Pi/poles from ``cur``; QP equation anchored at the mean-field e0."""
    nt = len(targets)
    static = np.zeros(nt) if static is None else np.asarray(static)
    if pairs is None:
        miP, abP = ov_only_blocks(B_ov, nocc, len(cur) - nocc)
        mn = assemble_mnP(B_ov, miP, abP, nocc, len(cur) - nocc, B_ov.shape[1])
        pairs = pair_list(mn, targets)
    pt, pm, Bp = pairs
    wnm = wmn_imag(cur, B_ov, nocc, Bp, freqs)
    ctx = CDContext(cur, e0, nocc, B_ov, np.asarray(targets), pt, pm, Bp,
                    np.asarray(freqs), np.asarray(wts), wnm, eta)
    ef = ctx.ef
    E = np.empty(nt)
    Z = np.empty(nt)
    S = np.empty(nt, dtype=np.complex128)
    for t, n in enumerate(targets):
        f = lambda om, t=t: sigma_cd_eval(ctx, t, om + ef)
        if solver == "linearized":
            ev = e0[n] if guesses is None else guesses[t]
            s0, s1 = f(ev - ef), f(ev - ef + 1e-6).real
            z = min(max(1.0 / (1.0 - (s1 - s0.real) / 1e-6), 0.1), 0.999)
            E[t], Z[t], S[t] = e0[n] + z * (static[t] + s0.real), z, s0
            continue
        g = linearized_guess(f, e0[n], static[t], ef) if guesses is None else guesses[t]
        r = NewtonQPSolver(e0[n], static[t], 0.0, ef, newton).solve(f, g)
        E[t], Z[t], S[t] = r.qp_energy, r.z_factor, r.sigma_c
    return dict(E=E, Z=Z, sigma=S, wnm=wnm)


# --------------------------------------------------------------------------------------
# A4/A7/A8. Analytic-continuation path (the reference's live Sigma method)
# --------------------------------------------------------------------------------------

def get_sigma_diag(mo_energy, lpq, orbs, nocc, nw=100, x0=0.5, iw_cutoff=5.0):
    """Sigma_c(i nu) on the AC grid.  Restates correlation_ac.rs:241-449.

    lpq is the full [naux, nmo, nmo] tensor.  Returns (sigma [norbs, nw_sigma],
    omega [norbs, nw_sigma]) as complex arrays.
    """
    e = np.asarray(mo_energy)
    naux, nmo, _ = lpq.shape
    if nocc == 0 or nocc >= nmo:
        raise ValueError(f"Invalid nocc={nocc}")
    freqs, wts = get_scaled_legendre_roots(nw, x0)
    idx = [k for k, f in enumerate(freqs) if f < iw_cutoff]
    nws = min(len(idx) + 1, nw + 1)
    ef = 0.5 * (e[nocc - 1] + e[nocc])
    om_occ = np.zeros(nws, dtype=np.complex128)
    om_vir = np.zeros(nws, dtype=np.complex128)
    for k, fi in enumerate(idx):
        if k + 1 < nws:
            om_occ[k + 1] = -1j * freqs[fi]
            om_vir[k + 1] = 1j * freqs[fi]
    emo_occ = om_occ[None, :] + ef - e[:, None]
    emo_vir = om_vir[None, :] + ef - e[:, None]
    lpq_ov = lpq[:, :nocc, nocc:]
    B_ov = lpq_ov.transpose(1, 2, 0).reshape(nocc * (nmo - nocc), naux)
    sigma = np.zeros((len(orbs), nws), dtype=np.complex128)
    omega = np.array([om_occ if o < nocc else om_vir for o in orbs])
    for w in range(nw):
        pi = get_rho_response(freqs[w], e, B_ov, nocc)
        eps = np.eye(naux) - pi
        try:
            c = sla.cho_factor(eps, lower=True)
            inv = sla.cho_solve(c, np.eye(naux))
        except (np.linalg.LinAlgError, sla.LinAlgError):
            try:
                inv = np.linalg.inv(eps)
            except np.linalg.LinAlgError:
                inv = np.eye(naux)
        pinv = inv - np.eye(naux)
        g_occ = wts[w] * emo_occ / (emo_occ ** 2 + freqs[w] ** 2)
        g_vir = wts[w] * emo_vir / (emo_vir ** 2 + freqs[w] ** 2)
        for i, n in enumerate(orbs):
            q = pinv.T @ lpq[:, n, :]
            wm = np.sum(q * lpq[:, :, n], axis=0)
            g = g_occ if n < nocc else g_vir
            sigma[i] += (-1.0 / math.pi) * (wm @ g.real + 1j * (wm @ g.imag))
    return sigma, omega


def thiele(fn, zn):
    """Thiele reciprocal-difference coefficients.  Restates correlation_ac.rs:470-509."""
    fn = np.asarray(fn, dtype=np.complex128)
    zn = np.asarray(zn, dtype=np.complex128)
    nfit = len(fn)
    if nfit == 0:
        return np.zeros(0, dtype=np.complex128)
    g = np.zeros((nfit, nfit), dtype=np.complex128)
    g[:, 0] = fn
    for j in range(1, nfit):
        for i in range(j, nfit):
            d = (zn[i] - zn[j - 1]) * g[i, j - 1]
            g[i, j] = 1e10 if abs(d) < 1e-14 else (g[j - 1, j - 1] - g[i, j - 1]) / d
    return np.diag(g).copy()


def pade_thiele(omega, zn, coeff):
    """Continued-fraction evaluation.  Restates correlation_ac.rs:530-566."""
    nfit = len(coeff)
    if nfit == 0:
        return 0j
    if nfit == 1:
        return complex(coeff[0])
    z = complex(omega, 0.0)
    x = coeff[nfit - 1] * (z - zn[nfit - 2])
    for i in range(nfit - 2, 0, -1):
        d = 1.0 + x
        x = 1e10 if abs(d) < 1e-14 else coeff[i] * (z - zn[i - 1]) / d
    d = 1.0 + x
    return complex(coeff[0]) if abs(d) < 1e-14 else complex(coeff[0] / d)


def ac_pade_thiele_diag(sigma, omega):
    """Point selection 1,7,..,37 then every 4th; even count.  Restates :587-654."""
    norbs, nw = sigma.shape
    idx = list(range(1, min(40, nw), 6))
    if idx:
        idx += list(range(idx[-1] + 4, nw, 4))
    nfit = (len(idx) // 2) * 2
    idx = idx[:nfit]
    coeff = np.zeros((nfit, norbs), dtype=np.complex128)
    ofit = omega[:, idx]
    for p in range(norbs):
        coeff[:, p] = thiele(sigma[p, idx], ofit[p])
    return coeff, ofit


def linearized_qp_ac(mo_energy, sigma_x, vxc, coeff, ofit, orbs, nocc, de=1e-6):
    """E = eps + Z (Sigma_x + ReSigma_c(eps) - Vxc), forward-difference Z (no clamp).

    Restates correlation_ac.rs:684-714.
    """
    e = np.asarray(mo_energy)
    ef = 0.5 * (e[nocc - 1] + e[nocc])
    out = np.empty(len(orbs))
    for i, p in enumerate(orbs):
        s = pade_thiele(e[p] - ef, ofit[i], coeff[:, i]).real
        sp = pade_thiele(e[p] - ef + de, ofit[i], coeff[:, i]).real
        z = 1.0 / (1.0 - (sp - s) / de)
        out[i] = e[p] + z * (sigma_x[p] + s - vxc[p])
    return out


def qp_step_ac(coeff, ofit, orbs, e_hf, nocc, sigma_x, vxc, solver, prev_qp, z_min=0.1,
               z_max=0.999, newton=None):
    """QP step of the evGW driver on the Pade model, quasix_core/src/qp/evgw.rs:705-820:
    linearised at prev_qp (or the mean-field energy at iteration 0) with a clamped forward-
    difference Z, or Newton from prev_qp / the linearised guess with the linearised value at
    the starting point as the fallback.  ef from the mean-field energies."""
    e = np.asarray(e_hf, dtype=np.float64)
    ef = 0.5 * (e[nocc - 1] + e[nocc])
    E, Z = np.empty(len(orbs)), np.empty(len(orbs))
    for i, p in enumerate(orbs):
        f = lambda om, i=i: pade_thiele(om, ofit[i], coeff[:, i])

        def lin(ev):
            s0, s1 = f(ev - ef).real, f(ev - ef + 1e-6).real
            z = min(max(1.0 / (1.0 - (s1 - s0) / 1e-6), z_min), z_max)
            return e[p] + z * (sigma_x[p] + s0 - vxc[p]), z

        start = e[p] if prev_qp is None else prev_qp[p]
        if solver == "linearized":
            E[i], Z[i] = lin(start)
            continue
        guess = lin(e[p])[0] if prev_qp is None else start
        try:
            r = NewtonQPSolver(e[p], sigma_x[p], vxc[p], ef, newton).solve(f, guess)
            E[i], Z[i] = r.qp_energy, r.z_factor
        except RuntimeError:
            E[i], Z[i] = lin(guess)
    return E, Z


def evgw_ac(mo_energy, lpq, nocc, orbs=None, *, nw=100, x0=0.5, iw_cutoff=5.0, n_iter=8,
            damping=0.5, tol=1e-4, diis=False, diis_space=6, solver="linearized",
            sigma_x=None, vxc=None, frontier_convergence=True, frontier_window=1.0,
            newton=None):
    """evGW on the AC self-energy (the reference's live driver, quasix_core/src/qp/evgw.rs:
    420-1049): get_sigma_diag on the previous QP energies (:603-690), ac_pade_thiele_diag,
    the QP step ``qp_step_ac`` and the mixing step ``evgw_step_mix``; ``orbs`` (default all,
    as the reference) are updated, others keep their mean-field energies.  Stop rule
    (:929-1027): after every iteration, including the first, the max |dE| over
    ``convergence_window`` of the mean-field energies below ``tol``; the history records
    that max and the RMS over the same window, and each iteration's Z."""
    e0 = np.asarray(mo_energy, dtype=np.float64)
    nmo = e0.size
    orbs = np.arange(nmo) if orbs is None else np.asarray(orbs)
    sx = np.zeros(nmo) if sigma_x is None else np.asarray(sigma_x, dtype=np.float64)
    vx = np.zeros(nmo) if vxc is None else np.asarray(vxc, dtype=np.float64)
    win = convergence_window(e0, nocc, frontier_convergence, frontier_window)
    cur = e0.copy()
    z = np.ones(len(orbs))
    mon = DIISMonitor(diis_space)
    hist = []
    converged = False
    for it in range(n_iter):
        sig, om = get_sigma_diag(cur, lpq, orbs, nocc, nw, x0, iw_cutoff)
        coeff, ofit = ac_pade_thiele_diag(sig, om)
        prev = None if it == 0 else cur
        En, Zn = qp_step_ac(coeff, ofit, orbs, e0, nocc, sx, vx, solver, prev, newton=newton)
        et, z = evgw_step_mix(it, mon, cur[orbs], z, En, Zn, damping, diis)
        new = cur.copy()
        new[orbs] = et
        dmax, drms = _window_changes(new, cur, win)
        zf = np.ones(nmo)
        zf[orbs] = z
        hist.append(dict(cycle=it, qp_energies=new.copy(), z_factors=zf, energy_change=dmax,
                         rms_change=drms))
        cur = new
        if dmax < tol:
            converged = True
            break
    return dict(E=cur, Z=z, history=hist, converged=converged, n_cycles=len(hist), window=win)


# --------------------------------------------------------------------------------------
# Adjoint of the G0W0 solve (d E_QP / d B, d E_QP / d eps) -- no reference counterpart
# --------------------------------------------------------------------------------------

def _g_and_dg(E, em, sign_eta, freqs, wts):
    """Green's-function weights g[k, j] = W_k emo / (emo^2 + w_k^2) and dg/dE (complex)."""
    emo = (E - em)[None, :] - 1j * sign_eta[None, :]
    w2 = (freqs ** 2)[:, None]
    den = emo ** 2 + w2
    g = wts[:, None] * emo / den
    dg = wts[:, None] * (w2 - emo ** 2) / den ** 2
    return g, dg


def g0w0_adjoint(mo_energy, B_ov, nocc, targets, freqs, wts, E, grad_E, *, eta=0.01):
    """Gradient of L = sum_t grad_E[t] * E_QP[t] w.r.t. B_ov and mo_energy (ov-only pairs,
    no residue poles at the solution; Sigma_x - V_xc held fixed).

    Implicit differentiation of f(E) = E - eps_n - s - Re Sigma_n(E) = 0:
      dE/dtheta = Z (d eps_n/dtheta + d Re Sigma_n/dtheta),  Z = 1 / (1 - d Re Sigma_n / dE);
    Sigma_n = -(1/pi) sum_k sum_m w_nm(k) g(E - eps_m, w_k);  w_nm(k) = b^T Wc_k b;
    dWc = eps^-1 dPi eps^-1;  Pi_k = sum_r s_k(r) b_r b_r^T,  s_k = 4 e / (w_k^2 + e^2).
    Returns (grad_B [N_ov, naux], grad_eps [nmo]).
    """
    e = np.asarray(mo_energy, dtype=np.float64)
    B = np.asarray(B_ov, dtype=np.float64)
    nmo = e.size
    nvir = nmo - nocc
    naux = B.shape[1]
    targets = np.asarray(targets)
    eia = (e[:nocc, None] - e[None, nocc:]).reshape(-1)
    nf = len(freqs)
    # forward pieces with checkpoints
    einv, wc = [], []
    for w in freqs:
        pi = get_rho_response(w, e, B, nocc)
        inv = np.linalg.inv(np.eye(naux) - pi)
        einv.append(inv)
        wc.append(inv - np.eye(naux))
    gB = np.zeros_like(B)
    geps = np.zeros(nmo)
    pibar = np.zeros((nf, naux, naux))
    for t, n in enumerate(targets):
        if n < nocc:
            rows = n * nvir + np.arange(nvir)
            m = nocc + np.arange(nvir)
        else:
            rows = np.arange(nocc) * nvir + (n - nocc)
            m = np.arange(nocc)
        bp = B[rows]                                              # [np, naux]
        wnm = np.stack([np.einsum("jp,jp->j", bp @ wc[k], bp) for k in range(nf)])
        sign = np.where(e[n] > e[m], 1.0, -1.0) * eta
        g, dg = _g_and_dg(E[t], e[m], sign, freqs, wts)
        dsig = -np.sum(wnm * dg.real) / math.pi
        zc = grad_E[t] / (1.0 - dsig)                             # c_t * Z_raw
        geps[n] += zc                                              # anchor eps_n
        geps[m] += zc * (np.sum(wnm * dg.real, axis=0) / math.pi)  # d/d eps_m = -d/dE
        abar = zc * (-1.0 / math.pi) * g.real                     # [nf, np]
        for k in range(nf):
            wb = (bp * abar[k][:, None]).T @ bp                   # W-bar_k
            pibar[k] += einv[k] @ wb @ einv[k]
            np.add.at(gB, rows, 2.0 * abar[k][:, None] * (bp @ wc[k]))
    for k, w in enumerate(freqs):
        s = 4.0 * eia / (w * w + eia * eia)
        ds = 4.0 * (w * w - eia * eia) / (w * w + eia * eia) ** 2
        Y = B @ pibar[k]
        gB += 2.0 * s[:, None] * Y
        ebar = ds * np.einsum("rp,rp->r", Y, B)
        geps[:nocc] += ebar.reshape(nocc, nvir).sum(axis=1)
        geps[nocc:] -= ebar.reshape(nocc, nvir).sum(axis=0)
    return gB, geps
