"""Brute-force RPA (Casida) G0W0 self-energy -- TEST INFRASTRUCTURE ONLY (oracle checker).

Independent of any frequency grid: diagonalises the spin-adapted RPA matrix
    M = Delta^2 + 4 Delta^(1/2) B B^T Delta^(1/2),   Delta_ia = eps_a - eps_i > 0
(B laid out [N_ov, naux], the convention of correlation_ac.rs:184-211 where
Pi(i w) = 4 sum_ia B_ia B_ia^T e_ia / (w^2 + e_ia^2)).  With M = X diag(Omega^2) X^T,
    Wc_nm(w) = sum_s |w^s_nm|^2 * 2 Omega_s / (w^2 - Omega_s^2),
    |w^s_nm|^2 = 4 (b_nm . B^T Delta^(1/2) x_s)^2 / (2 Omega_s),
and the exact G0W0 correlation self-energy (eta -> 0+) is
    Sigma_c_n(E) = sum_{m,s} |w^s_nm|^2 / (E - eps_m + sgn_m (Omega_s - i eta)),
sgn_m = +1 below the Fermi level, -1 above.  The CD self-energy (quadrature + residues)
must converge to this as the grid is refined and eta -> 0; the imaginary-axis Wc_nm(i w)
must equal sum_s |w^s_nm|^2 (-2 Omega_s) / (w^2 + Omega_s^2) at every w.
Sizes: N_ov up to a few hundred (dense eigh).
"""
from __future__ import annotations

import numpy as np


def casida_poles(mo_energy, B_ov, nocc):
    """Return (Omega [N_ov], A [N_ov, naux]) with A_s = 2 B^T Delta^(1/2) x_s / sqrt(2 Omega_s).

    Then Wc(w) in the auxiliary basis is sum_s A_s A_s^T * 2 Omega_s / (w^2 - Omega_s^2).
    """
    e = np.asarray(mo_energy, dtype=np.float64)
    delta = -(e[:nocc, None] - e[None, nocc:]).reshape(-1)
    sq = np.sqrt(delta)
    M = np.diag(delta ** 2) + 4.0 * (sq[:, None] * (B_ov @ B_ov.T)) * sq[None, :]
    om2, X = np.linalg.eigh(M)
    om = np.sqrt(om2)
    A = 2.0 * (B_ov.T @ (sq[:, None] * X)) / np.sqrt(2.0 * om)[None, :]
    return om, A.T.copy()


def wc_aux_exact(om, A, w_complex):
    """Wc(w) [naux, naux] from the pole expansion at a complex frequency."""
    f = 2.0 * om / (w_complex ** 2 - om ** 2)
    return (A.T * f[None, :]) @ A


def sigma_exact(mo_energy, nocc, om, A, b_rows, m_idx, E, eta=0.0):
    """Exact G0W0 Sigma_c_n(E) for one target given its pair rows b_rows[j] (partner m_idx[j])."""
    e = np.asarray(mo_energy)
    ef = 0.5 * (e[nocc - 1] + e[nocc])
    ws2 = (b_rows @ A.T) ** 2                        # [np, N_ov] = |w^s_nm|^2
    sgn = np.where(e[m_idx] < ef, 1.0, -1.0)
    den = (E - e[m_idx])[:, None] + sgn[:, None] * (om[None, :] - 1j * eta)
    return complex(np.sum(ws2 / den))
