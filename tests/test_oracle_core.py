"""Pins the FP64 oracle (CPU): grids, Pi, Wc, CD self-energy, against closed forms,
invariants, brute force and the reference's own known-answer tests.

Each check is chosen so that a plausible slip (dropped factor, wrong sign, transposed
operand, wrong eta placement) fails it.
"""
import os

import numpy as np
import pytest

from oracle import casida as Cs
from oracle import gw_oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden")


# ---- A1 grid --------------------------------------------------------------------------
def test_gl_two_point_known_answer():
    # correlation_ac.rs:742-752: nodes +-1/sqrt(3), weights 1
    x, w = O.gauss_legendre_nodes_weights(2)
    assert np.allclose(x, [-1 / np.sqrt(3), 1 / np.sqrt(3)], atol=1e-14)
    assert np.allclose(w, [1.0, 1.0], atol=1e-14)


@pytest.mark.parametrize("n", [1, 3, 20, 32, 48, 64, 100])
def test_gl_matches_leggauss(n):
    x, w = O.gauss_legendre_nodes_weights(n)
    xr, wr = np.polynomial.legendre.leggauss(n)
    assert np.max(np.abs(x - xr)) < 1e-14
    assert np.max(np.abs(w - wr)) < 1e-14
    assert abs(w.sum() - 2.0) < 1e-13                 # imaginary_grid.rs:276+ (sum w = 2)
    assert np.max(np.abs(x + x[::-1])) < 1e-14        # symmetric nodes


def test_scaled_roots_integrate_rational():
    # int_0^inf dw / (w^2 + a^2) = pi / (2a): the quadrature the CD integral relies on
    f, w = O.get_scaled_legendre_roots(64, 0.5)
    assert np.all(f > 0) and np.all(np.diff(f) > 0) and np.all(w > 0)   # correlation_ac.rs:722
    for a in (0.3, 1.0, 2.5):
        assert abs(np.sum(w / (f * f + a * a)) - np.pi / (2 * a)) < 1e-8


def test_pyscf_transforms_match_scaled_roots():
    x, w = O.gauss_legendre_nodes_weights(32)
    f = O.transform_to_imaginary_axis_pyscf(x, 0.5)
    ww = O.transform_weights_pyscf(w, x, 0.5)
    f2, w2 = O.get_scaled_legendre_roots(32, 0.5)
    assert np.allclose(f, f2, rtol=1e-15) and np.allclose(ww, w2, rtol=1e-15)
    assert O.transform_to_imaginary_axis_pyscf([1.0])[0] == 1e12
    assert O.transform_weights_pyscf([0.3], [1.0])[0] == 0.0


# ---- A2 polarizability ------------------------------------------------------------------
def _small(seed=0, nocc=3, nvir=5, naux=7):
    rng = np.random.default_rng(seed)
    e = np.concatenate([np.sort(rng.uniform(-1.2, -0.3, nocc)),
                        np.sort(rng.uniform(0.1, 2.0, nvir))])
    B = rng.standard_normal((nocc * nvir, naux)) * 0.1
    return e, B


def test_pi_brute_force_outer_products():
    e, B = _small()
    nocc, nvir = 3, 5
    w = 0.7
    ref = np.zeros((B.shape[1],) * 2)
    for i in range(nocc):
        for a in range(nvir):
            eia = e[i] - e[nocc + a]
            b = B[i * nvir + a]
            ref += 4.0 * eia / (w * w + eia * eia) * np.outer(b, b)
    assert np.allclose(O.get_rho_response(w, e, B, nocc), ref, rtol=1e-14, atol=1e-16)


def test_pi_reference_known_answer():
    # polarizability.rs:432+: occ 0.5, virt -0.5 (e_ia = +1), w = 1, B = 0.1 -> every element 0.02
    ia = np.full((1, 1, 2), 0.1)
    p0 = O.compute_polarizability_p0(ia, [0.5, -0.5], [2.0, 0.0], [1.0])
    assert np.allclose(p0, 0.02, atol=1e-15)


def test_pi_symmetric_nsd_and_asymptote():
    e, B = O.make_inputs(4, 9, 11, seed=2)
    for w in (0.0, 0.5, 3.0):
        p = O.get_rho_response(w, e, B, 4)
        assert np.max(np.abs(p - p.T)) < 1e-15                      # correlation_ac.rs:813
        assert np.max(np.linalg.eigvalsh(p)) < 1e-14                 # negative semidefinite
    # w -> inf: Pi ~ 4 B^T diag(e) B / w^2
    w = 1e4
    eia = (e[:4, None] - e[None, 4:]).reshape(-1)
    asym = 4.0 * (B * eia[:, None]).T @ B / w ** 2
    assert np.allclose(O.get_rho_response(w, e, B, 4), asym, rtol=1e-6)


def test_pi_complex_reduces_on_imaginary_axis():
    e, B = _small(1)
    for w in (0.1, 1.3):
        assert np.allclose(O.rho_response_complex(1j * w, e, B, 3),
                           O.get_rho_response(w, e, B, 3), rtol=1e-13, atol=1e-16)


def test_stage_p0_matches_live_path():
    e, B = _small(4)
    fr, _ = O.get_scaled_legendre_roots(5)
    p0 = O.compute_polarizability_p0(B.reshape(3, 5, 7), e, [2, 2, 2, 0, 0, 0, 0, 0], fr)
    for k, w in enumerate(fr):
        assert np.allclose(p0[k], O.get_rho_response(w, e, B, 3), rtol=1e-13, atol=1e-17)


# ---- A3 screened interaction --------------------------------------------------------------
def test_wc_matches_casida_pole_expansion():
    e, B = _small(5)
    om, A = Cs.casida_poles(e, B, 3)
    for w in (0.0, 0.4, 2.0):
        wc = O.screened_wc(O.get_rho_response(w, e, B, 3))
        assert np.allclose(wc, Cs.wc_aux_exact(om, A, 1j * w).real, atol=1e-14)


def test_wc_woodbury_and_properties():
    e, B = O.make_inputs(3, 6, 12, seed=7)            # rank 12 <= N_ov = 18
    w = 0.8
    eia = (e[:3, None] - e[None, 3:]).reshape(-1)
    D = 4.0 * (-eia) / (w * w + eia * eia)             # Pi = -B^T D B
    wc = O.screened_wc(O.get_rho_response(w, e, B, 3))
    sq = np.sqrt(D)
    A = sq[:, None] * B
    wood = -A.T @ np.linalg.solve(np.eye(A.shape[0]) + A @ A.T, A)
    assert np.allclose(wc, wood, atol=1e-14)
    assert np.max(np.abs(wc - wc.T)) < 1e-14
    assert np.max(np.linalg.eigvalsh(wc)) < 1e-14     # negative semidefinite
    assert np.max(np.abs(O.screened_wc(np.zeros((5, 5))))) == 0.0   # Pi = 0 -> Wc = 0


def test_wc_fallback_chain():
    rng = np.random.default_rng(3)
    Q = np.linalg.qr(rng.standard_normal((6, 6)))[0]
    pi = (Q * np.array([1.7, -0.5, -1, -2, 0.3, -0.1])) @ Q.T   # eps indefinite: LU path
    assert np.allclose(O.screened_wc(pi), np.linalg.inv(np.eye(6) - pi) - np.eye(6))
    # and the LU result really differs from a (failed) Cholesky attempt's domain: eps is
    # indefinite, so a Cholesky-only implementation would raise here
    with pytest.raises(np.linalg.LinAlgError):
        np.linalg.cholesky(np.eye(6) - pi)


def test_wc_identity_fallback_on_exactly_singular():
    pi = np.eye(4)                                     # eps = 0: both factorisations fail
    assert np.max(np.abs(O.screened_wc(pi))) == 0.0    # correlation_ac.rs:396: eps^-1 := I


# ---- A4-A6 CD self-energy --------------------------------------------------------------------
def test_cd_sigma_matches_casida_golden():
    g = np.load(os.path.join(GOLD, "casida_cd.npz"))
    assert np.max(np.abs(g["cd"].real - g["exact"].real)) < 1e-12


def test_cd_sigma_recomputes_golden():
    g = np.load(os.path.join(GOLD, "casida_cd.npz"))
    e, full = g["mo_energy"], g["full"]
    nocc, nmo = 3, 9
    B = full[:nocc, nocc:].reshape(-1, full.shape[2]).copy()
    targets = np.arange(nmo)
    pt, pm, Bp = O.pair_list(full, targets)
    fr, wt = O.get_scaled_legendre_roots(400, 0.5)
    wnm = O.wmn_imag(e, B, nocc, Bp, fr)
    ctx = O.CDContext(e, e, nocc, B, targets, pt, pm, Bp, fr, wt, wnm, eta=1e-7)
    for t in (0, 5, 8):
        for q in range(2):
            v = O.sigma_cd_eval(ctx, t, g["energies"][t, q])
            assert abs(v - g["cd"][t, q]) < 1e-13


def test_residues_are_needed():
    """Dropping the residue term breaks agreement with Casida where poles lie between
    the Fermi level and E (guards against a silently-zero R)."""
    g = np.load(os.path.join(GOLD, "casida_cd.npz"))
    e, full = g["mo_energy"], g["full"]
    B = full[:3, 3:].reshape(-1, full.shape[2]).copy()
    targets = np.arange(9)
    pt, pm, Bp = O.pair_list(full, targets)
    fr, wt = O.get_scaled_legendre_roots(400, 0.5)
    wnm = O.wmn_imag(e, B, 3, Bp, fr)
    ctx = O.CDContext(e, e, 3, B, targets, pt, pm, Bp, fr, wt, wnm, eta=1e-7)
    t, E = 8, g["energies"][8, 1]                     # above several virtual levels
    ef = ctx.ef
    sel = pt == t
    n_poles = np.count_nonzero((e[pm[sel]] > ef) & (e[pm[sel]] < E))
    assert n_poles > 0
    ctx_no_res = O.CDContext(e, e, 3, B, targets, pt, pm, Bp, fr, wt, wnm, eta=1e-7)
    ctx_no_res.mo_energy = e.copy()
    imag_only = O.sigma_cd_eval(ctx_no_res, t, E) - sum(
        O.wc_real_axis(ctx_no_res, complex(E - e[m], 1e-7), Bp[j])
        for j, m in zip(np.nonzero(sel)[0], pm[sel]) if ef < e[m] < E)
    assert abs(imag_only.real - g["exact"][t, 1].real) > 1e-4
    assert abs(O.sigma_cd_eval(ctx, t, E).real - g["exact"][t, 1].real) < 1e-12


def test_one_pole_closed_form():
    """Single excitation: Sigma_c = |w|^2/(E - eps_m - Omega) for a virtual partner."""
    e = np.array([-0.6, 0.4])
    b = np.array([[0.3]])                               # nocc 1, nvir 1, naux 1
    om, A = Cs.casida_poles(e, b, 1)
    Om = np.sqrt(1.0 ** 2 + 4 * 1.0 * 0.09)             # Delta = 1, sqrt(D^2 + 4 D |b|^2)
    assert abs(om[0] - Om) < 1e-14
    full = np.zeros((2, 2, 1))
    full[0, 1, 0] = full[1, 0, 0] = 0.3
    targets = np.array([0])
    pt, pm, Bp = O.pair_list(full, targets)
    fr, wt = O.get_scaled_legendre_roots(300, 0.5)
    wnm = O.wmn_imag(e, b, 1, Bp, fr)
    ctx = O.CDContext(e, e, 1, b, targets, pt, pm, Bp, fr, wt, wnm, eta=1e-8)
    w2 = 4 * (0.3 * 0.3 * 1.0) ** 2 / (2 * Om)           # |w|^2 = 4 (b . B^T D^1/2 x)^2 / 2 Om
    for E in (-0.7, -0.5):
        assert abs(O.sigma_cd_eval(ctx, 0, E).real - w2 / (E - 0.4 - Om)) < 1e-10


def test_sigma_zero_for_zero_tensor():
    # correlation_fixed.rs test_sigma_c_fixed_zero_p0: Pi = 0 -> Sigma = 0
    e = np.array([-0.5, -0.3, 0.2, 0.6])
    B = np.zeros((4, 3))
    fr, wt = O.get_scaled_legendre_roots(16)
    r = O.g0w0_cd(e, B, 2, np.arange(4), fr, wt, pairs=(np.array([0], np.int32),
                  np.array([2], np.int32), np.zeros((1, 3))))
    assert np.allclose(r["E"], e) and np.allclose(r["sigma"], 0.0)


def test_eta_sign_convention():
    """emo = E - eps_m - i eta sign(eps_n - eps_m): Im Sigma of an occupied target from
    virtual partners is negative in the small-eta regime far from poles."""
    e, B = O.make_inputs(2, 4, 6, seed=9)
    targets = np.array([1])
    fr, wt = O.get_scaled_legendre_roots(32)
    r = O.g0w0_cd(e, B, 2, targets, fr, wt, eta=0.05)
    assert r["sigma"][0].imag != 0.0
    r2 = O.g0w0_cd(e, B, 2, targets, fr, wt, eta=0.0)
    assert r2["sigma"][0].imag == 0.0
    assert np.sign(r["sigma"][0].imag) == -1.0


def test_exchange_diag_rank_one_closed_form():
    """mi[p, i, R] = a_p b_i c_R gives Sigma_x[p] = -a_p^2 |b|^2 |c|^2 (only the nocc
    occupied columns enter, each auxiliary index once)."""
    rng = np.random.default_rng(12)
    nmo, nocc, naux = 9, 3, 7
    a, b, c = rng.standard_normal(nmo), rng.standard_normal(nocc), rng.standard_normal(naux)
    mi = a[:, None, None] * b[None, :, None] * c[None, None, :]
    got = O.exchange_diag(mi.reshape(nmo * nocc, naux), nmo, nocc)
    assert np.allclose(got, -a ** 2 * (b @ b) * (c @ c), rtol=1e-13)
    assert np.all(got <= 0.0)


def test_mo_transform_identity_and_norm():
    """Identity coefficients pick AO pairs (i, nocc + a); with an orthogonal C = [C_occ C_vir]
    of the full space and a J symmetric in its AO pair, the sum over all (i, a), (i, j), (a, b)
    blocks keeps |J|_F -- checked here through the ov block plus its complement."""
    rng = np.random.default_rng(14)
    nao, nocc, naux = 9, 3, 5
    nvir = nao - nocc
    J = rng.standard_normal((nao, nao, naux))
    J = 0.5 * (J + J.transpose(1, 0, 2))
    eye = np.eye(nao)
    got = O.transform_mo_3center(J, eye[:, :nocc], eye[:, nocc:])
    assert np.array_equal(got.reshape(nocc, nvir, naux), J[:nocc, nocc:, :])
    Q = np.linalg.qr(rng.standard_normal((nao, nao)))[0]
    full = np.einsum("mp,mnP,nq->pqP", Q, J, Q)
    ov = O.transform_mo_3center(J, Q[:, :nocc], Q[:, nocc:]).reshape(nocc, nvir, naux)
    assert np.allclose(ov, full[:nocc, nocc:])
    assert np.isclose(np.sum(full ** 2), np.sum(J ** 2))
