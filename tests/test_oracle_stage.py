"""CPU pins of the oracle's stage restatements that the reference exposes to Python:
compute_screening_w (quasix_core/src/dielectric/screening.rs:164-420), the static CD
self-energy compute_sigma_c_pyscf_cd (quasix_core/src/selfenergy/correlation_pyscf_cd.rs:
60-298) and solve_qp_equation (quasix_core/src/qp/solver.rs:137-221)."""
import numpy as np
import pytest
import scipy.linalg as sla

from oracle import gw_oracle as O


def _metric(n, seed, cond=1e4):
    rng = np.random.default_rng(seed)
    Q = np.linalg.qr(rng.standard_normal((n, n)))[0]
    lam = np.geomspace(1.0, 1.0 / cond, n)
    return (Q * lam) @ Q.T


def _p0(e, B, nocc, freqs):
    return np.stack([O.get_rho_response(w, e, B, nocc) for w in freqs])


def test_metric_sqrt_is_the_principal_root():
    v = _metric(30, 1)
    vs = O.metric_sqrt(v)
    assert np.allclose(vs, sla.sqrtm(v).real, atol=1e-12)
    assert np.allclose(vs, vs.T, atol=0)


def test_metric_sqrt_rejects_indefinite():
    v = _metric(10, 2)
    v -= 2e-3 * np.eye(10)          # smallest eigenvalue 1e-4 - 2e-3 < 0
    with pytest.raises(ArithmeticError):
        O.metric_sqrt(v)


def test_screening_w_dyson_identity():
    """W = v + v P0 W for the symmetric P0 of the polarizability."""
    nocc, nvir, naux = 3, 8, 20
    e, B = O.make_inputs(nocc, nvir, naux, seed=3)
    fr, _ = O.get_scaled_legendre_roots(6, 0.5)
    p0 = _p0(e, B, nocc, fr)
    v = _metric(naux, 4)
    W = O.compute_screening_w(p0, v)
    for f in range(len(fr)):
        assert np.allclose(W[f], v + v @ p0[f] @ W[f], atol=1e-11 * np.abs(W[f]).max())
    # identity metric: W = (I - P0)^-1 = screened_wc + I
    Wi = O.compute_screening_w(p0, np.eye(naux))
    for f in range(len(fr)):
        assert np.allclose(Wi[f], O.screened_wc(p0[f]) + np.eye(naux), atol=1e-13)
    # v_sqrt given: used as is
    vs = sla.sqrtm(v).real
    assert np.allclose(O.compute_screening_w(p0, None, v_sqrt=vs), W, atol=1e-11)


def test_static_cd_equals_imaginary_axis_part_at_zero():
    """With occupied energies below -0.3 and virtual ones above 0.05, no residue pole lies
    between the Fermi level and E = 0, so the static formula equals the CD self-energy at 0."""
    nocc, nvir, naux, nf = 3, 7, 18, 24
    e, B = O.make_inputs(nocc, nvir, naux, seed=5)
    miP, abP = O.ov_only_blocks(B, nocc, nvir)
    rng = np.random.default_rng(6)
    miP = miP + 0.01 * rng.standard_normal(miP.shape)       # dense blocks: oo, vo, vv partners
    abP = abP + 0.01 * rng.standard_normal(abP.shape)
    mn = O.assemble_mnP(B, miP, abP, nocc, nvir, naux)
    Bov = np.ascontiguousarray(mn[:nocc, nocc:].reshape(nocc * nvir, naux))
    lpq = np.ascontiguousarray(mn.transpose(2, 0, 1))
    fr, wt = O.get_scaled_legendre_roots(nf, 0.5)
    sig = O.sigma_c_pyscf_cd(e, 0.01, _p0(e, Bov, nocc, fr), lpq, fr, wt)
    tg = np.arange(nocc + nvir)
    pt, pm, Bp = O.pair_list(mn, tg)
    ctx = O.CDContext(e, e, nocc, Bov, tg, pt, pm, Bp, fr, wt, O.wmn_imag(e, Bov, nocc, Bp, fr),
                      0.01)
    ref = np.array([O.sigma_cd_eval(ctx, t, 0.0) for t in tg])
    assert np.max(np.abs(sig - ref)) < 1e-12
    assert np.all(np.abs(sig) > 1e-6)
    # zero polarizability: zero self-energy
    assert np.all(O.sigma_c_pyscf_cd(e, 0.01, np.zeros((nf, naux, naux)), lpq, fr, wt) == 0)


def test_solve_qp_linearized_closed_form():
    e, sx, sc, vxc = (np.array(x) for x in ([-0.5, 0.2], [-0.3, -0.1], [0.05, -0.02],
                                            [-0.25, -0.15]))
    E, Z = O.solve_qp_linearized(e, sx, sc, vxc)
    assert np.array_equal(E, e + sx + sc - vxc) and np.array_equal(Z, np.ones(2))
