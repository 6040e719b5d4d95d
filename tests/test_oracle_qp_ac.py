"""Pins the oracle's Newton solver and analytic-continuation pieces with the reference's
known-answer tests (restated from quasix_core/src/qp/newton_solver.rs:609-732,
quasix_core/tests/test_zfactor_accuracy.rs and correlation_ac.rs:755-810)."""
import os

import numpy as np

from oracle import gw_oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def test_newton_constant_sigma():
    s = O.NewtonQPSolver(-0.5, -0.3, -0.4, 0.0)
    r = s.solve(lambda om: complex(-0.1, 0.0), -0.5)
    assert r.converged and abs(r.qp_energy - (-0.5)) < 1e-7


def test_newton_linear_sigma_residual():
    f = lambda om: complex(-0.1 - 0.2 * om, 0.0)
    s = O.NewtonQPSolver(-0.5, -0.3, -0.4, 0.0)
    r = s.solve(f, -0.5)
    res = r.qp_energy + 0.5 + 0.3 - f(r.qp_energy).real - 0.4
    assert r.converged and abs(res) < 1e-7
    # exact root of E = -0.5 - 0.3 - 0.1 - 0.2E + 0.4  ->  1.2 E = -0.5
    assert abs(r.qp_energy - (-0.5 / 1.2)) < 1e-8


def test_z_factor_known_answer():
    s = O.NewtonQPSolver(-0.5, 0.0, 0.0, 0.0)
    assert abs(s._z(lambda om: complex(-0.3 * om, 0.0), -0.5) - 1.0 / 1.3) < 1e-4


def test_richardson_derivative_cubic():
    s = O.NewtonQPSolver(-0.5, 0.0, 0.0, 0.0)
    assert abs(s._deriv(lambda om: complex(om ** 3, 0.0), 0.5) - 0.75) < 1e-10


def test_bisection_fallback_converges():
    cfg = O.NewtonConfig(max_iterations=3)
    s = O.NewtonQPSolver(-0.5, -0.3, -0.4, 0.0, cfg)
    f = lambda om: complex(-0.1 / (1.0 + om * om), 0.0)
    r = s.solve(f, -0.5)
    assert r.converged
    assert abs(r.qp_energy + 0.5 + 0.3 - f(r.qp_energy).real - 0.4) < 1e-7


def test_zfactor_pole_model():
    # test_zfactor_accuracy.rs: Sigma = a/(w - b) + c w, analytic derivative
    a, b, c = 0.05, 1.2, -0.1
    f = lambda om: complex(a / (om - b) + c * om, 0.0)
    s = O.NewtonQPSolver(-0.5, 0.0, 0.0, 0.0)
    for w in (-0.4, 0.1, 0.6):
        exact = -a / (w - b) ** 2 + c
        assert abs(s._deriv(f, w) - exact) < 1e-9


def test_thiele_constant():
    zn = np.array([0.1j, 0.2j, 0.3j])
    assert abs(O.thiele(np.ones(3), zn)[0] - 1.0) < 1e-10


def test_pade_known_answer():
    zn = np.array([0.5j, 1j, 2j, 4j])
    fn = 1.0 / (1.0 + zn)
    c = O.thiele(fn, zn)
    v = O.pade_thiele(0.1, zn, c)
    assert abs(v.real - 1.0 / 1.1) < 0.01 and abs(v.imag) < 0.1
    assert abs(O.pade_thiele(0.0, zn, c).real - 1.0) < 0.01


def test_thiele_reproduces_rational_function():
    # a degree-(1,1) rational function is represented exactly by 3 continued-fraction terms
    zn = np.array([0.3j, 0.9j, 1.7j, 2.2j])
    f = lambda z: (0.3 - 0.7 * z) / (1.0 + 0.4 * z)
    c = O.thiele(f(zn), zn)
    for x in (-0.3, 0.2, 0.8):
        assert abs(O.pade_thiele(x, zn, c) - f(x)) < 1e-12


def test_ac_point_selection():
    sigma = np.zeros((1, 60), dtype=complex)
    omega = np.zeros((1, 60), dtype=complex)
    coeff, ofit = O.ac_pade_thiele_diag(sigma + 1.0, omega + np.arange(60) * 1j)
    idx = [1, 7, 13, 19, 25, 31, 37, 41, 45, 49, 53, 57]
    assert np.allclose(ofit[0], np.array(idx) * 1j)


def test_ac_sigma_freq_cache_fixture():
    g = np.load(os.path.join(GOLD, "freq_cache_ac.npz"))
    sig, om = O.get_sigma_diag(g["mo_energy"], g["lpq"], list(range(6)), 3, nw=20)
    assert np.allclose(sig, g["sigma"], rtol=1e-12, atol=1e-15)
    assert np.allclose(om, g["omega"])
    # occupied orbitals live on -i w, virtual on +i w (correlation_ac.rs:289-301)
    assert np.all(om[0, 1:].imag < 0) and np.all(om[5, 1:].imag > 0)


def test_ac_sigma_matches_casida_on_imaginary_axis():
    """Sigma(i nu) of the AC path equals the exact pole sum at E = ef + i nu."""
    from oracle import casida as Cs
    rng = np.random.default_rng(4)
    nocc, nvir, naux = 2, 4, 5
    nmo = nocc + nvir
    e = np.concatenate([np.sort(rng.uniform(-1.0, -0.4, nocc)),
                        np.sort(rng.uniform(0.1, 1.0, nvir))])
    full = rng.standard_normal((nmo, nmo, naux)) * 0.1
    full = 0.5 * (full + full.transpose(1, 0, 2))
    lpq = full.transpose(2, 0, 1).copy()
    sig, om = O.get_sigma_diag(e, lpq, list(range(nmo)), nocc, nw=200)
    B = full[:nocc, nocc:].reshape(-1, naux)
    omg, A = Cs.casida_poles(e, B, nocc)
    ef = 0.5 * (e[nocc - 1] + e[nocc])
    for n in range(nmo):
        rows = full[n]
        for k in (1, 5, 12):
            z = ef + om[n, k]
            ws2 = (rows @ A.T) ** 2
            sgn = np.where(e < ef, 1.0, -1.0)
            exact = np.sum(ws2 / ((z - e)[:, None] + sgn[:, None] * omg[None, :]))
            assert abs(sig[n, k] - exact) < 1e-7


def test_linearized_qp_known_answer():
    # qp/solver.rs:314-400 style: constant Sigma_c gives E = eps + Z (Sx + Sc - Vxc), Z = 1
    zn = np.array([0.5j, 1j, 2j, 4j])
    c = O.thiele(np.full(4, -0.05 + 0j), zn)
    E = O.linearized_qp_ac(np.array([-0.5, 0.3]), np.array([-0.2, 0.0]), np.array([-0.1, 0.0]),
                           np.stack([c, c], axis=1), np.stack([zn, zn]), [0, 1], 1)
    assert np.allclose(E, [-0.5 - 0.2 - 0.05 + 0.1, 0.3 - 0.05], atol=1e-10)
