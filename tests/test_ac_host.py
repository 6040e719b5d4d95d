"""CPU: the host-side analytic continuation in libqxgw (Thiele fit, continued fraction,
per-orbital point selection, QP step on the Pade model) against the oracle's restatement of
correlation_ac.rs / newton_solver.rs, and the exactness property that pins both: a Thiele
fit through 2m points of a rational function of matching order reproduces it everywhere."""
import numpy as np
import pytest

from oracle import gw_oracle as O


@pytest.fixture(scope="module")
def core():
    from quasix_amd import _lib, core
    try:
        _lib.load()
    except ImportError:
        pytest.skip("libqxgw.so not built")
    return core


def _rational(z):
    return (0.3 - 0.2 * z) / ((z - 1.5) * (z + 0.7 + 0.4j)) + 0.1


def test_thiele_reproduces_rational_function(core):
    # a [2/2] rational function is fixed by 5 points: the fit must reproduce it off the axis
    zn = 1j * np.linspace(0.05, 4.0, 5)
    c = core.thiele(_rational(zn), zn)
    assert np.allclose(c, O.thiele(_rational(zn), zn), rtol=1e-12, atol=1e-14)
    for w in (-1.2, -0.3, 0.0, 0.4, 2.7):
        got = core.pade_thiele(w, zn, c)
        assert abs(got - _rational(w)) < 1e-10
        assert abs(got - O.pade_thiele(w, zn, c)) < 1e-13


def test_thiele_continuation_matches_oracle(core):
    """A non-rational function: reciprocal differences lose ~30x precision per level, so the
    late coefficients of two correct implementations differ in rounding (5e-3 relative at
    the 12th here); the leading ones and the continued fraction on the real axis agree."""
    zn = 1j * np.linspace(0.05, 4.0, 12)
    f = lambda z: np.exp(-0.3 * z) / (1.0 + z * z + 0.5 * z)
    c, c2 = core.thiele(f(zn), zn), O.thiele(f(zn), zn)
    assert np.allclose(c[:5], c2[:5], rtol=1e-11)
    for w in (-0.8, 0.2, 1.1):
        got = core.pade_thiele(w, zn, c)
        assert abs(got - O.pade_thiele(w, zn, c2)) < 1e-10
        assert abs(got - f(w)) < 1e-9


def test_ac_pade_thiele_diag_matches_oracle(core):
    rng = np.random.default_rng(4)
    norbs, nw = 5, 41
    nu = np.r_[0.0, np.sort(rng.uniform(0.01, 5.0, nw - 1))]
    omega = np.array([(-1j if p < 2 else 1j) * nu for p in range(norbs)])
    sigma = np.array([_rational(om * (1 + 0.1 * p)) for p, om in enumerate(omega)])
    c1, o1 = core.ac_pade_thiele_diag(sigma, omega)
    c2, o2 = O.ac_pade_thiele_diag(sigma, omega)
    assert c1.shape == c2.shape and np.array_equal(o1, o2)
    assert np.allclose(c1, c2, rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("solver", ["linearized", "newton"])
def test_qp_solve_ac_matches_oracle(core, solver):
    rng = np.random.default_rng(5)
    nocc, nmo = 3, 7
    e = np.sort(np.r_[rng.uniform(-1.2, -0.4, nocc), rng.uniform(0.1, 1.5, nmo - nocc)])
    ef = 0.5 * (e[nocc - 1] + e[nocc])
    nu = np.linspace(0.0, 5.0, 41)
    orbs = np.arange(nmo)
    # a smooth model self-energy with one pole on each side of the Fermi level
    omega = np.array([(-1j if p < nocc else 1j) * nu for p in orbs])
    sig = np.array([0.02 / (om - 0.9 + 0.05j) + 0.03 / (om + 1.1 - 0.05j) for om in omega])
    c, ofit = core.ac_pade_thiele_diag(sig, omega)
    sx = rng.uniform(-0.05, 0.05, nmo)
    vx = np.zeros(nmo)
    got = core.qp_solve_ac(c, ofit, orbs, e, nocc, sx, vx, solver=solver)
    for i, p in enumerate(orbs):
        f = lambda om, i=i: O.pade_thiele(om, ofit[i], c[:, i])
        s0, s1 = f(e[p] - ef).real, f(e[p] - ef + 1e-6).real
        z = min(max(1.0 / (1.0 - (s1 - s0) / 1e-6), 0.1), 0.999)
        lin = e[p] + z * (sx[p] + s0)
        if solver == "linearized":
            # the 1e-6 forward difference amplifies 1e-16 evaluation differences ~1e-10 x
            assert abs(got["qp_energies"][i] - lin) < 1e-9
        else:
            r = O.NewtonQPSolver(e[p], sx[p], 0.0, ef).solve(f, lin)
            assert abs(got["qp_energies"][i] - r.qp_energy) < 1e-9
            assert abs(got["z_factors"][i] - r.z_factor) < 1e-7
