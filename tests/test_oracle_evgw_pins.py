"""Pins of the oracle's evGW drivers (evgw_cd, evgw_ac) and of its AC QP step (qp_step_ac)
by facts the mathematics fixes, each chosen so that a plausible wiring mistake fails it:

1. Closed form.  nocc = nvir = naux = 1 with B = b: Pi(iw) = -4 b^2 D / (w^2 + D^2) with
   D = E_1 - E_0 of the energies Pi is built from, Wc_01(iw) = -4 b^4 D / (w^2 + W^2),
   W^2 = D^2 + 4 b^2 D, and the imaginary-axis integral in closed form gives
   Sigma_0(E) = -c / (W + E_1 - E),  Sigma_1(E) = c / (W + E - E_0),  c = 2 b^4 D / W.
   * evGW on the CD self-energy (eta -> 0): E_0 = e_0 + s_0 - c/(W + D),
     E_1 = e_1 + s_1 + c/(W + D), a scalar equation in D solved by bisection.
   * evGW on the AC self-energy (the reference's driver): Sigma(i nu) is sampled around the
     Fermi level of the CURRENT energies while the QP step evaluates the Pade model at
     E - ef of the MEAN-FIELD energies (quasix_core/src/qp/evgw.rs:626-633 vs :701), so
     E_0 = e_0 + s_0 - c/(W + D - d),  E_1 = e_1 + s_1 + c/(W + D + d),
     d = (E_0 + E_1 - e_0 - e_1) / 2, a 2-d root found by scipy.optimize.fsolve.
   Building Pi or the poles from the mean-field energies, anchoring the QP equation at the
   current energies, or dropping the Fermi-level shift each moves the answer by >= 1e-5 Ha.
2. Fixed point.  Converged evGW energies E* (all orbitals, tol 1e-12) satisfy the QP
   equation with Pi, the Green's function and the poles built from E* and the anchor at
   the mean-field energies, evaluated independently of the driver loop.
3. qp_step_ac on the Pade model of a known rational self-energy sampled around a shifted
   Fermi level: the returned energy solves the QP equation with the mean-field Fermi level
   and not with the shifted one (the quirk of evgw.rs:701).
4. The stop rule of evgw.rs:929-1027: the frontier window of evgw.rs:1121-1160 against
   hand-counted windows, and a case where the frontier and all-orbital criteria stop at
   different iterations.
"""
import math

import numpy as np
import pytest
from scipy.optimize import brentq, fsolve

from oracle import gw_oracle as O

TIGHT = O.NewtonConfig(energy_tolerance=1e-13)


# --------------------------------------------------------------------------- closed form

E0 = np.array([-0.6, 0.4])
BB = 0.3
STAT = np.array([-0.05, 0.01])


def _omega(D):
    return math.sqrt(D * D + 4.0 * BB * BB * D)


def _cd_closed_form():
    def eq(D):
        W = _omega(D)
        return D - (E0[1] - E0[0] + STAT[1] - STAT[0]) - 4.0 * BB ** 4 * D / (W * (W + D))
    D = brentq(eq, 1e-3, 10.0, xtol=1e-15)
    W = _omega(D)
    c = 2.0 * BB ** 4 * D / W
    return np.array([E0[0] + STAT[0] - c / (W + D), E0[1] + STAT[1] + c / (W + D)])


def _ac_closed_form(start):
    def eq(x):
        a, b = x
        D = b - a
        W = _omega(D)
        d = 0.5 * (a + b - E0[0] - E0[1])
        c = 2.0 * BB ** 4 * D / W
        return [a - E0[0] - STAT[0] + c / (W + D - d), b - E0[1] - STAT[1] - c / (W + D + d)]
    x = fsolve(eq, start, xtol=1e-15)
    assert np.max(np.abs(eq(x))) < 1e-15
    return x


@pytest.mark.parametrize("nf", [32, 64])
def test_evgw_cd_closed_form(nf):
    ref = _cd_closed_form()
    fr, wt = O.get_scaled_legendre_roots(nf, 0.5)
    r = O.evgw_cd(E0, np.array([[BB]]), 1, np.array([0, 1]), fr, wt, n_iter=200, tol=1e-13,
                  eta=1e-7, static=STAT, frontier_convergence=False, newton=TIGHT)
    assert r["converged"]
    err = np.max(np.abs(r["E"] - ref))
    print(f"evgw_cd closed form nf={nf}: {r['n_cycles']} cycles, err {err:.2e}")
    assert err < 1e-11


@pytest.mark.parametrize("nw", [60, 100])
def test_evgw_ac_closed_form_with_mean_field_fermi_level(nw):
    cd = _cd_closed_form()
    ref = _ac_closed_form(cd)
    # the Fermi-level shift matters at this size: the two closed forms differ
    assert np.max(np.abs(ref - cd)) > 1e-5
    lpq = np.zeros((1, 2, 2))
    lpq[0, 0, 1] = lpq[0, 1, 0] = BB
    r = O.evgw_ac(E0, lpq, 1, nw=nw, n_iter=200, tol=1e-13, solver="newton", sigma_x=STAT,
                  frontier_convergence=False, newton=TIGHT)
    assert r["converged"]
    err = np.max(np.abs(r["E"] - ref))
    print(f"evgw_ac closed form nw={nw}: {r['n_cycles']} cycles, err {err:.2e}")
    assert err < 1e-11


# --------------------------------------------------------------------------- fixed point

def _model(seed=3, nocc=3, nvir=9, naux=20):
    e, B = O.make_inputs(nocc, nvir, naux, seed=seed)
    return e, B, nocc, nvir, naux


@pytest.mark.parametrize("solver", ["newton", "linearized"])
def test_evgw_cd_fixed_point(solver):
    e, B, nocc, nvir, naux = _model()
    nmo = nocc + nvir
    tg = np.arange(nmo)
    stat = np.linspace(-0.05, 0.03, nmo)
    fr, wt = O.get_scaled_legendre_roots(24, 0.5)
    r = O.evgw_cd(e, B, nocc, tg, fr, wt, n_iter=300, tol=1e-12, static=stat, diis=True,
                  frontier_convergence=False, newton=TIGHT, solver=solver)
    assert r["converged"]
    Es = r["E"]
    # Sigma built from E* on its own (pairs, Wc_nm, CD context), anchor and eta signs from e
    miP, abP = O.ov_only_blocks(B, nocc, nvir)
    pt, pm, Bp = O.pair_list(O.assemble_mnP(B, miP, abP, nocc, nvir, naux), tg)
    ctx = O.CDContext(Es, e, nocc, B, tg, pt, pm, Bp, fr, wt,
                      O.wmn_imag(Es, B, nocc, Bp, fr), 0.01)
    def resid(c, t, n):
        s0 = O.sigma_cd_eval(c, t, Es[n]).real
        if solver == "newton":
            return Es[n] - e[n] - stat[n] - s0
        s1 = O.sigma_cd_eval(c, t, Es[n] + 1e-6).real
        z = min(max(1.0 / (1.0 - (s1 - s0) / 1e-6), 0.1), 0.999)
        return Es[n] - e[n] - z * (stat[n] + s0)

    res = np.array([resid(ctx, t, n) for t, n in enumerate(tg)])
    print(f"evgw_cd {solver} fixed point: {r['n_cycles']} cycles, max residual "
          f"{np.max(np.abs(res)):.2e}")
    assert np.max(np.abs(res)) < 1e-10
    # the same residual with Pi from the mean-field energies is far from zero
    ctx0 = O.CDContext(Es, e, nocc, B, tg, pt, pm, Bp, fr, wt,
                       O.wmn_imag(e, B, nocc, Bp, fr), 0.01)
    res0 = np.array([resid(ctx0, t, n) for t, n in enumerate(tg)])
    assert np.max(np.abs(res0)) > 1e-6


@pytest.mark.parametrize("solver", ["newton", "linearized"])
def test_evgw_ac_fixed_point(solver):
    e, B, nocc, nvir, naux = _model(seed=4)
    nmo = nocc + nvir
    miP, abP = O.ov_only_blocks(B, nocc, nvir)
    lpq = np.ascontiguousarray(O.assemble_mnP(B, miP, abP, nocc, nvir, naux).transpose(2, 0, 1))
    sx = O.exchange_diag(miP, nmo, nocc)
    vxc = 0.9 * sx
    r = O.evgw_ac(e, lpq, nocc, nw=60, n_iter=300, tol=1e-12, diis=True, solver=solver,
                  sigma_x=sx, vxc=vxc, frontier_convergence=False, newton=TIGHT)
    assert r["converged"]
    Es = r["E"]
    sig, om = O.get_sigma_diag(Es, lpq, np.arange(nmo), nocc, 60)
    coeff, ofit = O.ac_pade_thiele_diag(sig, om)
    ef = 0.5 * (e[nocc - 1] + e[nocc])
    res = []
    for p in range(nmo):
        s0 = O.pade_thiele(Es[p] - ef, ofit[p], coeff[:, p]).real
        if solver == "newton":
            res.append(Es[p] - e[p] - sx[p] - s0 + vxc[p])
        else:
            s1 = O.pade_thiele(Es[p] - ef + 1e-6, ofit[p], coeff[:, p]).real
            z = min(max(1.0 / (1.0 - (s1 - s0) / 1e-6), 0.1), 0.999)
            res.append(Es[p] - e[p] - z * (sx[p] + s0 - vxc[p]))
    err = float(np.max(np.abs(res)))
    print(f"evgw_ac {solver} fixed point: {r['n_cycles']} cycles, max residual {err:.2e}")
    assert err < 1e-10


# --------------------------------------------------------------------------- qp_step_ac

def test_qp_step_ac_uses_mean_field_fermi_level():
    """Sigma(w) = sum_k a_k / (w + ef_cur - p_k) sampled on the AC grid as get_sigma_diag
    lays it out around ef_cur (occupied: -i nu, virtual: +i nu)."""
    e = np.array([-0.9, -0.5, 0.3, 0.8])
    nocc = 2
    ef_mf = 0.5 * (e[1] + e[2])
    ef_cur = ef_mf + 0.07
    poles = [(-1.6, 0.04), (1.5, 0.05)]

    def sigma(z):
        return sum(a / (z - p) for p, a in poles)

    fr, _ = O.get_scaled_legendre_roots(100, 0.5)
    nu = fr[fr < 5.0]
    om = np.zeros((4, len(nu) + 1), dtype=np.complex128)
    for p in range(4):
        om[p, 1:] = (-1j if p < nocc else 1j) * nu
    sig = np.array([[sigma(w + ef_cur) for w in om[p]] for p in range(4)])
    coeff, ofit = O.ac_pade_thiele_diag(sig, om)
    sx = np.array([-0.4, -0.3, -0.1, -0.05])
    vxc = np.array([-0.35, -0.3, -0.2, -0.1])
    E, Z = O.qp_step_ac(coeff, ofit, np.arange(4), e, nocc, sx, vxc, "newton", None,
                        newton=TIGHT)
    good = [E[p] - e[p] - sx[p] + vxc[p] - sigma(E[p] - ef_mf + ef_cur).real for p in range(4)]
    bad = [E[p] - e[p] - sx[p] + vxc[p] - sigma(E[p]).real for p in range(4)]
    print(f"qp_step_ac residual with the mean-field ef {np.max(np.abs(good)):.2e}, "
          f"with the current ef {np.max(np.abs(bad)):.2e}")
    assert np.max(np.abs(good)) < 1e-10
    assert np.max(np.abs(bad)) > 1e-4
    # linearised: at prev_qp, forward-difference Z clamped to [0.1, 0.999]
    prev = e + 0.01
    El, Zl = O.qp_step_ac(coeff, ofit, np.arange(4), e, nocc, sx, vxc, "linearized", prev)
    for p in range(4):
        x = prev[p] - ef_mf + ef_cur
        s0, s1 = sigma(x).real, sigma(x + 1e-6).real
        z = min(max(1.0 / (1.0 - (s1 - s0) / 1e-6), 0.1), 0.999)
        assert Zl[p] == pytest.approx(z, abs=1e-8)
        assert El[p] == pytest.approx(e[p] + z * (sx[p] + s0 - vxc[p]), abs=1e-10)


# --------------------------------------------------------------------------- stop rule

def test_convergence_window_hand_counted():
    e = np.array([-3.0, -1.2, -0.9, -0.5, 0.1, 0.6, 1.05, 1.2])
    # HOMO -0.5: -0.9 and -1.2 within 1.0 (>= -1.5), -3.0 not; LUMO 0.1: 0.6 and 1.05
    # within 1.0 (<= 1.1), 1.2 not
    assert O.convergence_window(e, 4) == (1, 7)
    assert O.convergence_window(e, 4, frontier_window=0.1) == (3, 5)
    assert O.convergence_window(e, 4, frontier_convergence=False) == (0, 8)
    # the scan stops at the first orbital outside the window (sorted energies)
    assert O.convergence_window(e, 4, frontier_window=2.6) == (0, 8)


def test_frontier_and_all_orbital_stop_rules_differ():
    """V_xc of the frontier orbitals is set to Sigma_x + Re Sigma_c(eps) of iteration 0, so
    their first (linearised) update is exactly zero: the frontier rule stops after the first
    iteration (the reference checks convergence after every iteration, the first included),
    the all-orbital rule keeps going.  Both runs share their first iterate."""
    e, B, nocc, nvir, naux = _model(seed=6, nocc=4, nvir=12, naux=30)
    nmo = nocc + nvir
    miP, abP = O.ov_only_blocks(B, nocc, nvir)
    lpq = np.ascontiguousarray(O.assemble_mnP(B, miP, abP, nocc, nvir, naux).transpose(2, 0, 1))
    sx = O.exchange_diag(miP, nmo, nocc)
    lo, hi = O.convergence_window(e, nocc, frontier_window=0.1)
    assert (lo, hi) == (3, 5)
    sig, om = O.get_sigma_diag(e, lpq, np.arange(nmo), nocc, 40)
    coeff, ofit = O.ac_pade_thiele_diag(sig, om)
    ef = 0.5 * (e[nocc - 1] + e[nocc])
    vxc = 0.9 * sx
    for p in range(lo, hi):
        vxc[p] = sx[p] + O.pade_thiele(e[p] - ef, ofit[p], coeff[:, p]).real
    kw = dict(nw=40, n_iter=40, tol=1e-6, solver="linearized", sigma_x=sx, vxc=vxc)
    fr = O.evgw_ac(e, lpq, nocc, frontier_window=0.1, **kw)
    al = O.evgw_ac(e, lpq, nocc, frontier_convergence=False, **kw)
    print(f"window {lo}..{hi - 1}: frontier stops after {fr['n_cycles']}, all after "
          f"{al['n_cycles']}")
    assert fr["converged"] and fr["n_cycles"] == 1
    assert al["converged"] and al["n_cycles"] > 3
    h0, a0 = fr["history"][0], al["history"][0]
    assert np.array_equal(h0["qp_energies"], a0["qp_energies"])
    d = np.abs(h0["qp_energies"] - e)
    assert h0["energy_change"] == np.max(d[lo:hi]) < 1e-12
    assert h0["rms_change"] == pytest.approx(np.sqrt(np.mean(d[lo:hi] ** 2)), abs=1e-15)
    assert a0["energy_change"] == np.max(d) > 1e-3
    assert a0["rms_change"] == pytest.approx(np.sqrt(np.mean(d ** 2)), rel=1e-12)
