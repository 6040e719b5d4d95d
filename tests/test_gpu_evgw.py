"""GPU: evGW drivers (quasix_amd.gw.evgw / run_evgw / evgw_core / evgw_ac_core) against the
oracle's evGW on the same seeded inputs.  FP64: every updated QP energy within 1e-6 Ha of
the oracle, the same cycle count and the same window max / RMS changes; the mixed path
(BF16x3 Pi) within 1 meV of FP64.  Covers damping only and Pulay mixing (whose driver call
order falls back to damping from the fourth iteration on), both QP solvers, the frontier
and all-orbital stop rules, and dense blocks whose symmetrisation matters."""
import numpy as np
import pytest

from oracle import gw_oracle as O

pytestmark = pytest.mark.gpu
MEV = 1.0 / 27211.386245988
NF = 16


@pytest.mark.parametrize("diis,solver", [(False, "newton"), (True, "newton"),
                                         (False, "linearized")])
def test_evgw_core_matches_oracle(gpu, diis, solver):
    """CD driver: same iterates, cycle count, window max/RMS changes and Z as oracle.evgw_cd
    (frontier window 1 Ha, stop check after every iteration)."""
    from quasix_amd import gw
    nocc, nvir, naux = 5, 19, 84
    e, B = O.make_inputs(nocc, nvir, naux, seed=7)
    tg = gw.frontier_targets(nocc, nocc + nvir, 2, 2)
    fr, wt = O.get_scaled_legendre_roots(NF, 0.5)
    ref = O.evgw_cd(e, B, nocc, tg, fr, wt, n_iter=5, damping=0.5, tol=1e-7, diis=diis,
                    solver=solver)
    got = gw.evgw_core(e, B, nocc, tg, n_iter=5, damping=0.5, conv_tol=1e-7, diis=diis,
                       nfreq=NF, solver=solver)
    err = float(np.max(np.abs(got["qp_energies"] - ref["E"])))
    print(f"evgw cd diis={diis} {solver}: {got['n_cycles']} cycles, max |dE| vs oracle "
          f"{err:.3e}")
    assert got["n_cycles"] == ref["n_cycles"] and got["converged"] == ref["converged"]
    assert got["window"] == ref["window"]
    assert err < 1e-6
    for h, hr in zip(got["history"], ref["history"]):
        assert h["energy_change"] == pytest.approx(hr["energy_change"], abs=1e-8)
        assert h["rms_change"] == pytest.approx(hr["rms_change"], abs=1e-8)
        assert np.max(np.abs(h["z_factors"] - hr["z_factors"])) < 1e-6
    mixed = gw.evgw_core(e, B, nocc, tg, n_iter=5, damping=0.5, conv_tol=1e-7, diis=diis,
                         nfreq=NF, precision="bf16x3", solver=solver)
    dm = float(np.max(np.abs(mixed["qp_energies"] - got["qp_energies"])))
    print(f"mixed vs fp64: {dm / MEV:.3e} meV")
    assert dm < MEV


def _dense_blocks(nocc, nvir, naux, seed):
    """iaP, a miP whose vo block is not iaP transposed and whose oo block is not symmetric,
    and a non-symmetric abP: the symmetrisation of the reference's assembly matters for the
    polarizability, the pair rows and the exchange term."""
    e, B = O.make_inputs(nocc, nvir, naux, seed=seed)
    rng = np.random.default_rng(seed + 100)
    miP, abP = O.ov_only_blocks(B, nocc, nvir)
    miP = miP + 0.02 * rng.standard_normal(miP.shape)
    abP = 0.02 * rng.standard_normal(abP.shape)
    return e, B, miP, abP


def _ref_evgw_ac(e, ia, mi, ab, nocc, naux, vxc, **kw):
    nvir = e.size - nocc
    mn = O.assemble_mnP(ia, mi, ab, nocc, nvir, naux)               # [nmo, nmo, naux]
    lpq = np.ascontiguousarray(mn.transpose(2, 0, 1))
    sx = O.exchange_diag(np.ascontiguousarray(mn[:, :nocc, :]), e.size, nocc)
    return O.evgw_ac(e, lpq, nocc, sigma_x=sx, vxc=vxc, **kw), sx


def test_evgw_reference_interface(gpu):
    """quasix.quasix.evgw with the reference defaults (AC self-energy, linearised QP step,
    every orbital updated, frontier window 1 Ha) on dense, inconsistent blocks: result keys,
    Sigma_x from the symmetrised [nmo, nocc] block, iterates, n_cycles and the history
    against oracle.evgw_ac; ip / ea / gap in Hartree."""
    from quasix_amd import gw
    nocc, nvir, naux = 5, 19, 84
    e, B, miP, abP = _dense_blocks(nocc, nvir, naux, seed=8)
    occ = np.r_[np.full(nocc, 2.0), np.zeros(nvir)]
    vxc = np.linspace(-0.6, -0.2, nocc + nvir)
    r = gw.evgw(e, occ, B, miP, abP, np.eye(naux), vxc, max_cycle=6, conv_tol=1e-6)
    for k in ("qp_energies", "z_factors", "converged", "n_cycles", "final_error", "sigma_x",
              "sigma_c_re", "sigma_c_im", "iteration_history", "ip", "ea", "gap"):
        assert k in r
    ref, sx = _ref_evgw_ac(e, B, miP, abP, nocc, naux, vxc, nw=48, n_iter=6, tol=1e-6,
                           solver="linearized")
    assert np.max(np.abs(r["sigma_x"] - sx)) < 1e-13
    err = float(np.max(np.abs(r["qp_energies"] - ref["E"])))
    print(f"evgw defaults: {r['n_cycles']} cycles (oracle {ref['n_cycles']}), max |dE| "
          f"{err:.3e}")
    assert err < 1e-6
    assert r["n_cycles"] == ref["n_cycles"] and r["converged"] == ref["converged"]
    hist = r["iteration_history"]
    assert [h["cycle"] for h in hist] == list(range(1, r["n_cycles"] + 1))
    for h, hr in zip(hist, ref["history"]):
        assert h["energy_change"] == pytest.approx(hr["energy_change"], abs=1e-8)
        assert h["rms_change"] == pytest.approx(hr["rms_change"], abs=1e-8)
        assert np.max(np.abs(h["z_factors_cycle"] - hr["z_factors"])) < 1e-6
        assert np.array_equal(h["z_factors"], r["z_factors"])
    assert r["final_error"] == hist[-1]["energy_change"]
    assert r["gap"] == pytest.approx(r["qp_energies"][nocc] - r["qp_energies"][nocc - 1])
    assert r["ip"] == -r["qp_energies"][nocc - 1] and r["ea"] == -r["qp_energies"][nocc]
    with pytest.raises(ValueError):
        gw.evgw(e, occ, B, miP, abP, np.eye(naux), vxc, qp_solver="secant")


def test_run_evgw_defaults_match_oracle(gpu):
    """quasix.quasix.run_evgw with only the required keys: Pulay mixing on, frontier window
    0.1 Ha, linearised AC evGW over every orbital, eta fixed."""
    from quasix_amd import gw
    nocc, nvir, naux = 4, 16, 64
    e, B, miP, abP = _dense_blocks(nocc, nvir, naux, seed=11)
    occ = np.r_[np.full(nocc, 2.0), np.zeros(nvir)]
    vxc = np.linspace(-0.5, -0.3, nocc + nvir)
    cfg = dict(mo_energy=e, mo_occ=occ, iaP=B.reshape(nocc, nvir, naux), ijP=miP, abP=abP,
               chol_v=np.eye(naux), vxc_dft=vxc, max_cycle=8, conv_tol=1e-6, nfreq=40)
    r = gw.run_evgw(cfg)
    ref, _ = _ref_evgw_ac(e, B, miP, abP, nocc, naux, vxc, nw=40, n_iter=8, tol=1e-6,
                          solver="linearized", diis=True, frontier_window=0.1)
    err = float(np.max(np.abs(r["qp_energies"] - ref["E"])))
    print(f"run_evgw: {r['n_cycles']} cycles (oracle {ref['n_cycles']}), max |dE| {err:.3e}")
    assert r["n_cycles"] == ref["n_cycles"]
    assert err < 1e-6
    with pytest.raises(KeyError):
        gw.run_evgw({k: v for k, v in cfg.items() if k != "vxc_dft"})


def test_evgw_frontier_and_all_orbital_stop_rules(gpu):
    """A case where the frontier rule stops after the first iteration and the all-orbital
    rule does not (V_xc of the frontier orbitals cancels their first update): the product
    driver follows the oracle's cycle count under both rules."""
    from quasix_amd import gw
    nocc, nvir, naux = 4, 12, 30
    e, B = O.make_inputs(nocc, nvir, naux, seed=6)
    nmo = nocc + nvir
    miP, abP = O.ov_only_blocks(B, nocc, nvir)
    lpq = np.ascontiguousarray(O.assemble_mnP(B, miP, abP, nocc, nvir, naux).transpose(2, 0, 1))
    sx = O.exchange_diag(miP, nmo, nocc)
    lo, hi = O.convergence_window(e, nocc, frontier_window=0.1)
    sig, om = O.get_sigma_diag(e, lpq, np.arange(nmo), nocc, 40)
    coeff, ofit = O.ac_pade_thiele_diag(sig, om)
    ef = 0.5 * (e[nocc - 1] + e[nocc])
    vxc = 0.9 * sx
    for p in range(lo, hi):
        vxc[p] = sx[p] + O.pade_thiele(e[p] - ef, ofit[p], coeff[:, p]).real
    occ = np.r_[np.full(nocc, 2.0), np.zeros(nvir)]
    for front in (True, False):
        r = gw.evgw(e, occ, B, miP, abP, np.eye(naux), vxc, max_cycle=40, conv_tol=1e-6,
                    nfreq=40, frontier_convergence=front, frontier_window=0.1)
        ref = O.evgw_ac(e, lpq, nocc, nw=40, n_iter=40, tol=1e-6, solver="linearized",
                        sigma_x=sx, vxc=vxc, frontier_convergence=front, frontier_window=0.1)
        print(f"frontier={front}: {r['n_cycles']} cycles (oracle {ref['n_cycles']})")
        assert r["n_cycles"] == ref["n_cycles"] and r["converged"] == ref["converged"]
        assert (r["n_cycles"] == 1) == front
        assert np.max(np.abs(r["qp_energies"] - ref["E"])) < 1e-6


def test_sigma_ac_matches_oracle(gpu):
    """Sigma_c(i nu) on the AC grid (qx_sigma_ac_dev) against get_sigma_diag, all orbitals,
    reference default grid (100 points, iw_cutoff 5)."""
    from quasix_amd import core
    nocc, nvir, naux = 5, 19, 84
    e, B = O.make_inputs(nocc, nvir, naux, seed=9)
    miP, abP = O.ov_only_blocks(B, nocc, nvir)
    mn = O.assemble_mnP(B, miP, abP, nocc, nvir, naux)          # [nmo, nmo, naux]
    lpq = np.ascontiguousarray(mn.transpose(2, 0, 1))
    got = core.compute_sigma_c_ac(e, lpq, nocc)
    sig, om = O.get_sigma_diag(e, lpq, np.arange(nocc + nvir), nocc)
    assert got["sigma_c"].shape == sig.shape
    # grid nodes come from the library's Newton iteration on P_n vs the oracle's Golub-Welsch
    # eigenvalues: equal to the last bits
    assert np.allclose(got["omega"], om, rtol=1e-12, atol=0)
    err = float(np.max(np.abs(got["sigma_c"] - sig)))
    print(f"sigma_ac: max abs err {err:.3e}")
    assert err < 1e-10


@pytest.mark.parametrize("solver,diis", [("linearized", False), ("newton", True),
                                         ("Newton-Raphson", False)])
def test_evgw_ac_matches_oracle(gpu, solver, diis):
    from quasix_amd import gw
    nocc, nvir, naux = 4, 16, 64
    e, B, miP, abP = _dense_blocks(nocc, nvir, naux, seed=10)
    occ = np.r_[np.full(nocc, 2.0), np.zeros(nvir)]
    r = gw.evgw(e, occ, B, miP, abP, np.eye(naux), np.zeros(nocc + nvir), max_cycle=5,
                conv_tol=1e-7, diis=diis, nfreq=40, qp_solver=solver)
    ref, _ = _ref_evgw_ac(e, B, miP, abP, nocc, naux, np.zeros(nocc + nvir), nw=40, n_iter=5,
                          tol=1e-7, diis=diis, solver=gw.parse_qp_solver(solver))
    err = float(np.max(np.abs(r["qp_energies"] - ref["E"])))
    print(f"evgw ac {solver} diis={diis}: {r['n_cycles']} cycles, max |dE| vs oracle {err:.3e}")
    assert r["n_cycles"] == ref["n_cycles"]
    assert err < 1e-6


def test_evgw_cd_method_honours_solver(gpu):
    """sigma_method='cd' through the reference interface: Newton and linearised QP steps
    against oracle.evgw_cd with the same blocks (ov-only, no residue poles)."""
    from quasix_amd import gw
    nocc, nvir, naux = 5, 19, 84
    e, B = O.make_inputs(nocc, nvir, naux, seed=12)
    miP, abP = O.ov_only_blocks(B, nocc, nvir)
    occ = np.r_[np.full(nocc, 2.0), np.zeros(nvir)]
    vxc = np.linspace(-0.6, -0.2, nocc + nvir)
    sx = O.exchange_diag(miP, nocc + nvir, nocc)
    tg = np.arange(nocc + nvir)
    fr, wt = O.get_scaled_legendre_roots(NF, 0.5)
    for solver in ("linearized", "newton"):
        r = gw.evgw(e, occ, B, miP, abP, np.eye(naux), vxc, max_cycle=4, conv_tol=1e-7,
                    nfreq=NF, qp_solver=solver, sigma_method="cd")
        ref = O.evgw_cd(e, B, nocc, tg, fr, wt, n_iter=4, tol=1e-7, static=(sx - vxc),
                        solver=solver)
        err = float(np.max(np.abs(r["qp_energies"] - ref["E"])))
        print(f"evgw cd {solver}: {r['n_cycles']} cycles, max |dE| {err:.3e}")
        assert r["n_cycles"] == ref["n_cycles"]
        assert err < 1e-6


def test_exchange_diag_matches_oracle(gpu):
    from quasix_amd import core
    nocc, nvir, naux = 7, 40, 150
    _, B = O.make_inputs(nocc, nvir, naux, seed=15)
    miP, _ = O.ov_only_blocks(B, nocc, nvir)
    miP = miP + 0.01 * np.random.default_rng(2).standard_normal(miP.shape)   # dense block
    got = core.compute_exchange_diag(miP, nocc + nvir, nocc)
    ref = O.exchange_diag(miP, nocc + nvir, nocc)
    err = float(np.max(np.abs(got - ref)))
    print(f"exchange diag: max abs err {err:.3e}")
    assert err < 1e-13


@pytest.mark.parametrize("nao,nocc,naux", [(30, 5, 84), (97, 11, 130), (200, 20, 400)])
def test_mo_transform_matches_oracle(gpu, nao, nocc, naux):
    """Ragged sizes (not multiples of the 64 tile) included."""
    from quasix_amd import core
    rng = np.random.default_rng(nao)
    J = rng.standard_normal((nao, nao, naux))
    J = 0.5 * (J + J.transpose(1, 0, 2))
    C = np.linalg.qr(rng.standard_normal((nao, nao)))[0]
    got = core.transform_mo_3center(J, C[:, :nocc], C[:, nocc:])
    ref = O.transform_mo_3center(J, C[:, :nocc], C[:, nocc:])
    err = float(np.max(np.abs(got - ref)))
    print(f"mo transform {nao}/{nocc}/{naux}: max abs err {err:.3e}")
    assert err < 1e-11
