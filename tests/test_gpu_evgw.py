"""GPU: evGW driver (quasix_amd.gw.evgw / evgw_core) against the oracle's evGW on the same
seeded inputs.  FP64: every updated QP energy within 1e-6 Ha of the oracle after each run;
the mixed path (BF16x3 Pi) within 1 meV of FP64.  Covers damping only and DIIS (whose
driver call order falls back to damping from the fourth iteration on)."""
import numpy as np
import pytest

from oracle import gw_oracle as O

pytestmark = pytest.mark.gpu
MEV = 1.0 / 27211.386245988
NF = 16


@pytest.mark.parametrize("diis", [False, True])
def test_evgw_core_matches_oracle(gpu, diis):
    from quasix_amd import gw
    nocc, nvir, naux = 5, 19, 84
    e, B = O.make_inputs(nocc, nvir, naux, seed=7)
    tg = gw.frontier_targets(nocc, nocc + nvir, 2, 2)
    fr, wt = O.get_scaled_legendre_roots(NF, 0.5)
    ref = O.evgw_cd(e, B, nocc, tg, fr, wt, n_iter=5, damping=0.5, tol=1e-7, diis=diis)
    got = gw.evgw_core(e, B, nocc, tg, n_iter=5, damping=0.5, conv_tol=1e-7, diis=diis,
                       nfreq=NF)
    err = float(np.max(np.abs(got["qp_energies"] - ref["E"])))
    print(f"evgw diis={diis}: {len(got['history'])} cycles, max |dE| vs oracle {err:.3e}")
    assert len(got["history"]) == len(ref["history"])
    assert err < 1e-6
    mixed = gw.evgw_core(e, B, nocc, tg, n_iter=5, damping=0.5, conv_tol=1e-7, diis=diis,
                         nfreq=NF, precision="bf16x3")
    dm = float(np.max(np.abs(mixed["qp_energies"] - got["qp_energies"])))
    print(f"mixed vs fp64: {dm / MEV:.3e} meV")
    assert dm < MEV


def test_evgw_reference_interface(gpu):
    """quasix.quasix.evgw arguments (iaP/ijP/abP blocks, chol_v for naux, vxc) and result keys."""
    from quasix_amd import gw
    nocc, nvir, naux = 5, 19, 84
    e, B = O.make_inputs(nocc, nvir, naux, seed=8)
    miP, abP = O.ov_only_blocks(B, nocc, nvir)
    occ = np.r_[np.full(nocc, 2.0), np.zeros(nvir)]
    vxc = np.linspace(-0.6, -0.2, nocc + nvir)
    r = gw.evgw(e, occ, B, miP, abP, np.eye(naux), vxc, max_cycle=4,
                conv_tol=1e-7, diis=True, nfreq=NF, qp_solver="newton")
    for k in ("qp_energies", "z_factors", "converged", "n_cycles", "final_error", "sigma_x",
              "sigma_c_re", "sigma_c_im", "iteration_history", "ip", "ea", "gap"):
        assert k in r
    # Sigma_x from ijP on the GPU, as the reference driver computes it
    sx = O.exchange_diag(miP, nocc + nvir, nocc)
    assert np.max(np.abs(r["sigma_x"] - sx)) < 1e-13
    tg = gw.frontier_targets(nocc, nocc + nvir)
    fr, wt = O.get_scaled_legendre_roots(NF, 0.5)
    ref = O.evgw_cd(e, B, nocc, tg, fr, wt, n_iter=4, tol=1e-7, diis=True,
                    static=(sx - vxc)[tg])
    err = float(np.max(np.abs(r["qp_energies"] - ref["E"])))
    print(f"evgw interface: max |dE| vs oracle {err:.3e}")
    assert err < 1e-6
    assert r["gap"] == pytest.approx((r["qp_energies"][nocc] - r["qp_energies"][nocc - 1])
                                     * gw.HARTREE_TO_EV)


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


@pytest.mark.parametrize("solver,diis", [("linearized", False), ("newton", True)])
def test_evgw_ac_matches_oracle(gpu, solver, diis):
    from quasix_amd import gw
    nocc, nvir, naux = 4, 16, 64
    e, B = O.make_inputs(nocc, nvir, naux, seed=10)
    miP, abP = O.ov_only_blocks(B, nocc, nvir)
    occ = np.r_[np.full(nocc, 2.0), np.zeros(nvir)]
    r = gw.evgw(e, occ, B, miP, abP, np.eye(naux), np.zeros(nocc + nvir), max_cycle=5,
                conv_tol=1e-7, diis=diis, nfreq=40, qp_solver=solver, sigma_method="ac")
    mn = O.assemble_mnP(B, miP, abP, nocc, nvir, naux)
    lpq = np.ascontiguousarray(mn.transpose(2, 0, 1))
    ref = O.evgw_ac(e, lpq, nocc, nw=40, n_iter=5, tol=1e-7, diis=diis, solver=solver,
                    sigma_x=O.exchange_diag(miP, nocc + nvir, nocc))
    err = float(np.max(np.abs(r["qp_energies"] - ref["E"])))
    print(f"evgw ac {solver} diis={diis}: {r['n_cycles']} cycles, max |dE| vs oracle {err:.3e}")
    assert r["n_cycles"] == len(ref["history"])
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
