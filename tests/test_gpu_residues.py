"""GPU: contour-deformation residues with full oo / vv blocks at sizes where the real-axis
screened interaction is a real workload (naux 320 and 1000).  Targets below the HOMO and
above the LUMO have poles between their QP energy and the Fermi level, so every Newton
evaluation of theirs needs W(w + i eta) at those poles: complex Pi on the FP64 MFMA
polarizability kernel, blocked complex LU on the device.  Checked against the oracle's
dense complex solves (the CD self-energy the oracle pins by brute-force Casida): QP energies
to 1e-6 Ha, Z to 1e-5; the larger case prints the residue-phase time."""
import ctypes as C

import numpy as np
import pytest

from oracle import gw_oracle as O

pytestmark = pytest.mark.gpu


def _full_model(nocc, nvir, naux, seed, scale=0.05):
    e, B = O.make_inputs(nocc, nvir, naux, seed=seed)
    rng = np.random.default_rng(seed + 1)
    nmo = nocc + nvir
    miP, _ = O.ov_only_blocks(B, nocc, nvir)
    mi = miP.reshape(nmo, nocc, naux)
    mi[:nocc] = scale * rng.standard_normal((nocc, nocc, naux)) / np.sqrt(naux)
    ab = scale * rng.standard_normal((nvir, nvir, naux)) / np.sqrt(naux)
    return e, B, mi.reshape(nmo * nocc, naux), ab.reshape(nvir * nvir, naux)


def test_residues_full_blocks_naux320(gpu):
    from quasix_amd import gw
    nocc, nvir, naux, nf = 6, 30, 320, 24
    e, B, miP, abP = _full_model(nocc, nvir, naux, seed=50)
    mn = O.assemble_mnP(B, miP, abP, nocc, nvir, naux)
    Bs = np.ascontiguousarray(mn[:nocc, nocc:].reshape(nocc * nvir, naux))
    tg = np.arange(nocc - 4, nocc + 3)
    fr, wt = O.get_scaled_legendre_roots(nf, 0.5)
    ref = O.g0w0_cd(e, Bs, nocc, tg, fr, wt, pairs=O.pair_list(mn, tg), eta=0.01)
    pairs = gw.pairs_from_blocks(B, miP, abP, nocc, nvir, naux, tg)
    r = gw.g0w0(e, Bs, nocc, tg, nfreq=nf, eta=0.01, pairs=pairs, stats=True)
    st = r["stats"]
    dE = float(np.max(np.abs(r["qp_energies"] - ref["E"])))
    dZ = float(np.max(np.abs(r["z_factors"] - ref["Z"])))
    print(f"residues naux {naux}: {st['residue_targets']} targets, {st['residue_jobs']} jobs, "
          f"max |dE| {dE:.3e} Ha, max |dZ| {dZ:.3e}")
    assert st["residue_targets"] > 0 and st["residue_jobs"] > 0
    assert dE < 1e-6 and dZ < 1e-5


def test_residues_full_blocks_naux1000_timed(gpu):
    """The residue path at naux 1000 (npad 1024): mixed and FP64 imaginary-axis parts agree
    within 1 meV, and the residue phase time per job is printed."""
    import quasix_amd._lib as L
    from quasix_amd import gw
    nocc, nvir, naux, nf = 20, 120, 1000, 32
    e, B, miP, abP = _full_model(nocc, nvir, naux, seed=51)
    tg = np.array([nocc - 3, nocc - 2, nocc - 1, nocc, nocc + 1])
    Bs, _, pairs = gw.assemble_blocks(B, miP, abP, nocc, nvir, naux, tg, ov=True, mi=False)
    h = L.handle(0)
    h.check(h.lib.qx_set_profiling(h.h, 1))
    r = gw.g0w0(e, Bs, nocc, tg, nfreq=nf, eta=0.01, pairs=pairs, stats=True)
    ms, cnt = (C.c_double * 8)(), (C.c_longlong * 8)()
    h.check(h.lib.qx_get_profile(h.h, ms, cnt))
    h.check(h.lib.qx_set_profiling(h.h, 0))
    st = r["stats"]
    print(f"residues naux {naux}: {st['residue_targets']} targets, {st['residue_jobs']} jobs, "
          f"residue phase {ms[6]:.1f} ms ({ms[6] / max(st['residue_jobs'], 1):.2f} ms per job), "
          f"solve {st['seconds']:.2f} s, states {r['state'].tolist()}")
    assert st["residue_jobs"] > 0
    rm = gw.g0w0(e, Bs, nocc, tg, nfreq=nf, eta=0.01, pairs=pairs, precision="bf16x3")
    d = float(np.max(np.abs(rm["qp_energies"] - r["qp_energies"]))) * 27211.386245988
    print(f"mixed vs FP64 with residues: {d:.3e} meV")
    assert d < 1.0


def test_ov_path_residue_targets(gpu):
    """The ov-only entry (pairs gathered on the device) launches without downloading the
    energies or the pair list; a target whose Newton iterate needs the residue path makes it
    fetch them afterwards.  An occupied level inside the virtual range (HOMO at 2.0 Ha, Fermi
    level near 1.1 Ha) puts virtual poles between the HOMO's QP energy and the Fermi level.
    Checked against the oracle's G0W0 solve, and with static terms, guesses and reference
    energies against its evGW step."""
    from quasix_amd import gw
    nocc, nvir, naux, nf = 4, 20, 64, 16
    e, B = O.make_inputs(nocc, nvir, naux, seed=70)
    e[nocc - 1] = 2.0
    tg = np.arange(nocc - 2, nocc + 2)
    fr, wt = O.get_scaled_legendre_roots(nf, 0.5)
    ref = O.g0w0_cd(e, B, nocc, tg, fr, wt, eta=0.01)
    r = gw.g0w0(e, B, nocc, tg, nfreq=nf, eta=0.01, stats=True)
    st = r["stats"]
    dE = float(np.max(np.abs(r["qp_energies"] - ref["E"])))
    print(f"ov residues: {st['residue_targets']} targets, {st['residue_jobs']} jobs, "
          f"states {r['state'].tolist()}, max |dE| {dE:.3e} Ha")
    assert st["residue_targets"] > 0 and st["residue_jobs"] > 0
    assert dE < 1e-6
    assert np.max(np.abs(r["z_factors"] - ref["Z"])) < 1e-5
    rng = np.random.default_rng(71)
    static = 0.05 * rng.standard_normal(len(tg))
    eref = e + 0.01 * rng.standard_normal(e.shape)
    guess = ref["E"] + 0.002
    # the library's mo_energy_ref anchors the QP equation and sets the eta sign, as in the
    # oracle's evGW step (Green's function and poles from e, anchor at eref)
    ref2 = O._g0w0_iter(e, eref, B, nocc, tg, fr, wt, static, 0.01, guess)
    r2 = gw.g0w0(e, B, nocc, tg, nfreq=nf, eta=0.01, static=static, mo_energy_ref=eref,
                 guess=guess, stats=True)
    dE2 = float(np.max(np.abs(r2["qp_energies"] - ref2["E"])))
    print(f"ov residues with static/ref/guess: {r2['stats']['residue_targets']} targets, "
          f"max |dE| {dE2:.3e} Ha")
    assert r2["stats"]["residue_targets"] > 0
    assert dE2 < 1e-6


def test_residue_lu_panel_forms_agree(gpu, monkeypatch):
    """The 16-wave fused LU panel (default) and the four-wave panel with separate search,
    scale and update passes (QXGW_LU_PANEL=0) choose the same pivots and do the same
    arithmetic: bit-identical QP energies on a residue workload (full blocks, naux 320)."""
    from quasix_amd import gw
    nocc, nvir, naux, nf = 6, 30, 320, 24
    e, B, miP, abP = _full_model(nocc, nvir, naux, seed=52)
    mn = O.assemble_mnP(B, miP, abP, nocc, nvir, naux)
    Bs = np.ascontiguousarray(mn[:nocc, nocc:].reshape(nocc * nvir, naux))
    tg = np.arange(nocc - 4, nocc + 3)
    pairs = gw.pairs_from_blocks(B, miP, abP, nocc, nvir, naux, tg)
    monkeypatch.setenv("QXGW_LU_PANEL", "0")
    r0 = gw.g0w0(e, Bs, nocc, tg, nfreq=nf, eta=0.01, pairs=pairs, stats=True)
    monkeypatch.setenv("QXGW_LU_PANEL", "1")
    r1 = gw.g0w0(e, Bs, nocc, tg, nfreq=nf, eta=0.01, pairs=pairs, stats=True)
    print(f"lu panel forms: {r1['stats']['residue_jobs']} jobs, max |dE| "
          f"{np.max(np.abs(r1['qp_energies'] - r0['qp_energies'])):.3e} Ha")
    assert r1["stats"]["residue_jobs"] > 0
    assert np.array_equal(r1["qp_energies"], r0["qp_energies"])
    assert np.array_equal(r1["z_factors"], r0["z_factors"])


def test_residue_targets_in_lockstep(gpu, monkeypatch):
    """Several residue targets: their Newton iterations run in lockstep, every target's
    residue jobs of a round in one batch (default), or one target after the other
    (QXGW_RES_LOCKSTEP=0).  A job's value does not depend on its batch, so the results are
    bit-identical, with the same number of jobs."""
    from quasix_amd import gw
    nocc, nvir, naux, nf = 6, 30, 320, 24
    e, B, miP, abP = _full_model(nocc, nvir, naux, seed=53)
    mn = O.assemble_mnP(B, miP, abP, nocc, nvir, naux)
    Bs = np.ascontiguousarray(mn[:nocc, nocc:].reshape(nocc * nvir, naux))
    tg = np.arange(nocc - 4, nocc + 3)
    pairs = gw.pairs_from_blocks(B, miP, abP, nocc, nvir, naux, tg)
    monkeypatch.setenv("QXGW_RES_LOCKSTEP", "0")
    r0 = gw.g0w0(e, Bs, nocc, tg, nfreq=nf, eta=0.01, pairs=pairs, stats=True)
    monkeypatch.setenv("QXGW_RES_LOCKSTEP", "1")
    r1 = gw.g0w0(e, Bs, nocc, tg, nfreq=nf, eta=0.01, pairs=pairs, stats=True)
    print(f"lockstep: {r1['stats']['residue_targets']} residue targets, "
          f"{r1['stats']['residue_jobs']} jobs ({r0['stats']['residue_jobs']} sequential)")
    assert r1["stats"]["residue_targets"] > 1
    assert r1["stats"]["residue_jobs"] == r0["stats"]["residue_jobs"]
    for k in ("qp_energies", "z_factors", "state"):
        assert np.array_equal(r1[k], r0[k]), k
    assert np.array_equal(r1["sigma_c"], r0["sigma_c"])
