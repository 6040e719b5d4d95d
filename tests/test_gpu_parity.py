"""GPU parity: the HIP path through the C ABI against the FP64 oracle and golden fixtures.

Tolerances: FP64 path 1e-6 Ha on every QP energy (the acceptance bar; observed agreement is
~1e-10), FP64 Pi/Wc relative 1e-11; FP32 / BF16x3 Pi relative 3e-5 and max |dE_QP| <= 1 meV
(3.67e-5 Ha) against the FP64 path.
"""
import os

import numpy as np
import pytest

from oracle import gw_oracle as O

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")
MEV = 1.0 / 27211.386245988


def _c0():
    return np.load(os.path.join(GOLD, "c0_g0w0.npz"))


def _close(name, got, ref, tol, scale=None):
    """max |got - ref| <= tol * scale (scale defaults to 1); prints the observed error."""
    err = float(np.max(np.abs(np.asarray(got) - np.asarray(ref))))
    lim = tol * (1.0 if scale is None else float(scale))
    print(f"{name}: max abs err {err:.3e} (limit {lim:.3e})")
    assert err <= lim, f"{name}: {err:.3e} > {lim:.3e}"


def test_pi_fp64_matches_golden(gpu):
    from quasix_amd import core
    g = _c0()
    e, B = g["mo_energy"], g["B_ov"]
    nocc, naux = 5, B.shape[1]
    occ = np.array([2.0] * nocc + [0.0] * (e.size - nocc))
    p0 = core.compute_polarizability_p0(B.reshape(nocc, -1, naux), e, occ, g["freqs"])
    ref = g["pi"]
    got = p0[g["pi_idx"]]
    _close("pi fp64", got, ref, 1e-12, np.max(np.abs(ref)))
    _close("pi symmetry", p0, p0.transpose(0, 2, 1), 0.0)


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 3e-5), ("bf16x3", 3e-5)])
def test_pi_ragged_shape(gpu, prec, tol):
    from quasix_amd import core
    nocc, nvir, naux = 7, 45, 150
    e, B = O.make_inputs(nocc, nvir, naux, seed=11)
    fr, _ = O.get_scaled_legendre_roots(6, 0.5)
    occ = np.array([2.0] * nocc + [0.0] * nvir)
    p0 = core.compute_polarizability_p0(B.reshape(nocc, nvir, naux), e, occ, fr, precision=prec)
    for k, w in enumerate(fr):
        ref = O.get_rho_response(w, e, B, nocc)
        _close(f"pi {prec} k={k}", p0[k], ref, tol, np.max(np.abs(ref)))


def test_screened_wc(gpu):
    from quasix_amd import core
    g = _c0()
    pis = g["pi"]
    wc, nfb = core.compute_screening_wc(pis)
    assert nfb == 0
    for k in range(pis.shape[0]):
        ref = O.screened_wc(pis[k])
        _close(f"wc k={k}", wc[k], ref, 1e-11, max(1.0, np.max(np.abs(ref))))


@pytest.mark.parametrize("pair,fuse,n", [("1", "1", 300), ("0", "1", 300), ("1", "0", 300),
                                         ("0", "0", 300), ("1", "1", 250)])
def test_screened_wc_multi_block(gpu, pair, fuse, n, monkeypatch):
    """Five (n = 300, padded to 320) or four (n = 250) 64-blocks exercise every column-update
    tile kind (Cholesky NT tiles, inverse NN sums) with a ragged last block, in step pairs
    (wide update of two block columns, then the K = 64 update; an odd block count ends with a
    single step) or single steps (QXGW_POTRF_PAIR=0), with the diagonal factor fused into
    the update or launched on its own (QXGW_POTRF_FUSE=0)."""
    from quasix_amd import core
    monkeypatch.setenv("QXGW_POTRF_PAIR", pair)
    monkeypatch.setenv("QXGW_POTRF_FUSE", fuse)
    rng = np.random.default_rng(21)
    pis = []
    for lo in (-3.0, -0.2):
        Q = np.linalg.qr(rng.standard_normal((n, n)))[0]
        pi = (Q * rng.uniform(lo, 0.0, n)) @ Q.T
        pis.append(0.5 * (pi + pi.T))
    pis = np.stack(pis)
    wc, nfb = core.compute_screening_wc(pis)
    assert nfb == 0
    for k in range(pis.shape[0]):
        ref = O.screened_wc(pis[k])
        _close(f"wc n={n} pair={pair} fuse={fuse} k={k}", wc[k], ref, 1e-11, max(1.0, np.max(np.abs(ref))))


@pytest.mark.parametrize("pair", ["1", "0"])
def test_screened_wc_xcd_mapping(gpu, pair, monkeypatch):
    """27 frequencies in three stream groups of 9 (more frequencies than XCDs, not a
    multiple of 8) and four 64-blocks: the per-XCD workgroup mapping of the update and
    panel kernels computes every tile with the same operands in the same order as launch
    order, so Wc is bit-identical with QXGW_POTRF_XCD=0, and matches the oracle."""
    from quasix_amd import core
    monkeypatch.setenv("QXGW_POTRF_PAIR", pair)
    rng = np.random.default_rng(5)
    n, nf = 230, 27
    pis = []
    for k in range(nf):
        Q = np.linalg.qr(rng.standard_normal((n, n)))[0]
        pi = (Q * rng.uniform(-3.0 / (1 + k), 0.0, n)) @ Q.T
        pis.append(0.5 * (pi + pi.T))
    pis = np.stack(pis)
    monkeypatch.setenv("QXGW_POTRF_XCD", "1")
    wc1, nfb1 = core.compute_screening_wc(pis)
    monkeypatch.setenv("QXGW_POTRF_XCD", "0")
    wc0, nfb0 = core.compute_screening_wc(pis)
    assert nfb0 == nfb1 == 0
    assert np.array_equal(wc0, wc1)
    for k in (0, 8, 13, 26):
        ref = O.screened_wc(pis[k])
        _close(f"wc xcd k={k}", wc1[k], ref, 1e-11, max(1.0, np.max(np.abs(ref))))


def test_screened_wc_lu_fallback(gpu):
    """A Pi with an eigenvalue above 1 makes eps indefinite: Cholesky fails, LU is used."""
    from quasix_amd import core
    rng = np.random.default_rng(3)
    n = 70
    Q = np.linalg.qr(rng.standard_normal((n, n)))[0]
    lam = np.concatenate([[1.7], rng.uniform(-2.0, 0.0, n - 1)])
    pi = (Q * lam) @ Q.T
    pi = 0.5 * (pi + pi.T)
    wc, nfb = core.compute_screening_wc(pi[None])
    assert nfb == 1
    ref = O.screened_wc(pi)
    _close("wc lu", wc[0], ref, 1e-10, np.max(np.abs(ref)))


def test_wnm_matches_golden(gpu):
    from quasix_amd import core
    g = _c0()
    e, B = g["mo_energy"], g["B_ov"]
    mn = O.assemble_mnP(B, *O.ov_only_blocks(B, 5, 19), 5, 19, B.shape[1])
    pt, pm, Bp = O.pair_list(mn, g["targets"])
    wnm = core.compute_wmn_imag(e, B, 5, Bp, g["freqs"])
    ref = g["wnm"]
    _close("wnm", wnm, ref, 1e-12, np.max(np.abs(ref)))


def test_g0w0_c0_fp64_matches_golden(gpu):
    from quasix_amd import gw
    g = _c0()
    r = gw.g0w0(g["mo_energy"], g["B_ov"], 5, g["targets"], nfreq=32, stats=True)
    assert np.all(r["state"] == 0)
    _close("E_QP", r["qp_energies"], g["E"], 1e-6)
    _close("E_QP headroom", r["qp_energies"], g["E"], 1e-9)
    _close("Z", r["z_factors"], g["Z"], 1e-6)
    _close("Sigma", r["sigma_c"], g["sigma"], 1e-8)
    assert r["stats"]["residue_targets"] == 0


def test_g0w0_full_blocks_with_residues(gpu):
    """oo/vv blocks present: Newton iterates cross orbital energies, residues are evaluated."""
    from quasix_amd import gw
    g = np.load(os.path.join(GOLD, "casida_cd.npz"))
    e, full = g["mo_energy"], g["full"]
    nocc, nmo, naux = 3, 9, 8
    nvir = nmo - nocc
    B = np.ascontiguousarray(full[:nocc, nocc:].reshape(-1, naux))
    targets = np.arange(nmo)
    fr, wt = O.get_scaled_legendre_roots(32, 0.5)
    pt, pm, Bp = O.pair_list(full, targets)
    ref = O.g0w0_cd(e, B, nocc, targets, fr, wt, pairs=(pt, pm, Bp), eta=0.01)
    iaP = full[:nocc, nocc:].reshape(-1, naux)
    miP = full[:, :nocc].reshape(-1, naux)
    abP = full[nocc:, nocc:].reshape(-1, naux)
    pairs = gw.pairs_from_blocks(iaP, miP, abP, nocc, nvir, naux, targets)
    r = gw.g0w0(e, B, nocc, targets, nfreq=32, eta=0.01, pairs=pairs, stats=True)
    print("stats", r["stats"])
    assert r["stats"]["residue_targets"] > 0 and r["stats"]["residue_jobs"] > 0
    _close("E_QP", r["qp_energies"], ref["E"], 1e-6)
    _close("Z", r["z_factors"], ref["Z"], 1e-5)


def test_g0w0_ov_cholesky_fallback(gpu):
    """Single-instance ov solve with eps = I - Pi indefinite at 9 of 16 frequencies (an
    occupied level at 3.4 Ha above the virtuals, B scaled by 2.5; smallest |eigenvalue| 0.078
    per the oracle): the failure flags come back with the Newton states, Wc_nm is redone by
    the LU fallback and Newton relaunched.  FP64 against the oracle's Cholesky -> LU chain."""
    from quasix_amd import gw
    nocc, nvir, naux, nf = 4, 20, 64, 16
    e, B = O.make_inputs(nocc, nvir, naux, seed=72)
    e[nocc - 1] = 3.4
    B = B * 2.5
    tg = np.arange(nocc - 2, nocc + 2)
    fr, wt = O.get_scaled_legendre_roots(nf, 0.5)
    ref = O.g0w0_cd(e, B, nocc, tg, fr, wt, eta=0.01)
    r = gw.g0w0(e, B, nocc, tg, nfreq=nf, eta=0.01, stats=True)
    print("fallback stats", r["stats"], "states", r["state"].tolist())
    assert r["stats"]["chol_fallbacks"] == 9 and r["stats"]["singular_fallbacks"] == 0
    _close("g0w0 ov fallback", r["qp_energies"], ref["E"], 1e-6, 1.0)
    _close("g0w0 ov fallback Z", r["z_factors"], ref["Z"], 1e-5, 1.0)


def test_g0w0_c3_fp64_and_mixed(gpu):
    """nocc 20 / nvir 200 / naux 900 (a ragged naux), 32 frequencies, HOMO-5..LUMO+5."""
    from quasix_amd import gw
    nocc, nvir, naux = 20, 200, 900
    e, B = O.make_inputs(nocc, nvir, naux, seed=3)
    targets = O.frontier_targets(nocc, nocc + nvir)
    fr, wt = O.get_scaled_legendre_roots(32, 0.5)
    ref = O.g0w0_cd(e, B, nocc, targets, fr, wt)
    r64 = gw.g0w0(e, B, nocc, targets, nfreq=32)
    _close("E_QP fp64", r64["qp_energies"], ref["E"], 1e-6)
    for prec in ("fp32", "bf16x3"):
        rm = gw.g0w0(e, B, nocc, targets, nfreq=32, precision=prec)
        dE = np.max(np.abs(rm["qp_energies"] - r64["qp_energies"]))
        print(f"{prec}: max|dE_QP| vs FP64 = {dE / MEV:.3e} meV")
        assert dE <= 1.0 * MEV


@pytest.mark.parametrize("f2", ["0", "1"])
@pytest.mark.parametrize("nf", [7, 8])
def test_wnm_bf16x3_both_kernels(gpu, f2, nf, monkeypatch):
    """The one- and two-frequency mixed Wc_nm kernels (QXGW_WNM_F2), with an odd frequency
    count (the two-frequency kernel computes the last one twice and writes it once) and
    pair/auxiliary counts off the 128 tiles."""
    from quasix_amd import core
    monkeypatch.setenv("QXGW_WNM_F2", f2)
    nocc, nvir, naux = 7, 45, 200
    e, B = O.make_inputs(nocc, nvir, naux, seed=9)
    targets = O.frontier_targets(nocc, nocc + nvir, 3, 3)
    mn = O.assemble_mnP(B, *O.ov_only_blocks(B, nocc, nvir), nocc, nvir, naux)
    pt, pm, Bp = O.pair_list(mn, targets)
    fr, _ = O.get_scaled_legendre_roots(nf, 0.5)
    ref = O.wmn_imag(e, B, nocc, Bp, fr)
    got = core.compute_wmn_imag(e, B, nocc, Bp, fr, precision="bf16x3")
    _close(f"wnm bf16x3 f2={f2} nf={nf}", got, ref, 2e-5, np.max(np.abs(ref)))


@pytest.mark.parametrize("nf,blocked", [(7, "1"), (20, "1"), (20, "0"), (3, "1")])
def test_wnm_bf16x3_xcd_blocks(gpu, nf, blocked, monkeypatch):
    """The two-frequency mixed Wc_nm kernel with its work items cut into per-XCD blocks of
    pair tiles x frequency pairs (QXGW_WNM_BLOCKED=1) and in the plain grid order (0):
    5 pair tiles, 3 auxiliary row tiles (naux 300), and 4, 10 or 2 frequency pairs, so
    blocks are partial in the frequency direction, the block count is not a multiple of 8,
    and empty slots return at once."""
    from quasix_amd import core
    monkeypatch.setenv("QXGW_WNM_BLOCKED", blocked)
    nocc, nvir, naux = 7, 100, 300
    e, B = O.make_inputs(nocc, nvir, naux, seed=19)
    targets = O.frontier_targets(nocc, nocc + nvir, 4, 4)
    mn = O.assemble_mnP(B, *O.ov_only_blocks(B, nocc, nvir), nocc, nvir, naux)
    pt, pm, Bp = O.pair_list(mn, targets)
    assert 512 < Bp.shape[0] <= 640
    fr, _ = O.get_scaled_legendre_roots(nf, 0.5)
    ref = O.wmn_imag(e, B, nocc, Bp, fr)
    got = core.compute_wmn_imag(e, B, nocc, Bp, fr, precision="bf16x3")
    _close(f"wnm bf16x3 blocked={blocked} nf={nf}", got, ref, 2e-5, np.max(np.abs(ref)))


@pytest.mark.parametrize("prec,tol", [("fp32", 2e-5), ("bf16x3", 2e-5)])
def test_wnm_mixed_precision(gpu, prec, tol):
    """Wc_nm through the mixed kernels against the FP64 oracle (ragged naux = 150)."""
    from quasix_amd import core
    nocc, nvir, naux = 6, 40, 150
    e, B = O.make_inputs(nocc, nvir, naux, seed=5)
    targets = O.frontier_targets(nocc, nocc + nvir, 2, 2)
    mn = O.assemble_mnP(B, *O.ov_only_blocks(B, nocc, nvir), nocc, nvir, naux)
    pt, pm, Bp = O.pair_list(mn, targets)
    fr, _ = O.get_scaled_legendre_roots(8, 0.5)
    ref = O.wmn_imag(e, B, nocc, Bp, fr)
    got = core.compute_wmn_imag(e, B, nocc, Bp, fr, precision=prec)
    _close(f"wnm {prec}", got, ref, tol, np.max(np.abs(ref)))


@pytest.mark.parametrize("variant", [0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10])
def test_pi_bf16x3_every_variant(gpu, variant, monkeypatch):
    """Each BF16x3 Pi kernel instance the autotuner can pick, forced in turn."""
    from quasix_amd import core
    monkeypatch.setenv("QXGW_PI_VARIANT", str(variant))
    nocc, nvir, naux = 9, 70, 300                # naux not a multiple of 64/128/256
    e, B = O.make_inputs(nocc, nvir, naux, seed=13)
    fr, _ = O.get_scaled_legendre_roots(5, 0.5)
    occ = np.array([2.0] * nocc + [0.0] * nvir)
    p0 = core.compute_polarizability_p0(B.reshape(nocc, nvir, naux), e, occ, fr,
                                        precision="bf16x3")
    _close(f"pi symmetry v{variant}", p0, p0.transpose(0, 2, 1), 0.0)
    for k, w in enumerate(fr):
        ref = O.get_rho_response(w, e, B, nocc)
        _close(f"pi bf16x3 v{variant} k={k}", p0[k], ref, 3e-5, np.max(np.abs(ref)))


@pytest.mark.parametrize("variant", [1, 5, 7, 8, 9, 10])
@pytest.mark.parametrize("nocc,nvir", [(4, 15), (4, 45)])
def test_pi_bf16x3_odd_chunk_counts(gpu, variant, nocc, nvir, monkeypatch):
    """N_ov padded to 64 and 192: one and three K64 chunks (the producer loops of the
    double-buffered instances advance two chunks per pass), with full and partial tiles, and
    an odd frequency count (the frequency-pair instance duplicates the last one)."""
    from quasix_amd import core
    monkeypatch.setenv("QXGW_PI_VARIANT", str(variant))
    naux = 300
    e, B = O.make_inputs(nocc, nvir, naux, seed=31)
    fr, _ = O.get_scaled_legendre_roots(5, 0.5)
    occ = np.array([2.0] * nocc + [0.0] * nvir)
    p0 = core.compute_polarizability_p0(B.reshape(nocc, nvir, naux), e, occ, fr,
                                        precision="bf16x3")
    for k, w in enumerate(fr):
        ref = O.get_rho_response(w, e, B, nocc)
        _close(f"pi v{variant} nov {nocc * nvir} k={k}", p0[k], ref, 3e-5, np.max(np.abs(ref)))


@pytest.mark.parametrize("naux,nf", [(300, 7), (1000, 16), (1000, 17)])
def test_pi_pair_kernel_tail_items(gpu, naux, nf, monkeypatch):
    """The frequency-pair instance runs the items past the last full round of CUs as
    single-frequency workgroups (every item at naux 300 / 7 frequencies, 32 of 288 at
    naux 1000 / 16, with the odd last pair at 17): the same MFMA sequence per frequency, so
    Pi is bit-identical to the all-pairs launch (QXGW_PI_TAIL=0), and within 3e-5 of the
    oracle."""
    from quasix_amd import core
    monkeypatch.setenv("QXGW_PI_VARIANT", "8")
    nocc, nvir = 5, 40
    e, B = O.make_inputs(nocc, nvir, naux, seed=41)
    fr, _ = O.get_scaled_legendre_roots(nf, 0.5)
    occ = np.array([2.0] * nocc + [0.0] * nvir)
    args = (B.reshape(nocc, nvir, naux), e, occ, fr)
    p_tail = core.compute_polarizability_p0(*args, precision="bf16x3")
    monkeypatch.setenv("QXGW_PI_TAIL", "0")
    p_pairs = core.compute_polarizability_p0(*args, precision="bf16x3")
    _close(f"pi tail vs pairs naux {naux} nf {nf}", p_tail, p_pairs, 0.0)
    for k in (0, nf - 1):
        ref = O.get_rho_response(fr[k], e, B, nocc)
        _close(f"pi tail naux {naux} k={k}", p_tail[k], ref, 3e-5, np.max(np.abs(ref)))


@pytest.mark.parametrize("variant", [0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10])
def test_g0w0_bf16x3_every_variant(gpu, variant, monkeypatch):
    from quasix_amd import gw
    monkeypatch.setenv("QXGW_PI_VARIANT", str(variant))
    nocc, nvir, naux = 12, 108, 480
    e, B = O.make_inputs(nocc, nvir, naux, seed=17)
    targets = O.frontier_targets(nocc, nocc + nvir)
    r64 = gw.g0w0(e, B, nocc, targets, nfreq=24)
    rm = gw.g0w0(e, B, nocc, targets, nfreq=24, precision="bf16x3", stats=True)
    assert rm["stats"]["pi_variant"] == variant
    _close(f"E_QP bf16x3 v{variant}", rm["qp_energies"], r64["qp_energies"], MEV)


@pytest.mark.parametrize("variant", ["0", "1", "2", "3", "4"])
@pytest.mark.parametrize("naux,nf,nocc,nvir", [(150, 6, 7, 45), (300, 7, 7, 45), (64, 1, 7, 45),
                                               (200, 5, 3, 9)])
def test_pi_fp64_both_instances(gpu, variant, naux, nf, nocc, nvir, monkeypatch):
    """FP64 Pi, one-frequency (0) and frequency-pair instances (1: K16, 2: K32, 3: K32 and
    4: K16 with double-buffered LDS), against the oracle:
    ragged naux, N_ov not a multiple of the chunk (315) and below one chunk (27), an odd
    frequency count (the last pair holds one frequency) and a single frequency, in both
    output forms (full Pi, and eps = I - Pi inside a G0W0 solve)."""
    from quasix_amd import core, gw
    monkeypatch.setenv("QXGW_PI_VARIANT", variant)
    e, B = O.make_inputs(nocc, nvir, naux, seed=naux + nf)
    fr, wt = O.get_scaled_legendre_roots(nf, 0.5)
    occ = np.array([2.0] * nocc + [0.0] * nvir)
    p0 = core.compute_polarizability_p0(B.reshape(nocc, nvir, naux), e, occ, fr)
    for k, w in enumerate(fr):
        ref = O.get_rho_response(w, e, B, nocc)
        _close(f"pi fp64 v{variant} k={k}", p0[k], ref, 1e-12, np.max(np.abs(ref)))
    tg = np.arange(nocc - 2, nocc + 2)
    r = gw.g0w0(e, B, nocc, tg, nfreq=nf)
    _close(f"g0w0 fp64 v{variant}", r["qp_energies"], O.g0w0_cd(e, B, nocc, tg, fr, wt)["E"], 1e-6)
