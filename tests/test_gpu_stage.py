"""GPU: the quasix_core stage functions that take the reference's intermediates, and the
entry points no other GPU test calls.

* compute_sigma_c_pyscf_cd (qx_sigma_c_pyscf_cd_dev) against oracle.sigma_c_pyscf_cd
  (correlation_pyscf_cd.rs:60-298), symmetric and non-symmetric lpq, zero and singular P0;
* qx_sigma_cd_dev at E = 0 against the same static-CD formula (no residue pole lies between
  the Fermi level and 0 for these energies);
* compute_screening_w with a non-identity metric and with v_sqrt (qx_screening_w_dev)
  against oracle.compute_screening_w (screening.rs:164-420);
* the identity fallback of the Wc chain on an exactly singular eps (correlation_ac.rs:
  383-399);
* the host-pointer entry qx_g0w0_cd against the oracle and against qx_g0w0_cd_dev.
Tolerances: 1e-10 absolute on self-energies (Ha), 1e-9 relative on W."""
import ctypes as C

import numpy as np
import pytest

from oracle import gw_oracle as O

pytestmark = pytest.mark.gpu


def _p0(e, B, nocc, freqs):
    return np.stack([O.get_rho_response(w, e, B, nocc) for w in freqs])


def _metric(n, seed, cond=1e4):
    rng = np.random.default_rng(seed)
    Q = np.linalg.qr(rng.standard_normal((n, n)))[0]
    lam = np.geomspace(1.0, 1.0 / cond, n)
    return (Q * lam) @ Q.T


@pytest.mark.parametrize("symmetric", [True, False])
def test_sigma_c_pyscf_cd_matches_oracle(gpu, symmetric):
    from quasix_amd import core
    nocc, nvir, naux, nf = 5, 19, 84, 32
    e, B = O.make_inputs(nocc, nvir, naux, seed=21)
    rng = np.random.default_rng(3)
    miP, abP = O.ov_only_blocks(B, nocc, nvir)
    miP = miP + 0.02 * rng.standard_normal(miP.shape)
    abP = 0.02 * rng.standard_normal(abP.shape)
    nmo = nocc + nvir
    if symmetric:
        mn = O.assemble_mnP(B, miP, abP, nocc, nvir, naux)
    else:                                    # the blocks as given, no symmetrisation
        mn = np.zeros((nmo, nmo, naux))
        mn[:, :nocc] = miP.reshape(nmo, nocc, naux)
        mn[:nocc, nocc:] = B.reshape(nocc, nvir, naux)
        mn[nocc:, nocc:] = abP.reshape(nvir, nvir, naux)
    lpq = np.ascontiguousarray(mn.transpose(2, 0, 1))
    fr, wt = O.get_scaled_legendre_roots(nf, 0.5)
    p0 = _p0(e, B, nocc, fr)
    got = core.compute_sigma_c_pyscf_cd(e, 0.01, p0, lpq, fr, wt)
    ref = O.sigma_c_pyscf_cd(e, 0.01, p0, lpq, fr, wt)
    err = float(np.max(np.abs(got - ref)))
    print(f"static CD symmetric={symmetric}: max abs err {err:.3e}, |sigma| max "
          f"{np.max(np.abs(ref)):.3e}")
    assert got.shape == (nmo,) and err < 1e-10
    assert np.all(core.compute_sigma_c_pyscf_cd(e, 0.01, np.zeros_like(p0), lpq, fr, wt) == 0)


def test_sigma_c_pyscf_cd_errors(gpu):
    from quasix_amd import core
    nocc, nvir, naux, nf = 2, 3, 8, 4
    e, B = O.make_inputs(nocc, nvir, naux, seed=4)
    miP, abP = O.ov_only_blocks(B, nocc, nvir)
    lpq = np.ascontiguousarray(O.assemble_mnP(B, miP, abP, nocc, nvir, naux).transpose(2, 0, 1))
    fr, wt = O.get_scaled_legendre_roots(nf, 0.5)
    p0 = _p0(e, B, nocc, fr)
    with pytest.raises(RuntimeError, match="Dimension mismatch"):
        core.compute_sigma_c_pyscf_cd(e, 0.01, p0, lpq, fr, wt[:3])
    with pytest.raises(RuntimeError, match="Dimension mismatch"):
        core.compute_sigma_c_pyscf_cd(e[:4], 0.01, p0, lpq, fr, wt)
    p0[2] = np.eye(naux)                     # eps = I - P0 = 0 at one frequency
    with pytest.raises(RuntimeError, match="singular at frequency 2"):
        core.compute_sigma_c_pyscf_cd(e, 0.01, p0, lpq, fr, wt)


def test_sigma_cd_dev_at_zero_is_static_cd(gpu):
    """qx_sigma_cd_dev (I + R) at E = 0 for every orbital equals the reference's static CD
    self-energy when no residue pole lies between the Fermi level and 0."""
    from quasix_amd import core
    nocc, nvir, naux, nf = 5, 19, 84, 32
    e, B = O.make_inputs(nocc, nvir, naux, seed=22)
    nmo = nocc + nvir
    ef = 0.5 * (e[nocc - 1] + e[nocc])
    assert not np.any((np.minimum(ef, 0) < e) & (e < np.maximum(ef, 0)))
    tg = np.arange(nmo)
    got = core.compute_sigma_c_cd(e, B, nocc, tg, np.zeros((nmo, 1)), nfreq=nf)[:, 0]
    miP, abP = O.ov_only_blocks(B, nocc, nvir)
    lpq = np.ascontiguousarray(O.assemble_mnP(B, miP, abP, nocc, nvir, naux).transpose(2, 0, 1))
    fr, wt = O.get_scaled_legendre_roots(nf, 0.5)
    ref = O.sigma_c_pyscf_cd(e, 0.01, _p0(e, B, nocc, fr), lpq, fr, wt)
    err = float(np.max(np.abs(got - ref)))
    print(f"sigma_cd_dev(E=0) vs static CD: {err:.3e}")
    assert err < 1e-10


@pytest.mark.parametrize("naux,given_root", [(84, False), (130, False), (130, True)])
def test_screening_w_with_metric(gpu, naux, given_root):
    from quasix_amd import core
    nocc, nvir, nf = 4, 20, 6
    e, B = O.make_inputs(nocc, nvir, naux, seed=naux)
    fr, _ = O.get_scaled_legendre_roots(nf, 0.5)
    p0 = _p0(e, B, nocc, fr)
    v = _metric(naux, seed=naux + 1)
    vs = O.metric_sqrt(v)
    got = core.compute_screening_w(p0, v, fr, vs if given_root else None)
    ref = O.compute_screening_w(p0, v, vs if given_root else None)
    err = float(np.max(np.abs(got - ref)) / np.max(np.abs(ref)))
    print(f"W with metric naux={naux} root={given_root}: max rel err {err:.3e}")
    assert err < 1e-9
    assert np.array_equal(got, got.transpose(0, 2, 1))
    # identity metric: W = (I - P0)^-1
    wi = core.compute_screening_w(p0, np.eye(naux), fr)
    inv = np.stack([np.linalg.inv(np.eye(naux) - 0.5 * (p + p.T)) for p in p0])
    assert np.max(np.abs(wi - inv)) < 1e-10


def test_screening_w_errors(gpu):
    from quasix_amd import core
    naux, nf = 16, 3
    p0 = np.zeros((nf, naux, naux))
    with pytest.raises(RuntimeError, match="W\\(iω\\) computation error"):
        core.compute_screening_w(p0, np.eye(naux + 1), np.ones(nf))
    v = _metric(naux, 5)
    v[0, 0] = -1.0                           # indefinite metric: no real square root
    with pytest.raises(RuntimeError, match="W\\(iω\\) computation error"):
        core.compute_screening_w(p0, v, np.ones(nf))
    p0[1] = np.eye(naux)                     # with v_sqrt = I exactly: eps = 0 at frequency 1
    with pytest.raises(RuntimeError, match="singular at frequency 1"):
        core.compute_screening_w(p0, np.eye(naux), np.ones(nf), np.eye(naux))


def test_screened_wc_identity_fallback(gpu):
    """An exactly singular eps at one frequency: Cholesky fails there, every frequency is
    inverted by Gauss-Jordan, the singular one gives Wc = 0 (the identity fallback) and the
    others match the oracle's chain."""
    from quasix_amd import core
    nocc, nvir, naux, nf = 4, 20, 70, 5
    e, B = O.make_inputs(nocc, nvir, naux, seed=31)
    fr, _ = O.get_scaled_legendre_roots(nf, 0.5)
    p0 = _p0(e, B, nocc, fr)
    d = np.linspace(0.1, 0.6, naux)
    d[0] = 1.0                               # eps[0][0] = 0 exactly
    p0[3] = np.diag(d)
    wc, nfb = core.compute_screening_wc(p0)
    ref = np.stack([O.screened_wc(p) for p in p0])
    assert nfb >= 1
    assert np.all(wc[3] == 0.0) and np.all(ref[3] == 0.0)
    err = float(np.max(np.abs(wc - ref)))
    print(f"identity fallback: Cholesky failures {nfb}, max abs err {err:.3e}")
    assert err < 1e-10


def test_host_pointer_g0w0_cd(gpu):
    """qx_g0w0_cd with host arrays (the entry the Rust extern block binds): same numbers as
    qx_g0w0_cd_dev and the oracle."""
    import quasix_amd._lib as L
    from quasix_amd import gw
    nocc, nvir, naux, nf = 5, 19, 84, 32
    e, B = O.make_inputs(nocc, nvir, naux, seed=23)
    tg = np.arange(nocc - 3, nocc + 3, dtype=np.int32)
    miP, abP = O.ov_only_blocks(B, nocc, nvir)
    pt, pm, Bp = O.pair_list(O.assemble_mnP(B, miP, abP, nocc, nvir, naux), tg)
    static = np.linspace(-0.05, 0.05, len(tg))
    h = L.handle(0)
    d = gw.make_desc(nocc, nvir, naux, nf, ntarget=len(tg), npair=len(pt))
    arrs = [np.ascontiguousarray(a) for a in (e, B, Bp, pt, pm, tg, static)]
    out = [np.empty(len(tg)) for _ in range(4)]
    st = np.empty(len(tg), np.int32)
    s = L.Stats()
    vp = lambda a: a.ctypes.data_as(C.c_void_p)
    h.check(h.lib.qx_g0w0_cd(h.h, C.byref(d), vp(arrs[0]), None, vp(arrs[1]), vp(arrs[2]),
                             vp(arrs[3]), vp(arrs[4]), vp(arrs[5]), vp(arrs[6]), None,
                             *map(vp, out), vp(st), C.byref(s)))
    pairs = gw.Pairs(*[gw._i32(pt, gw._dev(0)), gw._i32(pm, gw._dev(0)),
                       gw._f64(Bp, gw._dev(0))])
    dev = gw.g0w0(e, B, nocc, tg, nfreq=nf, static=static, pairs=pairs)
    assert np.array_equal(out[0], dev["qp_energies"]) and np.array_equal(out[1], dev["z_factors"])
    fr, wt = O.get_scaled_legendre_roots(nf, 0.5)
    ref = O.g0w0_cd(e, B, nocc, tg, fr, wt, pairs=(pt, pm, Bp), static=static)
    err = float(np.max(np.abs(out[0] - ref["E"])))
    print(f"host-pointer qx_g0w0_cd: max |dE| vs oracle {err:.3e}, states {st.tolist()}")
    assert err < 1e-6
    assert np.max(np.abs(out[1] - ref["Z"])) < 1e-6
