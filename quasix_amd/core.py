"""Per-stage functions of the ``quasix_core`` Python module, on the MI355X path.

Mirrors quasix_core/src/python/g0w0.rs:721-751 and quasix_core/src/python/mod.rs:29-193
(same names, argument meaning, return shapes; errors raise RuntimeError like the PyO3
layer's mapping).  Arrays may be numpy arrays or torch tensors; results are numpy arrays.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from .gw import _dev, _f64, _i32, _ptr, make_desc, pairs_from_blocks


def _run(h, rc):
    try:
        h.check(rc)
    except _lib.QxError as e:
        raise RuntimeError(str(e)) from e


def gauss_legendre_grid(nfreq: int) -> dict:
    """quasix_core.gauss_legendre_grid (python/mod.rs:29): {'nodes', 'weights'} on [-1, 1]."""
    lib = _lib.load()
    x = np.empty(nfreq)
    w = np.empty(nfreq)
    if lib.qx_gauss_legendre_grid(int(nfreq), x.ctypes.data_as(C.c_void_p),
                                  w.ctypes.data_as(C.c_void_p)) != 0:
        raise RuntimeError(f"GL grid error: invalid nfreq={nfreq}")
    return {"nodes": x, "weights": w}


def get_scaled_legendre_roots(nw: int, x0: float = 0.5):
    """correlation_ac::get_scaled_legendre_roots (correlation_ac.rs:94-114)."""
    lib = _lib.load()
    f = np.empty(nw)
    w = np.empty(nw)
    if lib.qx_scaled_legendre_roots(int(nw), float(x0), f.ctypes.data_as(C.c_void_p),
                                    w.ctypes.data_as(C.c_void_p)) != 0:
        raise RuntimeError(f"invalid nw={nw}")
    return f, w


def transform_to_imaginary_axis_pyscf(nodes, x0: float = 0.5):
    """freq/imaginary_grid.rs:130-145: x0 (1+t)/(1-t), 1e12 at t -> 1."""
    t = np.asarray(nodes, dtype=np.float64)
    out = np.where(np.abs(1.0 - t) < 1e-14, 1e12, x0 * (1.0 + t) / np.where(
        np.abs(1.0 - t) < 1e-14, 1.0, 1.0 - t))
    return out


def transform_weights_pyscf(weights, nodes, x0: float = 0.5):
    """freq/imaginary_grid.rs:196-213: w 2 x0 / (1-t)^2, 0 at t -> 1."""
    t = np.asarray(nodes, dtype=np.float64)
    w = np.asarray(weights, dtype=np.float64)
    d = np.where(np.abs(1.0 - t) < 1e-14, 1.0, 1.0 - t)
    return np.where(np.abs(1.0 - t) < 1e-14, 0.0, w * 2.0 * x0 / (d * d))


def compute_polarizability_p0(ia_p, mo_energy, mo_occ, freqs, *, precision="fp64", device=0):
    """quasix_core.compute_polarizability_p0 (python/g0w0.rs:50 -> polarizability.rs:169).

    ia_p [nocc, nvir, naux] -> P0 [nf, naux, naux] (real, symmetric)."""
    dev = _dev(device)
    h = _lib.handle(device)
    ia = _f64(ia_p, dev)
    if ia.dim() != 3:
        raise RuntimeError("ia_p must be [nocc, nvir, naux]")
    nocc, nvir, naux = ia.shape
    occ = np.asarray(mo_occ)
    if int(np.count_nonzero(occ > 0.5)) != nocc or int(np.count_nonzero(occ < 0.5)) != nvir:
        raise RuntimeError("Occupation mismatch")
    e = np.asarray(mo_energy, dtype=np.float64)
    order = np.concatenate([np.nonzero(occ > 0.5)[0], np.nonzero(occ < 0.5)[0]])
    e_dev = _f64(e[order], dev)
    fr = _f64(freqs, dev)
    out = torch.empty((fr.numel(), naux, naux), dtype=torch.float64, device=dev)
    torch.cuda.synchronize(dev)
    _run(h, h.lib.qx_rho_response_dev(h.h, _lib.PREC[precision], nocc, nvir, naux, _ptr(e_dev),
                                      _ptr(ia.reshape(nocc * nvir, naux)), fr.numel(), _ptr(fr),
                                      _ptr(out)))
    return out.cpu().numpy()


def compute_screening_wc(p0, *, device=0):
    """Wc = (I - P0)^-1 - I per frequency (correlation_ac.rs:376-401, metric = I).

    Returns (wc [nf, naux, naux], n_fallback_frequencies)."""
    dev = _dev(device)
    h = _lib.handle(device)
    p = _f64(p0, dev)
    nf, naux, _ = p.shape
    out = torch.empty_like(p)
    nfb = C.c_int(0)
    torch.cuda.synchronize(dev)
    _run(h, h.lib.qx_screened_wc_dev(h.h, naux, nf, _ptr(p), _ptr(out), C.byref(nfb)))
    return out.cpu().numpy(), nfb.value


def compute_screening_w(p0, coulomb_matrix=None, freqs=None, v_sqrt=None, *, device=0):
    """quasix_core.compute_screening_w (python/g0w0.rs:103 -> screening.rs:334-420).

    With the Cholesky-normalised DF tensors of the live path the metric is the identity and
    W = eps^-1 = Wc + I.  A non-identity metric is outside this path and raises."""
    for m in (coulomb_matrix, v_sqrt):
        if m is not None and not np.allclose(np.asarray(m), np.eye(np.asarray(m).shape[0])):
            raise RuntimeError("only the identity metric (Cholesky-normalised B) is supported")
    wc, _ = compute_screening_wc(p0, device=device)
    return wc + np.eye(wc.shape[1])[None]


def compute_wmn_imag(mo_energy, B_ov, nocc, B_pair, freqs, *, precision="fp64", device=0):
    """Wc_nm(i w) = b_nm^T Wc(i w) b_nm for explicit pair rows (correlation_ac.rs:404-421)."""
    dev = _dev(device)
    h = _lib.handle(device)
    e = _f64(mo_energy, dev)
    B = _f64(B_ov, dev)
    bp = _f64(B_pair, dev)
    fr = _f64(freqs, dev)
    nvir = e.numel() - nocc
    out = torch.empty((fr.numel(), bp.shape[0]), dtype=torch.float64, device=dev)
    torch.cuda.synchronize(dev)
    _run(h, h.lib.qx_wmn_imag_dev(h.h, _lib.PREC[precision], nocc, nvir, B.shape[1], _ptr(e),
                                  _ptr(B), fr.numel(), _ptr(fr), bp.shape[0], _ptr(bp),
                                  _ptr(out)))
    return out.cpu().numpy()


def compute_sigma_c_cd(mo_energy, B_ov, nocc, targets, energies, *, pairs=None, nfreq=48,
                       x0=0.5, eta=0.01, precision="fp64", mo_energy_ref=None, device=0):
    """Contour-deformation Sigma_c_n(E) = I_n(E) + R_n(E) at arbitrary energies.

    Replaces quasix_core.compute_sigma_c_pyscf_cd (correlation_pyscf_cd.rs:60-298, E = 0 only,
    no residues there).  energies [nt, nq] -> complex [nt, nq]."""
    dev = _dev(device)
    h = _lib.handle(device)
    e = _f64(mo_energy, dev)
    B = _f64(B_ov, dev)
    nmo = e.numel()
    nvir = nmo - nocc
    naux = B.shape[1]
    tg = np.asarray(targets, dtype=np.int32)
    if pairs is None:
        from .gw import Pairs
        iaP = B
        miP = torch.zeros((nmo, nocc, naux), dtype=torch.float64, device=dev)
        miP[nocc:] = B.reshape(nocc, nvir, naux).permute(1, 0, 2)
        abP = torch.zeros((nvir * nvir, naux), dtype=torch.float64, device=dev)
        pairs = pairs_from_blocks(iaP, miP, abP, nocc, nvir, naux, tg, device)
    fr, wt = get_scaled_legendre_roots(nfreq, x0)
    wnm = torch.as_tensor(compute_wmn_imag(e, B, nocc, pairs.rows, fr, precision=precision,
                                           device=device), device=dev)
    E = _f64(np.asarray(energies, dtype=np.float64).reshape(len(tg), -1), dev)
    nq = E.shape[1]
    out = torch.empty((len(tg), nq, 2), dtype=torch.float64, device=dev)
    d = make_desc(nocc, nvir, naux, nfreq, x0, eta, precision, len(tg), pairs.rows.shape[0])
    er = None if mo_energy_ref is None else _f64(mo_energy_ref, dev)
    t_dev = _i32(tg, dev)
    fr_d, wt_d = _f64(fr, dev), _f64(wt, dev)
    torch.cuda.synchronize(dev)
    _run(h, h.lib.qx_sigma_cd_dev(h.h, C.byref(d), _ptr(e), _ptr(er), _ptr(B), _ptr(pairs.rows),
                                  _ptr(pairs.target), _ptr(pairs.m), _ptr(t_dev), _ptr(fr_d),
                                  _ptr(wt_d), _ptr(wnm), nq, _ptr(E), _ptr(out)))
    o = out.cpu().numpy()
    return o[..., 0] + 1j * o[..., 1]


# ---- Analytic continuation (the reference's live self-energy path) --------------------------

def _cplx(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.complex128))


def _vp(a):
    return a.ctypes.data_as(C.c_void_p)


def pairs_from_lpq(lpq, orbs):
    """(pair target, partner m, rows [npair, naux]) from a symmetric [naux, nmo, nmo] tensor:
    the rows lpq[:, n, m] of every listed orbital n that are not identically zero."""
    lpq = np.asarray(lpq, dtype=np.float64)
    pt, pm, rows = [], [], []
    for t, n in enumerate(orbs):
        blk = lpq[:, n, :].T                          # [nmo, naux]
        nz = np.flatnonzero(np.any(blk != 0.0, axis=1))
        pt += [t] * len(nz)
        pm += list(nz)
        rows.append(blk[nz])
    rows = np.concatenate(rows) if rows else np.zeros((0, lpq.shape[0]))
    return np.asarray(pt, np.int32), np.asarray(pm, np.int32), np.ascontiguousarray(rows)


def compute_sigma_c_ac(mo_energy, lpq, nocc, orbs=None, nw=100, x0=0.5, iw_cutoff=5.0, *,
                       precision="fp64", device=0):
    """quasix_core.compute_sigma_c_ac (python/g0w0.rs:721-751, correlation_ac.rs:241-449):
    Sigma_c(i nu) on the AC grid for ``orbs`` (default all) from the full symmetric DF tensor
    lpq [naux, nmo, nmo].  Returns {sigma_c [norbs, nw_sigma], omega [norbs, nw_sigma],
    freqs, wts, orbs, nocc}; Wc, the quadrature and the contraction run on the GPU."""
    dev = _dev(device)
    h = _lib.handle(device)
    lpq = np.asarray(lpq, dtype=np.float64)
    naux, nmo, _ = lpq.shape
    if nocc <= 0 or nocc >= nmo:
        raise RuntimeError(f"Invalid nocc={nocc}")
    orbs = np.arange(nmo, dtype=np.int32) if orbs is None else np.asarray(orbs, np.int32)
    nvir = nmo - nocc
    B = _f64(np.ascontiguousarray(lpq[:, :nocc, nocc:].transpose(1, 2, 0)).reshape(-1, naux), dev)
    e = _f64(mo_energy, dev)
    pt, pm, rows = pairs_from_lpq(lpq, orbs)
    d = make_desc(nocc, nvir, naux, nw, x0, 0.01, precision, len(orbs), len(pt))
    sig = torch.empty((len(orbs), nw + 1, 2), dtype=torch.float64, device=dev)
    om = torch.empty_like(sig)
    nws = C.c_int()
    args = [_f64(rows, dev), _i32(pt, dev), _i32(pm, dev), _i32(orbs, dev)]
    torch.cuda.synchronize(dev)
    _run(h, h.lib.qx_sigma_ac_dev(h.h, C.byref(d), _ptr(e), _ptr(B), *map(_ptr, args),
                                  float(iw_cutoff), _ptr(sig), _ptr(om), C.byref(nws)))
    n = nws.value
    s = sig.reshape(-1)[: len(orbs) * n * 2].cpu().numpy().reshape(len(orbs), n, 2)
    o = om.reshape(-1)[: len(orbs) * n * 2].cpu().numpy().reshape(len(orbs), n, 2)
    fr, wt = get_scaled_legendre_roots(nw, x0)
    return dict(sigma_c=s[..., 0] + 1j * s[..., 1], omega=o[..., 0] + 1j * o[..., 1],
                freqs=fr, wts=wt, orbs=orbs, nocc=nocc)


def thiele(fn, zn):
    """correlation_ac::thiele (correlation_ac.rs:470-509), host code in libqxgw."""
    fn, zn = _cplx(fn), _cplx(zn)
    out = np.empty(len(fn), np.complex128)
    if _lib.load().qx_thiele(len(fn), _vp(fn), _vp(zn), _vp(out)) != _lib.QX_OK:
        raise RuntimeError("qx_thiele failed")
    return out


def pade_thiele(omega, zn, coeff):
    """correlation_ac::pade_thiele (correlation_ac.rs:530-566) at a real omega."""
    zn, coeff = _cplx(zn), _cplx(coeff)
    out = np.empty(1, np.complex128)
    if _lib.load().qx_pade_thiele(len(coeff), float(omega), _vp(zn), _vp(coeff),
                                  _vp(out)) != _lib.QX_OK:
        raise RuntimeError("qx_pade_thiele failed")
    return complex(out[0])


def ac_pade_thiele_diag(sigma, omega):
    """correlation_ac::ac_pade_thiele_diag (correlation_ac.rs:587-654):
    (coeff [nfit, norbs], omega_fit [norbs, nfit])."""
    sigma, omega = _cplx(sigma), _cplx(omega)
    norbs, nw = sigma.shape
    coeff = np.empty((nw, norbs), np.complex128)
    ofit = np.empty((norbs, nw), np.complex128)
    nfit = C.c_int()
    if _lib.load().qx_ac_pade_thiele_diag(norbs, nw, _vp(sigma), _vp(omega), _vp(coeff),
                                          _vp(ofit), C.byref(nfit)) != _lib.QX_OK:
        raise RuntimeError("qx_ac_pade_thiele_diag failed")
    n = nfit.value
    return (np.ascontiguousarray(coeff.reshape(-1)[: n * norbs].reshape(n, norbs)),
            np.ascontiguousarray(ofit.reshape(-1)[: norbs * n].reshape(norbs, n)))


def qp_solve_ac(coeff, omega_fit, orbs, mo_energy, nocc, sigma_x, vxc, *, solver="linearized",
                prev_qp=None, newton=None):
    """QP step of the evGW driver on the Pade model (quasix_core/src/qp/evgw.rs:705-820)."""
    coeff, ofit = _cplx(coeff), _cplx(omega_fit)
    orbs = np.ascontiguousarray(orbs, np.int32)
    e = np.ascontiguousarray(mo_energy, np.float64)
    sx = np.ascontiguousarray(sigma_x, np.float64)
    vx = np.ascontiguousarray(vxc, np.float64)
    prev = None if prev_qp is None else np.ascontiguousarray(prev_qp, np.float64)
    cfg = _lib.newton_config(**(newton or {}))
    n = len(orbs)
    E, Z, sr, si = (np.empty(n) for _ in range(4))
    st = np.empty(n, np.int32)
    rc = _lib.load().qx_qp_solve_ac(
        n, coeff.shape[0], _vp(coeff), _vp(ofit), _vp(orbs), e.size, int(nocc), _vp(e), _vp(sx),
        _vp(vx), {"linearized": 0, "newton": 1}[solver],
        None if prev is None else _vp(prev), C.byref(cfg), _vp(E), _vp(Z), _vp(sr), _vp(si),
        _vp(st))
    if rc != _lib.QX_OK:
        raise RuntimeError(f"qx_qp_solve_ac failed ({_lib.STATUS_NAMES.get(rc, rc)})")
    return dict(qp_energies=E, z_factors=Z, sigma_c=sr + 1j * si, state=st)


def compute_exchange_diag(mi, nmo, nocc, *, device=0):
    """Diagonal exchange self-energy from the [nmo, nocc, naux] block (the reference's ijP /
    miP): sigma_x[p] = -sum_i sum_R mi[p, i, R]^2 (EvGWDriver::compute_exchange,
    quasix_core/src/qp/evgw.rs:1167-1193, whose diagonal the driver uses)."""
    dev = _dev(device)
    h = _lib.handle(device)
    m = _f64(mi, dev).reshape(nmo, nocc, -1)
    out = torch.empty(nmo, dtype=torch.float64, device=dev)
    torch.cuda.synchronize(dev)
    _run(h, h.lib.qx_exchange_diag_dev(h.h, int(nmo), int(nocc), int(m.shape[2]), _ptr(m),
                                       _ptr(out)))
    return out.cpu().numpy()


def transform_mo_3center(j3c_ao, c_occ, c_vir, *, device=0):
    """df::mo_transform::transform_mo_3center (quasix_core/src/df/mo_transform.rs:136-260):
    (ia|P) = sum_{mu nu} C_occ[mu, i] (mu nu|P) C_vir[nu, a] for Cholesky-normalised
    3-centre vectors j3c_ao [nao, nao, naux]; returns [nocc * nvir, naux] (the iaP layout),
    computed with FP64 MFMA GEMMs on the GPU."""
    dev = _dev(device)
    h = _lib.handle(device)
    J = _f64(j3c_ao, dev)
    co = _f64(c_occ, dev)
    cv = _f64(c_vir, dev)
    nao, _, naux = J.shape
    if co.shape[0] != nao or cv.shape[0] != nao or J.shape[1] != nao:
        raise RuntimeError("MO coefficient dimensions don't match the AO basis")
    nocc, nvir = co.shape[1], cv.shape[1]
    out = torch.empty((nocc * nvir, naux), dtype=torch.float64, device=dev)
    torch.cuda.synchronize(dev)
    _run(h, h.lib.qx_df_mo_transform_dev(h.h, nao, nocc, nvir, naux, _ptr(J), _ptr(co),
                                         _ptr(cv), _ptr(out)))
    return out.cpu().numpy()
