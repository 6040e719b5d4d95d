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


def compute_screening_w(p0, coulomb_matrix, freqs, v_sqrt=None, *, device=0):
    """quasix_core.compute_screening_w (python/g0w0.rs:102-130 -> screening.rs:164-420).

    W_f = vs (I - vs P0_f vs^T)^-1 vs^T (eps and W symmetrised), with vs = v_sqrt when given,
    else the square root of the Coulomb metric ``coulomb_matrix``; all of it on the GPU
    (qx_screening_w_dev).  ``freqs`` is only checked for length, as the reference ignores
    it.  Errors raise RuntimeError("W(iω) computation error: ...") as the PyO3 layer does."""
    dev = _dev(device)
    h = _lib.handle(device)
    p = _f64(p0, dev)
    try:
        if p.dim() != 3 or p.shape[1] != p.shape[2]:
            raise ValueError(f"P0 must be [nfreq, naux, naux], got {tuple(p.shape)}")
        nf, naux, _ = p.shape
        v = np.asarray(coulomb_matrix.cpu() if torch.is_tensor(coulomb_matrix)
                       else coulomb_matrix, dtype=np.float64)
        if v.shape != (naux, naux):
            raise ValueError(f"Dimension mismatch: v_aux must be {naux}x{naux}, got {v.shape}")
        vs = None
        if v_sqrt is not None:
            vs = _f64(v_sqrt, dev)
            if tuple(vs.shape) != (naux, naux):
                raise ValueError(f"Dimension mismatch: v_sqrt must be {naux}x{naux}, got "
                                 f"{tuple(vs.shape)}")
    except ValueError as e:
        raise RuntimeError(f"W(iω) computation error: {e}") from e
    vd = _f64(v, dev)
    out = torch.empty((nf, naux, naux), dtype=torch.float64, device=dev)
    torch.cuda.synchronize(dev)
    try:
        h.check(h.lib.qx_screening_w_dev(h.h, naux, nf, _ptr(p), _ptr(vd), _ptr(vs), _ptr(out)))
    except _lib.QxError as e:
        raise RuntimeError(f"W(iω) computation error: {e}") from e
    return out.cpu().numpy()


def compute_sigma_c_pyscf_cd(mo_energy, eta, p0_iw, lpq, freqs, weights, *, device=0):
    """quasix_core.compute_sigma_c_pyscf_cd (python/g0w0.rs:597-631 ->
    correlation_pyscf_cd.rs:60-298): the static CD self-energy Sigma_c_n(E = 0) of every
    orbital, without residues, from a given P0 [nfreq, naux, naux] and the full DF tensor
    lpq [naux, nmo, nmo], on the GPU (qx_sigma_c_pyscf_cd_dev).  Returns complex [nmo].
    Shape errors are the reference's DimensionMismatch cases, raised as RuntimeError."""
    dev = _dev(device)
    h = _lib.handle(device)
    e = _f64(mo_energy, dev)
    p = _f64(p0_iw, dev)
    L = _f64(lpq, dev)
    fr = _f64(freqs, dev)
    wt = _f64(weights, dev)
    nmo = e.numel()
    msg = None
    if p.dim() != 3 or p.shape[1] != p.shape[2]:
        msg = f"P₀ not square: {tuple(p.shape)}"
    elif L.dim() != 3 or L.shape[0] != p.shape[1]:
        msg = f"Lpq naux {L.shape[0] if L.dim() == 3 else '?'} != P₀ naux {p.shape[1]}"
    elif L.shape[1] != nmo or L.shape[2] != nmo:
        msg = f"Lpq MO dimensions ({L.shape[1]}, {L.shape[2]}) != nmo {nmo}"
    elif wt.numel() != p.shape[0]:
        msg = f"Weights length {wt.numel()} != nfreq {p.shape[0]}"
    elif fr.numel() != p.shape[0]:
        msg = f"Frequencies length {fr.numel()} != nfreq {p.shape[0]}"
    if msg:
        raise RuntimeError(f"Correlation self-energy (PySCF CD) error: Dimension mismatch: {msg}")
    out = torch.empty((nmo, 2), dtype=torch.float64, device=dev)
    torch.cuda.synchronize(dev)
    try:
        h.check(h.lib.qx_sigma_c_pyscf_cd_dev(h.h, nmo, p.shape[1], p.shape[0], float(eta),
                                              _ptr(e), _ptr(p), _ptr(L), _ptr(fr), _ptr(wt),
                                              _ptr(out)))
    except _lib.QxError as err:
        raise RuntimeError(f"Correlation self-energy (PySCF CD) error: {err}") from err
    o = out.cpu().numpy()
    return o[:, 0] + 1j * o[:, 1]


def solve_qp_equation(mo_energy, sigma_x, sigma_c, vxc):
    """quasix_core.solve_qp_equation (python/g0w0.rs:423-450 -> qp/solver.rs:137-221):
    E = eps + Sigma_x + Sigma_c - V_xc with Z = 1, host code in libqxgw.  Returns
    {'qp_energies', 'z_factors', 'converged': True}; a length mismatch or a non-finite
    energy raises RuntimeError("QP solver error: ...")."""
    arrs = [np.ascontiguousarray(np.asarray(a.cpu() if torch.is_tensor(a) else a,
                                            dtype=np.float64)).reshape(-1)
            for a in (mo_energy, sigma_x, sigma_c, vxc)]
    n = arrs[0].size
    for name, a in zip(("sigma_x", "sigma_c", "vxc"), arrs[1:]):
        if a.size != n:
            raise RuntimeError(f"QP solver error: Invalid input: {name} dimension mismatch: "
                               f"expected {n}, got {a.size}")
    qp, z = np.empty(n), np.empty(n)
    rc = _lib.load().qx_solve_qp_linearized(n, *(a.ctypes.data_as(C.c_void_p) for a in arrs),
                                            qp.ctypes.data_as(C.c_void_p),
                                            z.ctypes.data_as(C.c_void_p))
    if rc != _lib.QX_OK:
        bad = int(np.flatnonzero(~np.isfinite(qp))[0]) if rc == 3 else -1
        raise RuntimeError(f"QP solver error: Numerical error: QP energy[{bad}] is not finite"
                           if rc == 3 else f"QP solver error: {_lib.STATUS_NAMES.get(rc, rc)}")
    return {"qp_energies": qp, "z_factors": z, "converged": True}


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
