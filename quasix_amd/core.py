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
