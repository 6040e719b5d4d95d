"""G0W0 / evGW drivers on the MI355X path (the ``quasix.gw`` / ``quasix.quasix.evgw`` surface).

Every number comes from libqxgw.so (HIP kernels behind the C ABI in include/qxgw.h).
torch provides device buffers and the stream; no arithmetic of the path runs in Python.

Entry points
  * ``g0w0(...)``         one G0W0 solve (Pi -> Wc -> CD self-energy -> Newton) on a device
                          tensor set; the MI355X counterpart of the iteration-0 branch of
                          quasix_core/src/qp/evgw.rs:603-783.
  * ``evgw(...)``         eigenvalue self-consistency, same arguments, defaults and result
                          keys as quasix.quasix.evgw (quasix/src/evgw.rs:22-453).
  * ``run_evgw(config)``  dict front-end of quasix.quasix.run_evgw (quasix/src/evgw.rs:456-677).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib

HARTREE_TO_EV = 27.211386245988


def _dev(device):
    if not torch.cuda.is_available():
        raise RuntimeError("quasix_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
    return torch.device("cuda", device)


def _ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


def _f64(x, dev):
    return torch.as_tensor(np.ascontiguousarray(x, dtype=np.float64) if not torch.is_tensor(x)
                           else x, dtype=torch.float64, device=dev).contiguous()


def _i32(x, dev):
    return torch.as_tensor(np.ascontiguousarray(x, dtype=np.int32) if not torch.is_tensor(x)
                           else x, dtype=torch.int32, device=dev).contiguous()


def frontier_targets(nocc: int, nmo: int, below: int = 5, above: int = 5) -> np.ndarray:
    """HOMO-below .. LUMO+above, clipped to the orbital range."""
    return np.arange(max(0, nocc - 1 - below), min(nmo, nocc + above + 1), dtype=np.int32)


def make_desc(nocc, nvir, naux, nfreq=48, x0=0.5, eta=0.01, precision="fp64", ntarget=1,
              npair=0, **newton) -> _lib.G0W0Desc:
    d = _lib.G0W0Desc()
    _lib.load().qx_g0w0_desc_default(C.byref(d))
    d.nocc, d.nvir, d.naux = int(nocc), int(nvir), int(naux)
    d.nfreq, d.x0, d.eta = int(nfreq), float(x0), float(eta)
    d.precision = _lib.PREC[precision] if isinstance(precision, str) else int(precision)
    d.ntarget, d.npair = int(ntarget), int(npair)
    for k, v in newton.items():
        setattr(d.newton, k, v)
    return d


@dataclass
class Pairs:
    """Target-partner rows b_nm (symmetrised full tensor, zero rows dropped), on device."""
    target: torch.Tensor     # int32 [npair], sorted
    m: torch.Tensor          # int32 [npair]
    rows: torch.Tensor       # f64 [npair, naux]


def assemble_blocks(iaP, miP, abP, nocc, nvir, naux, targets=None, *, ov=True, mi=True,
                    device=0):
    """The symmetrised DF tensor of quasix/src/evgw.rs:161-205 (X[:, :nocc] = miP,
    X[:nocc, nocc:] = iaP, X[nocc:, nocc:] = abP, then (X[n, m] + X[m, n]) / 2), assembled on
    the device by qx_assemble_blocks_dev without forming the [nmo, nmo, naux] tensor.
    Returns (B_ov_sym [nocc*nvir, naux] or None, mi_sym [nmo, nocc, naux] or None,
    Pairs of the targets or None): the polarizability's operand, the exchange term's block
    (evgw.rs:305) and the target-partner rows, identically zero rows dropped, in (target,
    partner) order."""
    dev = _dev(device)
    h = _lib.handle(device)
    nmo = nocc + nvir
    ia = _f64(iaP, dev).reshape(nocc * nvir, naux)
    m = _f64(miP, dev).reshape(nmo * nocc, naux)
    ab = None if abP is None else _f64(abP, dev).reshape(nvir * nvir, naux)
    ov_out = torch.empty((nocc * nvir, naux), dtype=torch.float64, device=dev) if ov else None
    mi_out = torch.empty((nmo, nocc, naux), dtype=torch.float64, device=dev) if mi else None
    tg = None if targets is None else np.asarray(targets, dtype=np.int32)
    nt = 0 if tg is None else len(tg)
    cap = nt * nmo
    rows = torch.empty((max(cap, 1), naux), dtype=torch.float64, device=dev)
    pt = torch.empty(max(cap, 1), dtype=torch.int32, device=dev)
    pm = torch.empty(max(cap, 1), dtype=torch.int32, device=dev)
    t_dev = _i32(tg, dev) if nt else None
    npair = C.c_int(0)
    torch.cuda.synchronize(dev)
    h.check(h.lib.qx_assemble_blocks_dev(h.h, nocc, nvir, naux, _ptr(ia), _ptr(m), _ptr(ab), nt,
                                         _ptr(t_dev), _ptr(ov_out), _ptr(mi_out), cap,
                                         _ptr(rows), _ptr(pt), _ptr(pm), C.byref(npair)))
    n = npair.value
    pairs = Pairs(pt[:n], pm[:n], rows[:n]) if nt else None
    return ov_out, mi_out, pairs


def pairs_from_blocks(iaP, miP, abP, nocc, nvir, naux, targets, device=0) -> Pairs:
    """Target-partner rows of the symmetrised tensor (quasix/src/evgw.rs:161-205), rows that
    are identically zero dropped (they contribute exactly zero to Wc_nm)."""
    return assemble_blocks(iaP, miP, abP, nocc, nvir, naux, targets, ov=False, mi=False,
                           device=device)[2]


def symmetrized_ov(iaP, miP, nocc, nvir, naux, device=0):
    """Pi's ov block of the symmetrised tensor: (iaP[i,a] + miP[nocc+a, i]) / 2."""
    return assemble_blocks(iaP, miP, None, nocc, nvir, naux, ov=True, mi=False,
                           device=device)[0]


def parse_qp_solver(qp_solver) -> str:
    """'linearized' or 'newton' from the spellings quasix/src/evgw.rs:215-226 accepts
    (case-insensitive 'linearized' / 'linear', 'newton' / 'newton-raphson' / 'solved');
    anything else raises ValueError as there."""
    s = str(qp_solver).lower()
    if s in ("linearized", "linear"):
        return "linearized"
    if s in ("newton", "newton-raphson", "solved"):
        return "newton"
    raise ValueError(f"Invalid qp_solver '{qp_solver}'. Valid options: 'linearized' (default) "
                     "or 'newton'")


def g0w0(mo_energy, B_ov, nocc, targets=None, *, nfreq=48, x0=0.5, eta=0.01, precision="fp64",
         static=None, pairs: Pairs | None = None, mo_energy_ref=None, guess=None, device=0,
         newton=None, solver="newton", stats=False):
    """One G0W0 solve on the GPU.

    mo_energy [nmo]; B_ov [nocc*nvir, naux] (device tensor or array); targets: orbital indices
    (default HOMO-5..LUMO+5); static: Sigma_x - V_xc per target (default 0); pairs: explicit
    target-partner rows (default: the ov-only pairs gathered on the device); solver "newton"
    (NewtonQPSolver from the guess, else from the linearised value) or "linearized" (the
    linearised step at the guess, else at the mean-field energy).
    Returns dict(qp_energies, z_factors, sigma_c (complex), state, targets[, stats]).
    """
    dev = _dev(device)
    h = _lib.handle(device)
    e = _f64(mo_energy, dev)
    nmo = e.numel()
    B = _f64(B_ov, dev)
    naux = B.shape[1]
    nvir = nmo - nocc
    if B.shape[0] != nocc * nvir:
        raise ValueError(f"B_ov has {B.shape[0]} rows, expected nocc*nvir={nocc * nvir}")
    tg = frontier_targets(nocc, nmo) if targets is None else np.asarray(targets, dtype=np.int32)
    t_dev = _i32(tg, dev)
    nt = len(tg)
    d = make_desc(nocc, nvir, naux, nfreq, x0, eta, precision, nt, 0 if pairs is None
                  else pairs.rows.shape[0], **(newton or {}))
    d.qp_solver = _lib.SOLVER[parse_qp_solver(solver)]
    st = None if static is None else _f64(static, dev)
    er = None if mo_energy_ref is None else _f64(mo_energy_ref, dev)
    gs = None if guess is None else _f64(guess, dev)
    out = torch.empty((4, nt), dtype=torch.float64, device=dev)
    state = torch.empty(nt, dtype=torch.int32, device=dev)
    s = _lib.Stats()
    torch.cuda.synchronize(dev)
    if pairs is None:
        rc = h.lib.qx_g0w0_cd_ov_dev(h.h, C.byref(d), _ptr(e), _ptr(er), _ptr(B), _ptr(t_dev),
                                     _ptr(st), _ptr(gs), _ptr(out[0]), _ptr(out[1]),
                                     _ptr(out[2]), _ptr(out[3]), _ptr(state), C.byref(s))
    else:
        rc = h.lib.qx_g0w0_cd_dev(h.h, C.byref(d), _ptr(e), _ptr(er), _ptr(B), _ptr(pairs.rows),
                                  _ptr(pairs.target), _ptr(pairs.m), _ptr(t_dev), _ptr(st),
                                  _ptr(gs), _ptr(out[0]), _ptr(out[1]), _ptr(out[2]),
                                  _ptr(out[3]), _ptr(state), C.byref(s))
    h.check(rc)
    o = out.cpu().numpy()
    res = dict(qp_energies=o[0], z_factors=o[1], sigma_c=o[2] + 1j * o[3],
               state=state.cpu().numpy(), targets=tg)
    if stats:
        res["stats"] = {k: getattr(s, k) for k, _ in _lib.Stats._fields_}
    return res


def evgw_core(mo_energy, B_ov, nocc, targets=None, *, n_iter=8, damping=0.5, conv_tol=1e-4,
              conv_tol_z=1e-3, diis=False, diis_space=6, nfreq=64, x0=0.5, eta=0.01,
              precision="fp64", static=None, pairs=None, device=0, newton=None,
              solver="newton", frontier_convergence=True, frontier_window=1.0, progress=None):
    """evGW on the CD self-energy: Pi, poles and Green's function from the previous QP
    energies, QP equation anchored at the mean-field energies (quasix_core/src/qp/evgw.rs:
    603-1027 with the AC self-energy replaced by the CD one: iteration 0 direct, iteration 1
    damped, later iterations Pulay-or-damped through ``convergence.evgw_mix``; the QP step
    (``solver`` "newton" or "linearized") starts from the previous QP energy).  Orbitals
    outside ``targets`` keep their mean-field energies.  Stop rule (evgw.rs:929-1027): after
    every iteration, the first included, the largest |dE| over the frontier window of the
    mean-field energies (evgw.rs:1121-1160) is compared with ``conv_tol``.  B stays resident
    on the device across iterations.  ``progress(iteration, max_dE)`` is called after each
    iteration if given."""
    from .convergence import EvGWHistory, evgw_mix, frontier_window as fwin, window_change
    dev = _dev(device)
    e0 = _f64(mo_energy, dev)
    B = _f64(B_ov, dev)
    nmo = e0.numel()
    tg = frontier_targets(nocc, nmo) if targets is None else np.asarray(targets, dtype=np.int32)
    tgt = torch.as_tensor(tg, dtype=torch.long, device=dev)
    e0_h = e0.cpu().numpy()
    win = fwin(e0_h, nocc, frontier_convergence, frontier_window)
    cur = e0_h.copy()
    zc = np.ones(len(tg))
    mon = EvGWHistory(conv_tol, conv_tol_z, diis_space, use_diis=diis)
    history = []
    last = None
    converged = False
    for it in range(n_iter):
        guess = None if it == 0 else cur[tg]
        last = g0w0(cur, B, nocc, tg, nfreq=nfreq, x0=x0, eta=eta, precision=precision,
                    static=static, pairs=pairs, mo_energy_ref=e0, guess=guess, device=device,
                    newton=newton, solver=solver)
        e_old = cur[tg]
        et, zt, active = evgw_mix(it, mon, e_old, zc, np.asarray(last["qp_energies"]),
                                  np.asarray(last["z_factors"]), damping, use_diis=diis)
        new = cur.copy()
        new[tg] = et
        dmax, drms = window_change(new, cur, win)
        zf = np.ones(nmo)
        zf[tg] = zt
        history.append(dict(cycle=it, qp_energies=new.copy(), energy_change=dmax,
                            rms_change=drms, z_factors=zf, diis_active=active))
        cur, zc = new, zt
        if progress is not None:
            progress(it, dmax)
        if dmax < conv_tol:
            converged = True
            break
    return dict(qp_energies=cur, z_factors=zc, converged=converged, n_cycles=len(history),
                history=history, last=last, targets=tg, window=win)


def ov_pairs(B_ov, nocc, nvir, targets, device=0) -> Pairs:
    """Pair rows of an ov-only tensor: every virtual partner of an occupied target (rows
    n*nvir + a) and every occupied partner of a virtual target (rows i*nvir + n - nocc)."""
    dev = _dev(device)
    B = _f64(B_ov, dev)
    ts, ms, idx = [], [], []
    for t, n in enumerate(np.asarray(targets).tolist()):
        if n < nocc:
            r = n * nvir + np.arange(nvir)
            m = nocc + np.arange(nvir)
        else:
            r = np.arange(nocc) * nvir + (n - nocc)
            m = np.arange(nocc)
        ts.append(np.full(len(r), t))
        ms.append(m)
        idx.append(r)
    rows = B[torch.as_tensor(np.concatenate(idx), device=dev)].contiguous()
    return Pairs(_i32(np.concatenate(ts), dev), _i32(np.concatenate(ms), dev), rows)


def sigma_ac(mo_energy, B_ov, nocc, pairs: Pairs, targets, *, nw=100, x0=0.5, iw_cutoff=5.0,
             precision="fp64", device=0):
    """Sigma_c(i nu) of the targets on the AC grid (qx_sigma_ac_dev; correlation_ac.rs:
    241-449).  Returns (sigma [nt, nws], omega [nt, nws]) complex numpy arrays."""
    dev = _dev(device)
    h = _lib.handle(device)
    e = _f64(mo_energy, dev)
    B = _f64(B_ov, dev)
    nmo, naux = e.numel(), B.shape[1]
    tg = np.asarray(targets, dtype=np.int32)
    d = make_desc(nocc, nmo - nocc, naux, nw, x0, 0.01, precision, len(tg), pairs.rows.shape[0])
    sig = torch.empty((len(tg), nw + 1, 2), dtype=torch.float64, device=dev)
    om = torch.empty_like(sig)
    nws = C.c_int()
    t_dev = _i32(tg, dev)
    torch.cuda.synchronize(dev)
    h.check(h.lib.qx_sigma_ac_dev(h.h, C.byref(d), _ptr(e), _ptr(B), _ptr(pairs.rows),
                                  _ptr(pairs.target), _ptr(pairs.m), _ptr(t_dev),
                                  float(iw_cutoff), _ptr(sig), _ptr(om), C.byref(nws)))
    n = nws.value
    sv = sig.reshape(-1)[: len(tg) * n * 2].cpu().numpy().reshape(len(tg), n, 2)
    ov = om.reshape(-1)[: len(tg) * n * 2].cpu().numpy().reshape(len(tg), n, 2)
    return sv[..., 0] + 1j * sv[..., 1], ov[..., 0] + 1j * ov[..., 1]


def evgw_ac_core(mo_energy, B_ov, nocc, targets=None, *, pairs: Pairs | None = None, n_iter=8,
                 damping=0.5, conv_tol=1e-4, conv_tol_z=1e-3, diis=False, diis_space=6, nw=100,
                 x0=0.5, iw_cutoff=5.0, solver="linearized", sigma_x=None, vxc=None,
                 precision="fp64", device=0, newton=None, frontier_convergence=True,
                 frontier_window=1.0):
    """evGW on the AC self-energy, the reference's live driver (quasix_core/src/qp/evgw.rs:
    420-1049): Sigma_c(i nu) from the previous QP energies on the GPU, Thiele-Pade fit and
    the QP step on the host (libqxgw), the damping/Pulay step of ``convergence.evgw_mix``.
    ``targets`` (default all orbitals, as the reference) are updated; ef of the QP step is the
    mean-field one, as there.  Stop rule (evgw.rs:929-1027): after every iteration, the first
    included, the largest |dE| over the frontier window of the mean-field energies
    (evgw.rs:1121-1160) against ``conv_tol``; the history keeps that value, the RMS over the
    window and each iteration's Z."""
    from . import core
    from .convergence import EvGWHistory, evgw_mix, frontier_window as fwin, window_change
    dev = _dev(device)
    e0 = np.ascontiguousarray(np.asarray(mo_energy.cpu() if torch.is_tensor(mo_energy)
                                         else mo_energy, dtype=np.float64))
    B = _f64(B_ov, dev)
    nmo = e0.size
    tg = np.arange(nmo, dtype=np.int32) if targets is None else np.asarray(targets, np.int32)
    if pairs is None:
        pairs = ov_pairs(B, nocc, nmo - nocc, tg, device)
    sx = np.zeros(nmo) if sigma_x is None else np.asarray(sigma_x, dtype=np.float64)
    vx = np.zeros(nmo) if vxc is None else np.asarray(vxc, dtype=np.float64)
    win = fwin(e0, nocc, frontier_convergence, frontier_window)
    cur = e0.copy()
    zc = np.ones(len(tg))
    mon = EvGWHistory(conv_tol, conv_tol_z, diis_space, use_diis=diis)
    history, last, converged = [], None, False
    for it in range(n_iter):
        sig, om = sigma_ac(cur, B, nocc, pairs, tg, nw=nw, x0=x0, iw_cutoff=iw_cutoff,
                           precision=precision, device=device)
        coeff, ofit = core.ac_pade_thiele_diag(sig, om)
        last = core.qp_solve_ac(coeff, ofit, tg, e0, nocc, sx, vx, solver=solver,
                                prev_qp=None if it == 0 else cur, newton=newton)
        e_old = cur[tg]
        et, zt, active = evgw_mix(it, mon, e_old, zc, last["qp_energies"], last["z_factors"],
                                  damping, use_diis=diis)
        new = cur.copy()
        new[tg] = et
        dmax, drms = window_change(new, cur, win)
        zf = np.ones(nmo)
        zf[tg] = zt
        history.append(dict(cycle=it, qp_energies=new.copy(), energy_change=dmax,
                            rms_change=drms, z_factors=zf, diis_active=active))
        cur, zc = new, zt
        if dmax < conv_tol:
            converged = True
            break
    return dict(qp_energies=cur, z_factors=zc, converged=converged, n_cycles=len(history),
                history=history, last=last, targets=tg, window=win)


def evgw(mo_energy, mo_occ, iaP, ijP, abP, chol_v, vxc_dft, max_cycle=12, conv_tol=1e-4,
         conv_tol_z=1e-3, damping=0.5, damping_dynamic=False, diis=False, diis_space=6,
         nfreq=48, eta=0.01, verbose=1, energy_window=2.0, frontier_convergence=True,
         frontier_window=1.0, n_occ_active=None, n_vir_active=None, qp_solver="linearized",
         newton_max_iterations=50, newton_energy_tolerance=1e-8, *, precision="fp64",
         targets=None, sigma_x=None, sigma_method="ac", device=0):
    """quasix.quasix.evgw on the MI355X path (quasix/src/evgw.rs:22-453).

    Same positional arguments, keyword defaults, errors and result keys.  As the reference:
      * the DF tensor is assembled from iaP / ijP (the [nmo, nocc] miP block) / abP and
        symmetrised (evgw.rs:161-205); Pi uses its ov block, the self-energy its rows;
      * Sigma_x is the exchange diagonal of the symmetrised [nmo, nocc] block, computed on
        the GPU (quasix_core/src/qp/evgw.rs:501-502); ``sigma_x`` (per orbital) overrides it;
      * every orbital is updated (``orbs`` = all, quasix_core/src/qp/evgw.rs:572); the
        energy_window / n_occ_active / n_vir_active window is only reported there and so has
        no effect here; ``targets`` restricts the updated orbitals explicitly;
      * the self-energy is the AC one (Sigma(i nu) on nfreq points, Thiele-Pade, iw_cutoff 5)
        and ``qp_solver`` picks the linearised or the Newton QP step;
      * convergence: max |dE| over the frontier window (frontier_convergence,
        frontier_window) below conv_tol after any iteration, the first included.
    ``sigma_method="cd"`` switches to this library's contour-deformation self-energy (no
    reference counterpart on the live path), with the same driver and ``qp_solver``.
    Result keys as evgw.rs:339-452: ip, ea and gap in Hartree; iteration_history entries
    carry cycle (1-based), qp_energies, z_factors (the final Z in every entry, as the
    reference binding reports), z_factors_cycle (that iteration's Z), energy_change and
    rms_change (max and RMS over the window), damping_used (the damping argument) and
    converged.  ``eta`` is used by the CD self-energy only, as it is unused by the
    reference's AC driver; damping_dynamic and verbose are accepted and ignored as there.
    """
    solver = parse_qp_solver(qp_solver)
    e = np.ascontiguousarray(mo_energy, dtype=np.float64)
    occ = np.asarray(mo_occ)
    nocc = int(np.count_nonzero(occ > 0))
    nmo = e.size
    nvir = nmo - nocc
    naux = int(np.asarray(chol_v).shape[0])
    if not 0 < nocc < nmo:
        raise ValueError(f"need occupied and virtual orbitals (nocc={nocc}, nmo={nmo})")

    def as2d(x, rows, name):
        a = np.ascontiguousarray(x.cpu() if torch.is_tensor(x) else x, dtype=np.float64)
        if a.size != rows * naux:
            raise ValueError(f"Failed to reshape {name} tensor: {a.size} elements, expected "
                             f"{rows} x {naux}")
        return a.reshape(rows, naux)

    ia = as2d(iaP, nocc * nvir, "iaP")
    mi = as2d(ijP, nmo * nocc, "miP")
    ab = as2d(abP, nvir * nvir, "abP")
    tg = np.arange(nmo, dtype=np.int32) if targets is None else np.asarray(targets, np.int32)
    vxc = np.asarray(vxc_dft, dtype=np.float64)
    B, mi_sym, pairs = assemble_blocks(ia, mi, ab, nocc, nvir, naux, tg, ov=True,
                                       mi=sigma_x is None, device=device)
    if sigma_x is None:
        from .core import compute_exchange_diag
        sx = compute_exchange_diag(mi_sym, nmo, nocc, device=device)
        del mi_sym
    else:
        sx = np.asarray(sigma_x, dtype=np.float64)
    newton = dict(max_iterations=int(newton_max_iterations),
                  energy_tolerance=float(newton_energy_tolerance))
    common = dict(n_iter=int(max_cycle), damping=damping, conv_tol=conv_tol,
                  conv_tol_z=conv_tol_z, diis=bool(diis), diis_space=int(diis_space),
                  precision=precision, device=device, newton=newton, solver=solver,
                  frontier_convergence=bool(frontier_convergence),
                  frontier_window=float(frontier_window))
    if sigma_method == "ac":
        r = evgw_ac_core(e, B, nocc, tg, pairs=pairs, nw=int(nfreq), sigma_x=sx, vxc=vxc,
                         **common)
    elif sigma_method == "cd":
        r = evgw_core(e, B, nocc, tg, nfreq=int(nfreq), eta=eta, static=(sx - vxc)[tg],
                      pairs=pairs, **common)
    else:
        raise ValueError(f"sigma_method must be 'ac' or 'cd', got {sigma_method!r}")
    qp = r["qp_energies"]
    z = np.ones(nmo)
    z[tg] = r["z_factors"]
    sig = np.zeros(nmo, dtype=np.complex128)
    sig[tg] = r["last"]["sigma_c"]
    n = len(r["history"])
    hist = [dict(cycle=i + 1, qp_energies=h["qp_energies"], z_factors=z.copy(),
                 z_factors_cycle=h["z_factors"], energy_change=h["energy_change"],
                 rms_change=h["rms_change"], damping_used=float(damping),
                 converged=(i == n - 1) and r["converged"])
            for i, h in enumerate(r["history"])]
    homo, lumo = qp[nocc - 1], qp[nocc]
    return dict(qp_energies=qp, z_factors=z, converged=r["converged"], n_cycles=r["n_cycles"],
                final_error=r["history"][-1]["energy_change"] if n else 0.0, sigma_x=sx,
                sigma_c_re=sig.real, sigma_c_im=sig.imag, iteration_history=hist, ip=-homo,
                ea=-lumo, gap=lumo - homo)


_RUN_EVGW_REQUIRED = ("mo_energy", "mo_occ", "iaP", "ijP", "chol_v", "vxc_dft")


def run_evgw(config: dict):
    """Dict front-end with the keys of quasix.quasix.run_evgw (quasix/src/evgw.rs:456-677):
    a missing required key raises KeyError; iaP is the 3-D [nocc, nvir, naux] block; abP
    defaults to zeros; eta (0.01), conv_tol_z (1e-3) and damping_dynamic (off) are fixed;
    diis defaults to on, frontier_window to 0.1 Ha, n_occ_active / n_vir_active to 3; the
    AC self-energy and the driver of ``evgw`` as there (freq_int and mnP are unused)."""
    for k in _RUN_EVGW_REQUIRED:
        if k not in config:
            raise KeyError(k)
    ia = np.asarray(config["iaP"], dtype=np.float64)
    if ia.ndim != 3:
        raise ValueError(f"iaP must be [nocc, nvir, naux], got shape {ia.shape}")
    nocc, nvir, naux = ia.shape
    ab = config.get("abP")
    if ab is None:
        ab = np.zeros((nvir * nvir, naux))
    return evgw(np.asarray(config["mo_energy"], dtype=np.float64), np.asarray(config["mo_occ"]),
                ia.reshape(nocc * nvir, naux), config["ijP"], ab, config["chol_v"],
                config["vxc_dft"], max_cycle=config.get("max_cycle", 12),
                conv_tol=config.get("conv_tol", 1e-4), conv_tol_z=1e-3,
                damping=config.get("damping", 0.5), damping_dynamic=False,
                diis=config.get("diis", True), diis_space=config.get("diis_space", 6),
                nfreq=config.get("nfreq", 48), eta=0.01, verbose=config.get("verbose", 1),
                energy_window=config.get("energy_window", 2.0),
                frontier_convergence=config.get("frontier_convergence", True),
                frontier_window=config.get("frontier_window", 0.1),
                n_occ_active=config.get("n_occ_active", 3),
                n_vir_active=config.get("n_vir_active", 3),
                qp_solver=config.get("qp_solver", "linearized"),
                newton_max_iterations=config.get("newton_max_iterations", 50),
                newton_energy_tolerance=config.get("newton_energy_tolerance", 1e-8))


def _bind_current_stream(obj):
    """Orders the library's launches after the work torch has queued on its current stream:
    the handle is bound to that stream when it differs from the last one bound.  torch's
    default stream (handle 0, the legacy null stream) leaves the library on its own
    blocking stream, which the null stream already orders against."""
    s = torch.cuda.current_stream(obj.dev).cuda_stream
    h = obj.h
    if getattr(h, "bound_stream", 0) != s:
        h.check(h.lib.qx_set_stream(h.h, C.c_void_p(s) if s else None))
        h.bound_stream = s


class G0W0Solver:
    """Repeated G0W0 solves on device-resident inputs (no host round trips besides the
    library's own per-solve status check).  Outputs stay in ``self.out`` (device).

    Ordering: ``run()`` launches on torch's current stream of the device (the handle is
    rebound when that stream changes), so input updates queued there before ``run()`` are
    seen and work queued there after it sees the outputs."""

    def __init__(self, mo_energy, B_ov, nocc, targets=None, *, nfreq=48, x0=0.5, eta=0.01,
                 precision="bf16x3", device=0, newton=None):
        self.dev = _dev(device)
        self.h = _lib.handle(device)
        self.e = _f64(mo_energy, self.dev)
        self.B = _f64(B_ov, self.dev)
        nmo = self.e.numel()
        self.nocc = nocc
        self.targets = frontier_targets(nocc, nmo) if targets is None else np.asarray(targets,
                                                                                      np.int32)
        self.t_dev = _i32(self.targets, self.dev)
        nt = len(self.targets)
        self.desc = make_desc(nocc, nmo - nocc, self.B.shape[1], nfreq, x0, eta, precision, nt,
                              0, **(newton or {}))
        self.out = torch.empty((4, nt), dtype=torch.float64, device=self.dev)
        self.state = torch.empty(nt, dtype=torch.int32, device=self.dev)
        self.stats = _lib.Stats()
        self._args = None
        torch.cuda.synchronize(self.dev)

    def run(self):
        _bind_current_stream(self)
        if self._args is None:
            # the buffers live as long as the solver, so the call's arguments are built once
            o = self.out
            self._args = (self.h.h, C.byref(self.desc), _ptr(self.e), None, _ptr(self.B),
                          _ptr(self.t_dev), None, None, _ptr(o[0]), _ptr(o[1]), _ptr(o[2]),
                          _ptr(o[3]), _ptr(self.state), C.byref(self.stats))
        self.h.check(self.h.lib.qx_g0w0_cd_ov_dev(*self._args))

    def set_profiling(self, on: bool):
        self.h.check(self.h.lib.qx_set_profiling(self.h.h, 1 if on else 0))

    def profile(self):
        ms = (C.c_double * 8)()
        n = (C.c_longlong * 8)()
        self.h.check(self.h.lib.qx_get_profile(self.h.h, ms, n))
        return {name: (ms[i], n[i]) for i, name in enumerate(_lib.PHASES)}

    def result(self):
        o = self.out.cpu().numpy()
        return dict(qp_energies=o[0], z_factors=o[1], sigma_c=o[2] + 1j * o[3],
                    state=self.state.cpu().numpy(), targets=self.targets)


class G0W0BatchSolver:
    """Repeated batched G0W0 solves (``qx_g0w0_batch_ov_dev``) over device-resident
    instances of different sizes.  ``instances``: sequence of (mo_energy, B_ov, nocc,
    targets) with ov-only B_ov [nocc*nvir, naux]; targets None = HOMO-5..LUMO+5.  The first
    ``run()`` executes and tunes the grouped launches; later runs replay them as a HIP graph.
    Results: ``results()`` gives one dict per instance.  Ordering as ``G0W0Solver.run``:
    launches go to torch's current stream."""

    def __init__(self, instances, *, nfreq=32, x0=0.5, eta=0.01, precision="bf16x3", device=0,
                 newton=None):
        self.dev = _dev(device)
        self.h = _lib.handle(device)
        n = len(instances)
        if n == 0:
            raise ValueError("empty batch")
        self.arr = (_lib.BatchInstance * n)()
        self._keep = []
        self.targets = []
        off = 0
        self.offsets = []
        for q, (e, B, nocc, tg) in enumerate(instances):
            e = _f64(e, self.dev)
            B = _f64(B, self.dev)
            nmo = e.numel()
            tg = frontier_targets(nocc, nmo) if tg is None else np.ascontiguousarray(tg, np.int32)
            if B.dim() != 2 or B.shape[0] != nocc * (nmo - nocc):
                raise ValueError(f"instance {q}: B_ov must be [nocc*nvir, naux]")
            a = self.arr[q]
            a.nocc, a.nvir, a.naux, a.ntarget = int(nocc), int(nmo - nocc), int(B.shape[1]), len(tg)
            a.mo_energy, a.B_ov = e.data_ptr(), B.data_ptr()
            a.targets = tg.ctypes.data
            a.static_terms = None
            self._keep += [e, B, tg]
            self.targets.append(tg)
            self.offsets.append(off)
            off += len(tg)
        self.ntot = off
        self.desc = make_desc(1, 1, 1, nfreq, x0, eta, precision, 1, 0, **(newton or {}))
        self.out = torch.empty((4, off), dtype=torch.float64, device=self.dev)
        self.state = torch.empty(off, dtype=torch.int32, device=self.dev)
        self.stats = _lib.Stats()
        torch.cuda.synchronize(self.dev)

    def run(self):
        _bind_current_stream(self)
        o = self.out
        self.h.check(self.h.lib.qx_g0w0_batch_ov_dev(
            self.h.h, C.byref(self.desc), len(self.arr), self.arr, _ptr(o[0]), _ptr(o[1]),
            _ptr(o[2]), _ptr(o[3]), _ptr(self.state), C.byref(self.stats)))

    def set_profiling(self, on: bool):
        self.h.check(self.h.lib.qx_set_profiling(self.h.h, 1 if on else 0))

    def profile(self):
        ms = (C.c_double * 8)()
        n = (C.c_longlong * 8)()
        self.h.check(self.h.lib.qx_get_profile(self.h.h, ms, n))
        return {name: (ms[i], n[i]) for i, name in enumerate(_lib.PHASES)}

    def results(self):
        o = self.out.cpu().numpy()
        st = self.state.cpu().numpy()
        res = []
        for off, tg in zip(self.offsets, self.targets):
            sl = slice(off, off + len(tg))
            res.append(dict(qp_energies=o[0, sl], z_factors=o[1, sl],
                            sigma_c=o[2, sl] + 1j * o[3, sl], state=st[sl], targets=tg))
        return res


def g0w0_batch(instances, **kw):
    """One batched solve; returns a list of per-instance result dicts."""
    s = G0W0BatchSolver(instances, **kw)
    s.run()
    return s.results()
