"""torch.autograd.Function for G0W0 quasiparticle energies (FP64 forward, hand-written
HIP adjoint).  The reference has no derivatives; this is the differentiable form of the
CD G0W0 solve of libqxgw.so.

    E = g0w0_qp_energies(B_ov, mo_energy, nocc, targets)      # tensors on the GPU
    E.sum().backward()  ->  B_ov.grad, mo_energy.grad

Forward: qx_g0w0_cd_ov_dev in FP64.  Backward: qx_g0w0_adjoint_ov_dev, which re-forms the
FP64 factorisation, keeps X_k = L_k^-1 of every frequency as the checkpoint of Wc_k, and
propagates the implicit derivative of the Newton QP equation through the quadrature, Wc and
Pi.  Sigma_x - V_xc is held fixed (zero).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from .gw import _ptr, frontier_targets, make_desc


class G0W0QPFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, B_ov, mo_energy, nocc, targets, nfreq, x0, eta, newton_tol):
        if B_ov.dtype != torch.float64 or mo_energy.dtype != torch.float64:
            raise TypeError("B_ov and mo_energy must be float64")
        if not (B_ov.is_cuda and mo_energy.is_cuda):
            raise RuntimeError("B_ov and mo_energy must live on the GPU")
        dev = B_ov.device
        h = _lib.handle(dev.index or 0)
        B = B_ov.detach().contiguous()
        e = mo_energy.detach().contiguous()
        tg = torch.as_tensor(np.asarray(targets, dtype=np.int32), device=dev)
        nt = tg.numel()
        d = make_desc(nocc, e.numel() - nocc, B.shape[1], nfreq, x0, eta, "fp64", nt, 0,
                      energy_tolerance=newton_tol)
        out = torch.empty((4, nt), dtype=torch.float64, device=dev)
        state = torch.empty(nt, dtype=torch.int32, device=dev)
        s = _lib.Stats()
        torch.cuda.synchronize(dev)
        h.check(h.lib.qx_g0w0_cd_ov_dev(h.h, C.byref(d), _ptr(e), None, _ptr(B), _ptr(tg), None,
                                        None, _ptr(out[0]), _ptr(out[1]), _ptr(out[2]),
                                        _ptr(out[3]), _ptr(state), C.byref(s)))
        if s.residue_targets:
            raise RuntimeError("G0W0QPFunction: residue poles are active; adjoint unsupported")
        E = out[0].clone()
        ctx.save_for_backward(B, e, tg, E)
        ctx.cfg = (nocc, nfreq, x0, eta, newton_tol)
        return E

    @staticmethod
    def backward(ctx, gE):
        B, e, tg, E = ctx.saved_tensors
        nocc, nfreq, x0, eta, tol = ctx.cfg
        dev = B.device
        h = _lib.handle(dev.index or 0)
        d = make_desc(nocc, e.numel() - nocc, B.shape[1], nfreq, x0, eta, "fp64", tg.numel(), 0,
                      energy_tolerance=tol)
        g = gE.detach().to(torch.float64).contiguous()
        gB = torch.empty_like(B)
        ge = torch.empty_like(e)
        torch.cuda.synchronize(dev)
        h.check(h.lib.qx_g0w0_adjoint_ov_dev(h.h, C.byref(d), _ptr(e), _ptr(B), _ptr(tg),
                                             _ptr(E), _ptr(g), _ptr(gB), _ptr(ge)))
        return gB, ge, None, None, None, None, None, None


def g0w0_qp_energies(B_ov, mo_energy, nocc, targets=None, *, nfreq=32, x0=0.5, eta=0.01,
                     newton_tol=1e-8):
    """Differentiable QP energies of the targets (default HOMO, LUMO)."""
    nmo = mo_energy.numel()
    if targets is None:
        targets = np.array([nocc - 1, nocc], dtype=np.int32)
    return G0W0QPFunction.apply(B_ov, mo_energy, int(nocc), np.asarray(targets), int(nfreq),
                                float(x0), float(eta), float(newton_tol))
