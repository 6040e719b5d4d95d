"""Seeded synthetic inputs of the benchmark workloads.

Orbital energies eps_occ ~ U[-1.5, -0.3] and eps_vir ~ U[0.05, 3.0] (each sorted) and a DF
tensor B = U diag(s) V^T with Gaussian QR-orthonormalised U (N_ov x r) and V (naux x r),
r = min(N_ov, naux), s_k = 0.5 exp(-k/20), returned as [nocc*nvir, naux] (the reference's
iaP layout, quasix/src/evgw.rs:105-114).  Nothing of the G0W0 path is computed here.

* ``make_inputs``         numpy, draws in the order eps_occ, eps_vir, U, V from
                          ``numpy.random.default_rng(seed)``.
* ``make_inputs_device``  the same distributions drawn with torch's generator on the GPU
                          (QR and product on the device), for batches whose host-side
                          generation would take minutes.
* ``batch_shapes``        the batched workload: nocc ~ U{8..60}, nvir = 9 nocc,
                          naux = 4 (nocc + nvir).
"""
from __future__ import annotations

import numpy as np


def make_inputs(nocc: int, nvir: int, naux: int, seed: int = 0):
    rng = np.random.default_rng(seed)
    e_occ = np.sort(rng.uniform(-1.5, -0.3, nocc))
    e_vir = np.sort(rng.uniform(0.05, 3.0, nvir))
    nov = nocc * nvir
    r = min(nov, naux)
    U = np.linalg.qr(rng.standard_normal((nov, r)))[0]
    V = np.linalg.qr(rng.standard_normal((naux, r)))[0]
    s = 0.5 * np.exp(-np.arange(r) / 20.0)
    return np.concatenate([e_occ, e_vir]), np.ascontiguousarray((U * s) @ V.T)


def make_inputs_device(nocc: int, nvir: int, naux: int, seed: int, device):
    import torch
    g = torch.Generator(device=device)
    g.manual_seed(int(seed))
    f64 = dict(dtype=torch.float64, device=device)
    e_occ = torch.sort(torch.rand(nocc, generator=g, **f64) * 1.2 - 1.5).values
    e_vir = torch.sort(torch.rand(nvir, generator=g, **f64) * 2.95 + 0.05).values
    nov = nocc * nvir
    r = min(nov, naux)
    U = torch.linalg.qr(torch.randn(nov, r, generator=g, **f64))[0]
    V = torch.linalg.qr(torch.randn(naux, r, generator=g, **f64))[0]
    s = 0.5 * torch.exp(-torch.arange(r, **f64) / 20.0)
    B = (U * s) @ V.T.contiguous()
    return torch.cat([e_occ, e_vir]).contiguous(), B.contiguous()


def batch_shapes(n: int = 256, seed: int = 2):
    rng = np.random.default_rng(seed)
    nocc = rng.integers(8, 61, size=n)
    nvir = 9 * nocc
    naux = 4 * (nocc + nvir)
    return [(int(o), int(v), int(a)) for o, v, a in zip(nocc, nvir, naux)]
