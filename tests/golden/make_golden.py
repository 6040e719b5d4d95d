"""Writes the committed golden fixtures from the FP64 oracle (run in the build container).

    python tests/golden/make_golden.py

Fixtures (numpy .npz, allow_pickle=False):
  c0_g0w0.npz       nocc 5 / nvir 19 / naux 84 / 32 freqs, seed 0: inputs, Pi at 3
                    frequencies, Wc_nm for the ov-only pairs, E_QP / Z / Sigma for all 24 orbitals
  casida_cd.npz     nocc 3 / nvir 6 / naux 8 with full (oo/ov/vv) blocks: exact Casida
                    Sigma_c at energies that need residues, and the CD (I + R) value at 400
                    frequencies, eta 1e-7
  freq_cache_ac.npz the synthetic input of quasix_core/src/selfenergy/frequency_cache.rs:620-700
                    (nmo 6, nocc 3, naux 15, lpq[p,m,n] = 0.1 exp(-0.2 (m+n+p)), nw 20):
                    AC Sigma(i nu) from get_sigma_diag
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import gw_oracle as O  # noqa: E402
from oracle import casida as Cs  # noqa: E402


def c0():
    nocc, nvir, naux, nf = 5, 19, 84, 32
    e, B = O.make_inputs(nocc, nvir, naux, seed=0)
    fr, wt = O.get_scaled_legendre_roots(nf, 0.5)
    targets = np.arange(nocc + nvir)
    r = O.g0w0_cd(e, B, nocc, targets, fr, wt)
    pt, pm, Bp = r["pairs"]
    kidx = np.array([0, 15, 31])
    pis = np.stack([O.get_rho_response(fr[k], e, B, nocc) for k in kidx])
    np.savez_compressed(os.path.join(HERE, "c0_g0w0.npz"), mo_energy=e, B_ov=B, freqs=fr,
                        wts=wt, pi_idx=kidx, pi=pis, pair_t=pt, pair_m=pm, wnm=r["wnm"],
                        targets=targets, E=r["E"], Z=r["Z"], sigma=r["sigma"])


def casida():
    rng = np.random.default_rng(5)
    nocc, nvir, naux = 3, 6, 8
    nmo = nocc + nvir
    e = np.concatenate([np.sort(rng.uniform(-1.2, -0.4, nocc)),
                        np.sort(rng.uniform(0.1, 1.5, nvir))])
    full = rng.standard_normal((nmo, nmo, naux)) * 0.15
    full = 0.5 * (full + full.transpose(1, 0, 2))
    B = full[:nocc, nocc:, :].reshape(-1, naux).copy()
    om, A = Cs.casida_poles(e, B, nocc)
    targets = np.arange(nmo)
    pt, pm, Bp = O.pair_list(full, targets)
    fr, wt = O.get_scaled_legendre_roots(400, 0.5)
    wnm = O.wmn_imag(e, B, nocc, Bp, fr)
    ctx = O.CDContext(e, e, nocc, B, targets, pt, pm, Bp, fr, wt, wnm, eta=1e-7)
    Es = np.stack([e - 0.05, e + 0.04], axis=1)
    exact = np.empty(Es.shape, dtype=np.complex128)
    cd = np.empty(Es.shape, dtype=np.complex128)
    for t in range(nmo):
        sel = pt == t
        for q in range(2):
            exact[t, q] = Cs.sigma_exact(e, nocc, om, A, Bp[sel], pm[sel], Es[t, q], eta=1e-7)
            cd[t, q] = O.sigma_cd_eval(ctx, t, Es[t, q])
    np.savez_compressed(os.path.join(HERE, "casida_cd.npz"), mo_energy=e, full=full,
                        energies=Es, exact=exact, cd=cd, omega=om)


def freq_cache():
    nmo, nocc, naux = 6, 3, 15
    e = np.array([-0.5, -0.4, -0.3, 0.1, 0.2, 0.3])
    p, m, n = np.meshgrid(np.arange(naux), np.arange(nmo), np.arange(nmo), indexing="ij")
    lpq = 0.1 * np.exp(-0.2 * (m + n + p))
    sig, om = O.get_sigma_diag(e, lpq, list(range(nmo)), nocc, nw=20)
    np.savez_compressed(os.path.join(HERE, "freq_cache_ac.npz"), mo_energy=e, lpq=lpq,
                        sigma=sig, omega=om)


if __name__ == "__main__":
    c0()
    casida()
    freq_cache()
    print("fixtures written to", HERE)
