"""Pins the oracle adjoint of the G0W0 solve against central finite differences (CPU)."""
import numpy as np

from oracle import gw_oracle as O

TIGHT = O.NewtonConfig(energy_tolerance=1e-14)


def _solve(e, B, nocc, targets, fr, wt):
    return O.g0w0_cd(e, B, nocc, targets, fr, wt, newton=TIGHT)["E"]


def test_adjoint_matches_central_differences():
    nocc, nvir, naux = 3, 8, 12
    e, B = O.make_inputs(nocc, nvir, naux, seed=21)
    targets = np.array([nocc - 1, nocc])                       # HOMO, LUMO
    fr, wt = O.get_scaled_legendre_roots(16, 0.5)
    E = _solve(e, B, nocc, targets, fr, wt)
    rng = np.random.default_rng(0)
    gE = np.array([1.0, -0.7])
    gB, geps = O.g0w0_adjoint(e, B, nocc, targets, fr, wt, E, gE)
    h = 1e-5
    for trial in range(3):
        V = rng.standard_normal(B.shape) * float(np.std(B))
        fd = (gE @ _solve(e, B + h * V, nocc, targets, fr, wt)
              - gE @ _solve(e, B - h * V, nocc, targets, fr, wt)) / (2 * h)
        an = float(np.sum(gB * V))
        assert abs(an - fd) <= 1e-5 * max(abs(fd), 1e-8), (trial, an, fd)
    for i in (0, nocc - 1, nocc, nocc + nvir - 1):
        d = np.zeros_like(e)
        d[i] = 1.0
        fd = (gE @ _solve(e + h * d, B, nocc, targets, fr, wt)
              - gE @ _solve(e - h * d, B, nocc, targets, fr, wt)) / (2 * h)
        assert abs(geps[i] - fd) <= 1e-5 * max(abs(fd), 1e-8), (i, geps[i], fd)
