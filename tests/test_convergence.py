"""CPU: the evGW convergence monitor / DIIS (quasix_amd.convergence, host code of the product)
and the oracle's restatement, pinned by mathematics rather than by each other:

* the elimination solve equals numpy.linalg.solve on well-conditioned systems and rejects
  singular ones;
* the DIIS extrapolation is the constrained least-squares combination of the stored
  vectors (min |sum c_k r_k|, sum c_k = 1, closed form c = G^-1 1 / 1^T G^-1 1);
* the evGW driver's call order (record, then extrapolate) gives zero error vectors, so from
  the third iteration on the DIIS matrix is singular and the step is plain damping, while the
  third iteration itself takes the new values (one stored vector);
* both implementations give the same mixed iterates on the same inputs.
"""
import numpy as np
import pytest

from oracle import gw_oracle as O
from quasix_amd import convergence as Q


def _spd_border(n, rng):
    m = rng.standard_normal((n, n))
    bm = -np.ones((n + 1, n + 1))
    bm[n, n] = 0.0
    bm[:n, :n] = m @ m.T + n * np.eye(n)
    return bm


@pytest.mark.parametrize("solve", [Q.pulay_solve, O.diis_solve])
def test_solve_matches_numpy(solve):
    rng = np.random.default_rng(0)
    for n in (2, 3, 5, 7):
        bm = _spd_border(n, rng)
        rhs = np.r_[np.zeros(n), -1.0]
        x = solve(bm, rhs)
        assert np.allclose(x, np.linalg.solve(bm, rhs), rtol=1e-12, atol=1e-14)
        assert abs(x[:n].sum() - 1.0) < 1e-12


@pytest.mark.parametrize("solve,exc", [(Q.pulay_solve, Q.DIISError),
                                       (O.diis_solve, ArithmeticError)])
def test_solve_rejects_singular(solve, exc):
    bm = np.array([[0.0, 0.0, -1.0], [0.0, 0.0, -1.0], [-1.0, -1.0, 0.0]])
    with pytest.raises(exc):
        solve(bm, np.array([0.0, 0.0, -1.0]))


def _iterates(n, k, seed=1):
    rng = np.random.default_rng(seed)
    A = 0.5 * rng.standard_normal((n, n)) / np.sqrt(n)
    b = rng.standard_normal(n)
    xs = [rng.standard_normal(n)]
    for _ in range(k):
        xs.append(A @ xs[-1] + b)
    return xs


def _constrained_ls(xs):
    """argmin |sum c_k r_k| subject to sum c_k = 1 over the last len(xs)-1 iterates
    (r_k = x_k - x_{k-1}): c = G^-1 1 / (1^T G^-1 1); returns sum c_k x_k."""
    R = np.array([xs[k] - xs[k - 1] for k in range(1, len(xs))])
    G = R @ R.T
    g = np.linalg.solve(G, np.ones(len(R)))
    c = g / g.sum()
    return c @ np.array(xs[1:])


@pytest.mark.parametrize("impl", ["product", "oracle"])
def test_diis_is_constrained_least_squares(impl):
    n, m = 8, 4                                  # 4 stored vectors in 8 dimensions
    xs = _iterates(n, m)
    z = np.zeros(n)
    if impl == "product":
        mon = Q.EvGWHistory(space=m)
        mon.record(xs[0], z)
        for x in xs[1:]:
            e, _ = mon.extrapolate(x, z)         # error = x_k - x_{k-1}
            mon.record(x, z)
    else:
        mon = O.DIISMonitor(space=m)
        mon.record(xs[0], z)
        for x in xs[1:]:
            e, _ = mon.extrapolate(x, z)
            mon.record(x, z)
    assert np.allclose(e, _constrained_ls(xs), atol=1e-12)


def _drive(mix, mon_factory, seq, damping=0.5):
    mon = mon_factory()
    e, z = np.zeros(3), np.ones(3)
    out = []
    for it, (en, zn) in enumerate(seq):
        r = mix(it, mon, e, z, en, zn, damping)
        e, z = r[0], r[1]
        out.append((e.copy(), z.copy()))
    return out


def test_driver_order_falls_back_to_damping_and_both_agree():
    rng = np.random.default_rng(3)
    seq = [(rng.uniform(-1, 1, 3), rng.uniform(0.6, 0.95, 3)) for _ in range(6)]
    prod = _drive(lambda it, m, e, z, en, zn, d: Q.evgw_mix(it, m, e, z, en, zn, d,
                                                            use_diis=True),
                  lambda: Q.EvGWHistory(space=6), seq)
    orac = _drive(lambda it, m, e, z, en, zn, d: O.evgw_step_mix(it, m, e, z, en, zn, d, True),
                  lambda: O.DIISMonitor(6), seq)
    for (a, za), (b, zb) in zip(prod, orac):
        assert np.allclose(a, b, atol=1e-15) and np.allclose(za, zb, atol=1e-15)
    # iteration 2 takes the new values; iterations >= 3 are damped
    assert np.allclose(prod[2][0], seq[2][0])
    for it in range(3, 6):
        want = 0.5 * prod[it - 1][0] + 0.5 * seq[it][0]
        assert np.allclose(prod[it][0], want, atol=1e-15)


def test_frontier_window_product_matches_oracle():
    """The product's stop-rule window and the oracle's (hand-counted cases live in
    tests/test_oracle_evgw_pins.py) on random spectra and windows."""
    rng = np.random.default_rng(11)
    for _ in range(200):
        nocc, nvir = int(rng.integers(1, 8)), int(rng.integers(1, 8))
        e = np.r_[np.sort(rng.uniform(-3, -0.1, nocc)), np.sort(rng.uniform(0.0, 3, nvir))]
        w = float(rng.uniform(0.0, 2.5))
        for fc in (True, False):
            assert Q.frontier_window(e, nocc, fc, w) == O.convergence_window(e, nocc, fc, w)

