"""Host-only stage functions of libqxgw (no GPU needed): quasix_core.solve_qp_equation
(qx_solve_qp_linearized; quasix_core/src/qp/solver.rs:137-221) against the oracle's
restatement and its closed form, with the reference's error cases, and the qp_solver
spellings the evgw binding accepts (quasix/src/evgw.rs:215-226)."""
import numpy as np
import pytest

from oracle import gw_oracle as O


def test_solve_qp_equation_matches_oracle():
    from quasix_amd import core
    rng = np.random.default_rng(0)
    e, sx, sc, vxc = (rng.standard_normal(17) for _ in range(4))
    r = core.solve_qp_equation(e, sx, sc, vxc)
    E, Z = O.solve_qp_linearized(e, sx, sc, vxc)
    assert np.array_equal(r["qp_energies"], E) and np.array_equal(r["z_factors"], Z)
    assert r["converged"] is True
    # closed form on a hand-computed case
    r = core.solve_qp_equation([-0.5, 0.2], [-0.3, -0.1], [0.05, -0.02], [-0.25, -0.15])
    assert np.allclose(r["qp_energies"], [-0.5, 0.23], atol=1e-15)


def test_solve_qp_equation_errors():
    from quasix_amd import core
    with pytest.raises(RuntimeError, match="sigma_c dimension mismatch"):
        core.solve_qp_equation(np.zeros(3), np.zeros(3), np.zeros(2), np.zeros(3))
    with pytest.raises(RuntimeError, match=r"QP energy\[1\] is not finite"):
        core.solve_qp_equation(np.zeros(3), [0.0, np.inf, 0.0], np.zeros(3), np.zeros(3))


@pytest.mark.parametrize("spelling,parsed", [("linearized", "linearized"), ("Linear", "linearized"),
                                             ("newton", "newton"), ("NEWTON-RAPHSON", "newton"),
                                             ("solved", "newton")])
def test_qp_solver_spellings(spelling, parsed):
    from quasix_amd import gw
    assert gw.parse_qp_solver(spelling) == parsed


def test_qp_solver_rejects_unknown():
    from quasix_amd import gw
    with pytest.raises(ValueError, match="Invalid qp_solver 'secant'"):
        gw.parse_qp_solver("secant")
