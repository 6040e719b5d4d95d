"""GPU adjoint: hand-written HIP backward vs the oracle adjoint and vs central finite
differences of the GPU FP64 forward (h = 1e-5, agreement 1e-5 relative)."""
import numpy as np
import pytest

from oracle import gw_oracle as O

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return abs(a - b) / max(abs(b), 1e-12)


def test_adjoint_matches_oracle_small(gpu):
    import torch
    from quasix_amd.autograd import g0w0_qp_energies
    nocc, nvir, naux, nf = 4, 12, 40, 16
    e, B = O.make_inputs(nocc, nvir, naux, seed=31)
    targets = np.array([nocc - 2, nocc - 1, nocc, nocc + 1])
    Bt = torch.tensor(B, device="cuda", requires_grad=True)
    et = torch.tensor(e, device="cuda", requires_grad=True)
    E = g0w0_qp_energies(Bt, et, nocc, targets, nfreq=nf, newton_tol=1e-13)
    gE = torch.tensor([0.3, 1.0, -0.8, 0.5], device="cuda", dtype=torch.float64)
    (E * gE).sum().backward()
    fr, wt = O.get_scaled_legendre_roots(nf, 0.5)
    gB, geps = O.g0w0_adjoint(e, B, nocc, targets, fr, wt, E.detach().cpu().numpy(),
                              gE.cpu().numpy())
    errB = float(np.max(np.abs(Bt.grad.cpu().numpy() - gB)) / np.max(np.abs(gB)))
    errE = float(np.max(np.abs(et.grad.cpu().numpy() - geps)) / np.max(np.abs(geps)))
    print(f"grad_B rel err {errB:.2e}, grad_eps rel err {errE:.2e}")
    assert errB < 1e-9 and errE < 1e-9


def test_adjoint_c3_central_differences(gpu):
    """nocc 20 / nvir 200 / naux 900, 32 frequencies, d(E_HOMO, E_LUMO)/d(B, eps)."""
    import torch
    from quasix_amd import gw
    from quasix_amd.autograd import g0w0_qp_energies
    nocc, nvir, naux, nf = 20, 200, 900, 32
    e, B = O.make_inputs(nocc, nvir, naux, seed=3)
    targets = np.array([nocc - 1, nocc])
    Bt = torch.tensor(B, device="cuda", requires_grad=True)
    et = torch.tensor(e, device="cuda", requires_grad=True)
    E = g0w0_qp_energies(Bt, et, nocc, targets, nfreq=nf, newton_tol=1e-13)
    gE = torch.tensor([1.0, -1.0], device="cuda", dtype=torch.float64)   # d(gap)/d(.) sign
    (E * gE).sum().backward()
    gB = Bt.grad.cpu().numpy()
    ge = et.grad.cpu().numpy()

    def L(ee, BB):
        r = gw.g0w0(ee, BB, nocc, targets, nfreq=nf, newton=dict(energy_tolerance=1e-13))
        return float(r["qp_energies"] @ gE.cpu().numpy())

    h = 1e-5
    rng = np.random.default_rng(7)
    for trial in range(2):
        # direction with entries of B's size, so h = 1e-5 is a 1e-5 relative perturbation
        V = rng.standard_normal(B.shape) * float(np.std(B))
        fd = (L(e, B + h * V) - L(e, B - h * V)) / (2 * h)
        an = float(np.sum(gB * V))
        print(f"B direction {trial}: adjoint {an:.10e} fd {fd:.10e} rel {_rel(an, fd):.2e}")
        assert _rel(an, fd) < 1e-5
    for i in (nocc - 1, nocc, 0, nocc + nvir - 1):
        d = np.zeros_like(e)
        d[i] = 1.0
        fd = (L(e + h * d, B) - L(e - h * d, B)) / (2 * h)
        print(f"eps[{i}]: adjoint {ge[i]:.10e} fd {fd:.10e} rel {_rel(ge[i], fd):.2e}")
        assert abs(ge[i] - fd) <= 1e-5 * max(abs(fd), 1e-6)
