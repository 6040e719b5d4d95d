"""GPU: edge cases of the shapes -- single orbitals, one auxiliary function, auxiliary sizes
at and around the 64/128/256 tile boundaries, one frequency, a one-instance batch -- all
against the FP64 oracle (1e-6 Ha per QP energy, as the main parity tests) and, for the mixed
path, within 1 meV of FP64."""
import numpy as np
import pytest

from oracle import gw_oracle as O

pytestmark = pytest.mark.gpu
MEV = 1.0 / 27211.386245988


def _run(nocc, nvir, naux, nf, seed, targets=None):
    from quasix_amd import gw
    e, B = O.make_inputs(nocc, nvir, naux, seed=seed)
    tg = np.arange(nocc + nvir) if targets is None else np.asarray(targets)
    fr, wt = O.get_scaled_legendre_roots(nf, 0.5)
    ref = O.g0w0_cd(e, B, nocc, tg, fr, wt)
    r64 = gw.g0w0(e, B, nocc, tg, nfreq=nf)
    rmx = gw.g0w0(e, B, nocc, tg, nfreq=nf, precision="bf16x3")
    e64 = float(np.max(np.abs(r64["qp_energies"] - ref["E"])))
    emx = float(np.max(np.abs(rmx["qp_energies"] - r64["qp_energies"])))
    print(f"{nocc}/{nvir}/{naux} nf={nf}: fp64 vs oracle {e64:.2e} Ha, mixed vs fp64 {emx / MEV:.2e} meV")
    assert e64 < 1e-6 and emx < MEV


@pytest.mark.parametrize("nocc,nvir,naux", [(1, 1, 1), (1, 2, 3), (2, 3, 5), (3, 1, 7)])
def test_tiny_shapes(gpu, nocc, nvir, naux):
    _run(nocc, nvir, naux, 8, seed=nocc * 10 + nvir)


@pytest.mark.parametrize("naux", [63, 64, 65, 127, 128, 129, 255, 256, 257])
def test_aux_tile_boundaries(gpu, naux):
    _run(4, 20, naux, 6, seed=naux, targets=[2, 3, 4, 5])


def test_single_frequency(gpu):
    _run(3, 12, 40, 1, seed=3)


def test_one_instance_batch(gpu):
    import torch
    from quasix_amd import gw
    e, B = O.make_inputs(3, 11, 30, seed=4)
    dev = torch.device("cuda", 0)
    inst = [(torch.as_tensor(e, device=dev), torch.as_tensor(B, device=dev), 3, None)]
    got = gw.g0w0_batch(inst, nfreq=8, precision="fp64")[0]
    fr, wt = O.get_scaled_legendre_roots(8, 0.5)
    ref = O.g0w0_cd(e, B, 3, got["targets"], fr, wt)
    assert np.max(np.abs(got["qp_energies"] - ref["E"])) < 1e-6
