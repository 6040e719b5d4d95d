"""GPU: the no-allocation entry point qx_g0w0_cd_ov_noalloc_dev against qx_g0w0_cd_ov_dev,
called through the C ABI on the same inputs: FP64 bit-identical (same kernels, same
order), the BF16x3 path too once the handle has tuned the shape; two calls in a row with the
second on a smaller problem (the scratch keeps the first call's contents); a problem beyond
the scratch is rejected before anything runs."""
import ctypes as C

import numpy as np
import pytest
import torch

from oracle import gw_oracle as O

pytestmark = pytest.mark.gpu


def _solve(fn, h, d, e, B, tg):
    dev = e.device
    out = torch.empty((4, len(tg)), dtype=torch.float64, device=dev)
    st = torch.empty(len(tg), dtype=torch.int32, device=dev)
    t = torch.as_tensor(tg, dtype=torch.int32, device=dev)
    s = __import__("quasix_amd._lib", fromlist=["Stats"]).Stats()
    p = lambda x: C.c_void_p(x.data_ptr())
    torch.cuda.synchronize()
    rc = fn(h.h, C.byref(d), p(e), None, p(B), p(t), None, None, p(out[0]), p(out[1]),
            p(out[2]), p(out[3]), p(st), C.byref(s))
    h.check(rc)
    return out.cpu().numpy(), st.cpu().numpy()


@pytest.mark.parametrize("prec", ["fp64", "bf16x3"])
def test_noalloc_matches_allocating_entry(gpu, prec):
    import quasix_amd._lib as L
    from quasix_amd import gw
    h = L.handle(0)
    dev = torch.device("cuda", 0)
    for nocc, nvir, naux, nf, seed in [(8, 72, 320, 32, 1), (5, 19, 84, 16, 2)]:
        e, B = O.make_inputs(nocc, nvir, naux, seed=seed)
        e, B = torch.as_tensor(e, device=dev), torch.as_tensor(B, device=dev)
        tg = gw.frontier_targets(nocc, nocc + nvir, 2, 2)
        d = gw.make_desc(nocc, nvir, naux, nf, precision=prec, ntarget=len(tg))
        ref, st_ref = _solve(h.lib.qx_g0w0_cd_ov_dev, h, d, e, B, tg)
        got, st = _solve(h.lib.qx_g0w0_cd_ov_noalloc_dev, h, d, e, B, tg)
        diff = float(np.max(np.abs(got - ref)))
        print(f"noalloc {prec} {nocc}/{nvir}/{naux} nf={nf}: max diff {diff:.3e}, "
              f"states {st.tolist()}")
        assert np.array_equal(st, st_ref)
        assert np.array_equal(got, ref)


def test_noalloc_rejects_oversized_problem(gpu):
    import quasix_amd._lib as L
    from quasix_amd import gw
    h = L.handle(0)
    dev = torch.device("cuda", 0)
    nocc, nvir, naux, nf = 4, 40, 2000, 32          # eps and X alone need 2 x 1 GiB
    e, B = O.make_inputs(nocc, nvir, naux, seed=3)
    e, B = torch.as_tensor(e, device=dev), torch.as_tensor(B, device=dev)
    tg = np.array([nocc - 1, nocc], np.int32)
    d = gw.make_desc(nocc, nvir, naux, nf, ntarget=len(tg))
    with pytest.raises(L.QxError, match="too large for the no-allocation entry point"):
        _solve(h.lib.qx_g0w0_cd_ov_noalloc_dev, h, d, e, B, tg)


def test_noalloc_fallback_and_residues(gpu):
    """A problem whose Cholesky fails at 9 of 16 frequencies (an occupied level above the
    virtuals, B scaled by 2.5) and whose HOMO needs the residue path: in no-allocation mode
    the deferred failure check, the LU fallback, the relaunched Newton and the host residue
    solve run from the scratch and give the allocating entry's results bit for bit."""
    import quasix_amd._lib as L
    from quasix_amd import gw
    h = L.handle(0)
    dev = torch.device("cuda", 0)
    nocc, nvir, naux, nf = 4, 20, 64, 16
    e, B = O.make_inputs(nocc, nvir, naux, seed=72)
    e[nocc - 1] = 3.4
    B = B * 2.5
    e, B = torch.as_tensor(e, device=dev), torch.as_tensor(B, device=dev)
    tg = np.arange(nocc - 2, nocc + 2).astype(np.int32)
    d = gw.make_desc(nocc, nvir, naux, nf, precision="fp64", ntarget=len(tg))
    ref, st_ref = _solve(h.lib.qx_g0w0_cd_ov_dev, h, d, e, B, tg)
    got, st = _solve(h.lib.qx_g0w0_cd_ov_noalloc_dev, h, d, e, B, tg)
    print(f"noalloc fallback: states {st.tolist()}, max diff {np.max(np.abs(got - ref)):.3e}")
    assert np.array_equal(st, st_ref)
    assert np.array_equal(got, ref)
