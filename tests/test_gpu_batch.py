"""GPU: batched solves (qx_g0w0_batch_ov_dev) of instances with different sizes.

Each instance of a batch must reproduce the single-instance path (qx_g0w0_cd_ov_dev) with the
same precision: FP64 to 1e-10 Ha (same kernels; only the launch grouping differs); BF16x3 to
5e-7 Ha (the grouped Pi kernels and the tuner's pick among them differ from the
single-instance ones, so the FP32 accumulation order differs) and within 1 meV of FP64.
The first call runs the grouped launches directly, later calls replay them as a captured
graph; both must agree, and a replay must follow new contents of the input buffers.  One
FP64 instance is also checked against the FP64 oracle (1e-6 Ha).
"""
import numpy as np
import pytest

from oracle import gw_oracle as O

pytestmark = pytest.mark.gpu
MEV = 1.0 / 27211.386245988
SHAPES = [(5, 19, 84), (8, 72, 320), (6, 54, 240), (10, 90, 400), (4, 30, 136), (7, 63, 280)]
NF = 16


def _close(name, got, ref, tol):
    err = float(np.max(np.abs(np.asarray(got) - np.asarray(ref))))
    print(f"{name}: max abs err {err:.3e} (limit {tol:.3e})")
    assert err <= tol, f"{name}: {err:.3e} > {tol:.3e}"


def _instances(torch, shapes=SHAPES, seed0=40):
    dev = torch.device("cuda", 0)
    out = []
    for q, (o, v, a) in enumerate(shapes):
        e, B = O.make_inputs(o, v, a, seed=seed0 + q)
        out.append((torch.as_tensor(e, device=dev), torch.as_tensor(B, device=dev), o, None))
    return out


def _single(gw, inst, precision):
    res = []
    for e, B, nocc, tg in inst:
        s = gw.G0W0Solver(e, B, nocc, tg, nfreq=NF, precision=precision)
        s.run()
        res.append(s.result())
    return res


@pytest.mark.parametrize("precision", ["fp64", "bf16x3"])
def test_batch_matches_single_instance_path(gpu, precision):
    import torch
    from quasix_amd import gw
    inst = _instances(torch)
    single = _single(gw, inst, precision)
    bs = gw.G0W0BatchSolver(inst, nfreq=NF, precision=precision)
    bs.run()
    first = bs.results()
    assert bs.stats.graph_replay == 0 and bs.stats.resolved_instances == 0
    bs.run()
    assert bs.stats.graph_replay == 1
    second = bs.results()
    tol = 1e-10 if precision == "fp64" else 5e-7
    for q, (a, b, c) in enumerate(zip(first, second, single)):
        _close(f"{precision} inst {q} direct vs single", a["qp_energies"], c["qp_energies"], tol)
        _close(f"{precision} inst {q} graph vs direct", b["qp_energies"], a["qp_energies"], 1e-12)
        _close(f"{precision} inst {q} Z", b["z_factors"], c["z_factors"], 100 * tol)
        assert np.all(b["state"] == 0)


def test_batch_factorisation_xcd_mapping(gpu, monkeypatch):
    """The grouped factorisation kernels place the tiles of each (instance, frequency) on
    one XCD (instance offsets not a multiple of 8, more frequencies than XCDs); every tile
    is computed as in launch order, so the FP64 energies are bit-identical with
    QXGW_POTRF_XCD=0."""
    import torch
    from quasix_amd import gw
    inst = _instances(torch, seed0=300)
    monkeypatch.setenv("QXGW_POTRF_XCD", "1")
    r1 = gw.g0w0_batch(inst, nfreq=NF, precision="fp64")
    monkeypatch.setenv("QXGW_POTRF_XCD", "0")
    r0 = gw.g0w0_batch(inst, nfreq=NF, precision="fp64")
    for q, (a, b) in enumerate(zip(r1, r0)):
        assert np.array_equal(a["qp_energies"], b["qp_energies"]), q
        assert np.array_equal(a["z_factors"], b["z_factors"]), q


def test_batch_fp64_pi_instances(gpu, monkeypatch):
    """The FP64 batch runs its Pi with two frequencies per tile (grouped instance 1) when nf is
    even; the one-frequency grouped instance (QXGW_PI_VARIANT=0) gives the same energies to
    1e-10 Ha (the K chunking differs, so the FP64 sums differ in rounding only)."""
    import torch
    from quasix_amd import gw
    inst = _instances(torch, seed0=320)
    bs = gw.G0W0BatchSolver(inst, nfreq=NF, precision="fp64")
    bs.run()
    assert bs.stats.pi_variant == 1
    pair = bs.results()
    monkeypatch.setenv("QXGW_PI_VARIANT", "0")
    bs0 = gw.G0W0BatchSolver(inst, nfreq=NF, precision="fp64")
    bs0.run()
    assert bs0.stats.pi_variant == 0
    for q, (a, b) in enumerate(zip(pair, bs0.results())):
        _close(f"fp64 pi pair vs single inst {q}", a["qp_energies"], b["qp_energies"], 1e-10)


def test_batch_bf16x3_within_1mev_and_oracle(gpu):
    import torch
    from quasix_amd import gw
    inst = _instances(torch)
    r64 = gw.g0w0_batch(inst, nfreq=NF, precision="fp64")
    rmx = gw.g0w0_batch(inst, nfreq=NF, precision="bf16x3")
    worst = max(float(np.max(np.abs(a["qp_energies"] - b["qp_energies"])))
                for a, b in zip(r64, rmx))
    print(f"max |dE_QP| bf16x3 vs fp64 = {worst / MEV:.3e} meV")
    assert worst <= MEV
    # instance 0 against the oracle (ov-only blocks, same targets and grid)
    e, B = (t.cpu().numpy() for t in inst[0][:2])
    fr, wt = O.get_scaled_legendre_roots(NF, 0.5)
    ref = O.g0w0_cd(e, B, SHAPES[0][0], r64[0]["targets"], fr, wt)
    _close("fp64 inst 0 vs oracle", r64[0]["qp_energies"], ref["E"], 1e-6)


def test_batch_replay_follows_new_inputs(gpu):
    import torch
    from quasix_amd import gw
    inst = _instances(torch, SHAPES[:3], seed0=60)
    bs = gw.G0W0BatchSolver(inst, nfreq=NF, precision="fp64")
    bs.run()
    bs.run()                                      # captured now
    for e, B, _, _ in inst:
        B.mul_(0.9)
        e[0] -= 0.05
    torch.cuda.synchronize()
    bs.run()
    assert bs.stats.graph_replay == 1
    got = bs.results()
    ref = _single(gw, inst, "fp64")
    for q in range(len(inst)):
        _close(f"replay inst {q}", got[q]["qp_energies"], ref[q]["qp_energies"], 1e-10)


@pytest.mark.parametrize("precision", ["fp64", "bf16x3"])
def test_batch_waves_and_fallback(gpu, precision, monkeypatch):
    """Several waves through one workspace (the mixed batch also writes the BF16 split of
    M = X - I into each instance's eps block) and an instance re-solved alone."""
    import torch
    from quasix_amd import gw
    inst = _instances(torch, SHAPES[:4], seed0=80)
    # instance 2 gets an occupied level above the virtuals and a large B: eps = I - Pi is
    # indefinite at the 9 lowest frequencies (checked with the oracle: min eigenvalue -0.65),
    # so its Cholesky fails and it is re-solved alone
    e2, B2 = inst[2][0].clone(), inst[2][1] * 20.0
    e2[inst[2][2] - 1] = 3.4
    inst[2] = (e2, B2, inst[2][2], None)
    monkeypatch.setenv("QXGW_BATCH_GB", "0.02")   # forces several waves
    bs = gw.G0W0BatchSolver(inst, nfreq=NF, precision=precision)
    bs.run()
    print("waves", bs.stats.waves, "resolved", bs.stats.resolved_instances,
          "chol fallbacks", bs.stats.chol_fallbacks)
    assert bs.stats.waves > 1
    assert bs.stats.resolved_instances >= 1 and bs.stats.chol_fallbacks > 0
    got = bs.results()
    ref = _single(gw, inst, precision)
    tol = 1e-10 if precision == "fp64" else 5e-7
    for q in range(len(inst)):
        _close(f"{precision} waves inst {q}", got[q]["qp_energies"], ref[q]["qp_energies"], tol)


@pytest.mark.parametrize("variant", [0, 1, 2, 3, 4])
def test_batch_every_grouped_pi_variant(gpu, variant, monkeypatch):
    """Each grouped BF16x3 Pi kernel instance (K32, K64, double-buffered K32, K64 on 32x32x16
    MFMAs, producer/consumer with frequency pairs), forced."""
    import torch
    from quasix_amd import gw
    inst = _instances(torch, SHAPES[:4], seed0=100)
    r64 = gw.g0w0_batch(inst, nfreq=NF, precision="fp64")
    monkeypatch.setenv("QXGW_PI_VARIANT", str(variant))
    bs = gw.G0W0BatchSolver(inst, nfreq=NF, precision="bf16x3")
    bs.run()
    assert bs.stats.pi_variant == variant
    for q, (a, b) in enumerate(zip(bs.results(), r64)):
        _close(f"grouped v{variant} inst {q}", a["qp_energies"], b["qp_energies"], MEV / 100)


def test_batch_odd_frequency_count(gpu):
    """An odd number of frequencies: the autotuner must leave out the frequency-pair
    instance, and the mixed batch still agrees with FP64 within 1 meV."""
    import torch
    from quasix_amd import gw
    inst = _instances(torch, SHAPES[:3], seed0=200)
    r64 = gw.g0w0_batch(inst, nfreq=15, precision="fp64")
    bs = gw.G0W0BatchSolver(inst, nfreq=15, precision="bf16x3")
    bs.run()
    assert bs.stats.pi_variant != 4
    for q, (a, b) in enumerate(zip(bs.results(), r64)):
        _close(f"odd nf inst {q}", a["qp_energies"], b["qp_energies"], MEV)
