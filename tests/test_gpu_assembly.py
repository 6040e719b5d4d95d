"""GPU: assembly of the symmetrised DF tensor from the iaP / miP / abP blocks on the device
(qx_assemble_blocks_dev; quasix/src/evgw.rs:161-205) against oracle.assemble_mnP and
oracle.pair_list: the ov block, the [nmo, nocc] block and the target-partner rows with
identically zero rows dropped, all bit-identical (the same one addition and halving per
entry)."""
import numpy as np
import pytest

from oracle import gw_oracle as O

pytestmark = pytest.mark.gpu


def _blocks(nocc, nvir, naux, seed, ov_only):
    e, B = O.make_inputs(nocc, nvir, naux, seed=seed)
    miP, abP = O.ov_only_blocks(B, nocc, nvir)
    if not ov_only:
        rng = np.random.default_rng(seed)
        nmo = nocc + nvir
        mi = miP.reshape(nmo, nocc, naux) + 0.05 * rng.standard_normal((nmo, nocc, naux))
        mi[1, 2] = -mi[2, 1]                  # this symmetrised oo row is exactly zero
        ab = 0.05 * rng.standard_normal((nvir, nvir, naux))
        ab[3, 0] = -ab[0, 3]                  # and this vv row
        miP, abP = mi.reshape(nmo * nocc, naux), ab.reshape(nvir * nvir, naux)
    return B, miP, abP


@pytest.mark.parametrize("nocc,nvir,naux,ov_only", [(6, 23, 70, False), (6, 23, 70, True),
                                                    (4, 9, 1, False), (5, 19, 65, False)])
def test_assemble_blocks_bitwise(gpu, nocc, nvir, naux, ov_only):
    from quasix_amd import gw
    B, miP, abP = _blocks(nocc, nvir, naux, 40 + naux, ov_only)
    nmo = nocc + nvir
    mn = O.assemble_mnP(B, miP, abP, nocc, nvir, naux)
    tg = np.array(sorted({0, 1, 2, nocc - 1, nocc, nocc + 3, nmo - 1}), dtype=np.int32)
    ov, mi, pairs = gw.assemble_blocks(B, miP, abP, nocc, nvir, naux, tg)
    pt, pm, rows = O.pair_list(mn, tg)
    print(f"assembly {nocc}/{nvir}/{naux} ov_only={ov_only}: {len(pt)} of {len(tg) * nmo} rows")
    assert np.array_equal(pairs.target.cpu().numpy(), pt)
    assert np.array_equal(pairs.m.cpu().numpy(), pm)
    assert np.array_equal(pairs.rows.cpu().numpy(), rows)
    assert np.array_equal(ov.cpu().numpy(), mn[:nocc, nocc:].reshape(nocc * nvir, naux))
    assert np.array_equal(mi.cpu().numpy(), mn[:, :nocc])
    if not ov_only and nocc > 2 and nvir > 3:
        assert not np.any((pt == list(tg).index(1)) & (pm == 2))     # the zero row is gone
