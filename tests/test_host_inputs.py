"""CPU: the package's synthetic-input generator draws exactly what the oracle's does, and the
batched workload has the stated shape distribution."""
import numpy as np

from oracle import gw_oracle as O
from quasix_amd import synthetic


def test_make_inputs_matches_oracle():
    for shape, seed in [((5, 19, 84), 0), ((7, 45, 150), 11), ((3, 10, 40), 5)]:
        e1, B1 = synthetic.make_inputs(*shape, seed=seed)
        e2, B2 = O.make_inputs(*shape, seed=seed)
        assert np.array_equal(e1, e2) and np.array_equal(B1, B2)


def test_batch_shapes():
    s = synthetic.batch_shapes(256, seed=2)
    assert len(s) == 256
    nocc = np.array([o for o, _, _ in s])
    assert nocc.min() >= 8 and nocc.max() <= 60
    assert all(v == 9 * o and a == 4 * (o + v) for o, v, a in s)
    assert len(set(nocc)) > 30                       # uneven sizes
