"""The oracle's samples inside a pass (or_pass_sampled): the pass itself is the one or_pass
makes (same path, same draws), and the last of its samples is the sample a run that samples
after each pass takes.  The samples in between are or_sample on the path at those points, so
they differ from one another and from the pass-end sample."""
import numpy as np
import pytest

import oracle as O
import pimc_configs as pc

SEED = 985456376


def system(wl):
    S = O.System(wl["replicas"][0], wl["tables"], seed=SEED, replica=0)
    return S


@pytest.mark.parametrize("order,k", [(O.ORDER_SEGMENTS, 2), (O.ORDER_SEGMENTS, 8), (O.ORDER_REFERENCE, 4),
                                     (O.ORDER_REFERENCE, 16)])
def test_pass_sampled_is_the_pass_with_samples_inside(order, k):
    wl = pc.workload("tiny")  # P = 16, segment size 8
    A, B = system(wl), system(wl)
    got = np.concatenate([A.pass_sampled(p, order, k) for p in range(2)])
    want = B.run(0, 2, order, every=1)
    assert got.shape == (2 * k, 8)
    assert np.array_equal(got[k - 1::k], want)
    for x, y in zip(A.state(), B.state()):
        assert np.array_equal(x, y)
    assert np.array_equal(A.counters()[0], B.counters()[0]) and np.array_equal(A.counters()[1], B.counters()[1])
    assert len(np.unique(got[:, 1])) == 2 * k


def test_pass_sampled_rejects_a_grid_it_does_not_divide():
    S = system(pc.workload("tiny"))
    with pytest.raises(RuntimeError, match="divide"):
        S.pass_sampled(0, O.ORDER_SEGMENTS, 3)
    with pytest.raises(RuntimeError, match="divide"):
        S.pass_sampled(0, O.ORDER_REFERENCE, 5)
