"""Estimator samples inside a pass.  The reference samples when its count of time steps is a
multiple of MCSKIP_AVERG (mc_main.cc:348-386), so MCSKIP_AVERG below the slice count samples
several times per pass.  With the option samples_per_pass = k the device samples after every
P/k time steps in the reference's order, or after every S/k phases in segment order; FP64 runs
follow the oracle's chain, so every sample equals the oracle's at the same point of the pass."""
import numpy as np
import pytest

import oracle as O
import pimc_configs as pc
import pimc_hip as ph

pytestmark = pytest.mark.gpu
SEED = 985456376


def ctx_with(wl, flags, k):
    c = ph.Context(len(wl["replicas"]), flags=flags, seed=SEED)
    c.set_option("samples_per_pass", k)
    return pc.load_into_context(c, wl)


@pytest.mark.parametrize("ref_order,k", [(False, 2), (False, 8), (True, 4), (True, 16)])
def test_samples_inside_a_pass_follow_the_oracle(ref_order, k):
    wl = pc.workload("tiny")  # P = 16, segment size 8
    order = O.ORDER_REFERENCE if ref_order else O.ORDER_SEGMENTS
    ctx = ctx_with(wl, ph.FLAG_FP64 | (ph.FLAG_REF_ORDER if ref_order else 0), k)
    ctx.run(2, sample_every=1)
    got = ctx.samples(0)
    S = O.System(wl["replicas"][0], wl["tables"], seed=SEED, replica=0)
    want = np.concatenate([S.pass_sampled(p, order, k) for p in range(2)])
    assert got.shape == want.shape == (2 * k, 8)
    assert np.allclose(got, want, rtol=1e-9, atol=1e-9), np.abs(got - want).max()
    assert len(np.unique(got[:, 1])) == 2 * k  # the potential moves from sample to sample


def test_last_sample_of_a_pass_is_the_pass_sample():
    """FP32 production arithmetic: the k-th sample of a pass is the one a context with the
    default (one sample after the pass) takes on the same chain."""
    wl = pc.workload("c0")
    a = ctx_with(wl, 0, 8)
    b = ctx_with(wl, 0, 1)
    for c in (a, b):
        c.run(3, sample_every=1)
    sa, sb = a.samples(0), b.samples(0)
    assert sa.shape == (24, 8) and sb.shape == (3, 8)
    assert np.array_equal(sa[7::8], sb)


def test_samples_per_pass_must_divide_the_grid():
    wl = pc.workload("tiny")
    c = ph.Context(1, seed=SEED)
    c.set_option("samples_per_pass", 3)
    with pytest.raises(ph.PimcError, match="divide"):
        pc.load_into_context(c, wl)


def test_reading_more_samples_than_the_buffer_holds_fails():
    """A replica stores 4096 samples between clears; reading after more were taken raises
    instead of returning a truncated series, and a clear starts over."""
    wl = pc.workload("tiny")
    c = ctx_with(wl, ph.FLAG_REF_ORDER, 16)  # a sample after every time step of P = 16
    c.run(257, sample_every=1)  # 4112 samples
    with pytest.raises(ph.PimcError, match="4096"):
        c.samples(0)
    c.clear_samples()
    c.run(2, sample_every=1)
    assert c.samples(0).shape == (32, 8)
