"""The launch-shape autotuner: it keeps shapes from the allowed sets, and a tuned context samples
exactly the chains an untuned one does (the state is restored after the timing runs, and no draw
depends on a launch shape)."""
import numpy as np
import pytest

import pimc_configs as pc
import pimc_hip as ph

pytestmark = pytest.mark.gpu
SEED = 985456376


def ctx_for(wl, flags=0):
    return pc.load_into_context(ph.Context(len(wl["replicas"]), flags=flags, seed=SEED), wl)


@pytest.mark.parametrize("flags", [0, ph.FLAG_FP64])
def test_autotune_keeps_the_chains(flags):
    wl = pc.workload("c2", n_replicas=24)  # N from 1 to 100: all three bisection size classes
    tuned, plain = ctx_for(wl, flags), ctx_for(wl, flags)
    chosen = tuned.autotune(passes=1)
    print("chosen", chosen)
    for k in ("bisect_lanes_small", "bisect_lanes_medium", "bisect_lanes_large"):
        assert chosen[k] in (64, 128, 256)
    assert chosen["rotor_lds_kb"] in (40, 52, 64, 78)
    assert 0 < chosen["centroid_parts"] <= 1024
    for r in range(len(wl["replicas"])):
        assert tuned.pass_counter(r) == plain.pass_counter(r) == 0
        assert np.array_equal(tuned.counters(r)[0], plain.counters(r)[0])
    for c in (tuned, plain):
        c.run(3, sample_every=1)
    for r in range(len(wl["replicas"])):
        st, sp = tuned.get_state(r), plain.get_state(r)
        for a, b in zip(st, sp):
            assert np.array_equal(a, b), r
        tt, pt = tuned.counters(r), plain.counters(r)
        assert np.array_equal(tt[0], pt[0]) and np.array_equal(tt[1], pt[1])
        # sums over pairs may be added in another order under another lane count
        assert np.allclose(tuned.samples(r), plain.samples(r), rtol=1e-12, atol=1e-9)
