"""pimc_pot_energy_slice_noalloc: the PotEnergy(atom, pos, it) queries of pimc_pot_energy_slice
(mc_piqmc.cc:1796-1965) staged in the library's per-device array instead of per-call device
buffers.  Checked against the allocating entry point (same kernel: bit-identical results) and
against the FP64 oracle, on two calls in a row, the second smaller, so stale staging from the
first call cannot pass for the second's results; and the size limit is refused with an error."""
import numpy as np
import pytest

import oracle as O
import pimc_configs as pc
import pimc_hip as ph

pytestmark = pytest.mark.gpu
SEED = 985456376
MAXQ = 16384


def queries(rp, n, seed):
    rng = np.random.default_rng(seed)
    P, N = rp["P"], sum(rp["numb"])
    qa = rng.integers(0, N, n).astype(np.int32)
    qi = rng.integers(0, P, n).astype(np.int32)
    c = rp["coords"]
    pos = np.array([c[:, a * P + it] for a, it in zip(qa, qi)]) + rng.normal(0.0, 0.3, (n, 3))
    return np.zeros(n, np.int32), qa, qi, pos


def oracle_values(rp, tables, qa, qi, pos):
    S = O.System(rp, tables)
    c = rp["coords"]
    P = rp["P"]
    out = []
    for a, it, p in zip(qa, qi, pos):
        pb = c.copy()
        pb[:, a * P + it] = p
        out.append(S.L.or_pot_energy_slice(S.h, int(a), O.P(pb), int(it)))
    return np.array(out)


@pytest.mark.parametrize("fp64", [True, False])
def test_noalloc_matches_allocating_entry_point(fp64):
    wl = pc.workload("c0")
    rp = wl["replicas"][0]
    ctx = pc.load_into_context(ph.Context(1, flags=ph.FLAG_FP64 if fp64 else 0, seed=SEED), wl)
    for n, seed in ((3000, 1), (257, 2)):  # the second call is smaller than the first
        qr, qa, qi, pos = queries(rp, n, seed)
        got = ctx.pot_energy_slice(qr, qa, qi, pos, noalloc=True)
        ref = ctx.pot_energy_slice(qr, qa, qi, pos)
        assert got.shape == (n,)
        assert np.array_equal(got, ref)  # the same kernel on the same queries: bit for bit
        if fp64:  # the FP64 path against the oracle's PotEnergy (FP32 pair terms: per-config V bar)
            want = oracle_values(rp, wl["tables"], qa, qi, pos)
            assert np.all(np.abs(got - want) <= 1e-11 * (np.abs(want) + 1.0)), np.max(np.abs(got - want))
    ctx.close()


def test_noalloc_refuses_more_than_the_staging_holds():
    wl = pc.workload("tiny")
    ctx = pc.load_into_context(ph.Context(1, seed=SEED), wl)
    qr, qa, qi, pos = queries(wl["replicas"][0], MAXQ + 1, 3)
    with pytest.raises(ph.PimcError, match="at most 16384 queries"):
        ctx.pot_energy_slice(qr, qa, qi, pos, noalloc=True)
    # the limit itself is accepted, and agrees with the allocating path
    got = ctx.pot_energy_slice(qr[:MAXQ], qa[:MAXQ], qi[:MAXQ], pos[:MAXQ], noalloc=True)
    ref = ctx.pot_energy_slice(qr[:MAXQ], qa[:MAXQ], qi[:MAXQ], pos[:MAXQ])
    assert np.array_equal(got, ref)
    ctx.close()
