"""GPU parity at the edges of the input space: the smallest systems (one atom, two atoms, a
chain of two segments) and a dense system with many partners per atom, against the oracle's
chain (FP64, both pass orders).  MCMolecularMove mc_piqmc.cc:54-102, MCBisectionMove :194-294,
the estimators mc_estim.cc:500-687,876-937."""
import numpy as np
import pytest

import oracle as O
import pimc_configs as pc
import pimc_hip as ph
import pimc_tables as pt

pytestmark = pytest.mark.gpu
SEED = 985456376


def helium(N, P, spacing, seed, T=1.0, levels=3):
    """N He on a cubic lattice of the given spacing, beads scattered about their sites."""
    rng = np.random.default_rng(seed)
    n = int(np.ceil(N ** (1 / 3)))
    sites = np.array([[i, j, k] for i in range(n) for j in range(n) for k in range(n)], float)[:N] * spacing
    coords = np.zeros((3, N * P))
    for a in range(N):
        coords[:, a * P:(a + 1) * P] = sites[a][:, None] + rng.normal(0.0, 0.2, (3, P))
    r, v = pt.lj_table()
    rp = dict(ntypes=1, numb=[N], molecule=[0], levels=[levels], mcstep=[0.5], mass=[4.0026032497], P=P, Prot=0,
              T=T, rtstep=0.0, pot1d=[0], pot2d=[-1], rotdens=-1, coords=coords, angles=np.zeros_like(coords))
    return dict(tables=dict(pot1d=[(r, v)], pot2d=[], rot=[]), replicas=[rp])


@pytest.mark.parametrize("N,P,spacing,ref_order", [(1, 16, 3.5, False), (1, 16, 3.5, True), (2, 8, 3.5, False),
                                                   (2, 8, 3.5, True), (400, 16, 3.6, False)])
def test_edge_systems_trajectory_fp64(N, P, spacing, ref_order):
    """One atom (no partners: every level accepted on the free-particle action), two atoms,
    P = 8 with segments of 4 (two segments), and 400 atoms (the largest bisection class with
    399 partners per item row); counters, paths and the estimators after two passes."""
    wl = helium(N, P, spacing, seed=N)
    if P == 8:
        wl["replicas"][0]["levels"] = [2]
    rp = wl["replicas"][0]
    ctx = pc.load_into_context(ph.Context(1, flags=ph.FLAG_FP64 | (ph.FLAG_REF_ORDER if ref_order else 0),
                                          seed=SEED), wl)
    ctx.run(2)
    S = O.System(rp, wl["tables"], seed=SEED, replica=0)
    S.run(0, 2, order=O.ORDER_REFERENCE if ref_order else O.ORDER_SEGMENTS)
    cg, _, _ = ctx.get_state(0)
    co, _, _ = S.state()
    tg, acg = ctx.counters(0)
    to, aco = S.counters()
    assert np.array_equal(tg, to) and np.array_equal(acg, aco), (acg, aco)
    assert np.abs(cg - co).max() < 1e-9
    got = ctx.estimate()[0]
    ref = S.sample()
    for k in range(8):
        assert abs(got[k] - ref[k]) <= 1e-10 * max(1.0, abs(ref[k])), (k, got[k], ref[k])
