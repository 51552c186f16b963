"""GPU parity under the minimum image (MINIMAGE, mc_input.cc:298-300): the HIP path against the
oracle, whose minimum image is pinned to the compiled reference's (tests/test_oracle_golden.py).

Separations of atom pairs in PotEnergy (mc_piqmc.cc:1239-1240, 1832-1833), in the potential
estimator (mc_estim.cc:536-537) and the kinetic links (mc_estim.cc:916-917) take their nearest
periodic image; pairs with a non-linear rotor take the positions themselves, as vcord and caleng
do there.  FP64 runs follow the oracle's chain move for move; FP32 runs hold V of the reached
configuration to 1e-6 relative."""
import os

import numpy as np
import pytest

import oracle as O
import pimc_configs as pc
import pimc_hip as ph
import pimc_tables as pt
from golden import tables3d

pytestmark = pytest.mark.gpu
SEED = 985456376
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def helium_box(N, P, L, seed, T=1.5, levels=3):
    """N He on P slices anywhere in a cubic box of edge L (at least 2.6 Angstrom from every image
    of the others), helium.pot of the reference's examples."""
    rng = np.random.default_rng(seed)
    centres = []
    while len(centres) < N:
        x = rng.uniform(-0.5 * L, 0.5 * L, 3)
        if all(np.linalg.norm(x - c - L * np.rint((x - c) / L)) > 2.6 for c in centres):
            centres.append(x)
    coords = np.zeros((3, N * P))
    for a, c in enumerate(centres):
        coords[:, a * P:(a + 1) * P] = c[:, None] + rng.normal(0.0, 0.25, (3, P))
    r, v = pt.read_pot1d(os.path.join(GOLD, "helium.pot"))
    rp = dict(ntypes=1, numb=[N], molecule=[0], levels=[levels], mcstep=[0.6], mass=[4.0026032497], P=P, Prot=0,
              T=T, rtstep=0.0, pot1d=[0], pot2d=[-1], rotdens=-1, box=L, coords=coords,
              angles=np.zeros_like(coords))
    return dict(tables=dict(pot1d=[(r, v)], pot2d=[], rot=[]), replicas=[rp])


def gpu_ctx(wl, flags=0):
    return pc.load_into_context(ph.Context(len(wl["replicas"]), flags=flags, seed=SEED), wl)


def oracle_chain(wl, npass, order):
    S = O.System(wl["replicas"][0], wl["tables"], seed=SEED, replica=0)
    S.run(0, npass, order=order)
    return S


@pytest.mark.parametrize("N,P,L,ref_order", [(16, 16, 8.5, False), (16, 16, 8.5, True), (64, 64, 14.3, False),
                                             (100, 32, 16.6, False)])
def test_periodic_helium_trajectory_fp64(N, P, L, ref_order):
    """Liquid-density helium in a periodic box, every bisection size class: centroid and
    bisection moves under the minimum image follow the oracle's chain."""
    wl = helium_box(N, P, L, seed=N + P)
    ctx = gpu_ctx(wl, ph.FLAG_FP64 | (ph.FLAG_REF_ORDER if ref_order else 0))
    ctx.run(2)
    S = oracle_chain(wl, 2, O.ORDER_REFERENCE if ref_order else O.ORDER_SEGMENTS)
    cg, _, _ = ctx.get_state(0)
    co, _, _ = S.state()
    tg, acg = ctx.counters(0)
    to, aco = S.counters()
    assert np.array_equal(tg, to) and np.array_equal(acg, aco), (acg, aco)
    assert acg[0, 1] > 0  # some bisection moves were taken
    assert np.abs(cg - co).max() < 1e-9


@pytest.mark.parametrize("fp64", [True, False])
def test_periodic_helium_energies(fp64):
    """PotEnergy(atom, pos, it) across the box edges, and V, K and the g(r) bins of a reached
    configuration, against the oracle."""
    wl = helium_box(16, 16, 8.5, seed=5)
    rp = wl["replicas"][0]
    ctx = gpu_ctx(wl, ph.FLAG_FP64 if fp64 else 0)
    ctx.run(2)
    ctx.clear_samples()
    ctx.run(1, sample_every=1)  # one sample, of the configuration the pass ends on
    g1 = ctx.histograms(0)[0]
    cg, _, _ = ctx.get_state(0)
    got = ctx.estimate()[0]
    S = O.System(rp, wl["tables"])
    S.set_state(cg, np.zeros_like(cg))
    ref = S.sample()
    assert abs(got[0] - ref[0]) <= 1e-10 * abs(ref[0]), (got[0], ref[0])  # K: FP64 on both paths
    assert abs(got[1] - ref[1]) <= (1e-12 if fp64 else 1e-6) * abs(ref[1]), (got[1], ref[1])
    # g(r) bins the FP64 distances in both precisions
    gr1 = np.zeros(300)
    S.L.or_pot_energy_densities(S.h, gr1.ctypes.data_as(O._dp), None)
    assert np.array_equal(g1, gr1)
    # single-slice energies of displaced atoms, some pushed across a face of the box
    rng = np.random.default_rng(3)
    qa, qi, qp, want = [], [], [], []
    for a in range(16):
        for it in rng.choice(16, size=4, replace=False):
            pos = cg[:, a * 16 + it] + rng.normal(0.0, 0.3, 3)
            pos[int(rng.integers(3))] += 8.5 * rng.choice([-1.0, 1.0]) * (rng.uniform() < 0.3)
            pb = cg.copy()
            pb[:, a * 16 + it] = pos
            want.append(S.L.or_pot_energy_slice(S.h, a, O.P(pb), int(it)))
            qa.append(a); qi.append(int(it)); qp.append(pos)
    out = ctx.pot_energy_slice(np.zeros(len(qa), np.int32), qa, qi, np.array(qp))
    want = np.array(want)
    tol = 1e-11 if fp64 else 2e-5
    assert np.all(np.abs(out - want) <= tol * (np.abs(want) + 1.0)), np.max(np.abs(out - want) / (np.abs(want) + 1.0))


@pytest.fixture(scope="module")
def rot3d():
    return tables3d.rot3d()


@pytest.mark.parametrize("ref_order", [False, True])
def test_periodic_nonlinear_trajectory_fp64(rot3d, ref_order):
    """Para-H2 around an asymmetric top in a box of 11 Angstrom: the para-H2 pairs take their
    nearest image, the vcord pairs the positions; the FP64 chain follows the oracle's."""
    r, v = pt.h2h2_table()
    rng = np.random.default_rng(9)
    P, natoms = 8, 5
    coords = np.zeros((3, natoms * P))
    for a in range(natoms):
        centre = np.zeros(3) if a == natoms - 1 else np.array(
            [3.6 * np.cos(2 * np.pi * a / 4), 3.6 * np.sin(2 * np.pi * a / 4), 0.7 * (a % 2)])
        coords[:, a * P:(a + 1) * P] = centre[:, None] + rng.normal(0.0, 0.15, (3, P))
    angles = np.zeros((3, natoms * P))
    angles[0] = rng.uniform(0.0, 2.0 * np.pi, natoms * P)
    angles[1] = rng.uniform(-1.0, 1.0, natoms * P)
    angles[2] = rng.uniform(0.0, 2.0 * np.pi, natoms * P)
    rp = dict(ntypes=2, numb=[4, 1], molecule=[0, 2], levels=[2, 2], mcstep=[0.2, 0.15], mass=[2.01565, 18.010565],
              P=P, Prot=P, T=0.5, rtstep=0.25, pot1d=[0, -1], pot2d=[-1, -1], rotdens=-1, pot3d=0, rot3d=0,
              box=11.0, coords=coords, angles=angles)
    pot3d = tables3d.pot3d(nch=181)
    wl = dict(tables=dict(pot1d=[(r, v)], pot2d=[], rot=[], pot3d_list=[pot3d], rot3d_list=[rot3d]), replicas=[rp])
    ctx = gpu_ctx(wl, ph.FLAG_FP64 | (ph.FLAG_REF_ORDER if ref_order else 0))
    ctx.run(2)
    S = O.System(rp, dict(pot1d=[(r, v)], pot2d=[], rot=[], pot3d=pot3d, rot3d=rot3d), seed=SEED, replica=0)
    S.run(0, 2, order=O.ORDER_REFERENCE if ref_order else O.ORDER_SEGMENTS)
    cg, ag, _ = ctx.get_state(0)
    co, ao, _ = S.state()
    assert np.array_equal(ctx.counters(0)[1], S.counters()[1]), (ctx.counters(0)[1], S.counters()[1])
    assert np.abs(cg - co).max() < 1e-9
    rot = slice(4 * P, 5 * P)
    assert np.abs(ag[:, rot] - ao[:, rot]).max() < 1e-9


def test_minimum_image_refused_for_a_rotating_linear_rotor():
    """The reference's linear-rotor move stops under MINIMAGE (mc_piqmc.cc:2013-2016)."""
    wl = pc.workload("tiny")
    wl["replicas"][0]["box"] = 10.0
    with pytest.raises(RuntimeError, match="MIN IMAGE for orient pot"):
        gpu_ctx(wl)


@pytest.mark.parametrize("box", [0.0, -8.5, float("inf"), float("nan")])
def test_minimum_image_box_must_be_positive_and_finite(box):
    wl = helium_box(4, 16, 8.5, seed=1)
    wl["replicas"][0]["box"] = box
    with pytest.raises(RuntimeError, match="box edges must be positive"):
        gpu_ctx(wl)
