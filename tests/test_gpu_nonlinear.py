"""GPU parity for non-linear rotors: the HIP path against the oracle through the C ABI.

Systems: para-H2 around one asymmetric-top rotor (atom/rotor pairs through vcord on a
R x 181 x 181 table) and two rigid TIP4P rotors (rotor/rotor pairs through caleng).  The
density-matrix and potential tables are the deterministic stand-ins of tests/golden/tables3d.py
in the reference's layouts.  FP64 runs follow the oracle's chain move for move; FP32 runs hold
per-configuration energies to the workload tolerance (V within 1e-6 relative)."""
import numpy as np
import pytest

import oracle as O
import pimc_configs as pc
import pimc_hip as ph
import pimc_tables as pt
from golden import tables3d

pytestmark = pytest.mark.gpu
SEED = 985456376


@pytest.fixture(scope="module")
def rot3d():
    return tables3d.rot3d()


def nl_workload(kind, rot3d, n_h2=4, P=8, Prot=8, seed=7, nch=181):
    """One replica of the para-H2 + rotor system ("h2o_h2") or the two-rotor system ("water2");
    nch is the chi grid of the 3D potential (91, 181 or 361 points: vcord.f:50-85 folds chi by it)."""
    r, v = pt.h2h2_table()
    pot3d = tables3d.pot3d(nch=nch)
    if kind == "h2o_h2":
        rp = dict(ntypes=2, numb=[n_h2, 1], molecule=[0, 2], levels=[2, 2], mcstep=[0.2, 0.15],
                  mass=[2.01565, 18.010565], P=P, Prot=Prot, T=0.5, rtstep=0.25, pot1d=[0, -1], pot2d=[-1, -1],
                  rotdens=-1, pot3d=0, rot3d=0)
    else:
        rp = dict(ntypes=1, numb=[2], molecule=[2], levels=[2], mcstep=[0.15], mass=[18.010565], P=P, Prot=Prot,
                  T=1.0, rtstep=0.25, pot1d=[-1], pot2d=[-1], rotdens=-1, pot3d=-1, rot3d=0)
    natoms = sum(rp["numb"])
    rng = np.random.default_rng(seed)
    coords = np.zeros((3, natoms * P))
    for a in range(natoms):
        if kind == "h2o_h2" and a == natoms - 1:
            centre = np.zeros(3)
        else:
            ang = 2.0 * np.pi * a / natoms
            rad = 3.6 if kind == "h2o_h2" else 1.6
            centre = np.array([rad * np.cos(ang), rad * np.sin(ang), 0.5 * (a % 2)])
        coords[:, a * P:(a + 1) * P] = centre[:, None] + rng.normal(0.0, 0.12, (3, P))
    angles = np.zeros((3, natoms * P))
    angles[0] = rng.uniform(0.0, 2.0 * np.pi, natoms * P)
    angles[1] = rng.uniform(-1.0, 1.0, natoms * P)
    angles[2] = rng.uniform(0.0, 2.0 * np.pi, natoms * P)
    rp["coords"], rp["angles"] = coords, angles
    tables = dict(pot1d=[(r, v)], pot2d=[], rot=[], pot3d_list=[pot3d] if kind == "h2o_h2" else [],
                  rot3d_list=[rot3d])
    return dict(tables=tables, replicas=[rp])


def oracle_system(wl, r=0):
    rp = wl["replicas"][r]
    t = wl["tables"]
    tabs = dict(pot1d=t["pot1d"], pot2d=[], rot=[], pot3d=t["pot3d_list"][0] if t["pot3d_list"] else None,
                rot3d=t["rot3d_list"][0])
    S = O.System(rp, tabs, seed=SEED, replica=r)
    return S


def gpu_ctx(wl, flags=0):
    return pc.load_into_context(ph.Context(len(wl["replicas"]), flags=flags, seed=SEED), wl)


@pytest.mark.parametrize("fp64", [True, False])
@pytest.mark.parametrize("kind,nch", [("h2o_h2", 181), ("h2o_h2", 91), ("h2o_h2", 361), ("water2", 181)])
def test_nl_pot_energy_slice(rot3d, kind, nch, fp64):
    """PotEnergy(atom, pos, it) mc_piqmc.cc:1796-1965 with vcord / caleng partners, on each chi
    grid vcord folds to (91: C2v quadrant, 181: half turn, 361: full turn; vcord.f:76-85)."""
    wl = nl_workload(kind, rot3d, nch=nch)
    rp = wl["replicas"][0]
    ctx = gpu_ctx(wl, ph.FLAG_FP64 if fp64 else 0)
    S = oracle_system(wl)
    c = rp["coords"]
    P, N = rp["P"], S.natoms
    rng = np.random.default_rng(3)
    qa, qi, qp, ref = [], [], [], []
    for atom in range(N):
        for it in range(P):
            pos = c[:, atom * P + it] + rng.normal(0, 0.2, 3)
            pb = c.copy()
            pb[:, atom * P + it] = pos
            ref.append(S.L.or_pot_energy_slice(S.h, atom, O.P(pb), int(it)))
            qa.append(atom); qi.append(it); qp.append(pos)
    got = ctx.pot_energy_slice(np.zeros(len(qa), np.int32), qa, qi, np.array(qp))
    ref = np.array(ref)
    tol = 1e-11 if fp64 else 2e-5
    assert np.all(np.abs(got - ref) <= tol * (np.abs(ref) + 1.0)), np.max(np.abs(got - ref) / (np.abs(ref) + 1.0))


@pytest.mark.parametrize("fp64", [True, False])
@pytest.mark.parametrize("kind", ["h2o_h2", "water2"])
def test_nl_estimators(rot3d, kind, fp64):
    """MCGetAverage with GetRotE3D (mc_estim.cc:989-1096) and the 3D pair terms of V."""
    wl = nl_workload(kind, rot3d)
    ctx = gpu_ctx(wl, ph.FLAG_FP64 if fp64 else 0)
    S = oracle_system(wl)
    got = ctx.estimate()[0]
    ref = S.sample()
    tol = 1e-10 if fp64 else 1e-6
    for k in (0, 2, 3, 7):  # K and the rotational terms are FP64 on both paths
        assert abs(got[k] - ref[k]) <= 1e-10 * max(1.0, abs(ref[k])), (k, got[k], ref[k])
    assert abs(got[1] - ref[1]) <= tol * max(1.0, abs(ref[1])), (got[1], ref[1])


@pytest.mark.parametrize("ref_order", [False, True])
@pytest.mark.parametrize("kind,nch", [("h2o_h2", 181), ("h2o_h2", 91), ("h2o_h2", 361), ("water2", 181)])
def test_nl_trajectory_fp64(rot3d, kind, nch, ref_order):
    """Centroid, bisection and MCRotations3D moves on counter-based draws follow the oracle."""
    wl = nl_workload(kind, rot3d, nch=nch)
    flags = ph.FLAG_FP64 | (ph.FLAG_REF_ORDER if ref_order else 0)
    ctx = gpu_ctx(wl, flags)
    ctx.run(2)
    S = oracle_system(wl)
    S.run(0, 2, order=O.ORDER_REFERENCE if ref_order else O.ORDER_SEGMENTS)
    cg, ag, _ = ctx.get_state(0)
    co, ao, _ = S.state()
    tg, acg = ctx.counters(0)
    to, aco = S.counters()
    assert np.array_equal(tg, to), (tg, to)
    assert np.array_equal(acg, aco), (acg, aco)
    assert np.abs(cg - co).max() < 1e-9
    rot = rotor_columns(wl["replicas"][0])
    assert np.abs(ag[:, rot] - ao[:, rot]).max() < 1e-9


def rotor_columns(rp):
    """Entries of the rotors in [3][natoms*P] (atoms carry no orientation on the device)."""
    nat = sum(n for n, m in zip(rp["numb"], rp["molecule"]) if m == 0)
    return slice(nat * rp["P"], sum(rp["numb"]) * rp["P"])


def test_nl_larger_system_fp64(rot3d):
    """32 para-H2 and a rotor on 32 slices: more lane groups per slice and longer sums."""
    wl = nl_workload("h2o_h2", rot3d, n_h2=32, P=32, Prot=32, seed=11)
    ctx = gpu_ctx(wl, ph.FLAG_FP64)
    ctx.run(1)
    S = oracle_system(wl)
    S.run(0, 1, order=O.ORDER_SEGMENTS)
    cg, ag, _ = ctx.get_state(0)
    co, ao, _ = S.state()
    assert np.array_equal(ctx.counters(0)[1], S.counters()[1])
    assert np.abs(cg - co).max() < 1e-9
    rot = rotor_columns(wl["replicas"][0])
    assert np.abs(ag[:, rot] - ao[:, rot]).max() < 1e-9


def test_mixed_families_in_one_context(rot3d):
    """A linear-rotor replica and a non-linear one side by side keep their own chains."""
    lin = pc.workload("tiny")
    nl = nl_workload("h2o_h2", rot3d)
    t = dict(pot1d=lin["tables"]["pot1d"] + nl["tables"]["pot1d"], pot2d=lin["tables"]["pot2d"],
             rot=lin["tables"]["rot"], pot3d_list=nl["tables"]["pot3d_list"], rot3d_list=nl["tables"]["rot3d_list"])
    r_nl = dict(nl["replicas"][0])
    r_nl["pot1d"] = [1, -1]
    wl = dict(tables=t, replicas=[lin["replicas"][0], r_nl])
    ctx = gpu_ctx(wl, ph.FLAG_FP64)
    ctx.run(2)
    S0 = O.System(lin["replicas"][0], lin["tables"], seed=SEED, replica=0)
    S0.run(0, 2, order=O.ORDER_SEGMENTS)
    S1 = oracle_system(nl, 0)
    O.lib().or_set_rng(S1.h, SEED, 1)
    S1.run(0, 2, order=O.ORDER_SEGMENTS)
    for r, S in ((0, S0), (1, S1)):
        cg, ag, _ = ctx.get_state(r)
        co, ao, _ = S.state()
        assert np.array_equal(ctx.counters(r)[1], S.counters()[1]), r
        assert np.abs(cg - co).max() < 1e-9, r
        rot = rotor_columns(wl["replicas"][r])
        assert np.abs(ag[:2, rot] - ao[:2, rot]).max() < 1e-9, r


@pytest.mark.parametrize("kind", ["h2o_h2", "water2"])
def test_restore_state_continues_exactly(rot3d, kind):
    """pimc_get_state + pimc_get_pass into a fresh context, then more passes: bit-identical to
    one context running all passes (rotation matrices rebuilt on the device)."""
    wl = nl_workload(kind, rot3d)
    one = gpu_ctx(wl)
    one.run(4)
    two = gpu_ctx(wl)
    two.run(2)
    c, a, s = two.get_state(0)
    p = two.pass_counter(0)
    assert p == 2
    three = gpu_ctx(wl)
    three.restore_state(0, c, a, s, p)
    three.run(2)
    c1, a1, s1 = one.get_state(0)
    c3, a3, s3 = three.get_state(0)
    assert np.array_equal(c1, c3) and np.array_equal(a1, a3) and np.array_equal(s1, s3)


@pytest.mark.parametrize("ref_order", [False, True])
@pytest.mark.parametrize("kind", ["h2o_h2", "water2"])
def test_nl_trajectory_split_chain_fp64(rot3d, kind, ref_order):
    """P = Prot = 64: the rotor chain is split into parts of 16 slices that trade boundary frames
    each half-sweep; FP64 trajectories still follow the oracle move for move."""
    wl = nl_workload(kind, rot3d, n_h2=6, P=64, Prot=64, seed=5)
    flags = ph.FLAG_FP64 | (ph.FLAG_REF_ORDER if ref_order else 0)
    ctx = gpu_ctx(wl, flags)
    ctx.run(2)
    S = oracle_system(wl)
    S.run(0, 2, order=O.ORDER_REFERENCE if ref_order else O.ORDER_SEGMENTS)
    cg, ag, _ = ctx.get_state(0)
    co, ao, _ = S.state()
    tg, acg = ctx.counters(0)
    to, aco = S.counters()
    assert np.array_equal(tg, to) and np.array_equal(acg, aco), (acg, aco)
    assert np.abs(cg - co).max() < 1e-9
    rot = rotor_columns(wl["replicas"][0])
    assert np.abs(ag[:, rot] - ao[:, rot]).max() < 1e-9


@pytest.mark.parametrize("kind,P,ratio,ref_order", [("h2o_h2", 16, 2, False), ("h2o_h2", 16, 4, True),
                                                    ("water2", 16, 2, True), ("h2o_h2", 128, 2, False),
                                                    ("water2", 128, 4, False)])
def test_nl_trajectory_rotor_ratio_fp64(rot3d, kind, P, ratio, ref_order):
    """NumbRotTimes = NumbTimes / ratio (RotRatio, mc_setup.cc:379-384): a rotor slice's partner
    energy sums over its `ratio` translational slices (mc_piqmc.cc:1071-1072); at P = 128 the
    rotor chain is split into parts.  FP64 trajectories follow the oracle move for move
    (orientations compared on the NumbRotTimes rotor slices)."""
    wl = nl_workload(kind, rot3d, n_h2=4, P=P, Prot=P // ratio, seed=9)
    rp = wl["replicas"][0]
    flags = ph.FLAG_FP64 | (ph.FLAG_REF_ORDER if ref_order else 0)
    ctx = gpu_ctx(wl, flags)
    ctx.run(2)
    S = oracle_system(wl)
    S.run(0, 2, order=O.ORDER_REFERENCE if ref_order else O.ORDER_SEGMENTS)
    cg, ag, _ = ctx.get_state(0)
    co, ao, _ = S.state()
    tg, acg = ctx.counters(0)
    to, aco = S.counters()
    assert np.array_equal(tg, to) and np.array_equal(acg, aco), (acg, aco)
    assert np.abs(cg - co).max() < 1e-9
    nat = sum(n for n, m in zip(rp["numb"], rp["molecule"]) if m == 0)
    for q in range(sum(rp["numb"]) - nat):
        a0 = (nat + q) * P
        assert np.abs(ag[:, a0:a0 + rp["Prot"]] - ao[:, a0:a0 + rp["Prot"]]).max() < 1e-9, q


@pytest.mark.parametrize("kind,ratio", [("h2o_h2", 2), ("water2", 4)])
def test_nl_estimators_fp32_rotor_ratio(rot3d, kind, ratio):
    """Production arithmetic with fewer rotor slices (P = 128, split chains): after three FP32
    passes, V of the reached configuration equals the FP64 oracle's to 1e-6 relative, and K and
    the rotational terms to 1e-10."""
    P = 128
    wl = nl_workload(kind, rot3d, n_h2=4, P=P, Prot=P // ratio, seed=9)
    rp = wl["replicas"][0]
    ctx = gpu_ctx(wl, 0)
    ctx.run(3)
    got = ctx.estimate()[0]
    cg, ag, _ = ctx.get_state(0)
    angles = np.array(rp["angles"], copy=True)
    nat = sum(n for n, m in zip(rp["numb"], rp["molecule"]) if m == 0)
    for q in range(sum(rp["numb"]) - nat):
        a0 = (nat + q) * P
        angles[:, a0:a0 + rp["Prot"]] = ag[:, a0:a0 + rp["Prot"]]
    S = oracle_system(wl)
    S.set_state(cg, angles)
    ref = S.sample()
    for k in (0, 2, 3, 7):  # K and the rotational terms are FP64 on both paths
        assert abs(got[k] - ref[k]) <= 1e-10 * max(1.0, abs(ref[k])), (k, got[k], ref[k])
    assert abs(got[1] - ref[1]) <= 1e-6 * max(1.0, abs(ref[1])), (got[1], ref[1])


def test_split_chain_restore_continues_exactly(rot3d):
    wl = nl_workload("water2", rot3d, P=64, Prot=64, seed=5)
    one = gpu_ctx(wl)
    one.run(3)
    two = gpu_ctx(wl)
    two.run(1)
    c, a, s = two.get_state(0)
    three = gpu_ctx(wl)
    three.restore_state(0, c, a, s, two.pass_counter(0))
    three.run(2)
    for x, y in zip(one.get_state(0), three.get_state(0)):
        assert np.array_equal(x, y)


N3D = 300 * 50 * 100


@pytest.mark.parametrize("kind", ["h2o_h2", "water2"])
def test_nl_density_bins_fp64(rot3d, kind):
    """GetPotEnergy_Densities' (r, theta, chi) bins around the rotor and GetRotE3D's relative
    Euler angle bins, accumulated over three sampled passes, equal the oracle's on the same
    chain (FP64 runs follow it move for move)."""
    wl = nl_workload(kind, rot3d)
    ctx = gpu_ctx(wl, ph.FLAG_FP64)
    ctx.clear_samples()
    ctx.run(3, sample_every=1)
    g3, rel = ctx.histograms3d(0)
    S = oracle_system(wl)
    w3, wr = np.zeros(N3D), np.zeros(250)
    for p in range(3):
        S.run(p, 1)
        assert S.L.or_nl_densities(S.h, O.P(w3), O.P(wr)) == 0
    nrot = 1 if kind == "h2o_h2" else 2
    if kind == "h2o_h2":
        assert g3 is not None and np.array_equal(g3.reshape(-1), w3)
        assert w3.sum() == 3 * 4 * 8  # every H2 within 15 A at every slice
    else:
        assert g3 is None and not w3.any()
    assert rel is not None and np.array_equal(np.concatenate(rel), wr)
    assert wr[:50].sum() == 3 * nrot * 8


def test_nl_density_bins_fp32_configuration(rot3d):
    """An FP32 run bins from the FP64 coordinates and frames: one sampled pass, then the
    oracle's bins of the configuration the device holds."""
    wl = nl_workload("h2o_h2", rot3d)
    ctx = gpu_ctx(wl)
    ctx.run(2)
    ctx.clear_samples()
    ctx.run(1, sample_every=1)
    g3, rel = ctx.histograms3d(0)
    cg, ag, _ = ctx.get_state(0)
    rp = wl["replicas"][0]
    S = oracle_system(wl)
    angles = np.array(rp["angles"], copy=True)
    rot = rotor_columns(rp)
    angles[:, rot] = ag[:, rot]
    S.set_state(cg, angles)
    w3, wr = np.zeros(N3D), np.zeros(250)
    assert S.L.or_nl_densities(S.h, O.P(w3), O.P(wr)) == 0
    assert np.array_equal(g3.reshape(-1), w3) and np.array_equal(np.concatenate(rel), wr)
