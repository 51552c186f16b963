"""GPU parity: the HIP path against the oracle, through the C ABI.

FP64 runs (PIMC_FLAG_FP64) must follow the oracle's trajectory: same moves, same counter-based
random numbers, same accept/reject decisions, coordinates equal to round-off.  FP32 runs (the
production arithmetic) are held to the energy tolerance of the workload definition: total V of
a configuration within 1e-6 relative of the FP64 oracle."""
import numpy as np
import pytest

import oracle as O
import pimc_configs as pc
import pimc_hip as ph

pytestmark = pytest.mark.gpu
SEED = 985456376


def perturbed(rp, seed, spread=0.25):
    rng = np.random.default_rng(seed)
    c = rp["coords"] + rng.normal(0.0, spread, rp["coords"].shape)
    a = rp["angles"].copy()
    n = a.shape[1]
    a[0] = rng.uniform(0, 2 * np.pi, n)
    a[1] = rng.uniform(-1, 1, n)
    a[2] = 0.0
    return c, a


def gpu_ctx(wl, flags=0, seed=SEED, **options):
    ctx = ph.Context(len(wl["replicas"]), flags=flags, seed=seed)
    for k, v in options.items():
        ctx.set_option(k, v)
    return pc.load_into_context(ctx, wl)


@pytest.mark.parametrize("fp64", [True, False])
@pytest.mark.parametrize("name", ["tiny", "c0"])
def test_pot_energy_slice(name, fp64):
    """PotEnergy(atom, pos, it) mc_piqmc.cc:1796 for every atom on a sample of slices."""
    wl = pc.workload(name)
    rp = wl["replicas"][0]
    c, a = perturbed(rp, 3)
    ctx = gpu_ctx(wl, ph.FLAG_FP64 if fp64 else 0)
    ctx.set_state(0, c, a)
    S = O.System(rp, wl["tables"])
    S.set_state(c, a)
    P, N = rp["P"], S.natoms
    rng = np.random.default_rng(5)
    qa, qi, qp, ref = [], [], [], []
    for atom in range(N):
        for it in rng.choice(P, size=min(P, 12), replace=False):
            pos = c[:, atom * P + it] + rng.normal(0, 0.2, 3)
            pb = c.copy()
            pb[:, atom * P + it] = pos
            ref.append(S.L.or_pot_energy_slice(S.h, atom, O.P(pb), int(it)))
            qa.append(atom); qi.append(int(it)); qp.append(pos)
    got = ctx.pot_energy_slice(np.zeros(len(qa), np.int32), qa, qi, np.array(qp))
    ref = np.array(ref)
    tol = 1e-11 if fp64 else 2e-5
    assert np.all(np.abs(got - ref) <= tol * (np.abs(ref) + 1.0)), np.max(np.abs(got - ref) / (np.abs(ref) + 1.0))


@pytest.mark.parametrize("fp64", [True, False])
@pytest.mark.parametrize("name", ["tiny", "c0", "c1"])
def test_estimators(name, fp64):
    """K, V, E_rot and the Cv terms of MCGetAverage on a fixed path; V within 1e-6 relative in FP32."""
    wl = pc.workload(name)
    rp = wl["replicas"][0]
    c, a = perturbed(rp, 7, spread=0.15)
    ctx = gpu_ctx(wl, ph.FLAG_FP64 if fp64 else 0)
    ctx.set_state(0, c, a)
    S = O.System(rp, wl["tables"])
    S.set_state(c, a)
    got = ctx.estimate()[0]
    ref = S.sample()
    # kinetic and rotational terms are FP64 on both sides
    for k in (0, 2, 3, 7):
        assert abs(got[k] - ref[k]) <= 1e-10 * (abs(ref[k]) + 1.0), (k, got[k], ref[k])
    tolv = 1e-12 if fp64 else 1e-6
    assert abs(got[1] - ref[1]) <= tolv * abs(ref[1]), (got[1], ref[1], abs(got[1] - ref[1]) / abs(ref[1]))


@pytest.mark.parametrize("ref_order", [False, True])
@pytest.mark.parametrize("name,npass", [("tiny", 3), ("c0", 1)])
def test_trajectory_fp64(name, npass, ref_order):
    """Moves on identical counter-based draws: the GPU path follows the oracle's chain."""
    wl = pc.workload(name)
    rp = wl["replicas"][0]
    flags = ph.FLAG_FP64 | (ph.FLAG_REF_ORDER if ref_order else 0)
    ctx = gpu_ctx(wl, flags)
    ctx.run(npass)
    cg, ag, _ = ctx.get_state(0)
    S = O.System(rp, wl["tables"], seed=SEED, replica=0)
    S.run(0, npass, order=O.ORDER_REFERENCE if ref_order else O.ORDER_SEGMENTS)
    co, ao, _ = S.state()
    tg, acg = ctx.counters(0)
    to, aco = S.counters()
    assert np.array_equal(tg, to)
    assert np.array_equal(acg, aco), (acg, aco)
    assert np.abs(cg - co).max() < 1e-9
    assert np.abs(ag[:2] - ao[:2]).max() < 1e-9


def test_batch_trajectory_fp64():
    """Uneven replicas in one launch, each on its own generator key."""
    wl = pc.workload("mini")
    ctx = gpu_ctx(wl, ph.FLAG_FP64)
    ctx.run(2)
    for r, rp in enumerate(wl["replicas"]):
        S = O.System(rp, wl["tables"], seed=SEED, replica=r)
        S.run(0, 2, order=O.ORDER_SEGMENTS)
        cg, ag, _ = ctx.get_state(r)
        co, ao, _ = S.state()
        assert np.array_equal(ctx.counters(r)[1], S.counters()[1]), r
        assert np.abs(cg - co).max() < 1e-9, r
        assert np.abs(ag[:2] - ao[:2]).max() < 1e-9, r


@pytest.mark.parametrize("name,slices,ref_order", [("tiny", 4, False), ("tiny", 6, True), ("c0", 6, False),
                                                   ("mini", 4, False)])
def test_trajectory_fp64_rotor_parts(name, slices, ref_order):
    """Rotor chains split over several workgroups (boundary slices traded through memory
    between half-sweeps) follow the oracle's chain like the single-workgroup sweep."""
    wl = pc.workload(name)
    flags = ph.FLAG_FP64 | (ph.FLAG_REF_ORDER if ref_order else 0)
    ctx = gpu_ctx(wl, flags, rotor_part_slices=slices)
    ctx.run(2)
    for r, rp in enumerate(wl["replicas"]):
        S = O.System(rp, wl["tables"], seed=SEED, replica=r)
        S.run(0, 2, order=O.ORDER_REFERENCE if ref_order else O.ORDER_SEGMENTS)
        cg, ag, _ = ctx.get_state(r)
        co, ao, _ = S.state()
        assert np.array_equal(ctx.counters(r)[0], S.counters()[0]), r
        assert np.array_equal(ctx.counters(r)[1], S.counters()[1]), r
        assert np.abs(cg - co).max() < 1e-9, r
        assert np.abs(ag[:2] - ao[:2]).max() < 1e-9, r


def ratio_workload(name, ratio):
    """Workload `name` with NumbRotTimes = NumbTimes / ratio (RotRatio, mc_setup.cc:379-384): a
    rotor slice spans `ratio` translational slices, and its partner energy sums over all of them
    (mc_piqmc.cc:854-855); the density table is the one of (T, NumbRotTimes)."""
    wl = pc.workload(name)
    rp = wl["replicas"][0]
    rp["Prot"] = rp["P"] // ratio
    wl["tables"]["rot"] = [pc.rot_table(rp["T"], rp["Prot"])]
    rp["rotdens"] = 0
    return wl


@pytest.mark.parametrize("name,ratio,slices,ref_order", [("tiny", 2, 0, False), ("tiny", 4, 0, True),
                                                         ("tiny", 2, 4, False), ("c0", 4, 0, False),
                                                         ("c0", 2, 8, True)])
def test_trajectory_fp64_rotor_ratio(name, ratio, slices, ref_order):
    """Fewer rotor slices than translational slices, whole and split rotor chains: the GPU chain
    follows the oracle's (orientations compared on the NumbRotTimes rotor slices)."""
    wl = ratio_workload(name, ratio)
    rp = wl["replicas"][0]
    flags = ph.FLAG_FP64 | (ph.FLAG_REF_ORDER if ref_order else 0)
    ctx = gpu_ctx(wl, flags, **({"rotor_part_slices": slices} if slices else {}))
    ctx.run(2)
    S = O.System(rp, wl["tables"], seed=SEED, replica=0)
    S.run(0, 2, order=O.ORDER_REFERENCE if ref_order else O.ORDER_SEGMENTS)
    cg, ag, _ = ctx.get_state(0)
    co, ao, _ = S.state()
    assert np.array_equal(ctx.counters(0)[0], S.counters()[0])
    assert np.array_equal(ctx.counters(0)[1], S.counters()[1]), (ctx.counters(0)[1], S.counters()[1])
    assert np.abs(cg - co).max() < 1e-9
    rot = sum(rp["numb"]) - 1  # the rotor is the last atom
    cols = slice(rot * rp["P"], rot * rp["P"] + rp["Prot"])
    assert np.abs(ag[:2, cols] - ao[:2, cols]).max() < 1e-9


@pytest.mark.parametrize("name,ratio,slices", [("tiny", 2, 0), ("c0", 4, 8)])
def test_estimators_fp32_rotor_ratio(name, ratio, slices):
    """Production arithmetic with fewer rotor slices: after three FP32 passes, K and E_rot of the
    reached configuration equal the FP64 oracle's to 1e-10 and V to 1e-6 relative."""
    wl = ratio_workload(name, ratio)
    rp = wl["replicas"][0]
    ctx = gpu_ctx(wl, 0, **({"rotor_part_slices": slices} if slices else {}))
    ctx.run(3)
    got = ctx.estimate()[0]
    cg, ag, _ = ctx.get_state(0)
    rot = sum(rp["numb"]) - 1
    cols = slice(rot * rp["P"], rot * rp["P"] + rp["Prot"])
    angles = np.array(rp["angles"], copy=True)
    angles[:2, cols] = ag[:2, cols]
    S = O.System(rp, wl["tables"])
    S.set_state(cg, angles)
    ref = S.sample()
    for k in (0, 2, 3, 7):
        assert abs(got[k] - ref[k]) <= 1e-10 * (abs(ref[k]) + 1.0), (k, got[k], ref[k])
    assert abs(got[1] - ref[1]) <= 1e-6 * abs(ref[1]), (got[1], ref[1])


def test_rotor_parts_fp32_match_single_part():
    """Production arithmetic: the split changes nothing but the launch shape."""
    wl = pc.workload("c0")
    a = gpu_ctx(wl, 0)
    b = gpu_ctx(wl, 0, rotor_part_slices=8)
    a.run(3)
    b.run(3)
    ca, aa, _ = a.get_state(0)
    cb, ab, _ = b.get_state(0)
    assert np.array_equal(a.counters(0)[1], b.counters(0)[1])
    assert np.array_equal(aa, ab) and np.array_equal(ca, cb)


def test_trajectory_fp32_decisions():
    """FP32 pair terms: on a short run the accept/reject decisions still match FP64."""
    wl = pc.workload("tiny")
    ctx = gpu_ctx(wl, 0)
    ctx.run(2)
    S = O.System(wl["replicas"][0], wl["tables"], seed=SEED, replica=0)
    S.run(0, 2, order=O.ORDER_SEGMENTS)
    assert np.array_equal(ctx.counters(0)[1], S.counters()[1])
    cg, _, _ = ctx.get_state(0)
    co, _, _ = S.state()
    assert np.abs(cg - co).max() < 1e-9


def test_negative_density_is_an_error():
    """A .rot table with a negative lobe stops the run with the reference's message."""
    wl = pc.workload("tiny")
    c, rho, er, es = wl["tables"]["rot"][0]
    rho = rho.copy()
    rho[c < 0.5] = -1e-3
    wl["tables"]["rot"][0] = (c, rho, er, es)
    rp = wl["replicas"][0]
    _, a = perturbed(rp, 11)
    ctx = gpu_ctx(wl, 0)
    ctx.set_state(0, rp["coords"], a)
    with pytest.raises(ph.PimcError, match="Negative rot density"):
        ctx.run(1)


def test_rotational_correlation_of_a_sample():
    """GetRCF (mc_estim.cc:1098-1139) on the state after one sampled pass: row 0 sums
    n(t0).n(t0+t) over time origins, rows l sum P_l of it (numpy's Legendre series here)."""
    from numpy.polynomial import legendre as L
    wl = pc.workload("tiny")
    ctx = gpu_ctx(wl, ph.FLAG_FP64, SEED)
    ctx.run(2)
    ctx.clear_samples()
    ctx.run(1, sample_every=1)
    got = ctx.rcf(0)
    rp = wl["replicas"][0]
    _, _, cos = ctx.get_state(0)
    P, Prot = rp["P"], rp["Prot"]
    rot = sum(rp["numb"]) - 1
    n = cos[:, rot * P: rot * P + Prot]
    want = np.zeros((10, Prot))
    for t in range(Prot):
        p = np.sum(n * np.roll(n, -t, axis=1), axis=0)
        want[0, t] = p.sum()
        for l in range(1, 10):
            want[l, t] = L.legval(p, [0] * l + [1]).sum()
    assert np.allclose(got, want, atol=1e-10 * Prot), np.abs(got - want).max()
    assert abs(got[0, 0] - Prot) < 1e-9
