"""GPU parity at the production sizes: the HIP path against the oracle, move for move, at the
chain lengths and atom counts of the benchmark workloads.

* 64 He + linear rotor at P = 512 (the single-chain GPU workload), both pass orders.
* One replica per bisection size class and class boundary (N = 20, 30, 63, 64, 100 He; P 256
  and 512), under several lane counts per segment and rotor-part LDS budgets: the multi-wave
  level sums, the batched draws with fewer lanes than items, and rotor chains split into 16 to
  40 parts all follow the oracle's chain.
* The asymmetric top in 32 para-H2 at P = 256 on its 91-point chi grid, and two TIP4P rotors
  at P = 512 (both orders) and P = 2048.
* FP32 (the production arithmetic) on the same systems: total V of the configuration the run
  reaches within 1e-6 relative of the FP64 oracle on that configuration.

Oracle runs are cached per module, so each (system, order, passes) is restated once and the
device variants are compared with the same chain.  MCBisectionMove mc_piqmc.cc:194-294,
MCMolecularMove :54-102, MCRotationsMove :421-517, MCRotations3D :692-779 with MCRot3Dstep
:938-1199, vcord.f:76-85."""
import numpy as np
import pytest

import oracle as O
import pimc_configs as pc
import pimc_hip as ph
from golden import tables3d

pytestmark = pytest.mark.gpu
SEED = 985456376

_cache = {}


def oracle_run(tag, wl, r, npass, order, tabs=None):
    """State and counters of replica r of wl after npass oracle passes (cached by tag)."""
    key = (tag, r, npass, order)
    if key not in _cache:
        S = O.System(wl["replicas"][r], tabs or wl["tables"], seed=SEED, replica=r)
        S.run(0, npass, order=order)
        _cache[key] = (S.state(), S.counters())
    return _cache[key]


def gpu_ctx(wl, flags=0, **options):
    ctx = ph.Context(len(wl["replicas"]), flags=flags, seed=SEED)
    for k, v in options.items():
        ctx.set_option(k, v)
    return pc.load_into_context(ctx, wl)


def rotor_columns(rp):
    nat = sum(n for n, m in zip(rp["numb"], rp["molecule"]) if m == 0)
    return slice(nat * rp["P"], sum(rp["numb"]) * rp["P"])


def assert_same_chain(ctx, r, ostate, ocounters, rp, angle_rows=slice(0, 2)):
    cg, ag, _ = ctx.get_state(r)
    co, ao, _ = ostate
    tg, acg = ctx.counters(r)
    to, aco = ocounters
    assert np.array_equal(tg, to), (r, tg, to)
    assert np.array_equal(acg, aco), (r, acg, aco)
    assert np.abs(cg - co).max() < 1e-9, (r, np.abs(cg - co).max())
    rot = rotor_columns(rp)
    assert np.abs(ag[angle_rows, rot] - ao[angle_rows, rot]).max() < 1e-9, r


def assert_fp32_v(ctx, wl, tabs=None, angle_rows=slice(0, 2)):
    """V of every replica's current configuration (FP32 pair terms, FP64 sums) within 1e-6 of
    the FP64 oracle's V on that configuration."""
    got = ctx.estimate()
    for r, rp in enumerate(wl["replicas"]):
        cg, ag, _ = ctx.get_state(r)
        S = O.System(rp, tabs or wl["tables"])
        angles = np.array(rp["angles"], copy=True)
        rot = rotor_columns(rp)
        angles[angle_rows, rot] = ag[angle_rows, rot]
        S.set_state(cg, angles)
        ref = S.sample()
        assert abs(got[r][1] - ref[1]) <= 1e-6 * abs(ref[1]), (r, got[r][1], ref[1])


# ------------------------------------------------------------------------- linear family
@pytest.fixture(scope="module")
def c1():
    return pc.workload("c1")


@pytest.fixture(scope="module")
def classes():
    return pc.workload("classes")


@pytest.mark.parametrize("ref_order", [False, True])
def test_c1_trajectory_fp64(c1, ref_order):
    """64 He + rotor, P = Prot = 512: one pass, every move equal to the oracle's."""
    order = O.ORDER_REFERENCE if ref_order else O.ORDER_SEGMENTS
    ctx = gpu_ctx(c1, ph.FLAG_FP64 | (ph.FLAG_REF_ORDER if ref_order else 0))
    ctx.run(1)
    st, cn = oracle_run("c1", c1, 0, 1, order)
    assert_same_chain(ctx, 0, st, cn, c1["replicas"][0])


def test_c1_fp32_total_v(c1):
    ctx = gpu_ctx(c1, 0)
    ctx.run(2)
    assert_fp32_v(ctx, c1)


SHAPES = {
    "default": {},
    "lanes256": dict(bisect_lanes_small=256, bisect_lanes_medium=256, bisect_lanes_large=256),
    "lanes64_rotor40": dict(bisect_lanes_small=64, bisect_lanes_medium=64, bisect_lanes_large=64, rotor_lds_kb=40),
    "mixed_rotor78": dict(bisect_lanes_small=128, bisect_lanes_medium=64, bisect_lanes_large=128, rotor_lds_kb=78),
}


@pytest.mark.parametrize("shape", list(SHAPES))
def test_size_classes_trajectory_fp64(classes, shape):
    """Replicas in every bisection size class, in one context, under several launch shapes."""
    ctx = gpu_ctx(classes, ph.FLAG_FP64, **SHAPES[shape])
    ctx.run(1)
    for r, rp in enumerate(classes["replicas"]):
        st, cn = oracle_run("classes", classes, r, 1, O.ORDER_SEGMENTS)
        assert_same_chain(ctx, r, st, cn, rp)


def test_size_classes_trajectory_fp64_ref_order(classes):
    ctx = gpu_ctx(classes, ph.FLAG_FP64 | ph.FLAG_REF_ORDER)
    ctx.run(1)
    for r, rp in enumerate(classes["replicas"]):
        st, cn = oracle_run("classes", classes, r, 1, O.ORDER_REFERENCE)
        assert_same_chain(ctx, r, st, cn, rp)


def test_size_classes_fp32_total_v(classes):
    ctx = gpu_ctx(classes, 0)
    ctx.run(2)
    assert_fp32_v(ctx, classes)


# --------------------------------------------------------------------- non-linear family
def _nl_tabs(wl):
    t = wl["tables"]
    return dict(pot1d=t["pot1d"], pot2d=[], rot=[], pot3d=t["pot3d_list"][0] if t["pot3d_list"] else None,
                rot3d=t["rot3d_list"][0])


@pytest.fixture(scope="module")
def generator():
    """A context used only for the device density-matrix generator (pimc_asymrho)."""
    g = ph.Context(1)
    yield g
    g.close()


def _asymrho(gen):
    return lambda T, P, A, B, C, maxj: gen.asymrho(T, P, A, B, C, maxj)


@pytest.fixture(scope="module")
def c3(generator):
    return pc.nl_workload("c3", n_replicas=1, asymrho=_asymrho(generator))


def test_c3_trajectory_fp64(c3):
    """Asymmetric top in 32 para-H2 at P = Prot = 256 on the production 100 x 181 x 91 table."""
    assert c3["tables"]["pot3d_list"][0][2] == 91
    ctx = gpu_ctx(c3, ph.FLAG_FP64)
    ctx.run(1)
    st, cn = oracle_run("c3", c3, 0, 1, O.ORDER_SEGMENTS, _nl_tabs(c3))
    assert_same_chain(ctx, 0, st, cn, c3["replicas"][0], angle_rows=slice(0, 3))


def test_c3_fp32_total_v(c3):
    ctx = gpu_ctx(c3, 0)
    ctx.run(2)
    assert_fp32_v(ctx, c3, _nl_tabs(c3), angle_rows=slice(0, 3))


@pytest.fixture(scope="module")
def water512():
    """Two TIP4P rotors at P = Prot = 512 on the stand-in density tables."""
    rng = np.random.default_rng(17)
    P = 512
    coords = np.zeros((3, 2 * P))
    for a, cx in enumerate((-1.45, 1.45)):
        coords[:, a * P:(a + 1) * P] = np.array([cx, 0.0, 0.3 * a])[:, None] + rng.normal(0.0, 0.05, (3, P))
    angles = np.zeros((3, 2 * P))
    angles[0] = rng.uniform(0.0, 2.0 * np.pi, 2 * P)
    angles[1] = rng.uniform(-1.0, 1.0, 2 * P)
    angles[2] = rng.uniform(0.0, 2.0 * np.pi, 2 * P)
    rp = dict(ntypes=1, numb=[2], molecule=[2], levels=[3], mcstep=[0.05], mass=[pc.MASS_H2O], P=P, Prot=P, T=1.0,
              rtstep=0.3, pot1d=[-1], pot2d=[-1], rotdens=-1, pot3d=-1, rot3d=0, coords=coords, angles=angles)
    return dict(tables=dict(pot1d=[], pot2d=[], rot=[], pot3d_list=[], rot3d_list=[tables3d.rot3d()]), replicas=[rp])


@pytest.mark.parametrize("ref_order", [False, True])
def test_water_p512_trajectory_fp64(water512, ref_order):
    order = O.ORDER_REFERENCE if ref_order else O.ORDER_SEGMENTS
    ctx = gpu_ctx(water512, ph.FLAG_FP64 | (ph.FLAG_REF_ORDER if ref_order else 0))
    ctx.run(1)
    st, cn = oracle_run("water512", water512, 0, 1, order, _nl_tabs(water512))
    assert_same_chain(ctx, 0, st, cn, water512["replicas"][0], angle_rows=slice(0, 3))


@pytest.fixture(scope="module")
def c4(generator):
    return pc.nl_workload("c4", n_replicas=1, asymrho=_asymrho(generator))


def test_c4_trajectory_fp64(c4):
    """Two TIP4P rotors at P = Prot = 2048 on the asymmetric-top density of T = 1 K: one pass."""
    ctx = gpu_ctx(c4, ph.FLAG_FP64)
    ctx.run(1)
    st, cn = oracle_run("c4", c4, 0, 1, O.ORDER_SEGMENTS, _nl_tabs(c4))
    assert_same_chain(ctx, 0, st, cn, c4["replicas"][0], angle_rows=slice(0, 3))


def test_c4_fp32_total_v(c4):
    ctx = gpu_ctx(c4, 0)
    ctx.run(1)
    assert_fp32_v(ctx, c4, _nl_tabs(c4), angle_rows=slice(0, 3))


@pytest.mark.parametrize("nrep", [128, 256])
def test_c4_many_replicas_trajectory_fp64(generator, nrep):
    """The full two-water workload (128 chains of 2048 slices: four parts of 512 slices per chain,
    two parts per CU, one lane per move) and twice as many chains (two parts of 1024 slices, so
    each lane makes two moves per half-sweep).  A replica's chain after one pass against the
    oracle's."""
    wl = pc.nl_workload("c4", n_replicas=nrep, asymrho=_asymrho(generator))
    ctx = gpu_ctx(wl, ph.FLAG_FP64)
    ctx.run(1)
    r = 77
    st, cn = oracle_run(f"c4_{nrep}", wl, r, 1, O.ORDER_SEGMENTS, _nl_tabs(wl))
    assert_same_chain(ctx, r, st, cn, wl["replicas"][r], angle_rows=slice(0, 3))
