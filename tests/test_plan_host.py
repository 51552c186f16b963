"""Host-side plans and checks of the C ABI that need no device: the validation of replica
descriptions (pimc_check_replica) and the split of linear-rotor chains over workgroups
(pimc_plan_rotor_parts), run from the library on the CPU."""
import numpy as np
import pytest

import pimc_configs as pc
import pimc_hip as ph


def desc(**kw):
    base = dict(ntypes=2, numb=[8, 1], molecule=[0, 1], levels=[3, 3], mcstep=[0.15, 0.08], mass=[4.0026, 44.0],
                P=128, Prot=128, T=0.37, rtstep=0.3, pot1d=[0, -1], pot2d=[-1, 0], rotdens=0)
    base.update(kw)
    return ph.replica_desc(**base)


def test_check_accepts_the_workload_decks():
    for name in ("c0", "c1", "c2", "classes"):
        for rp in pc.workload(name, n_replicas=16)["replicas"]:
            d = ph.replica_desc(rp["ntypes"], rp["numb"], rp["molecule"], rp["levels"], rp["mcstep"], rp["mass"],
                                rp["P"], rp["Prot"], rp["T"], rp["rtstep"], rp["pot1d"], rp["pot2d"], rp["rotdens"])
            assert ph.check_replica(d) == ""


@pytest.mark.parametrize("molecule,extra", [(1, {}), (2, dict(pot3d=0, rot3d=0, pot2d=[-1, -1], rotdens=-1))])
def test_odd_rotor_slices_are_refused(molecule, extra):
    """Both checkerboard sweeps need an even number of rotor slices: with an odd count slices
    0 and Prot-1 would be moved at once although they are neighbours."""
    ok = desc(molecule=[0, molecule], P=126, Prot=126, **extra)
    assert ph.check_replica(ok) == ""
    odd = desc(molecule=[0, molecule], P=63 * 2, Prot=63, **extra)
    assert "even number of rotor slices" in ph.check_replica(odd)
    # a fixed (non-rotating) rotor has no sweep and no constraint
    assert ph.check_replica(desc(molecule=[0, molecule], P=126, Prot=0, **extra)) == ""


@pytest.mark.parametrize("kw,msg", [
    (dict(P=100, Prot=64), "NumbTimes is not proportional to NumbRotTimes"),
    (dict(levels=[7, 3]), "Segment size is larger then a number of time slices"),
    (dict(molecule=[1, 0]), "Molecules should follow atoms in input file"),
    (dict(numb=[8, 2]), "No more than one linear dopant molecule so far"),
    (dict(Prot=1, P=128), "rotational sampling needs at least two rotational slices"),
    (dict(numb=[0, 1]), "every type needs at least one atom"),
])
def test_reference_input_checks(kw, msg):
    assert ph.check_replica(desc(**kw)) == msg


def chains(wl):
    reps = wl["replicas"]
    return ([rp["Prot"] for rp in reps], [rp["P"] // rp["Prot"] for rp in reps], [sum(rp["numb"]) for rp in reps])


def covers(prot, W, ns):
    return W * ns >= prot and (W - 1) * ns < prot and ns % 2 == 0


@pytest.fixture(scope="module")
def c2_chains():
    return chains(pc.workload("c2"))


@pytest.mark.parametrize("fp64", [False, True])
@pytest.mark.parametrize("lds_kb", [40, 52, 64, 78])
def test_plan_of_the_256_replica_workload(c2_chains, fp64, lds_kb):
    """Every chain is covered by even parts within the budget; no chain has more parts than one
    XCD holds at the launch's LDS size (32 CUs, here up to 8 resident by registers)."""
    prot, rr, nat = c2_chains
    W, ns, lds = ph.plan_rotor_parts(prot, rr, nat, fp64=fp64, lds_kb=lds_kb)
    assert lds <= lds_kb * 1024
    cap = 32 * min(8, 160 * 1024 // lds)
    for p, w, s in zip(prot, W, ns):
        assert covers(p, w, s), (p, w, s)
        assert w <= cap
    # the budget alone sets these plans: N = 100 at P = 512 in FP32 takes 16 parts of 32 slices at 52 KB
    if not fp64 and lds_kb == 52:
        i = next(k for k in range(len(prot)) if nat[k] == 101 and prot[k] == 512)
        assert (W[i], ns[i]) == (16, 32)


def test_plan_caps_parts_by_residency():
    """300 partners per slice in FP64: the 52 KB budget alone cuts a P = 512 chain into 128 parts
    of 4 slices (38.8 KB each).  With 8 blocks per CU by registers, 4 per CU fit by LDS and one XCD
    holds all 128; with fewer resident blocks the plan takes fewer, larger parts."""
    W, ns, lds = ph.plan_rotor_parts([512], [1], [301], fp64=True, lds_kb=52, blocks_per_cu=8)
    assert (W[0], ns[0], lds) == (128, 4, 38752)
    W, ns, lds = ph.plan_rotor_parts([512], [1], [301], fp64=True, lds_kb=52, blocks_per_cu=2)
    assert (W[0], ns[0]) == (64, 8) and lds > 52 * 1024  # 2 per CU, 64 per XCD
    W, ns, lds = ph.plan_rotor_parts([512], [1], [301], fp64=True, lds_kb=52, blocks_per_cu=1)
    assert (W[0], ns[0]) == (32, 16) and lds <= 160 * 1024
    for bpc in (8, 4, 2, 1):
        W, ns, lds = ph.plan_rotor_parts([512, 256, 64], [1, 2, 1], [301, 120, 40], fp64=True, blocks_per_cu=bpc)
        cap = 32 * min(bpc, 160 * 1024 // lds)
        assert all(covers(p, w, s) and w <= cap for p, w, s in zip([512, 256, 64], W, ns))


@pytest.mark.parametrize("natoms,fp64", [(1000, False), (401, True)])
def test_plan_fails_with_a_message_when_nothing_fits(natoms, fp64):
    """Too many partners at P = 512: every plan needs parts beyond 160 KB or more parts than an
    XCD holds, so the commit fails with a message instead of spinning at run time."""
    with pytest.raises(ph.PimcError, match="exceeds LDS"):
        ph.plan_rotor_parts([512], [1], [natoms], fp64=fp64, lds_kb=52, blocks_per_cu=8)


def test_plan_respects_part_slices_max():
    W, ns, _ = ph.plan_rotor_parts([512, 128], [1, 1], [9, 9], part_slices_max=6)
    assert list(ns) == [6, 6] and list(W) == [86, 22]
    assert all(covers(p, w, s) for p, w, s in zip([512, 128], W, ns))


def test_rotor3d_split_fits_one_xcd():
    """One chain of two rotors at P = 2048: 64 parts of 32 slices when one XCD holds two parts per
    CU (64), but only 32 parts when it holds one per CU (the FP64 kernel's register budget) --
    the 64 parts of a chain then could not all be resident on their XCD and would wait on each
    other for ever."""
    assert ph.plan_rotor3d_split(2048, 1, 2, blocks_per_cu=2) == 64
    assert ph.plan_rotor3d_split(2048, 1, 2, blocks_per_cu=1) == 32
    # 128 chains: parts of 512 slices need 76 KB of LDS, so two share a CU when the registers allow
    # two waves per SIMD, and 4 x 128 parts fill the device's 256 CUs twice; at one block per CU
    # 2 x 128 parts of 1024 slices (92 KB) fill it once
    assert ph.plan_rotor3d_split(2048, 128, 2, blocks_per_cu=2) == 4
    assert ph.plan_rotor3d_split(2048, 128, 2, blocks_per_cu=1) == 2
    # short chains: parts keep at least 16 slices
    assert ph.plan_rotor3d_split(64, 1, 1, blocks_per_cu=2) == 4
    assert ph.plan_rotor3d_split(16, 1, 1, blocks_per_cu=2) == 1
