"""The host driver on the GPU against the reference program, statistically.

tests/golden/reference_runs.json holds the .eng block lines of the reference program (compiled
from the upstream sources) on the decks of tests/golden/decks.py.  The driver runs the same
deck files with several independent chains (default production path: segment order, FP32 pair
terms); the pooled block means of K, V, E_rot and E must agree with the reference's within two
combined standard errors of the mean.  Standard errors are blocking-corrected: the largest
over bins of 1 to 32 consecutive blocks that leave at least 8 bins.  The output files keep the
reference's layouts."""
import json
import os
import subprocess

import numpy as np
import pytest

from golden import decks

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "moribs-pimc_amd", "pimc")
with open(os.path.join(ROOT, "tests", "golden", "reference_runs.json")) as f:
    REF = json.load(f)

GPU_BLOCKS = {"co2_100K": 45, "he4_co2": 130, "c0_8he": 60, "he16_pbc": 130, "h2o_ph2": 40}
CHAINS = 8
PREFIX = {"co2_100K": "CO2_monomer", "he4_co2": "he4co2", "c0_8he": "c0he8", "he16_pbc": "he16pbc",
          "h2o_ph2": "h2oph2"}
NEQ = {"co2_100K": 5, "he4_co2": 10, "c0_8he": 10, "he16_pbc": 10, "h2o_ph2": 10}
SYMMETRIC = ("he4_co2", "c0_8he")  # linear-rotor decks: .grt compared through its symmetric part


def mean_se(chains):
    """Mean and blocking-corrected standard error of block values from independent chains: the
    largest estimate over bins of 1, 2, 4, ... 32 consecutive blocks that leave at least 8 bins
    (the blocking plateau; on the 8-He deck the estimate still grows up to 32 blocks)."""
    x = np.concatenate(chains)
    se = 0.0
    for b in (1, 2, 4, 8, 16, 32):
        bins = np.concatenate([c[: len(c) // b * b].reshape(-1, b).mean(axis=1) for c in chains])
        if len(bins) >= 8:
            se = max(se, bins.std(ddof=1) / np.sqrt(len(bins)))
    return x.mean(), se


def run_driver(name, d):
    """The driver on deck `name` in directory d with CHAINS independent chains: (d, [eng per chain])."""
    decks.materialise(name, d, blocks=GPU_BLOCKS[name])
    r = subprocess.run([DRIVER, "--replicas", str(CHAINS)], cwd=d, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-2000:]
    engs = []
    for c in range(CHAINS):
        engs.append(decks.read_eng(os.path.join(d, PREFIX[name] + (f"_r{c}" if c else "") + ".eng")))
    return d, engs


@pytest.fixture(scope="module")
def driver_runs(tmp_path_factory):
    """The driver on each deck with CHAINS independent chains: {name: (directory, [eng per chain])}."""
    return {name: run_driver(name, str(tmp_path_factory.mktemp(name)))
            for name in ("co2_100K", "he4_co2", "c0_8he", "he16_pbc")}


@pytest.fixture(scope="module")
def nl_runs(tmp_path_factory):
    """The driver on the rotating-H2O deck (its density tables, 1.4 GB of text, are written here
    and parsed by the driver first): {"h2o_ph2": (directory, [eng per chain])}."""
    if "h2o_ph2" not in REF:
        pytest.skip("reference_runs.json holds no runs of the rotating-H2O deck (make_reference_runs.py h2o_ph2)")
    return {"h2o_ph2": run_driver("h2o_ph2", str(tmp_path_factory.mktemp("h2o_ph2")))}


def test_atoms_only_deck_writes_the_reference_files(driver_runs):
    """Without a molecule the reference writes no block densities, only the accumulated g(r)
    (mc_main.cc:446-451,716-735)."""
    d, _ = driver_runs["he16_pbc"]
    names = os.listdir(d)
    assert "he16pbc_sum.gra" in names and "he16pbc.eng" in names
    assert not [n for n in names if n.startswith("he16pbc0") and n.endswith(".gra")]
    gra = np.loadtxt(os.path.join(d, "he16pbc_sum.gra"))
    assert gra.shape == (300, 2) and gra[:, 1].max() > 0.0


@pytest.mark.parametrize("name", ["co2_100K", "he4_co2", "c0_8he", "he16_pbc"])
def test_driver_matches_reference_statistics(name, driver_runs):
    compare_energies(name, *driver_runs[name])


@pytest.mark.timeout(900)
def test_nonlinear_driver_matches_reference_statistics(nl_runs):
    """2 para-H2 around a rotating H2O (translations of both, rotor moves of MCRot3Dstep,
    GetRotE3D): K, V, E_rot and E against the reference program's eight chains."""
    compare_energies("h2o_ph2", *nl_runs["h2o_ph2"])


def compare_energies(name, d, engs):
    neq = NEQ[name]
    for e in engs:
        assert e.shape == (GPU_BLOCKS[name] - neq, 10)
        assert np.array_equal(e[:, 0], np.arange(neq + 1, GPU_BLOCKS[name] + 1))
    # the reference's independent chains where the fixture holds several (its rotor moves
    # draw from a wall-clock seeded generator, so concurrent runs are distinct chains)
    refs = [np.array(c) for c in REF[name].get("eng_chains", [REF[name]["eng"]])]
    cols = REF[name]["columns"]
    report = []
    for col in ("kin", "pot", "rot", "total"):
        j = cols.index(col)
        mg, sg = mean_se([e[:, j] for e in engs])
        mr, sr = mean_se([r[:, j] for r in refs])
        report.append((col, mg, sg, mr, sr))
        if sr == 0 and sg == 0:
            assert mg == mr == 0.0
            continue
        assert abs(mg - mr) <= 2.0 * np.hypot(sg, sr), report
    print(name, report)


def block_columns(d, prefix, ext, blocks):
    """Column 1 of the block files <prefix><block><ext>, per chain: [chains, blocks, bins]."""
    out = []
    for c in range(CHAINS):
        stem = prefix + (f"_r{c}" if c else "")
        rows = []
        for b in blocks:
            with open(os.path.join(d, f"{stem}{b:03d}{ext}")) as f:
                text = f.read()
            if ext == ".rcf":
                text = text.split("\n\n")[0]
            rows.append(np.loadtxt(text.splitlines(), ndmin=2)[:, 1])
        out.append(rows)
    return np.array(out)


def grouped_mean_se(v, group):
    """Mean over [chains, blocks, bins] and the standard error from averages over groups of
    `group` consecutive blocks of each chain (slow modes make single blocks correlated)."""
    c, b, n = v.shape
    g = v[:, : b // group * group].reshape(c, -1, group, n).mean(axis=2).reshape(-1, n)
    return v.reshape(-1, n).mean(axis=0), g.std(axis=0, ddof=1) / np.sqrt(len(g))


@pytest.mark.parametrize("name,ext", [("he4_co2", ".gra"), ("he4_co2", ".gri"), ("he4_co2", ".grt"),
                                      ("he4_co2", ".rcf"), ("co2_100K", ".rcf"), ("c0_8he", ".gra"),
                                      ("c0_8he", ".gri"), ("c0_8he", ".grt"), ("c0_8he", ".rcf")])
def test_distributions_match_reference(name, ext, driver_runs):
    """He-He g(r), the radial and angular He density around the rotor, and the rotor's
    orientational correlation <n(tau).n(0)>: per-bin means of the driver's block files against
    the reference program's, in units of the combined standard error (both from groups of 10
    consecutive blocks).  Over the bins that hold density the sum of z^2 stays below the 0.999
    quantile of chi^2 with that many degrees of freedom, few bins pass 2 and none passes 4.5.

    The angular density is compared through its symmetric part, bins theta and pi - theta
    averaged in every block (bin i with bin 48 - i; bin 49 holds only theta = pi).  The rotor-He potential has
    only even Legendre terms, so the exact density is symmetric; its antisymmetric part is a
    slow mode (He crossing from one end of the rotor to the other) that neither run converges,
    and the two runs show it with opposite signs."""
    compare_distribution(name, ext, driver_runs[name][0])


@pytest.mark.timeout(900)
@pytest.mark.parametrize("ext", [".gra", ".gri", ".grt", ".grc", ".rcf"])
def test_nonlinear_distributions_match_reference(ext, nl_runs):
    """Around the rotating H2O: H2-H2 g(r), the H2 density in r, theta and chi of the rotor's
    frame (SaveRho1D mc_estim.cc:1233-1286) and the correlation of the rotor's axis, by the
    criterion of test_distributions_match_reference (.grt whole: the potential holds cos theta)."""
    compare_distribution("h2o_ph2", ext, nl_runs["h2o_ph2"][0])


def compare_distribution(name, ext, d):
    """Per-bin z of the driver's block files against the fixture's, and the acceptance criterion."""
    neq = NEQ[name]
    ref = REF[name]["distributions"][ext]
    v = block_columns(d, PREFIX[name], ext, range(neq + 1, GPU_BLOCKS[name] + 1))
    if ext == ".grt" and name in SYMMETRIC:  # the symmetric part of every block
        v = 0.5 * (v[..., :49] + v[..., 48::-1])
    mg, sg = grouped_mean_se(v, ref["group"])
    mr, sr = np.array(ref["mean"]), np.array(ref["se"])
    assert mg.shape == mr.shape
    se = np.hypot(sg, sr)
    keep = (np.maximum(np.abs(mg), np.abs(mr)) > 1e-3 * np.abs(mr).max()) & (se > 0)
    z = (mg - mr)[keep] / se[keep]
    print(name, ext, "bins", keep.sum(), "mean z^2", round(float(np.mean(z * z)), 3), "max |z|",
          round(float(np.abs(z).max()), 2))
    from scipy.stats import chi2
    k = int(keep.sum())
    assert k >= 3  # a P = 4 rotor has 3 free offsets: n(0).n(0) = 1 on both sides
    assert np.sum(z * z) < chi2.ppf(0.999, k)
    assert np.mean(np.abs(z) > 2) < 0.12 or k < 10
    assert np.abs(z).max() < 4.5


def test_chain_files_do_not_depend_on_the_other_chains(tmp_path):
    """Each chain's block and accumulated files are written on a worker thread of their own:
    chain 0 of a four-chain run writes, byte for byte, the files of a one-chain run (around a
    non-linear rotor, the 1.5-million-line _sum.g3d included), and the other chains differ."""
    out = {}
    for n in (1, 4):
        d = str(tmp_path / f"c{n}")
        decks.materialise_nl(d, blocks=3)
        r = subprocess.run([DRIVER, "--replicas", str(n)], cwd=d, capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stderr
        out[n] = d
    names = sorted(f for f in os.listdir(out[1]) if f.startswith("h2oh2"))
    assert "h2oh2_sum.g3d" in names and "h2oh2.eng" in names and len(names) > 20
    for f in names:
        with open(os.path.join(out[1], f), "rb") as a, open(os.path.join(out[4], f), "rb") as b:
            assert a.read() == b.read(), f
    for c in (1, 2, 3):
        with open(os.path.join(out[4], "h2oh2.eng")) as a, open(os.path.join(out[4], f"h2oh2_r{c}.eng")) as b:
            assert a.read() != b.read()


def test_driver_output_layouts(tmp_path):
    """.eng: block number and nine columns per measured block; _sum.eng: one line per
    MCSKIP_TOTAL steps; .xyz: natoms*P bead lines; the accept-ratio line on stdout."""
    d = str(tmp_path)
    decks.materialise("he4_co2", d, blocks=12)
    r = subprocess.run([DRIVER], cwd=d, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    eng = decks.read_eng(os.path.join(d, "he4co2.eng"))
    assert eng.shape == (2, 10)
    assert np.allclose(eng[:, 3], eng[:, 1] + eng[:, 2], rtol=1e-5)          # kin + pot
    assert np.allclose(eng[:, 6], eng[:, 1] + eng[:, 2] + eng[:, 4], rtol=1e-5)  # + rot
    sums = np.loadtxt(os.path.join(d, "he4co2_sum.eng"), ndmin=2)
    assert sums.shape == (2, 10) and np.array_equal(sums[:, 0], [1.0, 2.0])  # 12800 steps a block
    assert np.allclose(sums[-1, 1:7], eng[:, 1:7].mean(axis=0), rtol=1e-5)
    with open(os.path.join(d, "he4co2.xyz")) as f:
        lines = f.read().splitlines()
    assert int(lines[0]) == 5 * 64 and len(lines) == 2 + 5 * 64
    assert lines[2].split()[0] == "He41" and lines[-1].split()[0] == "CO21"
    ratio = [ln for ln in r.stdout.splitlines() if ln.startswith("BLOCK:")]
    assert len(ratio) == 12 and "Rot:" in ratio[0] and "He4" in ratio[0]
    # distributions of the measured blocks 11, 12 and the accumulated ones
    for b in ("011", "012"):
        for ext in (".gra", ".gri", ".grt", ".g2d", ".rcf"):
            assert os.path.exists(os.path.join(d, f"he4co2{b}{ext}")), b + ext
    gra = np.loadtxt(os.path.join(d, "he4co2_sum.gra"))
    assert gra.shape == (300, 2) and np.allclose(gra[:, 0], (np.arange(300) + 0.5) * 0.05)
    # sum over bins of the _sum.gra column times dr: each of the N(N-1)/2 pairs once per slice
    # within 15 A (the cluster is bound, so all of them)
    assert abs(gra[:, 1].sum() * 0.05 - 1.0) < 1e-6
    gri = np.loadtxt(os.path.join(d, "he4co2012.gri"))
    grt = np.loadtxt(os.path.join(d, "he4co2012.grt"))
    r = gri[:, 0]
    assert abs(np.sum(gri[:, 1] * 4 * np.pi * r * r * 0.05) - 4.0) < 1e-6   # 4 He around the rotor
    assert abs(np.sum(grt[:, 1]) * np.pi / 49 - 1.0) < 1e-6
    g2d = np.loadtxt(os.path.join(d, "he4co2_sum.g2d"))
    assert g2d.shape == (300 * 50, 3)
    with open(os.path.join(d, "he4co2_sum.rcf")) as f:
        blocks = f.read().split("\n\n\n#\n")
    c0 = np.loadtxt(blocks[0].splitlines())
    cl = np.loadtxt(blocks[1].splitlines())
    assert c0.shape == (65, 2) and cl.shape == (65, 10)
    assert abs(c0[0, 1] - 1.0) < 1e-6 and np.allclose(cl[0, 1:], 1.0, atol=1e-6)   # n(0).n(0) = 1
    assert np.allclose(c0[:, 1], cl[:, 1], atol=1e-9)                             # P_1(p) = p
    assert np.allclose(c0[:, 1], c0[::-1, 1], atol=1e-9)                          # symmetric in tau
    assert np.allclose(c0[:, 0], np.arange(65) / (0.37 * 64), rtol=1e-5)


def test_driver_nonlinear_densities(tmp_path):
    """Around a non-linear rotor (4 para-H2 + H2O, no rotation moves): the block files .gri,
    .grt, .grc (SaveRho1D), .gtc (SaveRhoThetaChi) and the accumulated _sum.g3d
    (SaveDensities3D), _sum.gri/.grt/.grc and _sum.eulphi/.eulchi/.eulthe.  Each normalisation
    is checked by integration: every file gives the same number of H2 within 15 A of the rotor
    per slice (at most 4), the H2O column is empty, and the accumulated files give the mean of
    the two measured blocks (equal sample counts)."""
    d = str(tmp_path)
    decks.materialise_nl(d, blocks=3)
    r = subprocess.run([DRIVER], cwd=d, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    stem = os.path.join(d, "h2oh2")
    dr, dth, dch, deg = 0.05, np.pi / 49, 2 * np.pi / 99, 180.0 / np.pi

    def counts(prefix, gtc=True):
        gri, grt, grc = (np.loadtxt(prefix + e) for e in (".gri", ".grt", ".grc"))
        assert gri.shape == (300, 3) and grt.shape == (50, 3) and grc.shape == (100, 3)
        assert not gri[:, 2].any() and not grt[:, 2].any() and not grc[:, 2].any()
        assert np.allclose(grc[:, 0], (np.arange(100) + 0.5) * dch * deg)
        n = gri[:, 1].sum() * dr
        assert abs(grt[:, 1].sum() * dth * deg * 4 - n) < 1e-5 * n
        assert abs(grc[:, 1].sum() * dch * deg * 4 - n) < 1e-5 * n
        if gtc:
            t = np.loadtxt(prefix + ".gtc")
            assert t.shape == (5000, 4) and not t[:, 3].any()
            assert abs(t[:, 2].sum() * dth * dch * deg * deg * 4 - n) < 1e-5 * n
        return n

    per_block = [counts(stem + b) for b in ("002", "003")]
    assert all(0.5 < n <= 4.0 + 1e-9 for n in per_block), per_block
    n_sum = counts(stem + "_sum", gtc=False)
    assert abs(n_sum - np.mean(per_block)) < 1e-5 * n_sum
    g3d = np.fromfile(stem + "_sum.g3d", sep=" ").reshape(-1, 8)
    assert g3d.shape[0] == 300 * 50 * 100 and not g3d[:, 7].any()
    ir, it, ic = np.unravel_index(np.arange(g3d.shape[0]), (300, 50, 100))
    assert np.allclose(g3d[:, 0], (ir + 0.5) * dr, rtol=1e-6) and np.allclose(g3d[:, 2], (ic + 0.5) * dch * deg, rtol=1e-6)
    th = (it + 0.5) * dth
    assert np.allclose(g3d[:, 1], th * deg, rtol=1e-6)
    n3 = np.sum(g3d[:, 3] * ((ir + 0.5) * dr) ** 2 * np.sin(th)) * dr * dth * dch
    assert abs(n3 - n_sum) < 1e-5 * n_sum
    for ext, nb in ((".eulphi", 100), (".eulchi", 100), (".eulthe", 50)):
        e = np.loadtxt(stem + "_sum" + ext)
        assert e.shape == (nb, 2) and not e[:, 1].any()  # no rotation moves, no rotor links binned


def test_restart_continues_exactly(tmp_path):
    """Four blocks in one run against two blocks, then RESTART for two more: the resumed run's
    .eng lines and final configuration are identical to the uninterrupted run's."""
    a, b = tmp_path / "a", tmp_path / "b"
    for d, blocks in ((a, 4), (b, 2)):
        decks.materialise("he4_co2", str(d), blocks=blocks)
        deck = (d / "qmc.input").read_text().replace(f"NUMBEROFBLOCKS     {blocks}  10", f"NUMBEROFBLOCKS     {blocks}  1")
        (d / "qmc.input").write_text(deck)
        r = subprocess.run([DRIVER], cwd=d, capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stderr
    assert (b / "yw001.stat").read_text().startswith("STARTBLOCK 2")
    deck = (b / "qmc.input").read_text() + "RESTART\n"
    (b / "qmc.input").write_text(deck)
    r = subprocess.run([DRIVER], cwd=b, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    eng_a = (a / "he4co2.eng").read_text().splitlines()
    eng_b = (b / "he4co2.eng").read_text().splitlines()
    assert [ln.split()[0] for ln in eng_a] == ["2", "3", "4"]
    assert eng_b == eng_a[1:]
    assert (b / "he4co2.xyz").read_text() == (a / "he4co2.xyz").read_text()
    assert (b / "yw001.stat.old").exists()
