"""The host driver moribs-pimc_amd/pimc on the CPU: qmc.input parsing, refusals of methods
outside the sampling path, the initial lattice, checkpoint guards, and a loud failure without
a device."""
import os
import subprocess

import numpy as np
import pytest

from golden import decks

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "moribs-pimc_amd", "pimc")


def run(d, *args):
    return subprocess.run([DRIVER, *args], cwd=d, capture_output=True, text=True, timeout=120)


def read_xyz(path):
    with open(path) as f:
        lines = f.read().splitlines()
    n = int(lines[0])
    assert lines[1].startswith("#   xyz format")
    rows = [ln.split() for ln in lines[2:]]
    assert len(rows) == n
    labels = [r[0] for r in rows]
    vals = np.array([[float(x) for x in r[1:]] for r in rows])
    return labels, vals[:, 0::2], vals[:, 1::2]  # coordinates, direction cosines


DECK = """OUTPUTDIR ./
FILENAMEPREFIX t
DIMENSION 3
DENSITY 0.02
TEMPERATURE 0.37
ATOM He4 {nhe} BOLTZMANN 0.3 3 He PRIMITIVE   # comment after the fields
MOLECULE CO2 1 BOLTZMANN 0.08 3 CO2He PRIMITIVE
ROTATION CO2 0.3 {P}
NUMBEROFSLICES {P}
NUMBEROFPASSES 10
NUMBEROFBLOCKS 2 1
{extra}
"""


def write(d, nhe=8, P=4, extra=""):
    with open(os.path.join(d, "qmc.input"), "w") as f:
        f.write(DECK.format(nhe=nhe, P=P, extra=extra))


def test_initial_lattice_of_eight_atoms(tmp_path):
    """8 atoms fill the box on a 2x2x2 cube at +-L/4, x fastest; the lone molecule sits at the
    centre (initLattice_config mc_setup.cc:491-650, box from DENSITY mc_setup.cc:343-367)."""
    write(tmp_path, nhe=8, P=4)
    r = run(tmp_path, "--init-only")
    assert r.returncode == 0, r.stderr
    labels, xyz, cos = read_xyz(tmp_path / "t.xyz")
    L = (9 / 0.02) ** (1.0 / 3.0)
    assert labels[0] == "He41" and labels[4] == "He42" and labels[-1] == "CO21"
    for a in range(8):
        want = np.array([a % 2, (a // 2) % 2, a // 4]) * L / 2 - L / 4
        assert np.allclose(xyz[4 * a:4 * a + 4], want, atol=5e-6 * L)  # every slice of the atom
    assert np.allclose(xyz[32:], 0.0)
    assert np.allclose(cos, [0.0, 0.0, 1.0])


def test_initial_lattice_equal_counts(tmp_path):
    """With as many atoms as molecules the molecules are shifted by one box length."""
    write(tmp_path, nhe=1, P=4)
    assert run(tmp_path, "--init-only").returncode == 0
    _, xyz, _ = read_xyz(tmp_path / "t.xyz")
    L = (2 / 0.02) ** (1.0 / 3.0)
    assert np.allclose(xyz[:4], 0.0, atol=1e-6) and np.allclose(xyz[4:], L, rtol=1e-6)


def test_xyz_init_is_read(tmp_path):
    write(tmp_path, nhe=1, P=2, extra="READMCCOORDS")
    with open(tmp_path / "xyz.init", "w") as f:
        f.write("0\ncomment\n")
        for i in range(4):
            f.write(f"X {i} 0.5 {i + 10} 0.25 {i + 20} 1.5\n")
    assert run(tmp_path, "--init-only").returncode == 0
    _, xyz, cos = read_xyz(tmp_path / "t.xyz")
    assert np.allclose(xyz[:, 0], [0, 1, 2, 3]) and np.allclose(xyz[:, 2], [20, 21, 22, 23])
    st = np.sqrt(1 - 0.25 ** 2)
    assert np.allclose(cos[0], [st * np.cos(0.5), st * np.sin(0.5), 0.25], atol=1e-6)


@pytest.mark.parametrize("extra,msg", [("WORM He4 0.13 16", "worm"), ("MINIMAGE", "MIN IMAGE for orient pot"),
                                       ("CRYSTAL FCC 1 1 1", "crystal"),
                                       ("REFLECTY 1", "REFLECTY"), ("ROTDENSI 1 -1 0 1 1 1 1", "ROTDENSI")])
def test_methods_outside_the_path_are_refused(tmp_path, extra, msg):
    write(tmp_path, extra=extra)
    r = run(tmp_path, "--init-only")
    assert r.returncode == 1 and msg in r.stderr, r.stderr


def test_minimum_image_accepted_without_a_rotating_linear_rotor(tmp_path):
    """MINIMAGE is read (mc_input.cc:298-300) for a deck whose rotor does not rotate; with the
    rotation line it is refused as the reference's rotor move stops (above)."""
    with open(tmp_path / "qmc.input", "w") as f:
        f.write(DECK.format(nhe=2, P=4, extra="MINIMAGE").replace("ROTATION CO2 0.3 4\n", ""))
    r = run(tmp_path, "--init-only")
    assert r.returncode == 0, r.stderr


def test_bose_statistics_refused(tmp_path):
    with open(tmp_path / "qmc.input", "w") as f:
        f.write(DECK.format(nhe=2, P=4, extra="").replace("He4 2 BOLTZMANN", "He4 2 BOSE"))
    r = run(tmp_path, "--init-only")
    assert r.returncode == 1 and "Bose" in r.stderr


def test_unknown_type_refused(tmp_path):
    with open(tmp_path / "qmc.input", "w") as f:
        f.write(DECK.format(nhe=2, P=4, extra="").replace("He4", "Xe"))
    r = run(tmp_path, "--init-only")
    assert r.returncode == 1 and "Unknown atom/molecule type" in r.stderr


def test_reference_example_deck_parses(tmp_path):
    decks.materialise("co2_100K", str(tmp_path), blocks=3)
    r = run(tmp_path, "--init-only")
    assert r.returncode == 0, r.stderr
    _, xyz, _ = read_xyz(tmp_path / "CO2_monomer.xyz")
    assert xyz.shape == (4, 3) and np.allclose(xyz, 0.0)


def test_no_device_fails_loudly(tmp_path):
    import torch
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    decks.materialise("co2_100K", str(tmp_path), blocks=3)
    r = run(tmp_path)
    assert r.returncode == 1 and "pimc_create" in r.stderr


def test_restart_needs_a_checkpoint(tmp_path):
    write(tmp_path, extra="RESTART")
    r = run(tmp_path)
    assert r.returncode == 1 and "yw001.stat" in r.stderr, r.stderr


def test_new_run_keeps_an_existing_checkpoint(tmp_path):
    """As the reference refuses to start over its status files (mc_main.cc:126-136)."""
    write(tmp_path)
    (tmp_path / "yw001.stat").write_text("STARTBLOCK 3\n")
    r = run(tmp_path)
    assert r.returncode == 1 and "RESTART" in r.stderr, r.stderr


FORMAT_CHECK = r"""
#include "sci_format.hpp"
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
int main() {
    long n = 0, bad = 0;
    auto check = [&](double x) {
        char a[64] = {0}, b[64] = {0};
        const int na = sci_format::e14_6(x, a);
        const int nb = std::snprintf(b, sizeof b, "%14.6e", x);
        ++n;
        if (na != nb || std::memcmp(a, b, (size_t)nb)) {
            if (bad++ < 5) std::printf("%.17g [%.*s] [%s]\n", x, na, a, b);
        }
    };
    std::mt19937_64 g(7);
    std::uniform_real_distribution<double> u(0.0, 1.0);
    for (long i = 0; i < 3000000; ++i) check((u(g) < 0.5 ? -1.0 : 1.0) * std::pow(10.0, 44.0 * u(g) - 22.0));
    for (int e = -18; e <= 18; ++e)  // seven-digit ties, decades and their neighbours
        for (long m = 999990; m <= 1000010; ++m)
            for (double h : {0.0, 0.5}) {
                const double x = (m + h) * std::pow(10.0, e);
                for (double y : {x, std::nextafter(x, 0.0), std::nextafter(x, 1e300)}) {
                    check(y);
                    check(-y);
                }
            }
    for (long m = 9999990; m <= 10000010; ++m)
        for (int e = -12; e <= 12; ++e) check((m + 0.5) * std::pow(10.0, e - 7));
    const double special[] = {0.0, -0.0, 5e-324, 1e-300, 1e300, 1e100, 9.9999995e99, 1e16, 1e-16,
                              1234567.5, 0.5, 2.5, INFINITY, -INFINITY, NAN};
    for (double x : special) {
        check(x);
        check(-x);
    }
    std::printf("%ld %ld\n", n, bad);
    return bad != 0;
}
"""


def test_density_field_format_equals_printf(tmp_path):
    """The 3D density file's fast "%14.6e" (host/sci_format.hpp) against snprintf on random
    magnitudes, seven-digit ties, decades, signed zeros and non-finite values."""
    src = tmp_path / "check.cc"
    src.write_text(FORMAT_CHECK)
    exe = tmp_path / "check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "moribs-pimc_amd", "host"), str(src), "-o",
                    str(exe)], check=True, timeout=120)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout
    n, bad = map(int, r.stdout.split()[-2:])
    assert bad == 0 and n > 3000000
