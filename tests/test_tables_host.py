"""Host-side logic on the CPU: table generators and file formats, workload definitions, the
bead-move count, and the C ABI's exported symbols."""
import ctypes
import math
import os
import re

import numpy as np
import pytest

import oracle as O
import pimc_configs as pc
import pimc_hip as ph
import pimc_tables as pt

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")


# ----------------------------------------------------------------------------- .rot tables
@pytest.mark.parametrize("fname,T,P,B", [("CO2_T100t4.rot", 100.0, 4, 0.39021),
                                         ("N2O_T0.5t128.rot", 0.5, 128, 0.419)])
def test_linden_reproduces_reference_rot_files(fname, T, P, B):
    """The .rot files shipped with the reference's examples are reproduced by the linden sum
    (linden.f:88-122) to the rounding of their E15.8 columns."""
    c, rho, er, es = pt.read_rot(os.path.join(GOLD, fname))
    c2, r2, e2, s2 = pt.linden(T, P, B, npt=len(c))
    assert np.max(np.abs(c2 - c)) < 1e-8
    for got, ref in ((r2, rho), (e2, er), (s2, es)):
        assert np.max(np.abs(got - ref)) <= 5e-9 * np.max(np.abs(ref))


def test_linden_matches_oracle():
    L = O.lib()
    T, P, B = 0.37, 128, 0.39
    tau = pt.TAU_UNIT / (T * P)
    lmax = pt.linden_lmax(tau, B)
    assert lmax == L.or_linden_lmax(tau, B)
    c, rho, er, es = pt.linden(T, P, B, npt=201)
    e1, e2 = ctypes.c_double(), ctypes.c_double()
    for k in range(0, 201, 7):
        r = L.or_linden_rho(c[k], tau, B, lmax, -1, P, ctypes.byref(e1), ctypes.byref(e2))
        assert abs(r - rho[k]) <= 1e-12 * max(1.0, abs(rho[k]))
        assert abs(e1.value - er[k]) <= 1e-12 * max(1.0, abs(er[k]))
        assert abs(e2.value - es[k]) <= 1e-12 * max(1.0, abs(es[k]))


def test_linden_cutoff_of_the_c0_table():
    """l cutoff exp(-tau B l(l+1)) < 1e-16 at T = 0.37 K, Prot = 128, B = 0.39 cm^-1: l = 56."""
    tau = pt.TAU_UNIT / (0.37 * 128)
    assert pt.linden_lmax(tau, 0.39) == 56


def test_rotor_density_normalisation():
    """Only l = 0 survives the integral over cos(gamma): int rho d(cos) = 2/(4 pi)."""
    c, rho, _, _ = pt.linden(0.37, 128, 0.39, npt=20001)
    assert abs(np.trapezoid(rho, c) - 1.0 / (2.0 * math.pi)) < 1e-6


def test_free_rotor_energy_limits():
    """Classical limit E -> T - B/3 (Kelvin) at high T; ground state E -> 0 at low T."""
    B_K = 0.39 / pt.BOLTZ_WNO
    e, _ = pt.linden_exact_energy(100.0, 0.39)
    assert abs(e - (100.0 - B_K / 3.0)) < 2e-3
    e, _ = pt.linden_exact_energy(0.05, 0.39)
    assert 0.0 <= e < 1e-8


# ----------------------------------------------------------------------------- file formats
def test_pot1d_round_trip(tmp_path):
    r, v = pt.lj_table()
    p = tmp_path / "He.pot"
    pt.write_pot1d(str(p), r, v)
    r2, v2 = pt.read_pot1d(str(p))
    assert np.array_equal(r, r2) and np.array_equal(v, v2)


def test_reference_pot1d_fixture_reads():
    r, v = pt.read_pot1d(os.path.join(GOLD, "helium.pot"))
    assert len(r) == 500 and np.all(np.diff(r) > 0)


def test_pot2d_round_trip(tmp_path):
    rg, cg, V = pt.morse_legendre_2d()
    p = tmp_path / "CO2He.pot"
    pt.write_pot2d(str(p), rg, cg, V)
    rg2, cg2, dr, dc, V2 = pt.read_pot2d(str(p))
    assert np.array_equal(rg, rg2) and np.array_equal(cg, cg2) and np.array_equal(V, V2)
    assert abs(dr - 0.02) < 1e-12 and abs(dc - 0.01) < 1e-12


def test_rot_round_trip(tmp_path):
    c, rho, er, es = pc.rot_table(0.5, 16, npt=101)
    p = tmp_path / "x.rot"
    pt.write_rot(str(p), c, rho, er, es)
    back = pt.read_rot(str(p))
    for a, b in zip((c, rho, er, es), back):
        assert np.array_equal(a, b)  # rot_table already rounds to E15.8


def test_e15_8_format():
    assert pt._e15_8(-1.0) == "-1.00000000E+00"
    assert pt._e15_8(1.20634723e-15) == " 1.20634723E-15"
    assert len(pt._e15_8(123456.789)) == 15


def test_temperature_format_is_cpp_default():
    """std::ostream << double with precision 6 (mc_poten.cc:517-524 file names)."""
    assert pt.fmt_temperature(0.37) == "0.37"
    assert pt.fmt_temperature(100.0) == "100"
    assert pt.fmt_temperature(1.2345678) == "1.23457"
    assert pt.fmt_temperature(0.000123456789) == "0.000123457"
    assert pt.rot_filename("CO2", 0.5, 128) == "CO2_T0.5t128.rot"


# ----------------------------------------------------------------------------- workloads
def test_bead_moves_per_pass_of_c0():
    """8 He + 1 rotor, P = Prot = 128, L = 3: 8*128*7 + 8*128 + 128*7 + 128 + 128*128."""
    rp = pc.workload("c0")["replicas"][0]
    assert pc.bead_moves_per_pass(rp) == 25600


def test_c2_workload_shape():
    wl = pc.workload("c2")
    reps = wl["replicas"]
    assert len(reps) == 256
    n = np.array([r["numb"][0] for r in reps])
    P = np.array([r["P"] for r in reps])
    T = np.array([r["T"] for r in reps])
    assert n.min() >= 1 and n.max() <= 100
    assert set(P.tolist()) <= {64, 128, 256, 512}
    assert T.min() >= 0.3 and T.max() <= 2.0
    # one .rot table per distinct (T, Prot), shared by replicas with the same key
    assert len(wl["tables"]["rot"]) == len({(r["T"], r["Prot"]) for r in reps})
    again = pc.workload("c2")["replicas"]
    assert all(np.array_equal(a["coords"], b["coords"]) for a, b in zip(reps[:8], again[:8]))


def test_initial_path_exclusion():
    rp = pc.workload("c0")["replicas"][0]
    P = rp["P"]
    pts = rp["coords"][:, ::P].T[:-1]
    d = np.linalg.norm(pts[:, None, :] - pts[None, :, :], axis=-1)
    assert np.all(d[np.triu_indices(len(pts), 1)] > 2.6)
    assert np.all(np.linalg.norm(pts, axis=1) >= 3.0)


# ----------------------------------------------------------------------------- C ABI
def declared_functions():
    with open(os.path.join(ROOT, "include", "moribs_pimc.h")) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pimc_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(ph.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 20
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) == set(ph.SIGNATURES), set(names) ^ set(ph.SIGNATURES)


def test_context_without_device_fails_loudly():
    """No HIP device here: creating a context reports an error instead of falling back."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    with pytest.raises(ph.PimcError):
        ph.Context(1)


# ----------------------------------------------------------------------------- asymmetric top
def test_asymrho_reproduces_reference_fixture():
    """The asymmetric-top propagator and estimators reproduce the table the reference's own
    asymrho run left in nmv_prop/ (rho.den010_rho/_eng/_esq: T 0.37 K, 128 slices, no spin
    symmetry, theta 10 deg, A B C 0.6666525 0.2306476 0.1769383 cm^-1, maxj 66; every 10th phi
    row kept) to the precision of its E15.8 columns, wherever rho is above the summation noise."""
    z = np.load(os.path.join(GOLD, "asymrho_theta010.npz"))
    T, P, iod, th, A, B, C, maxj = z["params"]
    rho, eng, esq, _ = pt.asymrho_tables(T, int(P), A, B, C, int(maxj), int(iod), [int(th)])
    rows = z["phi"]
    got = (rho[0][rows], eng[0][rows], esq[0][rows])
    ref = (z["rho"], z["eng"], z["esq"])
    mask = np.abs(ref[0]) > 1e-8 * np.abs(ref[0]).max()
    assert mask.sum() > 3000
    assert np.max(np.abs(got[0] - ref[0])[mask] / np.abs(ref[0][mask])) < 1e-7
    for g, r in zip(got[1:], ref[1:]):
        assert np.max(np.abs(g - r)[mask] / np.maximum(np.abs(r[mask]), 1.0)) < 1e-7


def test_wigner_d_matches_zare_sum():
    """d^J_mk(theta) from J_y eigenvectors equals the factorial sum of Zare eq. 3.57 at small J."""
    from math import factorial, cos, sin, sqrt

    def zare(j, m, k, t):
        s = 0.0
        for nu in range(max(0, k - m), min(j + k, j - m) + 1):
            den = factorial(j - m - nu) * factorial(j + k - nu) * factorial(nu + m - k) * factorial(nu) * (-1) ** nu
            s += sqrt(factorial(j + k) * factorial(j - k) * factorial(j + m) * factorial(j - m)) / den * \
                cos(t / 2) ** (2 * j + k - m - 2 * nu) * (-sin(t / 2)) ** (m - k + 2 * nu)
        return s
    for j in (0, 1, 3, 6):
        d = pt._wigner_d(j, 0.83)
        for m in range(-j, j + 1):
            for k in range(-j, j + 1):
                assert abs(d[m + j, k + j] - zare(j, m, k, 0.83)) < 1e-12


def test_asymmetric_top_normalisation():
    """Only J = 0 survives the integral over SO(3): sum over the degree grid of
    rho sin(theta) dtheta dphi dchi approaches 1."""
    thetas = list(range(0, 181, 2))
    rho, _, _, _ = pt.asymrho_tables(20.0, 4, 27.88, 14.51, 9.28, 14, -1, thetas)
    d = np.radians(1.0)
    w = np.sin(np.radians(thetas))[:, None, None]
    integral = np.sum(rho[:, :-1, :-1] * w) * np.radians(2.0) * d * d
    assert abs(integral - 1.0) < 2e-3
