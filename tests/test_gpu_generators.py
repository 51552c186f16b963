"""GPU density-matrix generators through the C ABI: linden (linear_prop/linden.f:88-122) and
asymrho (nmv_prop/asymrho.f:34-731).

Pins: the .rot files and the asymrho output the reference ships (tests/golden), the host
restatements of pimc_tables (numpy; the Wigner d there comes from J_y eigenvectors, the GPU's
from the recursion in J, so the two share no code), and the SO(3) normalisation."""
import os

import numpy as np
import pytest

import pimc_hip as ph
import pimc_tables as pt

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def ctx():
    return ph.Context(1)


@pytest.mark.parametrize("fname,T,P,B", [("CO2_T100t4.rot", 100.0, 4, 0.39021),
                                         ("N2O_T0.5t128.rot", 0.5, 128, 0.419)])
def test_linden_reproduces_reference_rot_files(ctx, fname, T, P, B):
    """The reference's own .rot files, to the rounding of their E15.8 columns."""
    c, rho, er, es = pt.read_rot(os.path.join(GOLD, fname))
    c2, r2, e2, s2, _ = ctx.linden(T, P, B, npt=len(c))
    assert np.max(np.abs(c2 - c)) < 1e-8
    for got, ref in ((r2, rho), (e2, er), (s2, es)):
        assert np.max(np.abs(got - ref)) <= 5e-9 * np.max(np.abs(ref))


@pytest.mark.parametrize("iodevn", [-1, 0, 1])
def test_linden_matches_host(ctx, iodevn):
    tau = pt.TAU_UNIT / (0.37 * 128)
    c, rho, er, es, lmax = ctx.linden(0.37, 128, 0.39, npt=2001, iodevn=iodevn)
    assert lmax == pt.linden_lmax(tau, 0.39) == 56
    c2, r2, e2, s2 = pt.linden(0.37, 128, 0.39, npt=2001, iodevn=iodevn)
    assert np.array_equal(c, c2), np.abs(c - c2).max()
    for got, ref in ((rho, r2), (er, e2), (es, s2)):
        assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref))


def test_asymrho_reproduces_reference_fixture(ctx):
    """The table the reference's own asymrho run left in nmv_prop/ (T 0.37 K, 128 slices, no
    spin symmetry, A B C 0.6666525 0.2306476 0.1769383 cm^-1, maxj 66; theta 10 deg, every 10th
    phi row), to the precision of its E15.8 columns wherever rho is above the summation noise."""
    z = np.load(os.path.join(GOLD, "asymrho_theta010.npz"))
    T, P, iod, th, A, B, C, maxj = z["params"]
    rho, eng, esq = ctx.asymrho(T, int(P), A, B, C, int(maxj), int(iod))
    rows = z["phi"]
    got = (rho[int(th)][rows], eng[int(th)][rows], esq[int(th)][rows])
    ref = (z["rho"], z["eng"], z["esq"])
    mask = np.abs(ref[0]) > 1e-8 * np.abs(ref[0]).max()
    assert mask.sum() > 3000
    assert np.max(np.abs(got[0] - ref[0])[mask] / np.abs(ref[0][mask])) < 1e-7
    for g, r in zip(got[1:], ref[1:]):
        assert np.max(np.abs(g - r)[mask] / np.maximum(np.abs(r[mask]), 1.0)) < 1e-7


@pytest.mark.parametrize("iodevn", [-1, 0, 1])
def test_asymrho_matches_host(ctx, iodevn):
    """Every grid point of six theta planes, including the theta = 0 and 180 degree edges,
    against the host restatement."""
    thetas = [0, 1, 37, 90, 179, 180]
    args = (20.0, 4, 27.88, 14.51, 9.28, 14, iodevn)
    rho, eng, esq = ctx.asymrho(*args)
    r2, e2, s2, _ = pt.asymrho_tables(*args, thetas=thetas)
    scale = np.abs(r2).max()
    for i, th in enumerate(thetas):
        assert np.max(np.abs(rho[th] - r2[i])) <= 1e-11 * scale, th
        big = np.abs(r2[i]) > 1e-6 * scale
        assert np.max(np.abs(eng[th] - e2[i])[big] / np.maximum(np.abs(e2[i][big]), 1.0)) < 1e-9, th
        assert np.max(np.abs(esq[th] - s2[i])[big] / np.maximum(np.abs(s2[i][big]), 1.0)) < 1e-9, th


def test_asymrho_normalisation_and_refusal(ctx):
    """Only J = 0 survives the integral over SO(3): the degree-grid sum of rho sin(theta)
    approaches 1.  A maxj too small for the temperature is refused as asymrho refuses it."""
    rho, _, _ = ctx.asymrho(20.0, 4, 27.88, 14.51, 9.28, 14, -1)
    d = np.radians(1.0)
    w = np.sin(np.radians(np.arange(181)))[:, None, None]
    assert abs(np.sum(rho[:, :-1, :-1] * w) * d ** 3 - 1.0) < 2e-3
    with pytest.raises(ph.PimcError, match="raise maxj"):
        ctx.asymrho(2000.0, 4, 27.88, 14.51, 9.28, 3, -1)
