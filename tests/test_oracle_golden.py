"""The oracle against the golden vectors of the compiled reference (tests/golden/).

Every quantity is compared at 1e-12 relative or tighter: the oracle restates the reference's
arithmetic in the same order, so the two agree to the last bits or to a few ulps where the
reference's Fortran and the oracle's C round a sum differently.  Moves are replayed on the
reference's own random numbers (tapes), so trajectories and acceptance counters must match.
"""
import json
import os

import numpy as np
import pytest

import oracle as O
import pimc_configs as pc
import pimc_tables as pt
from golden import tables3d

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def G():
    with open(os.path.join(GOLD, "ref_vectors.json")) as f:
        return json.load(f)


@pytest.fixture(scope="module")
def A():
    with np.load(os.path.join(GOLD, "ref_arrays.npz")) as z:
        return {k: z[k] for k in z.files}


@pytest.fixture(scope="module")
def rot3d():
    return tables3d.rot3d()


def close(got, ref, rtol=1e-12, atol=1e-300):
    got, ref = np.asarray(got, dtype=np.float64), np.asarray(ref, dtype=np.float64)
    err = np.abs(got - ref) / np.maximum(np.abs(ref), 1.0)
    assert got.shape == ref.shape
    assert np.all(err <= rtol) or np.all(np.abs(got - ref) <= atol), float(err.max())


def tiny_system():
    wl = pc.workload("tiny")
    return O.System(wl["replicas"][0], wl["tables"]), wl


def tapes_of(A, prefix):
    return {k: A[f"{prefix}_tape{k}"] for k in (1, 2, 3, 8, 9, 15)}


# ----------------------------------------------------------------------------- Philox
def test_philox_known_answers():
    """Random123 known-answer vectors for philox4x32-10 (kat_vectors)."""
    assert O.philox([0, 0, 0, 0], [0, 0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    m = 0xFFFFFFFF
    assert O.philox([m, m, m, m], [m, m]) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert O.philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0]) == \
        [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


# ----------------------------------------------------------------------------- linear tables
def test_spot1d(G):
    S, _ = tiny_system()
    g = G["linear"]
    close([S.L.or_spot1d(S.h, r, 0) for r in g["spot1d_r"]], g["spot1d"], rtol=0.0)


def test_lpot2d(G):
    S, _ = tiny_system()
    g = G["linear"]
    close([S.L.or_lpot2d(S.h, r, c, 1) for r, c in zip(g["lpot2d_r"], g["lpot2d_c"])], g["lpot2d"], rtol=0.0)


def test_srotdens(G):
    S, _ = tiny_system()
    g = G["linear"]
    for f, key in [(S.L.or_srotdens, "srotdens"), (S.L.or_srotdens_deriv, "srotdens_deriv"),
                   (S.L.or_srotdens_esqrt, "srotdens_esqrt")]:
        close([f(S.h, x, 1) for x in g["srot_g"]], g[key], rtol=0.0)


def test_table_node_identities():
    """Closed forms: the spline passes through its knots; LPot2D equals V at grid nodes."""
    S, wl = tiny_system()
    r, v = wl["tables"]["pot1d"][0]
    got = np.array([S.L.or_spot1d(S.h, x, 0) for x in r[1:-1]])
    assert np.array_equal(got, v[1:-1])
    rg, cg, _, _, V = wl["tables"]["pot2d"][0]
    for i in range(0, len(rg) - 1, 50):
        for j in range(0, len(cg) - 1, 20):
            assert abs(S.L.or_lpot2d(S.h, rg[i], cg[j], 1) - V[i, j]) <= 1e-12 * max(1.0, abs(V[i, j]))


def test_spot1d_tails():
    """Below the first knot U0 exp(-alpha r) through the first two knots, above the last
    -C6/r^6 through the last two (mc_poten.cc:422-437)."""
    S, wl = tiny_system()
    r, v = wl["tables"]["pot1d"][0]
    alpha = np.log(v[0] / v[1]) / (r[1] - r[0])
    u0 = v[0] * np.exp(alpha * r[0])
    for x in (0.3, 0.7, 0.95):
        assert abs(S.L.or_spot1d(S.h, x, 0) - u0 * np.exp(-alpha * x)) <= 1e-10 * abs(u0 * np.exp(-alpha * x))
    c6 = (v[-1] - v[-2]) / (1.0 / r[-2] ** 6 - 1.0 / r[-1] ** 6)  # slope of V against 1/r^6
    for x in (30.5, 40.0):
        assert abs(S.L.or_spot1d(S.h, x, 0) - (-c6 / x ** 6)) <= 1e-12 * abs(c6 / x ** 6)


# ----------------------------------------------------------------------------- linear energies
def test_linear_energies(G, A):
    S, _ = tiny_system()
    g = G["linear"]
    S.set_state(A["linear_coords"], A["linear_angles"])
    natoms, P = S.natoms, S.rp["P"]
    newc = np.ascontiguousarray(A["linear_newcoords"])
    got = [[S.L.or_pot_energy_slice(S.h, a, None, it) for it in range(P)] for a in range(natoms)]
    close(got, g["pot_slice"])
    got = [[S.L.or_pot_energy_slice(S.h, a, O.P(newc), it) for it in range(P)] for a in range(natoms)]
    close(got, g["pot_slice_new"])
    close([S.L.or_pot_energy_path(S.h, a, None) for a in range(natoms)], g["pot_path"])
    close([S.L.or_pot_energy_path(S.h, a, O.P(newc)) for a in range(natoms)], g["pot_path_new"])
    _, _, cs = S.state()
    close([S.L.or_pot_rot_energy(S.h, natoms - 1, O.P(cs), it) for it in range(P)], g["pot_rot"])


def estimators(S, nonlinear=False):
    import ctypes
    gr1, gr2 = np.zeros(300), np.zeros(300 * 50)
    e1, e2 = ctypes.c_double(), ctypes.c_double()
    out = dict(kin=S.L.or_kin_energy(S.h), pot=S.L.or_pot_energy_densities(S.h, O.P(gr1) if False else gr1.ctypes.data_as(O._dp),
                                                                          gr2.ctypes.data_as(O._dp)))
    f = S.L.or_rot_e3d if nonlinear else S.L.or_rot_energy
    out["rot"] = f(S.h, ctypes.byref(e1), ctypes.byref(e2))
    out["rotsq"], out["rot_termsq"] = e1.value, e2.value
    return out, gr1, gr2


def check_estimators(got, ref):
    for k in ("kin", "pot", "rot", "rotsq", "rot_termsq"):
        close(got[k], ref[k])


def test_linear_estimators(G, A):
    S, _ = tiny_system()
    S.set_state(A["linear_coords"], A["linear_angles"])
    est, gr1, gr2 = estimators(S)
    check_estimators(est, G["linear"]["estimators"])
    assert np.array_equal(gr1, A["linear_gr1d"])
    assert np.array_equal(gr2, A["linear_gr2d"])


def test_linear_reference_order_passes(G, A):
    """Two passes of mc_main.cc:348-381 on the reference's random numbers: the same path,
    the same acceptance counters and the same consumption of every stream."""
    S, wl = tiny_system()
    g = G["linear"]
    S.set_tapes(tapes_of(A, "linear"))
    S.L.or_reset_counters(S.h)
    for p in range(2):
        assert S.L.or_pass(S.h, p, O.ORDER_REFERENCE) == 0
    assert S.L.or_tape_short(S.h) == 0
    for k, n in g["tape_used"].items():
        assert S.L.or_tape_pos(S.h, int(k)) == n
    c, a, _ = S.state()
    close(c.reshape(-1), A["linear_pass_coords"], rtol=1e-13)
    close(a.reshape(-1), A["linear_pass_angles"], rtol=1e-13)
    t, acc = S.counters()
    assert t.reshape(-1).tolist() == g["pass_total"] and acc.reshape(-1).tolist() == g["pass_accep"]
    est, _, _ = estimators(S)
    check_estimators(est, g["pass_estimators"])


# ----------------------------------------------------------------------------- helium fixture
def helium_system(A):
    r, v = pt.read_pot1d(os.path.join(GOLD, "helium.pot"))
    rp = dict(ntypes=1, numb=[8], molecule=[0], levels=[3], mcstep=[0.2], mass=[4.0026032497], P=512, Prot=0,
              T=0.37, rtstep=0.0, pot1d=[0], pot2d=[-1], rotdens=-1)
    S = O.System(rp, dict(pot1d=[(r, v)], pot2d=[], rot=[]))
    S.set_state(A["helium_coords"], A["helium_angles"])
    return S


def test_helium_fixture(G, A):
    """helium.pot and the 8-He path of the reference's examples: SPot1D, V and K."""
    S = helium_system(A)
    g = G["helium"]
    close([S.L.or_spot1d(S.h, x, 0) for x in g["spot1d_r"]], g["spot1d"], rtol=0.0)
    close([S.L.or_pot_energy_path(S.h, a, None) for a in range(8)], g["pot_path"])
    close(S.L.or_kin_energy(S.h), g["kin"])
    gr1 = np.zeros(300)
    close(S.L.or_pot_energy_densities(S.h, gr1.ctypes.data_as(O._dp), None), g["pot"])
    assert np.array_equal(gr1, A["helium_gr1d"])


# ----------------------------------------------------------------------------- minimum image
def periodic_system(A, **extra):
    r, v = pt.read_pot1d(os.path.join(GOLD, "helium.pot"))
    rp = dict(ntypes=1, numb=[16], molecule=[0], levels=[3], mcstep=[0.6], mass=[4.0026032497], P=16, Prot=0,
              T=1.5, rtstep=0.0, pot1d=[0], pot2d=[-1], rotdens=-1, box=8.5, **extra)
    S = O.System(rp, dict(pot1d=[(r, v)], pot2d=[], rot=[]))
    S.set_state(A["periodic_coords"], np.zeros_like(A["periodic_coords"]))
    return S, rp, (r, v)


def test_periodic_energies(G, A):
    """MINIMAGE (mc_piqmc.cc:1239-1240,1832-1833; mc_estim.cc:536-537,916-917): PotEnergy on the
    path and on newcoords, V with its g(r) bins and K equal the compiled reference's."""
    S, rp, _ = periodic_system(A)
    g = G["periodic"]
    new = A["periodic_newcoords"]
    close([[S.L.or_pot_energy_slice(S.h, a, None, it) for it in range(16)] for a in range(16)], g["pot_slice"])
    close([[S.L.or_pot_energy_slice(S.h, a, O.P(new), it) for it in range(16)] for a in range(16)], g["pot_slice_new"])
    close([S.L.or_pot_energy_path(S.h, a, None) for a in range(16)], g["pot_path"])
    close([S.L.or_pot_energy_path(S.h, a, O.P(new)) for a in range(16)], g["pot_path_new"])
    close(S.L.or_kin_energy(S.h), g["kin"])
    gr1 = np.zeros(300)
    close(S.L.or_pot_energy_densities(S.h, gr1.ctypes.data_as(O._dp), None), g["pot"])
    assert np.array_equal(gr1, A["periodic_gr1d"])


def test_periodic_minimum_image_brute_force(A):
    """V under the minimum image equals a direct sum over pairs and slices that takes, for every
    pair, the closest of its 27 nearest periodic images; without the minimum image V differs."""
    S, rp, (r, v) = periodic_system(A)
    L, P, N = 8.5, 16, 16
    c = A["periodic_coords"].reshape(3, N, P)
    shifts = np.array([[i, j, k] for i in (-1, 0, 1) for j in (-1, 0, 1) for k in (-1, 0, 1)], float) * L
    V = 0.0
    for a in range(N):
        for b in range(a + 1, N):
            for it in range(P):
                d = c[:, a, it] - c[:, b, it]
                rr = np.min(np.linalg.norm(d[None, :] + shifts, axis=1))
                V += S.L.or_spot1d(S.h, rr, 0)
    V /= P
    gr1 = np.zeros(300)
    close(S.L.or_pot_energy_densities(S.h, gr1.ctypes.data_as(O._dp), None), V, rtol=1e-12)
    plain, _, _ = periodic_system(A)
    plain.L.or_set_minimage(plain.h, 0, None)
    assert abs(plain.L.or_pot_energy_densities(plain.h, gr1.ctypes.data_as(O._dp), None) - V) > 1.0


def test_periodic_reference_order_passes(G, A):
    """Two passes under MINIMAGE on the reference's random numbers: the same path, counters and
    stream consumption, and the same V and K afterwards."""
    S, rp, _ = periodic_system(A)
    g = G["periodic"]
    S.set_tapes(tapes_of(A, "periodic"))
    S.L.or_reset_counters(S.h)
    for p in range(2):
        assert S.L.or_pass(S.h, p, O.ORDER_REFERENCE) == 0
    assert S.L.or_tape_short(S.h) == 0
    for k, n in g["tape_used"].items():
        assert S.L.or_tape_pos(S.h, int(k)) == n
    c, _, _ = S.state()
    close(c.reshape(-1), A["periodic_pass_coords"], rtol=1e-13)
    t, acc = S.counters()
    assert t.reshape(-1).tolist() == g["pass_total"] and acc.reshape(-1).tolist() == g["pass_accep"]
    close(S.L.or_kin_energy(S.h), g["pass_kin"])
    gr1 = np.zeros(300)
    close(S.L.or_pot_energy_densities(S.h, gr1.ctypes.data_as(O._dp), None), g["pass_pot"])
    assert np.array_equal(gr1, A["periodic_pass_gr1d"])


def test_helium_potential_brute_force(A):
    """V of the estimator equals a direct double loop over pairs and slices."""
    S = helium_system(A)
    c = A["helium_coords"].reshape(3, 8, 512)
    tot = 0.0
    for i in range(8):
        for j in range(i + 1, 8):
            d = np.sqrt(((c[:, i, :] - c[:, j, :]) ** 2).sum(axis=0))
            tot += sum(S.L.or_spot1d(S.h, x, 0) for x in d)
    v = S.L.or_pot_energy_densities(S.h, None, None)
    assert abs(v - tot / 512) <= 1e-11 * abs(v)


# ----------------------------------------------------------------------------- non-linear functions
def test_rotden(G, rot3d):
    import ctypes
    L = O.lib()
    rho, er, es = rot3d
    g = G["nonlinear"]
    for a, b, ref in zip(g["rotden_e1"], g["rotden_e2"], g["rotden"]):
        erel = np.zeros(3)
        r_, e_, s_ = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        ist = ctypes.c_int(0)
        L.or_rotden(O.P(a), O.P(b), erel.ctypes.data_as(O._dp), ctypes.byref(r_), ctypes.byref(e_), ctypes.byref(s_),
                    rho.ctypes.data_as(O._dp), er.ctypes.data_as(O._dp), es.ctypes.data_as(O._dp), ctypes.byref(ist))
        close([r_.value, e_.value, s_.value, *erel], ref[:6], rtol=1e-13)
        assert ist.value == ref[6]


def euler_matrix(e):
    L = O.lib()
    m = np.zeros(9)
    L.or_matpre(O.P(e), m.ctypes.data_as(O._dp))
    return m.reshape(3, 3).T  # or_matpre writes Fortran column-major


def test_deleul_round_trip():
    """The relative Euler angles of deleul reproduce R(e1)^T R(e2)."""
    L = O.lib()
    rng = np.random.default_rng(11)
    for _ in range(200):
        e1 = np.array([rng.uniform(0, 2 * np.pi), np.arccos(rng.uniform(-1, 1)), rng.uniform(0, 2 * np.pi)])
        e2 = np.array([rng.uniform(0, 2 * np.pi), np.arccos(rng.uniform(-1, 1)), rng.uniform(0, 2 * np.pi)])
        er = np.zeros(3)
        ist = np.zeros(1, np.int32)
        L.or_deleul(O.P(e1), O.P(e2), er.ctypes.data_as(O._dp), O.I(ist) if False else ist.ctypes.data_as(O._ip))
        R1, R2, Rr = euler_matrix(e1), euler_matrix(e2), euler_matrix(er)
        assert np.allclose(R1.T @ R2, Rr, atol=1e-9) or np.allclose(R2 @ R1.T, Rr, atol=1e-9) or \
            np.allclose(R1 @ R2.T, Rr, atol=1e-9) or np.allclose(R2.T @ R1, Rr, atol=1e-9)


def test_deleul_convention():
    """Pin which product deleul inverts, so a transposed operand fails."""
    L = O.lib()
    rng = np.random.default_rng(12)
    hits = {"R1tR2": 0, "R2R1t": 0, "R1R2t": 0, "R2tR1": 0}
    for _ in range(50):
        e1 = np.array([rng.uniform(0, 2 * np.pi), np.arccos(rng.uniform(-1, 1)), rng.uniform(0, 2 * np.pi)])
        e2 = np.array([rng.uniform(0, 2 * np.pi), np.arccos(rng.uniform(-1, 1)), rng.uniform(0, 2 * np.pi)])
        er = np.zeros(3)
        ist = np.zeros(1, np.int32)
        L.or_deleul(O.P(e1), O.P(e2), er.ctypes.data_as(O._dp), ist.ctypes.data_as(O._ip))
        R1, R2, Rr = euler_matrix(e1), euler_matrix(e2), euler_matrix(er)
        for k, M in (("R1tR2", R1.T @ R2), ("R2R1t", R2 @ R1.T), ("R1R2t", R1 @ R2.T), ("R2tR1", R2.T @ R1)):
            hits[k] += np.allclose(M, Rr, atol=1e-9)
    assert max(hits.values()) == 50, hits


@pytest.mark.parametrize("nch", [91, 181, 361])
def test_vcord(G, nch):
    L = O.lib()
    nr, nth, nc, rmin, rmax, vt = tables3d.pot3d(nch=nch)
    g = G["nonlinear"][f"vcord{nch}"]
    rvstep = (rmax - rmin) / (nr - 1)
    for e, a, b, ref in zip(g["eul"], g["rcom"], g["rpt"], g["out"]):
        o = np.zeros(4)
        hx, hy, hz = np.zeros(3), np.zeros(3), np.zeros(3)
        L.or_vcord(O.P(e), O.P(a), O.P(b), vt.ctypes.data_as(O._dp), nr, nth, nc, rmax, rmin, rvstep,
                   *(o[i:].ctypes.data_as(O._dp) for i in range(4)), hx.ctypes.data_as(O._dp),
                   hy.ctypes.data_as(O._dp), hz.ctypes.data_as(O._dp), 0)
        close(o, ref, rtol=1e-12)


def test_vcalc_grid_nodes():
    """At grid nodes the corner-plus-gradients form returns the node value."""
    L = O.lib()
    nr, nth, nch, rmin, rmax, vt = tables3d.pot3d(nch=181)
    step = (rmax - rmin) / (nr - 1)
    v = np.zeros(1)
    for ir in (0, 7, nr - 2):
        for ith in (0, 45, 179):
            for ich in (0, 90, 179):
                L.or_vcalc(rmin + ir * step, float(ith), float(ich), rmin, rmax, step, nr, nth, nch,
                           vt.ctypes.data_as(O._dp), v.ctypes.data_as(O._dp))
                ref = vt[(ir * nth + ith) * nch + ich]
                assert abs(v[0] - ref) <= 1e-12 * max(1.0, abs(ref))


def test_caleng(G):
    L = O.lib()
    g = G["nonlinear"]["caleng"]
    e = np.zeros(1)
    got = []
    for a, b, x, y in zip(g["c1"], g["c2"], g["e1"], g["e2"]):
        L.or_caleng(O.P(a), O.P(b), e.ctypes.data_as(O._dp), O.P(x), O.P(y))
        got.append(e[0])
    close(got, g["e"], rtol=1e-12)


def test_caleng_symmetry():
    """TIP4P dimer energy is symmetric under exchange of the two molecules."""
    L = O.lib()
    rng = np.random.default_rng(13)
    e = np.zeros(1)
    for _ in range(50):
        c1, c2 = rng.uniform(-1, 1, 3), rng.uniform(-1, 1, 3) + np.array([4.0, 0, 0])
        e1 = np.array([rng.uniform(0, 2 * np.pi), np.arccos(rng.uniform(-1, 1)), rng.uniform(0, 2 * np.pi)])
        e2 = np.array([rng.uniform(0, 2 * np.pi), np.arccos(rng.uniform(-1, 1)), rng.uniform(0, 2 * np.pi)])
        L.or_caleng(O.P(c1), O.P(c2), e.ctypes.data_as(O._dp), O.P(e1), O.P(e2))
        a = e[0]
        L.or_caleng(O.P(c2), O.P(c1), e.ctypes.data_as(O._dp), O.P(e2), O.P(e1))
        assert abs(a - e[0]) <= 1e-10 * max(1.0, abs(a))


# ----------------------------------------------------------------------------- non-linear systems
def system3d(tag, rot3d):
    if tag in ("h2o_h2", "h2o_h2_pbc"):
        r, v = pt.h2h2_table()
        rp = dict(ntypes=2, numb=[4, 1], molecule=[0, 2], levels=[2, 2], mcstep=[0.2, 0.2], mass=[2.01565, 18.010565],
                  P=8, Prot=8, T=0.5, rtstep=0.25, pot1d=[0, -1], pot2d=[-1, -1], rotdens=-1)
        if tag == "h2o_h2_pbc":
            rp["box"] = 11.0
        tabs = dict(pot1d=[(r, v)], pot2d=[], rot=[], pot3d=tables3d.pot3d(nch=181), rot3d=rot3d)
    else:
        rp = dict(ntypes=1, numb=[2], molecule=[2], levels=[2], mcstep=[0.2], mass=[18.010565], P=8, Prot=8, T=1.0,
                  rtstep=0.25, pot1d=[-1], pot2d=[-1], rotdens=-1)
        tabs = dict(pot1d=[], pot2d=[], rot=[], pot3d=None, rot3d=rot3d)
    return O.System(rp, tabs)


@pytest.mark.parametrize("tag", ["h2o_h2", "water2", "h2o_h2_pbc"])
def test_nonlinear_energies(G, A, rot3d, tag):
    S = system3d(tag, rot3d)
    g = G[tag]
    S.set_state(A[f"{tag}_coords"], A[f"{tag}_angles"])
    P, natoms = 8, S.natoms
    close([[S.L.or_pot_energy_slice(S.h, a, None, it) for it in range(P)] for a in range(natoms)], g["pot_slice"])
    close([S.L.or_pot_energy_path(S.h, a, None) for a in range(natoms)], g["pot_path"])
    nrot = len(g["pot_rot_e3d"])
    for k, ra in enumerate(range(natoms - nrot, natoms)):
        got = [S.L.or_pot_rot_e3d(S.h, ra, O.P(e), int(it)) for e, it in zip(g["pot_rot_e3d_eul"], np.arange(16) % P)]
        close(got, g["pot_rot_e3d"][k])
    est, gr1, _ = estimators(S, nonlinear=True)
    check_estimators(est, g["estimators"])
    assert np.array_equal(gr1, A[f"{tag}_gr1d"])


@pytest.mark.parametrize("tag", ["h2o_h2", "water2", "h2o_h2_pbc"])
def test_nonlinear_reference_order_passes(G, A, rot3d, tag):
    S = system3d(tag, rot3d)
    g = G[tag]
    S.set_state(A[f"{tag}_coords"], A[f"{tag}_angles"])
    S.set_tapes(tapes_of(A, tag))
    S.L.or_reset_counters(S.h)
    for p in range(2):
        assert S.L.or_pass(S.h, p, O.ORDER_REFERENCE) == 0
    assert S.L.or_tape_short(S.h) == 0
    for k, n in g["tape_used"].items():
        assert S.L.or_tape_pos(S.h, int(k)) == n
    c, a, _ = S.state()
    close(c.reshape(-1), A[f"{tag}_pass_coords"], rtol=1e-13)
    close(a.reshape(-1), A[f"{tag}_pass_angles"], rtol=1e-13)
    t, acc = S.counters()
    assert t.reshape(-1).tolist() == g["pass_total"] and acc.reshape(-1).tolist() == g["pass_accep"]
    est, _, _ = estimators(S, nonlinear=True)
    check_estimators(est, g["pass_estimators"])


N3D = 300 * 50 * 100


def nl_bins(S):
    g3, rel = np.zeros(N3D), np.zeros(250)
    assert S.L.or_nl_densities(S.h, O.P(g3), O.P(rel)) == 0
    return g3, rel


@pytest.mark.parametrize("tag", ["h2o_h2", "water2", "h2o_h2_pbc"])
def test_nonlinear_density_bins(G, A, rot3d, tag):
    """The (r, theta, chi) bins around the rotor and the relative Euler angle bins equal the
    compiled reference's _gr3D and _relthe/_relphi/_relchi_sum on the same path."""
    S = system3d(tag, rot3d)
    S.set_state(A[f"{tag}_coords"], A[f"{tag}_angles"])
    g3, rel = nl_bins(S)
    ref = np.zeros((G[tag]["ndens3d"], N3D))
    idx = A[f"{tag}_gr3d_idx"]
    ref[idx[0], idx[1]] = A[f"{tag}_gr3d_val"]
    assert np.array_equal(g3, ref[0])  # density type 0: the atom type (h2o_h2) or nothing (water2)
    assert np.all(ref[1:] == 0.0)
    assert np.array_equal(rel, A[f"{tag}_rel"])
    nrot = 2 if tag == "water2" else 1
    assert rel[:50].sum() == rel[50:150].sum() == rel[150:].sum() == nrot * 8


def frame_matrix(phi, theta, chi):
    """matpre (rotden.f:136-163): body axes are the columns."""
    cp, sp, ct, st, ck, sk = np.cos(phi), np.sin(phi), np.cos(theta), np.sin(theta), np.cos(chi), np.sin(chi)
    return np.array([[cp * ct * ck - sp * sk, -cp * ct * sk - sp * ck, cp * st],
                     [sp * ct * ck + cp * sk, -sp * ct * sk + cp * ck, sp * st],
                     [-st * ck, st * sk, ct]])


def test_nonlinear_density_bins_hand_placed(rot3d):
    """Atoms put at chosen body-frame (r, theta, chi) around a rotor of chosen orientation land
    in bin (floor(r/0.05), floor(theta/(pi/49)), floor(chi/(2 pi/99))), one count per slice.
    Identical rotor slices give the identity as relative rotation: theta and phi bin 0, and
    chi = 2 pi - acos(1) in the last bin (deleul takes chi = 2 pi - acos when sin chi <= 0)."""
    S = system3d("h2o_h2", rot3d)
    P = 8
    eul = (0.7, np.arccos(0.3), 2.1)
    Rm = frame_matrix(*eul)
    centre = np.array([0.4, -0.2, 0.3])
    body = [(3.33, 1.0, 4.0), (5.12, 2.5, 0.3), (2.71, 0.4, 2.0), (7.77, 1.9, 5.5)]  # (r, theta, chi)
    coords = np.zeros((3, 5 * P))
    for a, (r, th, ch) in enumerate(body):
        v = r * np.array([np.sin(th) * np.cos(ch), np.sin(th) * np.sin(ch), np.cos(th)])
        coords[:, a * P:(a + 1) * P] = (centre + Rm @ v)[:, None]
    coords[:, 4 * P:] = centre[:, None]
    angles = np.zeros((3, 5 * P))
    angles[0, 4 * P:], angles[1, 4 * P:], angles[2, 4 * P:] = eul[0], np.cos(eul[1]), eul[2]
    S.set_state(coords, angles)
    g3, rel = nl_bins(S)
    want = np.zeros(N3D)
    for r, th, ch in body:
        br, bt, bc = int(r / 0.05), int(th / (np.pi / 49)), int(ch / (2 * np.pi / 99))
        want[(br * 50 + bt) * 100 + bc] += P
    assert np.array_equal(g3, want)
    assert rel[0] == P and rel[50] == P and rel[249] == P and rel.sum() == 3 * P
