#!/usr/bin/env python3
"""Writes the golden vectors of tests/golden/ from the compiled reference.

The reference routines run from oracle/_ref/libref.so, which oracle/ref.mk builds from the
upstream MoRiBS-PIMC sources where they lie (see oracle/ref_harness.cc).  This script runs in
the build container only, where those sources exist:

    make -C oracle ref && python3 tests/golden/make_golden.py

Each section runs in a child process of its own, since the reference keeps one system in
process globals.  Output: ref_vectors.json (scalars and short vectors) and ref_arrays.npz
(paths, tapes and histograms; plain arrays, no pickle).

Sections
  linear     3 He + 1 linear rotor, P = 16 (the 'tiny' deck of pimc_configs): SPot1D,
             LPot2D, SRotDens*, PotEnergy (slice, path, on MCCoords and on newcoords),
             PotRotEnergy, the estimators, and two reference-order passes on tapes
  helium     8 He on the path of examples/MF_8He_0.37K_512_128/xyz.init with
             examples/MF_1He_0.37K_512_128/helium.pot: SPot1D on the table and the
             estimators V and K
  nonlinear  rotden/deleul/rotpro, vcord/vcalc (CHgrd 91, 181, 361) and caleng on seeded
             inputs, with the stand-in tables of tables3d.py
  h2o_h2     4 para-H2 + 1 non-linear rotor, P = Prot = 8: PotEnergy, PotRotE3D, GetRotE3D,
             GetPotEnergy_Densities (with its 3D bins and the relative Euler angle bins) and
             two reference-order passes on tapes
  water2     2 non-linear rotors (TIP4P via caleng), P = Prot = 8: the same quantities
  periodic   16 He in a cubic box of 8.5 Angstrom under MINIMAGE, P = 16, on helium.pot:
             PotEnergy (slice, path, on MCCoords and on newcoords), the estimators V (with its
             g(r) bins) and K, and two reference-order passes on tapes
  h2o_h2_pbc the h2o_h2 system under MINIMAGE in a box of 11 Angstrom (para-H2 pairs take
             their nearest image, the vcord pairs the positions): the h2o_h2 quantities
"""
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REFTREE = "/root/reference"
sys.path.insert(0, os.path.join(ROOT, "moribs-pimc_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, HERE)

SEED = 20240611
TAPE_STREAMS = (1, 2, 3, 8, 9, 15)


def _lib():
    import oracle as O
    import ctypes
    R = O.ref()
    R.h_set_pot3d.restype = None
    R.h_set_pot3d.argtypes = [ctypes.c_int] * 3 + [ctypes.c_double] * 2 + [O._dp]
    return O, R


def _dp(a):
    import oracle as O
    return O.P(a)


def make_tapes(rng, n):
    t = {k: rng.uniform(0.0, 1.0, n) for k in TAPE_STREAMS}
    t[8] = 1.0 - t[8]  # gauss' first uniform enters log(); keep it in (0, 1]
    return t


def run_passes(R, P, ntypes, npasses):
    for _ in range(npasses):
        for t in range(P):
            for ty in range(ntypes):
                R.h_pimc_pass(ty, t)


def tape_record(R, tapes, prefix, arrays):
    """Keep only the consumed part of each tape."""
    used = {}
    for k in TAPE_STREAMS:
        pos = int(R.h_tape_pos(k))
        used[k] = pos
        arrays[f"{prefix}_tape{k}"] = tapes[k][:pos].copy()
    assert R.h_tape_short() == 0, "a tape ran short"
    return used


def get_state(R, n):
    c, a, cs = np.empty(3 * n), np.empty(3 * n), np.empty(3 * n)
    R.h_get_state(c.ctypes.data_as(_typ()), a.ctypes.data_as(_typ()), cs.ctypes.data_as(_typ()))
    return c, a, cs


def _typ():
    import ctypes
    return ctypes.POINTER(ctypes.c_double)


def counters(R, ntypes):
    t, a = np.zeros(3 * ntypes), np.zeros(3 * ntypes)
    R.h_get_counters(t.ctypes.data_as(_typ()), a.ctypes.data_as(_typ()))
    return t.tolist(), a.tolist()


def estimators(R, nonlinear=False):
    import ctypes
    e1, e2 = ctypes.c_double(), ctypes.c_double()
    gr1 = np.zeros(300)
    gr2 = np.zeros(300 * 50)
    out = dict(kin=R.h_kin_energy(), pot=R.h_pot_energy_densities(gr1.ctypes.data_as(_typ()), gr2.ctypes.data_as(_typ())))
    if nonlinear:
        out["rot"] = R.h_rot_e3d(ctypes.byref(e1), ctypes.byref(e2))
    else:
        out["rot"] = R.h_rot_energy(ctypes.byref(e1), ctypes.byref(e2))
    out["rotsq"], out["rot_termsq"] = e1.value, e2.value
    return out, gr1, gr2


# ----------------------------------------------------------------------------- linear
def section_linear(arrays):
    import ctypes
    import pimc_configs as pc
    import pimc_tables as pt
    O, R = _lib()
    wl = pc.workload("tiny")
    rp, tabs = wl["replicas"][0], wl["tables"]
    d = tempfile.mkdtemp(prefix="golden_lin_")
    r, v = tabs["pot1d"][0]
    pt.write_pot1d(os.path.join(d, "He.pot"), r, v)
    rg, cg, dr, dc, V = tabs["pot2d"][0]
    pt.write_pot2d(os.path.join(d, "CO2He.pot"), rg, cg, V, dr, dc)
    c, rho, er, es = tabs["rot"][0]
    pt.write_rot(os.path.join(d, pc.rot_file_name(rp)), c, rho, er, es)
    rc = R.h_setup(2, b"He4\0CO2\0", O.I(rp["numb"]), O.I(rp["molecule"]), O.I(rp["levels"]), O.P(rp["mcstep"]),
                   O.P(rp["mass"]), rp["P"], rp["Prot"], rp["T"], rp["rtstep"], d.encode(), b"He\0CO2He\0")
    assert rc == 0
    rng = np.random.default_rng(SEED)
    out = {}
    # pair tables: knots, the exponential and -C6/r^6 branches, interior points
    x1 = np.concatenate([r[::23], [r[0], r[-1], 0.5, 0.9, 30.5, 45.0], rng.uniform(0.6, 32.0, 300)])
    out["spot1d_r"] = x1.tolist()
    out["spot1d"] = [R.h_spot1d(x, 0) for x in x1]
    xr = np.concatenate([rg[::97], rng.uniform(1.0, 25.0, 300), [1.2, 22.0, 30.0]])
    xc = np.concatenate([cg[::20][: len(rg[::97])], rng.uniform(-1.2, 1.2, 300), [1.1, -1.05, 0.0]])
    n2 = min(len(xr), len(xc))
    out["lpot2d_r"], out["lpot2d_c"] = xr[:n2].tolist(), xc[:n2].tolist()
    out["lpot2d"] = [R.h_lpot2d(a, b, 1) for a, b in zip(xr[:n2], xc[:n2])]
    g = np.concatenate([c[::101], [-1.0, 1.0, -1.004, 1.003], rng.uniform(-1.01, 1.01, 300)])
    out["srot_g"] = g.tolist()
    out["srotdens"] = [R.h_srotdens(x, 1) for x in g]
    out["srotdens_deriv"] = [R.h_srotdens_deriv(x, 1) for x in g]
    out["srotdens_esqrt"] = [R.h_srotdens_esqrt(x, 1) for x in g]
    # a perturbed state
    n = (rp["numb"][0] + 1) * rp["P"]
    coords = rp["coords"] + rng.normal(0.0, 0.35, rp["coords"].shape)
    angles = np.zeros_like(rp["angles"])
    angles[0] = rng.uniform(0.0, 2.0 * np.pi, angles.shape[1])
    angles[1] = rng.uniform(-1.0, 1.0, angles.shape[1])
    angles[2] = rng.uniform(0.0, 2.0 * np.pi, angles.shape[1])
    newc = coords + rng.normal(0.0, 0.2, coords.shape)
    arrays["linear_coords"], arrays["linear_angles"], arrays["linear_newcoords"] = coords, angles, newc
    R.h_set_state(O.P(coords), O.P(angles))
    R.h_set_newcoords(O.P(newc))
    natoms = rp["numb"][0] + 1
    out["pot_slice"] = [[R.h_pot_energy_slice(a, 0, it) for it in range(rp["P"])] for a in range(natoms)]
    out["pot_slice_new"] = [[R.h_pot_energy_slice(a, 1, it) for it in range(rp["P"])] for a in range(natoms)]
    out["pot_path"] = [R.h_pot_energy_path(a, 0) for a in range(natoms)]
    out["pot_path_new"] = [R.h_pot_energy_path(a, 1) for a in range(natoms)]
    out["pot_rot"] = [R.h_pot_rot_energy(natoms - 1, 0, it) for it in range(rp["P"])]
    est, gr1, gr2 = estimators(R)
    out["estimators"] = est
    arrays["linear_gr1d"], arrays["linear_gr2d"] = gr1, gr2
    # two reference-order passes on tapes from the lattice start
    tapes = make_tapes(rng, 20000)
    for k, t in tapes.items():
        R.h_set_tape(k, O.P(t), len(t))
    R.h_set_state(O.P(rp["coords"]), O.P(rp["angles"]))
    R.h_reset_counters()
    run_passes(R, rp["P"], 2, 2)
    out["tape_used"] = tape_record(R, tapes, "linear", arrays)
    c1, a1, _ = get_state(R, n)
    arrays["linear_pass_coords"], arrays["linear_pass_angles"] = c1, a1
    out["pass_total"], out["pass_accep"] = counters(R, 2)
    est, _, _ = estimators(R)
    out["pass_estimators"] = est
    del ctypes
    return out


# ----------------------------------------------------------------------------- helium
def read_xyz_init(path, natoms, P):
    """Lines 3.. of xyz.init: name x phi y cos(theta) z chi (mc_input.cc:690-720 layout)."""
    with open(path) as f:
        f.readline()
        f.readline()
        rows = [f.readline().split() for _ in range(natoms * P)]
    a = np.array([[float(t) for t in row[1:7]] for row in rows])
    coords = np.stack([a[:, 0], a[:, 2], a[:, 4]])
    angles = np.stack([a[:, 1], a[:, 3], a[:, 5]])
    return coords, angles


def section_helium(arrays):
    import pimc_tables as pt
    O, R = _lib()
    pot = os.path.join(REFTREE, "examples/MF_1He_0.37K_512_128/helium.pot")
    xyz = os.path.join(REFTREE, "examples/MF_8He_0.37K_512_128/xyz.init")
    P, N = 512, 8
    coords, angles = read_xyz_init(xyz, N, P)
    arrays["helium_coords"], arrays["helium_angles"] = coords, angles
    d = tempfile.mkdtemp(prefix="golden_he_")
    r, v = pt.read_pot1d(pot)
    with open(pot) as fi, open(os.path.join(d, "helium.pot"), "w") as fo:
        fo.write(fi.read())
    rc = R.h_setup(1, b"He4\0", O.I([N]), O.I([0]), O.I([3]), O.P([0.2]), O.P([4.0026032497]), P, 0, 0.37, 0.0,
                   d.encode(), b"helium\0")
    assert rc == 0
    R.h_set_state(O.P(coords), O.P(angles))
    rng = np.random.default_rng(SEED + 1)
    x = np.concatenate([r[::17], [r[0] - 0.3, r[-1] + 2.0], rng.uniform(r[0] - 0.5, r[-1] + 5.0, 200)])
    out = {"spot1d_r": x.tolist(), "spot1d": [R.h_spot1d(q, 0) for q in x]}
    est, gr1, _ = estimators(R)
    out["kin"], out["pot"] = est["kin"], est["pot"]
    arrays["helium_gr1d"] = gr1
    out["pot_path"] = [R.h_pot_energy_path(a, 0) for a in range(N)]
    return out


# ----------------------------------------------------------------------------- non-linear functions
def random_euler(rng, n):
    e = np.empty((n, 3))
    e[:, 0] = rng.uniform(0.0, 2.0 * np.pi, n)
    e[:, 1] = np.arccos(rng.uniform(-1.0, 1.0, n))
    e[:, 2] = rng.uniform(0.0, 2.0 * np.pi, n)
    return e


def section_nonlinear(arrays):
    import ctypes
    import tables3d
    O, R = _lib()
    rng = np.random.default_rng(SEED + 2)
    rho, erot, esq = tables3d.rot3d()
    R.h_set_rot3d(O.P(rho), O.P(erot), O.P(esq))
    e1 = random_euler(rng, 400)
    e2 = random_euler(rng, 400)
    # degenerate relative orientations: equal, nearly equal, flipped about x, theta = 0 and pi
    e2[:20] = e1[:20]
    e2[20:40] = e1[20:40] + rng.normal(0.0, 1e-9, (20, 3))
    e1[40:50, 1] = 0.0
    e2[50:60, 1] = np.pi
    e2[60:70] = e1[60:70]
    e2[60:70, 1] = np.pi - e1[60:70, 1]
    e2[60:70, 0] = e1[60:70, 0] + np.pi
    out = {"rotden_e1": e1.tolist(), "rotden_e2": e2.tolist(), "rotden": []}
    for a, b in zip(e1, e2):
        a = np.ascontiguousarray(a)
        b = np.ascontiguousarray(b)
        erel = np.zeros(3)
        r_, e_, s_ = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        ist = ctypes.c_int(0)
        R.h_rotden(O.P(a), O.P(b), erel.ctypes.data_as(_typ()), ctypes.byref(r_), ctypes.byref(e_),
                   ctypes.byref(s_), ctypes.byref(ist))
        out["rotden"].append([r_.value, e_.value, s_.value, erel[0], erel[1], erel[2], ist.value])
    # vcord on three chi foldings
    for nch in (91, 181, 361):
        nr, nth, nc, rmin, rmax, vt = tables3d.pot3d(nch=nch)
        R.h_set_pot3d(nr, nth, nc, rmin, rmax, O.P(vt))
        eul = random_euler(rng, 300)
        rcom = rng.uniform(-3.0, 3.0, (300, 3))
        dirs = rng.normal(size=(300, 3))
        dirs /= np.linalg.norm(dirs, axis=1)[:, None]
        dist = rng.uniform(1.5, 14.0, 300)
        dist[:10] = [1.0, 1.9, 2.1, 12.3, 12.5, 20.0, 3.0, 3.0, 3.0, 3.0]
        rpt = rcom + dirs * dist[:, None]
        rpt[6] = rcom[6] + np.array([0.0, 0.0, 3.0])   # along the body axes of an identity orientation
        eul[6] = 0.0
        rpt[7] = rcom[7] + np.array([3.0, 0.0, 0.0])
        eul[7] = 0.0
        res = []
        for a, b, c in zip(eul, rcom, rpt):
            o = np.zeros(4)
            R.h_vcord(O.P(a), O.P(b), O.P(c), o.ctypes.data_as(_typ()))
            res.append(o.tolist())
        out[f"vcord{nch}"] = dict(eul=eul.tolist(), rcom=rcom.tolist(), rpt=rpt.tolist(), out=res)
    # caleng on water pairs
    c1 = rng.uniform(-1.0, 1.0, (300, 3))
    dirs = rng.normal(size=(300, 3))
    dirs /= np.linalg.norm(dirs, axis=1)[:, None]
    c2 = c1 + dirs * rng.uniform(2.4, 9.0, 300)[:, None]
    ea, eb = random_euler(rng, 300), random_euler(rng, 300)
    out["caleng"] = dict(c1=c1.tolist(), c2=c2.tolist(), e1=ea.tolist(), e2=eb.tolist(),
                         e=[R.h_caleng(O.P(a), O.P(b), O.P(x), O.P(y)) for a, b, x, y in zip(c1, c2, ea, eb)])
    return out


# ----------------------------------------------------------------------------- non-linear systems
def _system3d(arrays, tag, ntypes, names, numb, molecule, pots, P, Prot, T, seed, with_pot3d, box=None):
    import ctypes
    import pimc_tables as pt
    import tables3d
    O, R = _lib()
    d = tempfile.mkdtemp(prefix=f"golden_{tag}_")
    if with_pot3d:
        tables3d.write_pot3d(os.path.join(d, "H2Opot.pot"), *tables3d.pot3d(nch=181))
        r, v = pt.h2h2_table()
        pt.write_pot1d(os.path.join(d, "pH2.pot"), r, v)
    levels = [2] * ntypes
    mcstep = [0.2] * ntypes
    mass = [2.01565 if m == 0 else 18.010565 for m in molecule]
    rc = R.h_setup(ntypes, names, O.I(numb), O.I(molecule), O.I(levels), O.P(mcstep), O.P(mass), P, Prot, T, 0.25,
                   d.encode(), pots)
    assert rc == 0
    rho, erot, esq = tables3d.rot3d()
    R.h_set_rot3d(O.P(rho), O.P(erot), O.P(esq))
    if box is not None:
        R.h_set_minimage(1, O.P([box] * 3))
    rng = np.random.default_rng(seed)
    natoms = sum(numb)
    n = natoms * P
    coords = np.zeros((3, n))
    for a in range(natoms):
        if molecule[-1] == 2 and a == natoms - 1 and with_pot3d:
            centre = np.zeros(3)
        else:
            ang = 2.0 * np.pi * a / natoms
            centre = np.array([3.6 * np.cos(ang), 3.6 * np.sin(ang), 0.7 * (a % 2)])
        coords[:, a * P:(a + 1) * P] = centre[:, None] + rng.normal(0.0, 0.15, (3, P))
    angles = np.zeros((3, n))
    angles[0] = rng.uniform(0.0, 2.0 * np.pi, n)
    angles[1] = rng.uniform(-1.0, 1.0, n)
    angles[2] = rng.uniform(0.0, 2.0 * np.pi, n)
    arrays[f"{tag}_coords"], arrays[f"{tag}_angles"] = coords, angles
    R.h_set_state(O.P(coords), O.P(angles))
    out = {"pot_slice": [[R.h_pot_energy_slice(a, 0, it) for it in range(P)] for a in range(natoms)],
           "pot_path": [R.h_pot_energy_path(a, 0) for a in range(natoms)]}
    rot_atoms = [a for a in range(natoms) if a >= natoms - numb[-1]]
    eul = random_euler(rng, 16)
    out["pot_rot_e3d_eul"] = eul.tolist()
    out["pot_rot_e3d"] = [[R.h_pot_rot_e3d(ra, O.P(e), it) for e, it in zip(eul, np.arange(16) % P)] for ra in rot_atoms]
    est, gr1, _ = estimators(R, nonlinear=True)
    out["estimators"] = est
    arrays[f"{tag}_gr1d"] = gr1
    # bins around the non-linear rotor (_gr3D and the relative Euler angle bins), kept sparse
    n3 = 300 * 50 * 100
    g3, rel = np.zeros(2 * n3), np.zeros(250)
    nd = R.h_nl_densities(O.P(g3), O.P(rel))
    g3 = g3[:nd * n3].reshape(nd, n3)
    nz = np.nonzero(g3)
    arrays[f"{tag}_gr3d_idx"] = np.stack(nz).astype(np.int64)
    arrays[f"{tag}_gr3d_val"] = g3[nz]
    arrays[f"{tag}_rel"] = rel
    out["ndens3d"] = nd
    tapes = make_tapes(rng, 40000)
    for k, t in tapes.items():
        R.h_set_tape(k, O.P(t), len(t))
    R.h_reset_counters()
    run_passes(R, P, ntypes, 2)
    out["tape_used"] = tape_record(R, tapes, tag, arrays)
    c1, a1, _ = get_state(R, n)
    arrays[f"{tag}_pass_coords"], arrays[f"{tag}_pass_angles"] = c1, a1
    out["pass_total"], out["pass_accep"] = counters(R, ntypes)
    est, _, _ = estimators(R, nonlinear=True)
    out["pass_estimators"] = est
    del ctypes
    return out


def section_h2o_h2(arrays):
    return _system3d(arrays, "h2o_h2", 2, b"H2\0H2O\0", [4, 1], [0, 2], b"pH2\0H2Opot\0", 8, 8, 0.5, SEED + 3, True)


def section_water2(arrays):
    return _system3d(arrays, "water2", 1, b"H2O\0", [2], [2], b"none\0", 8, 8, 1.0, SEED + 4, False)


def section_h2o_h2_pbc(arrays):
    return _system3d(arrays, "h2o_h2_pbc", 2, b"H2\0H2O\0", [4, 1], [0, 2], b"pH2\0H2Opot\0", 8, 8, 0.5, SEED + 3,
                     True, box=11.0)


# ----------------------------------------------------------------------------- periodic helium
def section_periodic(arrays):
    import pimc_tables as pt
    O, R = _lib()
    pot = os.path.join(REFTREE, "examples/MF_1He_0.37K_512_128/helium.pot")
    P, N, L = 16, 16, 8.5
    d = tempfile.mkdtemp(prefix="golden_pbc_")
    with open(pot) as fi, open(os.path.join(d, "helium.pot"), "w") as fo:
        fo.write(fi.read())
    rc = R.h_setup(1, b"He4\0", O.I([N]), O.I([0]), O.I([3]), O.P([0.6]), O.P([4.0026032497]), P, 0, 1.5, 0.0,
                   d.encode(), b"helium\0")
    assert rc == 0
    R.h_set_minimage(1, O.P([L, L, L]))
    rng = np.random.default_rng(SEED + 5)
    centres = []
    while len(centres) < N:  # atoms anywhere in the box, at least 2.6 Angstrom from every image of the others
        x = rng.uniform(-0.5 * L, 0.5 * L, 3)
        if all(np.linalg.norm(x - c - L * np.rint((x - c) / L)) > 2.6 for c in centres):
            centres.append(x)
    coords = np.zeros((3, N * P))
    for a, c in enumerate(centres):
        coords[:, a * P:(a + 1) * P] = c[:, None] + rng.normal(0.0, 0.25, (3, P))
    angles = np.zeros_like(coords)
    newc = coords + rng.normal(0.0, 0.2, coords.shape)
    arrays["periodic_coords"], arrays["periodic_newcoords"] = coords, newc
    R.h_set_state(O.P(coords), O.P(angles))
    R.h_set_newcoords(O.P(newc))
    out = {"box": L,
           "pot_slice": [[R.h_pot_energy_slice(a, 0, it) for it in range(P)] for a in range(N)],
           "pot_slice_new": [[R.h_pot_energy_slice(a, 1, it) for it in range(P)] for a in range(N)],
           "pot_path": [R.h_pot_energy_path(a, 0) for a in range(N)],
           "pot_path_new": [R.h_pot_energy_path(a, 1) for a in range(N)]}
    est, gr1, _ = estimators(R)
    out["kin"], out["pot"] = est["kin"], est["pot"]
    arrays["periodic_gr1d"] = gr1
    tapes = make_tapes(rng, 40000)
    for k, t in tapes.items():
        R.h_set_tape(k, O.P(t), len(t))
    R.h_reset_counters()
    run_passes(R, P, 1, 2)
    out["tape_used"] = tape_record(R, tapes, "periodic", arrays)
    c1, _, _ = get_state(R, N * P)
    arrays["periodic_pass_coords"] = c1
    out["pass_total"], out["pass_accep"] = counters(R, 1)
    est, gr1, _ = estimators(R)
    out["pass_kin"], out["pass_pot"] = est["kin"], est["pot"]
    arrays["periodic_pass_gr1d"] = gr1
    return out


SECTIONS = dict(linear=section_linear, helium=section_helium, nonlinear=section_nonlinear,
                h2o_h2=section_h2o_h2, water2=section_water2, periodic=section_periodic,
                h2o_h2_pbc=section_h2o_h2_pbc)


def main():
    if len(sys.argv) > 2 and sys.argv[1] == "--section":
        name, dst = sys.argv[2], sys.argv[3]
        arrays = {}
        out = SECTIONS[name](arrays)
        np.savez(dst + ".npz", **arrays)
        with open(dst + ".json", "w") as f:
            json.dump(out, f)
        return
    if not os.path.isdir(REFTREE):
        sys.exit("the reference tree is not present; golden vectors are made in the build container")
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    tmp = tempfile.mkdtemp(prefix="golden_")
    allj, alla = {}, {}
    for name in SECTIONS:
        dst = os.path.join(tmp, name)
        env = dict(os.environ, OMP_NUM_THREADS="1")
        subprocess.run([sys.executable, os.path.abspath(__file__), "--section", name, dst], check=True, env=env)
        with open(dst + ".json") as f:
            allj[name] = json.load(f)
        with np.load(dst + ".npz") as z:
            alla.update({k: z[k] for k in z.files})
        print("section", name, "done", flush=True)
    with open(os.path.join(HERE, "ref_vectors.json"), "w") as f:
        json.dump(allj, f, separators=(",", ":"))
    np.savez_compressed(os.path.join(HERE, "ref_arrays.npz"), **alla)


if __name__ == "__main__":
    main()
