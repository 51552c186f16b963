"""Workload definitions of the linear-rotor decks (one linear rotor in a He cluster).

  c0  8 He + 1 linear rotor (B = 0.39 cm^-1), T = 0.37 K, P = Prot = 128, LJ He-He
      (eps 10.22 K, sigma 2.556 A, 500 points on 1-30 A), Morse-Legendre rotor-He
      (D = 30, 8, 2 K, a = 1.2 1/A, R0 = 3.5 A), rotor density on 2001 cos(gamma) points
  c1  as c0 with 64 He and P = Prot = 512
  c2  256 replicas from a seed: N ~ U{1..100} He, P in {64,128,256,512}, T ~ U(0.3, 2) K

Tables are generated from the closed forms, passed through the reference text formats
(so the GPU and the compiled reference read identical numbers), and the initial path of
each replica is a seeded classical start: the rotor at the origin along z, He atoms at
random points of a sphere with a 2.6 A exclusion, every bead of an atom at one point
(the reference starts all beads of an atom at one point too, mc_setup.cc:700-707).
"""
import numpy as np

from pimc_tables import (lj_table, morse_legendre_2d, linden, rot_filename, _e15_8, fmt_temperature)

MASS_HE4 = 4.0026032497                 # mc_const.h:44
MASS_CO2 = 12.0 + 2.0 * 15.994915       # mc_setup.cc:275-279 (CO2 type)
B_ROTOR_CM = 0.39

STEP_HE = 0.15      # centroid step, Angstrom
STEP_ROTOR = 0.08
ROT_STEP = 0.3      # rotational step (cos theta and phi)
LEVELS = 3


def _round_e15(a):
    return np.array([float(_e15_8(x)) for x in a])


def rot_table(T, P, npt=2001):
    """The four .rot columns at (T, Prot = P), rounded as written to the file."""
    c, rho, er, es = linden(T, P, B_ROTOR_CM, npt=npt)
    return tuple(_round_e15(x) for x in (c, rho, er, es))


def initial_path(n_he, P, rng):
    """[3][natoms*P] coordinates and angles: He atoms first, the rotor last."""
    natoms = n_he + 1
    radius = 3.5 + 1.6 * max(n_he, 1) ** (1.0 / 3.0)
    pts = []
    while len(pts) < n_he:
        x = rng.uniform(-radius, radius, 3)
        rr = np.linalg.norm(x)
        if rr > radius or rr < 3.0:
            continue
        if all(np.linalg.norm(x - p) > 2.6 for p in pts):
            pts.append(x)
    pts.append(np.zeros(3))
    coords = np.zeros((3, natoms * P))
    for a, p in enumerate(pts):
        coords[:, a * P:(a + 1) * P] = p[:, None]
    angles = np.zeros((3, natoms * P))
    angles[1, :] = 1.0  # cos(theta) = 1, phi = chi = 0
    return coords, angles


def replica(n_he, P, T, seed, Prot=None):
    rng = np.random.default_rng(seed)
    coords, angles = initial_path(n_he, P, rng)
    return dict(
        ntypes=2, type_names=["He4", "CO2"], numb=[n_he, 1], molecule=[0, 1], levels=[LEVELS, LEVELS],
        mcstep=[STEP_HE, STEP_ROTOR], mass=[MASS_HE4, MASS_CO2], P=P, Prot=Prot or P, T=T, rtstep=ROT_STEP,
        coords=coords, angles=angles,
    )


def workload(name, seed=985456376, n_replicas=256):
    """Tables and replicas of workload `name` ('c0', 'c1', 'c2', 'tiny', 'mini' or 'classes': one
    replica per bisection size class and boundary, N in {20, 30, 63, 64, 100} He, P 256-512)."""
    r1, v1 = lj_table()
    rg, cg, V = morse_legendre_2d()
    tabs = dict(pot1d=[(r1, v1)], pot2d=[(rg, cg, rg[1] - rg[0], cg[1] - cg[0], V)], rot=[])
    reps = []
    if name == "c0":
        reps.append(replica(8, 128, 0.37, seed))
    elif name == "c1":
        reps.append(replica(64, 512, 0.37, seed))
    elif name == "tiny":
        reps.append(replica(3, 16, 0.5, seed))
    elif name == "mini":  # a small uneven batch for replica-level parity checks
        rng = np.random.default_rng(seed)
        for r in range(n_replicas if n_replicas < 256 else 6):
            n = int(rng.integers(1, 13))
            P = int(rng.choice([16, 32, 64]))
            T = float(fmt_temperature(float(rng.uniform(0.3, 2.0))))
            reps.append(replica(n, P, T, seed + 1 + r))
    elif name == "classes":  # every bisection size class of the linear family, long chains
        for r, (n, P, T) in enumerate([(20, 512, 0.41), (30, 512, 0.93), (63, 256, 1.6), (64, 512, 0.75),
                                       (100, 512, 1.2)]):
            reps.append(replica(n, P, float(fmt_temperature(T)), seed + 1 + r))
    elif name == "c2":
        rng = np.random.default_rng(seed)
        for r in range(n_replicas):
            n = int(rng.integers(1, 101))
            P = int(rng.choice([64, 128, 256, 512]))
            T = float(rng.uniform(0.3, 2.0))
            T = float(fmt_temperature(T))  # the value the .rot file name encodes
            reps.append(replica(n, P, T, seed + 1 + r))
    else:
        raise ValueError(name)
    keys = {}
    for rp in reps:
        key = (rp["T"], rp["Prot"])
        if key not in keys:
            keys[key] = len(tabs["rot"])
            tabs["rot"].append(rot_table(*key))
        rp["rotdens"] = keys[key]
        rp["pot1d"] = [0, -1]
        rp["pot2d"] = [-1, 0]
    return dict(name=name, tables=tabs, replicas=reps)


# ---------------------------------------------------------------- non-linear rotor workloads
MASS_H2 = 2.015650642                   # mc_const.h:45
MASS_H2O = 2.0 * 1.0078 + 15.994915     # mc_setup.cc:303-306
ABC_H2O = (27.88, 14.51, 9.28)          # asymmetric-top rotational constants, cm^-1


def nl_initial_path(n_solvent, n_rotors, P, rng, shell=(5.0, 7.5), sep=3.0):
    """Rotors near the origin (sep apart), solvent molecules on a shell around them; every bead of
    a molecule at one point; random orientations."""
    natoms = n_solvent + n_rotors
    pts = []
    while len(pts) < n_solvent:
        x = rng.normal(size=3)
        x *= rng.uniform(*shell) / np.linalg.norm(x)
        if all(np.linalg.norm(x - p) > 2.9 for p in pts):
            pts.append(x)
    for q in range(n_rotors):
        pts.append(np.array([sep * (q - 0.5 * (n_rotors - 1)), 0.0, 0.0]))
    coords = np.zeros((3, natoms * P))
    for a, p in enumerate(pts):
        coords[:, a * P:(a + 1) * P] = p[:, None]
    angles = np.zeros((3, natoms * P))
    angles[0] = np.repeat(rng.uniform(0, 2 * np.pi, natoms), P)
    angles[1] = np.repeat(rng.uniform(-1, 1, natoms), P)
    angles[2] = np.repeat(rng.uniform(0, 2 * np.pi, natoms), P)
    return coords, angles


def nl_workload(name, seed=985456376, n_replicas=None, asymrho=None):
    """'c3': an asymmetric top (A, B, C = 27.88, 14.51, 9.28 cm^-1) in 32 para-H2 (12-6-8-9 pair
    potential), P = P_rot = 256, T = 0.5 K, rotor-H2 potential from a seeded Legendre expansion on
    a 100 x 181 x 91 grid, one replica.  'c4': two rigid TIP4P water rotors (caleng), T = 1 K,
    P = P_rot = 2048, 128 replicas.  Density tables from asymrho_tables at (T, P_rot), or from
    asymrho(T, P_rot, A, B, C, maxj) when given (the device generator, pimc_hip.Context.asymrho)."""
    from pimc_tables import asymrho_tables, h2h2_table, legendre_3d_table
    rng = np.random.default_rng(seed)
    if name == "c3":
        T, P, maxj, nrep = 0.5, 256, 20, n_replicas or 1
        tabs = dict(pot1d=[h2h2_table()], pot2d=[], rot=[], pot3d_list=[legendre_3d_table()])
        base = dict(ntypes=2, type_names=["pH2", "H2O"], numb=[32, 1], molecule=[0, 2], levels=[3, 3],
                    mcstep=[0.3, 0.1], mass=[MASS_H2, MASS_H2O], P=P, Prot=P, T=T, rtstep=0.4,
                    pot1d=[0, -1], pot2d=[-1, -1], rotdens=-1, pot3d=0, rot3d=0)
        nsolv, nrot = 32, 1
    elif name == "c4":
        T, P, maxj, nrep = 1.0, 2048, 44, n_replicas or 128
        tabs = dict(pot1d=[], pot2d=[], rot=[], pot3d_list=[])
        base = dict(ntypes=1, type_names=["H2O"], numb=[2], molecule=[2], levels=[3], mcstep=[0.05],
                    mass=[MASS_H2O], P=P, Prot=P, T=T, rtstep=0.3, pot1d=[-1], pot2d=[-1], rotdens=-1,
                    pot3d=-1, rot3d=0)
        nsolv, nrot = 0, 2
    else:
        raise ValueError(name)
    if asymrho is not None:
        rho, eng, esq = asymrho(T, P, *ABC_H2O, maxj)
    else:
        rho, eng, esq, _ = asymrho_tables(T, P, *ABC_H2O, maxj, -1)
    tabs["rot3d_list"] = [(rho.reshape(-1), eng.reshape(-1), esq.reshape(-1))]
    reps = []
    for r in range(nrep):
        rp = dict(base)
        rp["coords"], rp["angles"] = nl_initial_path(nsolv, nrot, P, rng, sep=3.0 if nsolv else 2.9)
        reps.append(rp)
    return dict(name=name, tables=tabs, replicas=reps)


def bead_moves_per_pass(rp):
    """Loop-structure count: sum n P (2^L - 1) + sum n P + P Prot n_rot (mc_main.cc:348-381)."""
    s = 0
    for t in range(rp["ntypes"]):
        s += rp["numb"][t] * rp["P"] * ((1 << rp["levels"][t]) - 1) + rp["numb"][t] * rp["P"]
    nrot = sum(rp["numb"][t] for t in range(rp["ntypes"]) if rp["molecule"][t] != 0)
    if rp["Prot"]:
        s += rp["P"] * rp["Prot"] * nrot
    return s


def load_into_context(ctx, wl):
    """Upload tables, replicas and initial paths of a workload into a pimc_hip.Context.
    Optional tables: "pot3d" (list of (nr, nth, nch, rmin, rmax, v)) and "rot3d" (list of
    (rho, erot, esq)); replicas then name them with "pot3d" / "rot3d" indices."""
    t = wl["tables"]
    for r, v in t["pot1d"]:
        ctx.add_pot1d(r, v)
    for rg, cg, dr, dc, V in t["pot2d"]:
        ctx.add_pot2d(rg, cg, dr, dc, V)
    for c, rho, er, es in t["rot"]:
        ctx.add_rotdens(c, rho, er, es)
    for tab in t.get("pot3d_list", []):
        ctx.add_pot3d(*tab)
    for tab in t.get("rot3d_list", []):
        ctx.add_rot3d(*tab)
    for i, rp in enumerate(wl["replicas"]):
        ctx.set_replica(i, rp["ntypes"], rp["numb"], rp["molecule"], rp["levels"], rp["mcstep"], rp["mass"], rp["P"],
                        rp["Prot"], rp["T"], rp["rtstep"], pot1d=rp["pot1d"], pot2d=rp["pot2d"], rotdens=rp["rotdens"],
                        pot3d=rp.get("pot3d", -1), rot3d=rp.get("rot3d", -1))
        if rp.get("box") is not None:  # MINIMAGE with this box
            ctx.set_minimum_image(i, rp["box"])
    ctx.commit()
    for i, rp in enumerate(wl["replicas"]):
        ctx.set_state(i, rp["coords"], rp["angles"])
    return ctx


def rot_file_name(rp):
    return rot_filename(rp["type_names"][1], rp["T"], rp["Prot"])
