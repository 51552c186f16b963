"""Potential and density-matrix tables in MoRiBS-PIMC's file formats.

Generators (closed forms the workloads are defined on) and readers/writers that keep the
reference layouts, so the same files drive the reference program and this build:
  1D pair potential  '.pot'  two columns r [Angstrom], V [K]; lines whose first token is '#'
                             are skipped (get_filesize/read_datafile, mc_poten.cc:731-788)
  2D rotor potential '.pot'  'nr nc / dr dc / r grid / cos grid / V[r][cos]' (init_pot2D,
                             mc_poten.cc:305-377)
  linear-rotor density '<TYPE>_T<T>t<Prot>.rot'  cos(gamma), rho, E-numerator, E^2-numerator
                             (init_rotdens mc_poten.cc:508-546, written as linden.f:60-70)
"""
import math

import numpy as np

TAU_UNIT = 1.4387752224  # cm K: tau [cm] = TAU_UNIT / (T * slices)   (linden.f:5,24)
BOLTZ_WNO = 0.69503476   # cm^-1 -> K used by linden's energy columns (linden.f:91)


# ---------------------------------------------------------------- generators
def lj_table(eps=10.22, sigma=2.556, rmin=1.0, rmax=30.0, n=500):
    """He-He Lennard-Jones on a uniform grid (the pair table of the linear-rotor workloads)."""
    r = np.linspace(rmin, rmax, n)
    sr6 = (sigma / r) ** 6
    return r, 4.0 * eps * (sr6 * sr6 - sr6)


def h2h2_table(eps=34.2, sigma=2.96, rmin=1.5, rmax=30.0, n=1000):
    """12-6-8-9 form for para-H2 pairs: 4 eps [(s/r)^12 - (s/r)^6] - eps [(s/r)^8 - (s/r)^9]."""
    r = np.linspace(rmin, rmax, n)
    x = sigma / r
    return r, 4.0 * eps * (x ** 12 - x ** 6) - eps * (x ** 8 - x ** 9)


def legendre(lmax, x):
    """P_0..P_lmax at x by the upward recurrence."""
    x = np.asarray(x, dtype=np.float64)
    p = [np.ones_like(x), x.copy()]
    for l in range(1, lmax):
        p.append(((2 * l + 1) * x * p[l] - l * p[l - 1]) / (l + 1))
    return p[: lmax + 1]


def morse_legendre_2d(D=(30.0, 8.0, 2.0), a=1.2, R0=3.5, rgrid=None, cgrid=None):
    """V(R, cos t) = sum_{l=0,2,4} D_l [(1 - e^{-a(R-R0)})^2 - 1] P_l(cos t), V[r][cos]."""
    if rgrid is None:
        rgrid = np.round(np.linspace(1.5, 21.5, 1001), 10)
    if cgrid is None:
        cgrid = np.round(np.linspace(-1.0, 1.0, 201), 10)
    R = np.asarray(rgrid)[:, None]
    c = np.asarray(cgrid)[None, :]
    morse = (1.0 - np.exp(-a * (R - R0))) ** 2 - 1.0
    P = legendre(4, c)
    V = D[0] * morse * P[0] + D[1] * morse * P[2] + D[2] * morse * P[4]
    return np.asarray(rgrid), np.asarray(cgrid), V


def linden_lmax(tau, bconst, maxl=500, eps=1.0e-16):
    """First l with exp(-tau B l(l+1)) < eps (linden.f:27-36)."""
    for l in range(maxl + 1):
        if math.exp(-tau * bconst * l * (l + 1)) < eps:
            return l
    return maxl


def linden(T, nslice, bconst_cm, npt=2001, iodevn=-1, lmax=None):
    """Free linear-rotor density matrix on a uniform cos(gamma) grid (exarho, linden.f:88-122).

    Returns (cos, rho, erot, erotsq) exactly as the four .rot columns, before rounding to the
    file's E15.8 format."""
    tau = TAU_UNIT / (T * nslice)
    if lmax is None:
        lmax = linden_lmax(tau, bconst_cm)
    cstep = 2.0 / (npt - 1)
    cost = np.array([(ic - 1) * cstep - 1.0 for ic in range(1, npt + 1)])
    P = legendre(max(lmax, 1), cost)
    rho = np.zeros(npt)
    er = np.zeros(npt)
    es = np.zeros(npt)
    for l in range(lmax + 1):
        if l % 2 == iodevn or iodevn == -1:
            tmp = (2 * l + 1) * P[l] * math.exp(-tau * bconst_cm * l * (l + 1))
            rho += tmp
            er += tmp * l * (l + 1) * bconst_cm
            es += tmp * l * (l + 1) * bconst_cm * l * (l + 1) * bconst_cm
    er = er / (nslice * BOLTZ_WNO)
    es = es / (nslice * BOLTZ_WNO) ** 2
    return cost, rho / (4.0 * math.pi), er / (4.0 * math.pi), es / (4.0 * math.pi)


def linden_exact_energy(T, bconst_cm, iodevn=-1, lmax=400):
    """<E_rot> and Cv of the free linear rotor at temperature T (linden.f:72-82 closed form)."""
    beta_cm = TAU_UNIT / T
    e = 0.0
    z = 0.0
    e2 = 0.0
    for l in range(lmax + 1):
        if l % 2 == iodevn or iodevn == -1:
            w = (2 * l + 1) * math.exp(-beta_cm * bconst_cm * l * (l + 1))
            el = bconst_cm * l * (l + 1) / BOLTZ_WNO
            z += w
            e += w * el
            e2 += w * el * el
    e /= z
    e2 /= z
    return e, (e2 - e * e) / T ** 2


# ---------------------------------------------------------------- file formats
def fmt_temperature(T):
    """C++ default ostream formatting of a double (precision 6, %g), as used to build the
    density-matrix file name (mc_poten.cc:517-524)."""
    return "%g" % T


def rot_filename(type_name, T, nrot):
    return f"{type_name}_T{fmt_temperature(T)}t{nrot}.rot"


def write_pot1d(path, r, v, comment="pair potential"):
    with open(path, "w") as f:
        f.write(f"# {comment}\n# (\\AA)   V(r) (K)\n")
        for a, b in zip(r, v):
            f.write(f"{a:.17g} {b:.17g}\n")


def read_pot1d(path):
    """read_datafile semantics: a line counts when its first token is not exactly '#'."""
    xs, ys = [], []
    with open(path) as f:
        for line in f:
            tok = line.split()
            if not tok or tok[0] == "#":
                continue
            xs.append(float(tok[0]))
            ys.append(float(tok[1]))
    return np.array(xs), np.array(ys)


def write_pot2d(path, rgrid, cgrid, V, dr=None, dc=None):
    nr, nc = len(rgrid), len(cgrid)
    dr = dr if dr is not None else (rgrid[-1] - rgrid[0]) / (nr - 1)
    dc = dc if dc is not None else (cgrid[-1] - cgrid[0]) / (nc - 1)
    with open(path, "w") as f:
        f.write(f"{nr} {nc}\n{dr:.17g} {dc:.17g}\n")
        for x in rgrid:
            f.write(f"{x:.17g}\n")
        for x in cgrid:
            f.write(f"{x:.17g}\n")
        np.savetxt(f, np.asarray(V).reshape(-1), fmt="%.17g")


def read_pot2d(path):
    with open(path) as f:
        tok = f.read().split()
    nr, nc = int(tok[0]), int(tok[1])
    dr, dc = float(tok[2]), float(tok[3])
    vals = np.array(tok[4:4 + nr + nc + nr * nc], dtype=np.float64)
    return vals[:nr], vals[nr:nr + nc], dr, dc, vals[nr + nc:].reshape(nr, nc)


def write_rot(path, cost, rho, erot, esq):
    """Fortran '(1p,7(1x,E15.8))' rows, as linden writes them."""
    with open(path, "w") as f:
        for row in zip(cost, rho, erot, esq):
            f.write("".join(" " + _e15_8(x) for x in row) + "\n")


def _e15_8(x):
    s = f"{x:.8E}"
    mant, exp = s.split("E")
    e = int(exp)
    return f"{mant}E{'+' if e >= 0 else '-'}{abs(e):02d}".rjust(15)


def read_rot(path):
    cols = [[], [], [], []]
    with open(path) as f:
        for line in f:
            tok = line.split()
            if not tok or tok[0] == "#":
                continue
            for k in range(4):
                cols[k].append(float(tok[k]))
    return tuple(np.array(c) for c in cols)
