"""Potential and density-matrix tables in MoRiBS-PIMC's file formats.

Generators (closed forms the workloads are defined on) and readers/writers that keep the
reference layouts, so the same files drive the reference program and the host driver here:
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


# ---------------------------------------------------------------- asymmetric-top density matrix
BOLTZ_ASYM = 0.6950356  # cm^-1 per K in asymrho (asymrho.f:42)
ROT3D_SHAPE = (181, 361, 361)  # theta, phi, chi in whole degrees (rotpro_sub.f:7-26)


def _asym_block(J, ks, A, B, C):
    """Rigid-rotor Hamiltonian of one k-parity block in the |J k> basis (rotmat, asymrho.f:792-813):
    diagonal (A+C)/2 J(J+1) + (B - (A+C)/2) k^2, k -> k-2 coupling (A-C)/4 c+(J,k-2) c+(J,k-1)."""
    def cplus(j, k):
        return 0.0 if (k >= j or k < -j) else math.sqrt(j * (j + 1.0) - k * (k + 1.0))
    n = len(ks)
    H = np.zeros((n, n))
    for i, k in enumerate(ks):
        H[i, i] = 0.5 * (A + C) * J * (J + 1) + (B - 0.5 * (A + C)) * k * k
        if i > 0:
            H[i, i - 1] = H[i - 1, i] = 0.25 * (A - C) * cplus(J, k - 2) * cplus(J, k - 1)
    return H


def _wigner_d(J, theta):
    """d^J_{mk}(theta) = <J m| exp(-i theta J_y) |J k> (the convention of Zare eq. 3.57 used by
    wigd, asymrho.f:1006-1039), from the eigenvectors of J_y: stable at large J where the
    alternating factorial sum loses its digits.  Rows m, columns k, both -J..J."""
    ms = np.arange(-J, J + 1)
    Jy = np.zeros((2 * J + 1, 2 * J + 1), complex)
    for i, m in enumerate(ms[:-1]):
        cp = math.sqrt(J * (J + 1.0) - m * (m + 1.0))
        Jy[i + 1, i] += cp / 2j
        Jy[i, i + 1] -= cp / 2j
    lam, V = np.linalg.eigh(Jy)
    return ((V * np.exp(-1j * theta * lam)) @ V.conj().T).real


def asymrho_tables(T, nslice, A, B, C, maxj, iodevn=-1, thetas=None):
    """Free asymmetric-top propagator, energy and energy-square estimators on the degree grid of
    relative Euler angles (asymrho, asymrho.f:34-731; J. Chem. Phys. 134, 054117 (2011) eqs 15-16).

    rho(theta, phi, chi) = sum_J (2J+1)/(8 pi^2) sum_s exp(-tau E_Js) sum_{m,k} c_m c_k d^J_mk(theta)
    cos(m phi + k chi), the sums running over the eigenstates s of each k-parity block (even k
    for iodevn 0, odd k for 1, both for -1); E and E^2 are the energy-weighted sums divided by rho
    (cm^-1 and cm^-2; zero where |rho| < 1e-16).  Each theta is a product of complex matrices
    Re[Phi M X^T] with Phi[phi, m] = e^{i m phi}, X[chi, k] = e^{i k chi}.  Values for chi beyond
    the symmetry-unique range are copied as asymrho.f:665-709 does.

    Returns rho, eng, esq of shape (len(thetas), 361, 361) and the list of thetas."""
    if iodevn not in (-1, 0, 1):
        raise ValueError("iodevn can only be -1 0 1")
    thetas = list(range(181)) if thetas is None else list(thetas)
    beta = 1.0 / (BOLTZ_ASYM * T)
    tau = beta / nslice
    eps = 1.0e-16
    blocks = {0: [], 1: []}  # parity of k -> list of (J, ks, energies, vectors)
    for J in range(maxj + 1):
        kev = list(range(-J if J % 2 == 0 else -J + 1, J + 1, 2))
        kod = list(range(-J + 1 if J % 2 == 0 else -J, J + 1, 2))
        for par, ks in ((0, kev), (1, kod)):
            if not ks:
                continue
            e, v = np.linalg.eigh(_asym_block(J, ks, A, B, C))
            blocks[par].append((J, ks, e, v))
    emax = max(e.max() for par in blocks for (_, _, e, _) in blocks[par])
    if math.exp(-tau * emax) > 1.0e-8:
        raise ValueError("too large contribution from emax: raise maxj")
    ang = np.radians(np.arange(361.0))
    n = 2 * maxj + 1
    idx = np.arange(-maxj, maxj + 1)
    E = np.exp(1j * np.outer(ang, idx))  # [angle, m] -> e^{i m angle}
    out = [np.zeros((len(thetas), 361, 361)) for _ in range(3)]
    pars = {-1: (0, 1), 0: (0,), 1: (1,)}[iodevn]
    for ti, ith in enumerate(thetas):
        th = math.radians(ith)
        acc = {}
        for par in pars:
            M = [np.zeros((n, n)) for _ in range(3)]
            for J, ks, e, v in blocks[par]:
                pre = (2 * J + 1) / (8.0 * math.pi * math.pi)
                d = _wigner_d(J, th)
                kk = np.asarray(ks) + J
                dsub = d[np.ix_(kk, kk)]
                w = np.exp(-tau * e)
                keep = w * pre >= eps
                if not keep.any():
                    continue
                cw = v[:, keep]
                for q, p in enumerate((0, 1, 2)):
                    wt = w[keep] * e[keep] ** p
                    G = (cw * wt) @ cw.T  # sum_s w_s E_s^p c_m c_k
                    sl = np.asarray(ks) + maxj
                    M[q][np.ix_(sl, sl)] += pre * G * dsub
            acc[par] = [np.real(E @ Mq @ E.T) for Mq in M]  # [phi, chi]
        if iodevn == -1:
            r = acc[0][0] + acc[1][0]
            with np.errstate(divide="ignore", invalid="ignore"):
                # E of each parity is divided by its own rho first (zero where |rho| < 1e-16)
                ee = np.where(np.abs(acc[0][0]) > eps, acc[0][1] / acc[0][0], 0.0)
                eo = np.where(np.abs(acc[1][0]) > eps, acc[1][1] / acc[1][0], 0.0)
                se = np.where(np.abs(acc[0][0]) > eps, acc[0][2] / acc[0][0], 0.0)
                so = np.where(np.abs(acc[1][0]) > eps, acc[1][2] / acc[1][0], 0.0)
                eng = (eo * acc[1][0] + ee * acc[0][0]) / r
                esq = (so * acc[1][0] + se * acc[0][0]) / r
            small = np.abs(r) < eps
            eng[small] = 0.0
            esq[small] = 0.0
        else:
            a = acc[pars[0]]
            r = a[0]
            with np.errstate(divide="ignore", invalid="ignore"):
                eng = np.where(np.abs(r) > eps, a[1] / r, 0.0)
                esq = np.where(np.abs(r) > eps, a[2] / r, 0.0)
        for q, x in enumerate((r, eng, esq)):
            out[q][ti] = _asym_symmetric_fill(x)
    return out[0], out[1], out[2], thetas


def _asym_symmetric_fill(x):
    """Copy the symmetry-unique (phi, chi) triangle onto the rest of the grid (asymrho.f:665-709)."""
    x = x.copy()
    for iphi in range(90, 181):
        for ichi in range(180 - iphi, iphi + 1):
            x[iphi, ichi] = x[180 - ichi, 180 - iphi]
    for iphi in range(180, 271):
        for ichi in range(iphi - 180, 360 - iphi + 1):
            x[iphi, ichi] = x[180 + ichi, iphi - 180]
    for iphi in range(180, 361):
        for ichi in range(360 - iphi, iphi + 1):
            x[iphi, ichi] = x[360 - ichi, 360 - iphi]
    for iphi in range(0, 361):
        for ichi in range(iphi, 361):
            x[iphi, ichi] = x[ichi, iphi]
    return x


def legendre_3d_table(seed=20110201, nr=100, rmin=4.0, rmax=24.0, nch=91, depth=80.0):
    """Rotor-particle potential V(R, theta, chi) on the vcord/vcalc grid: R in bohr (nr points),
    theta 0..180 deg, chi 0..nch-1 deg (91: the C2v-folded quadrant), V[(ir*181 + ith)*nch + ich]
    in Kelvin.  A seeded Legendre-type expansion, l <= 4 and even m <= min(l, 4):
    V = sum_lm a_lm f_lm(R) P_l^m(cos theta) cos(m chi) with an isotropic Morse well of `depth`
    (minimum near 6.3 bohr) and anisotropic terms of random sign and size on bounded Gaussian
    radial shapes around the well, so the repulsive wall stays repulsive in every direction.
    Returns (nr, 181, nch, rmin, rmax, v)."""
    rng = np.random.default_rng(seed)
    R = np.linspace(rmin, rmax, nr)[:, None, None]
    th = np.radians(np.arange(181.0))[None, :, None]
    ch = np.radians(np.arange(float(nch)))[None, None, :]
    x = np.cos(th)
    # associated Legendre P_l^m(x) for l <= 4 by explicit forms (no Condon-Shortley phase)
    s = np.sqrt(np.maximum(1.0 - x * x, 0.0))
    P = {(0, 0): np.ones_like(x), (1, 0): x, (2, 0): 0.5 * (3 * x * x - 1), (2, 2): 3 * s * s,
         (3, 0): 0.5 * (5 * x ** 3 - 3 * x), (3, 2): 15 * x * s * s,
         (4, 0): (35 * x ** 4 - 30 * x * x + 3) / 8, (4, 2): 7.5 * (7 * x * x - 1) * s * s, (4, 4): 105 * s ** 4}
    norm = {(l, m): 1.0 / max(1.0, float(np.abs(P[(l, m)]).max())) for (l, m) in P}

    def morse(re, a):
        return (1.0 - np.exp(-a * (R - re))) ** 2 - 1.0
    v = depth * morse(6.3, 0.9) * P[(0, 0)]
    for (l, m) in P:
        if (l, m) == (0, 0):
            continue
        amp = rng.uniform(0.05, 0.25) * depth * rng.choice([-1.0, 1.0])
        shape = np.exp(-((R - rng.uniform(5.8, 7.0)) / rng.uniform(1.0, 2.0)) ** 2)
        v = v + amp * shape * norm[(l, m)] * P[(l, m)] * np.cos(m * ch)
    v = np.broadcast_to(v, (nr, 181, nch))
    return nr, 181, nch, float(rmin), float(rmax), np.ascontiguousarray(v.reshape(-1))


def read_asym_column(path):
    """One E15.8 column of an asymrho output file (rho.denXXX_rho/_eng/_esq)."""
    with open(path) as f:
        return np.array([float(t) for t in f.read().split()])


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


def write_pot3d(path, nr, nth, nch, rmin, rmax, v):
    """The 3D rotor-partner potential as init_pot3D reads it (mc_poten.cc:283-297): the header
    'Rgrd THgrd CHgrd Rvmin Rvmax', then one value a line in the table's order."""
    with open(path, "w") as f:
        f.write(f"{nr} {nth} {nch} {rmin!r} {rmax!r}\n")
        np.asarray(v, dtype=np.float64).reshape(-1).tofile(f, sep="\n", format="%.17g")
        f.write("\n")


def write_rot3d(stem, rho, erot, esq):
    """An asymmetric top's density, energy and energy-square tables as init_rot3D reads them
    (mc_poten.cc:440-500): stem + '.rho', '.eng', '.esq', 181*361*361 values each in the
    (theta, phi, chi) order of rotpro_sub.f:7-26, whitespace-separated."""
    for ext, a in ((".rho", rho), (".eng", erot), (".esq", esq)):
        with open(stem + ext, "w") as f:
            np.asarray(a, dtype=np.float64).reshape(-1).tofile(f, sep="\n", format="%.17g")
            f.write("\n")


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
