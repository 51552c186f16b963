"""Deterministic stand-in 3D tables for golden vectors and parity tests.

They are smooth functions of the grid indices in the reference's layouts, not physical
densities or potentials; they exist so that the table look-ups, the Euler-angle algebra
and the moves of the non-linear-rotor path can be compared value for value.

  rot3d   rho, E, E^2 on the 181 x 361 x 361 grid of rotpro_sub.f:7-26: entry
          (ith*361 + iph)*361 + ich holds the value at theta = ith, phi = iph, chi = ich
          degrees.  rho dips below zero at large theta, which exercises the |rho| rule of
          MCRot3Dstep (mc_piqmc.cc:1067,1133).
  pot3d   V on the Rgrd x THgrd x CHgrd grid of vcalc.f:16-60 (R in bohr from Rvmin to
          Rvmax, theta in whole degrees, chi in whole degrees over the folded range that
          CHgrd implies, vcord.f:50-85): entry (ir*THgrd + ith)*CHgrd + ich.
"""
import numpy as np

NTH, NPH, NCH = 181, 361, 361


def rot3d(kappa=4.0):
    th = np.radians(np.arange(NTH, dtype=np.float64))[:, None, None]
    ph = np.radians(np.arange(NPH, dtype=np.float64))[None, :, None]
    ch = np.radians(np.arange(NCH, dtype=np.float64))[None, None, :]
    rho = np.exp(kappa * (np.cos(th) - 1.0)) * (1.0 + 0.3 * np.cos(ph + ch)) - 2.0e-3
    erot = 20.0 + 6.0 * np.cos(th) + 2.0 * np.sin(ph) * np.cos(2.0 * ch)
    esq = erot * erot + 15.0 + np.cos(ph - ch)
    return (np.ascontiguousarray(rho.reshape(-1)), np.ascontiguousarray(erot.reshape(-1)),
            np.ascontiguousarray(esq.reshape(-1)))


def pot3d(nch=91, nr=40, rmin=4.0, rmax=23.5):
    """Returns (nr, nth, nch, rmin, rmax, v) with R in bohr."""
    R = np.linspace(rmin, rmax, nr)[:, None, None]
    th = np.radians(np.arange(NTH, dtype=np.float64))[None, :, None]
    ch = np.radians(np.arange(nch, dtype=np.float64))[None, None, :]
    morse = 40.0 * ((1.0 - np.exp(-0.6 * (R - 6.6))) ** 2 - 1.0)
    v = morse * (1.0 + 0.35 * np.cos(th) + 0.1 * np.cos(2.0 * th)) + 4.0 * np.exp(-0.3 * R) * np.cos(2.0 * ch)
    return nr, NTH, nch, rmin, rmax, np.ascontiguousarray(v.reshape(-1))


def write_pot3d(path, nr, nth, nch, rmin, rmax, v):
    """The text file init_pot3D reads (mc_poten.cc:283-297): header, then one value a line."""
    with open(path, "w") as f:
        f.write(f"{nr} {nth} {nch} {rmin!r} {rmax!r}\n")
        np.savetxt(f, v, fmt="%.17g")
