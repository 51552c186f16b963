"""Small qmc.input decks with their table files, shared by the fixture generator
(make_reference_runs.py, which runs the reference program on them) and the statistical parity
tests of the host driver (which run moribs-pimc_amd/pimc on the same files).

  co2_100K     the reference's own example examples/CO2_100K_4_4 (one CO2 rotor, 100 K,
               P = Prot = 4) with fewer blocks; its .rot file is the fixture CO2_T100t4.rot
  he4_co2      4 He + CO2 at 0.37 K, P = Prot = 64, lattice start; LJ He-He, Morse-Legendre
               CO2-He and the linden CO2 density, all written in the reference's formats
  c0_8he       the 8-He workload: 8 He + CO2 at 0.37 K, P = Prot = 128, the same tables
  he16_pbc     16 He at liquid density (0.0218 Angstrom^-3, a periodic box of 9 Angstrom) under
               MINIMAGE at 2 K, P = 16, LJ He-He
  h2o_ph2      2 para-H2 + one rotating H2O at 1 K, P = Prot = 16, blocks of 500 passes (both
               programs rewrite a 1.5-million-line 3D density file after every block): the H2-H2
               table of pimc_tables.h2h2_table, the stand-in 3D potential and density tables of
               tables3d (H2Opot.pot, H2O_T1t16.rho/.eng/.esq, 1.4 GB of text written once per
               directory or shared between directories).  With 4 para-H2 at 0.5 K each chain of
               the reference program stayed in its own configuration for 60,000 passes (K per
               chain from 33 to 52 K); with 2 at 1 K chains started apart agree within their
               blocking errors.
"""
import os
import shutil

import numpy as np

import pimc_configs as pc
import pimc_tables as pt

HERE = os.path.dirname(os.path.abspath(__file__))

DECKS = {
    "co2_100K": """MASTERDIR        ./
OUTPUTDIR        ./
FILENAMEPREFIX   CO2_monomer
DIMENSION     3
DENSITY       0.0450   #  Angstrom^-3  Used to define the BoxLength
TEMPERATURE   100       #  Kelvin
MOLECULE     CO2    1   BOLTZMANN     0.18      1         CO2_fake  PRIMITIVE    #
ROTATION   CO2   0.15  4
ROTDENSI 0  -1 0.0 0.6666525 0.1769383 0.2306476 1
NUMBEROFSLICES      4       #  number of time slices
NUMBEROFPASSES     3000         #  number of passes per block
NUMBEROFBLOCKS     {blocks}  5   #  number of blocks and number of equil blocks
MCSKIP_RATIO     3000
MCSKIP_TOTAL     3000
MCSKIP_AVERG     4
""",
    "he4_co2": """MASTERDIR        ./
OUTPUTDIR        ./
FILENAMEPREFIX   he4co2
DIMENSION     3
DENSITY       0.02
TEMPERATURE   0.37
ATOM       He4  4    BOLTZMANN     0.30     3         He      PRIMITIVE
MOLECULE   CO2  1    BOLTZMANN     0.08     3         CO2He   PRIMITIVE
ROTATION   CO2   0.3  64
NUMBEROFSLICES     64
NUMBEROFPASSES     200
NUMBEROFBLOCKS     {blocks}  10
MCSKIP_RATIO     12800
MCSKIP_TOTAL     12800
MCSKIP_AVERG     64
""",
    "c0_8he": """MASTERDIR        ./
OUTPUTDIR        ./
FILENAMEPREFIX   c0he8
DIMENSION     3
DENSITY       0.02
TEMPERATURE   0.37
ATOM       He4  8    BOLTZMANN     0.15     3         He      PRIMITIVE
MOLECULE   CO2  1    BOLTZMANN     0.08     3         CO2He   PRIMITIVE
ROTATION   CO2   0.3  128
NUMBEROFSLICES     128
NUMBEROFPASSES     100
NUMBEROFBLOCKS     {blocks}  10
MCSKIP_RATIO     12800
MCSKIP_TOTAL     12800
MCSKIP_AVERG     128
""",
    "he16_pbc": """MASTERDIR        ./
OUTPUTDIR        ./
FILENAMEPREFIX   he16pbc
DIMENSION     3
DENSITY       0.0218
TEMPERATURE   2.0
ATOM       He4  16   BOLTZMANN     0.40     3         He      PRIMITIVE
MINIMAGE
NUMBEROFSLICES     16
NUMBEROFPASSES     100
NUMBEROFBLOCKS     {blocks}  10
MCSKIP_RATIO     1600
MCSKIP_TOTAL     1600
MCSKIP_AVERG     16
""",
    "h2o_ph2": """MASTERDIR        ./
OUTPUTDIR        ./
FILENAMEPREFIX   h2oph2
DIMENSION     3
DENSITY       0.02
TEMPERATURE   1.0
ATOM       H2   2    BOLTZMANN     0.30     2         pH2      PRIMITIVE
NONLINEAR  H2O  1    BOLTZMANN     0.05     2         H2Opot   PRIMITIVE
ROTATION   H2O   0.3  16
NUMBEROFSLICES     16
NUMBEROFPASSES     500
NUMBEROFBLOCKS     {blocks}  10
MCSKIP_RATIO     8000
MCSKIP_TOTAL     8000
MCSKIP_AVERG     16
""",
}
BLOCKS = {"co2_100K": 205, "he4_co2": 1010, "c0_8he": 310, "he16_pbc": 1010, "h2o_ph2": 130}
ROT3D_STEM = "H2O_T1t16"  # init_rot3D's file stem for H2O at 1 K with 16 rotor slices


def write_nl_tables(d, rotation):
    """The H2-H2 table, the stand-in 3D potential and, with rotation moves, the stand-in density,
    energy and energy-square tables of tables3d, as the reference reads them."""
    try:
        from golden import tables3d
    except ImportError:  # run from tests/golden itself
        import tables3d
    os.makedirs(d, exist_ok=True)
    r, v = pt.h2h2_table()
    pt.write_pot1d(os.path.join(d, "pH2.pot"), r, v)
    tables3d.write_pot3d(os.path.join(d, "H2Opot.pot"), *tables3d.pot3d(nch=181))
    if rotation:
        pt.write_rot3d(os.path.join(d, ROT3D_STEM), *tables3d.rot3d())


def materialise(name, d, blocks=None, shared=None):
    """Writes the deck and its table files into directory d.  For h2o_ph2 the tables are
    written once into `shared` when given (and linked from d), else into d."""
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, "qmc.input"), "w") as f:
        f.write(DECKS[name].format(blocks=blocks or BLOCKS[name]))
    if name == "h2o_ph2":
        if shared is None:
            write_nl_tables(d, rotation=True)
            return
        if not os.path.exists(os.path.join(shared, ROT3D_STEM + ".esq")):
            write_nl_tables(shared, rotation=True)
        for f in ("pH2.pot", "H2Opot.pot", ROT3D_STEM + ".rho", ROT3D_STEM + ".eng", ROT3D_STEM + ".esq"):
            os.symlink(os.path.join(shared, f), os.path.join(d, f))
    elif name == "co2_100K":
        shutil.copy(os.path.join(HERE, "CO2_T100t4.rot"), d)
        with open(os.path.join(d, "CO2_fake.pot"), "w") as f:  # the example's empty 2D table
            f.write("0 0\n0 0\n")
    elif name == "he16_pbc":
        r, v = pt.lj_table()
        pt.write_pot1d(os.path.join(d, "He.pot"), r, v)
    else:
        P = 128 if name == "c0_8he" else 64
        r, v = pt.lj_table()
        pt.write_pot1d(os.path.join(d, "He.pot"), r, v)
        rg, cg, V = pt.morse_legendre_2d()
        pt.write_pot2d(os.path.join(d, "CO2He.pot"), rg, cg, V)
        c, rho, er, es = pc.rot_table(0.37, P)
        pt.write_rot(os.path.join(d, pt.rot_filename("CO2", 0.37, P)), c, rho, er, es)


# A non-linear rotor deck for the driver's own tests (the reference program is not run on it):
# 4 para-H2 around one H2O without rotation moves, the H2-H2 table of pimc_tables.h2h2_table
# and the stand-in 3D table of tables3d.pot3d (40 x 181 x 181).
NL_DECK = """MASTERDIR        ./
OUTPUTDIR        ./
FILENAMEPREFIX   h2oh2
DIMENSION     3
DENSITY       0.02
TEMPERATURE   0.5
ATOM       H2   4    BOLTZMANN     0.30     2         pH2      PRIMITIVE
NONLINEAR  H2O  1    BOLTZMANN     0.05     2         H2Opot   PRIMITIVE
NUMBEROFSLICES     16
NUMBEROFPASSES     100
NUMBEROFBLOCKS     {blocks}  1
MCSKIP_RATIO     1600
MCSKIP_TOTAL     1600
MCSKIP_AVERG     16
"""


def materialise_nl(d, blocks=3):
    """Writes NL_DECK and its tables into directory d."""
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, "qmc.input"), "w") as f:
        f.write(NL_DECK.format(blocks=blocks))
    write_nl_tables(d, rotation=False)


ENG_COLUMNS = ("kin", "pot", "kin+pot", "rot", "rotsq", "total", "Cv", "Cv_trans", "Cv_rot")


def read_eng(path):
    """Block lines of a .eng file: [n_blocks, 10] (block number, then ENG_COLUMNS)."""
    return np.loadtxt(path, ndmin=2)
