// Host driver: a drop-in for the reference's `pimc` program on the sampling path.
//
// It reads the reference's qmc.input (keyword set of mc_input.cc:18-57 and IOReadParams
// mc_input.cc:71-399), the table files named as InitPotentials / InitRotDensity name them
// (mc_poten.cc:93-119,150-158,283-297,305-360,379-438,440-506,508-546), builds the initial
// configuration (xyz.init as initconf.f reads it, else the lattice of initLattice_config
// mc_setup.cc:491-650 and replInitial_config), runs blocks of passes through the C ABI of
// moribs_pimc.h, and writes the reference's outputs:
//   <OUTPUTDIR><FILENAMEPREFIX>.eng       one line per block after equilibration (SaveEnergy mc_main.cc:764-790)
//   <OUTPUTDIR><FILENAMEPREFIX>_sum.eng   accumulated averages every MCSKIP_TOTAL steps (SaveSumEnergy :792-836)
//   <OUTPUTDIR><FILENAMEPREFIX>.xyz       the configuration at the end of every block (IOxyz mc_input.cc:690-746)
//   <...><block>.gra/.gri/.grt/.g2d/.rcf  block distributions (MCSaveBlockAverages mc_main.cc:703-735,
//                                         SaveDensities1D/2D mc_estim.cc:1327-1444,1666-1727, SaveRCF :1141-1184)
//   <...><block>.grc/.gtc                 around a non-linear rotor, .gri/.grt/.grc from the 3D bins
//                                         (SaveRho1D mc_estim.cc:1445-1590, SaveRhoThetaChi :1930-1995)
//   <...>_sum.gra/_sum.g2d/_sum.rcf       accumulated distributions after every block (mc_main.cc:446-462)
//   <...>_sum.g3d/_sum.gri/.grt/.grc      around a non-linear rotor (SaveDensities3D mc_estim.cc:1728-1820,
//   <...>_sum.eulphi/.eulchi/.eulthe      SaveRho1D with the relative Euler angles :1591-1662)
//   stdout                                the accept-ratio line every MCSKIP_RATIO steps (MCSaveAcceptRatio :879-943)
//   yw001.stat, yw001.conf, yw001.rand    the checkpoint after every block, the previous one kept as
//                                         *.old (mc_main.cc:471-478, StatusIO/ConfigIO mc_input.cc:517-609)
// RESTART in qmc.input resumes from the checkpoint: the start block, every chain's coordinates,
// angles and direction cosines, and its generator state (the pass counter that addresses every
// draw), so the resumed run continues exactly as the uninterrupted one.  The reference's
// yw001.conf holds one coordinate row and no angles; here it holds all three rows of
// coordinates, cosines and angles per chain: an int64 byte count, then the three arrays.
// The atom-atom g(r), the atom density around a linear rotor (300 radial bins to 15 A, 50
// polar bins), the atom density around a non-linear rotor in its body frame (300 x 50 x 100
// chi bins) and the relative Euler angles of neighbouring rotor slices are binned on the device.
//
// Time steps and passes.  The reference counts "steps" (one time slice of one pass) and
// samples when the step count is a multiple of MCSKIP_AVERG (mc_main.cc:340-484).  A pass on
// the device moves every slice at once; sampling is taken every max(1, MCSKIP_AVERG /
// NUMBEROFSLICES) passes, and, when MCSKIP_AVERG is below NUMBEROFSLICES, k times inside each
// pass (library option samples_per_pass: after every P/k time steps in the reference's order,
// every S/k segment phases otherwise).  Sum lines and ratio lines are taken at pass granularity.
//
// Usage:  pimc [qmc.input] [--seed N] [--fp64] [--ref-order] [--replicas R] [--autotune] [--init-only]
// --autotune picks launch shapes by timing them on the deck before the first block (results
// are the same as without it)
// --init-only writes the initial configuration to <OUTPUTDIR><FILENAMEPREFIX>.xyz and stops
// before any device work.  With R > 1 the same deck runs as R independent chains in one context; chain r > 0 writes
// its files under the prefix <OUTPUTDIR><FILENAMEPREFIX>_r<r>.
#include "moribs_pimc.h"
#include "sci_format.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

namespace {

[[noreturn]] void die(const std::string& where, const std::string& what) {
    std::cerr << "pimc: " << where << ": " << what << std::endl;
    std::exit(1);
}

// ------------------------------------------------------------------------------- the deck
struct Species {
    std::string type, fpot;
    int numb = 0, molecule = 0, levels = 0;
    double mcstep = 0.0, mass = 0.0, rtstep = 0.0;
};

struct Deck {
    std::string master = "./", outdir = "./", prefix = "pimc";
    int ndim = 3;
    double T = 0.0, density = 0.0;
    std::vector<Species> types;
    int P = 0, Prot = 0;
    long passes = 0, blocks = 0, eqblocks = 0, startblock = 0;
    bool rotation = false, read_coords = false, restart = false, minimage = false;
    std::string rot_type;
    double rot_step = 0.0;
    long skip_ratio = 100000, skip_total = 10000, skip_averg = 1;  // mc_setup.cc:91-93
    int natoms() const { int n = 0; for (auto& s : types) n += s.numb; return n; }
};

// Masses by type name, MCInitParams mc_setup.cc:244-319 with the constants of mc_const.h:44-50.
double mass_of(const std::string& t) {
    const double H1 = 1.0078, H2 = 2.015650642, HE4 = 4.0026032497, C12 = 12.0, N14 = 14.003,
                 O16 = 15.994915, S32 = 31.972;
    if (t == "He4") return HE4;
    if (t == "H2") return H2;
    if (t == "OCS") return O16 + C12 + S32;
    if (t == "N2O") return 2.0 * N14 + O16;
    if (t == "CO2") return C12 + 2.0 * O16;
    if (t == "CO") return C12 + O16;
    if (t == "HCN") return H1 + C12 + N14;
    if (t == "HCCCN") return H1 + 3.0 * C12 + N14;
    if (t == "H2O") return 2.0 * H1 + O16;
    if (t == "SO2") return 2.0 * O16 + S32;
    if (t == "HCOOCH3") return 4.0 * H1 + 2.0 * O16 + 2.0 * C12;
    die("MCInitParams", "Unknown atom/molecule type " + t);
}

// Keywords as IOReadParams reads them: one keyword and its fields per line, the rest of the
// line is a comment.  Keywords that select methods outside the sampling path are refused.
Deck read_deck(const std::string& path) {
    std::ifstream in(path);
    if (!in.good()) die("IOReadParams", "cannot open " + path);
    Deck d;
    std::string key, rest;
    bool impurity = false;
    int rtypes = 0;
    auto need = [&](std::istream& s, const std::string& k) {
        if (!s) die("IOReadParams", "missing or bad field after " + k);
    };
    while (in >> key) {
        if (key == "RESTART") d.restart = true;
        else if (key == "MASTERDIR") { in >> d.master; need(in, key); }
        else if (key == "OUTPUTDIR") { in >> d.outdir; need(in, key); }
        else if (key == "FILENAMEPREFIX") { in >> d.prefix; need(in, key); }
        else if (key == "TEMPERATURE") { in >> d.T; need(in, key); }
        else if (key == "DENSITY") { in >> d.density; need(in, key); }
        else if (key == "DIMENSION") { in >> d.ndim; need(in, key); }
        else if (key == "ATOM" || key == "MOLECULE" || key == "NONLINEAR") {
            Species s;
            std::string stat, model;
            in >> s.type >> s.numb >> stat >> s.mcstep >> s.levels >> s.fpot >> model;
            need(in, key);
            if (s.numb < 0) die("IOReadParams", "spherical models (negative N) are not supported");
            if (stat == "BOSE") die("IOReadParams", "Bose statistics (exchange moves) are not supported; use BOLTZMANN");
            if (stat != "BOLTZMANN") die("IOReadParams", "Unknown statistics");
            if (model == "EFFECTIVE") die("IOReadParams", "EFFECTIVE pair actions are not supported; use PRIMITIVE");
            if (model != "PRIMITIVE") die("IOReadParams", "Unknown model of interaction");
            s.molecule = key == "MOLECULE" ? 1 : key == "NONLINEAR" ? 2 : 0;
            if (s.molecule == 0 && impurity) die("IOReadParams", "Molecules should follow atoms in input file");
            if (s.numb > 0) {
                if (s.molecule) impurity = true;
                d.types.push_back(s);
            }
        }
        else if (key == "NUMBEROFSLICES") { in >> d.P; need(in, key); }
        else if (key == "NUMBEROFPASSES") { in >> d.passes; need(in, key); }
        else if (key == "NUMBEROFBLOCKS") { in >> d.blocks >> d.eqblocks; need(in, key); }
        else if (key == "INITIALBLOCK") { in >> d.startblock; need(in, key); }
        else if (key == "ROTATION") {
            in >> d.rot_type >> d.rot_step >> d.Prot;
            need(in, key);
            d.rotation = true;
            ++rtypes;
        }
        else if (key == "ROTDENSI") {
            int kind = 0;
            in >> kind;
            need(in, key);
            if (kind != 0) die("IOReadParams", "ROTDENSI: only the tabulated density (0) is supported");
        }
        else if (key == "REFLECTX" || key == "REFLECTY" || key == "REFLECTZ") {
            int on = 0;
            in >> on;
            need(in, key);
            if (on) die("IOReadParams", key + " 1 is not supported");
        }
        else if (key == "ROTSYM") die("IOReadParams", "ROTSYM is not supported");
        else if (key == "WORM") die("IOReadParams", "the worm algorithm is not supported");
        else if (key == "MINIMAGE") d.minimage = true;  // mc_input.cc:298-300
        else if (key == "CRYSTAL") die("IOReadParams", "crystal lattices are not supported");
        else if (key == "READMCCOORDS") d.read_coords = true;
        else if (key == "MCSKIP_RATIO") { in >> d.skip_ratio; need(in, key); }
        else if (key == "MCSKIP_TOTAL") { in >> d.skip_total; need(in, key); }
        else if (key == "MCSKIP_AVERG") { in >> d.skip_averg; need(in, key); }
        std::getline(in, rest);  // the rest of the line is a comment
    }
    if (d.ndim != 3) die("IOReadParams", "DIMENSION must be 3");
    if (d.types.empty()) die("IOReadParams", "no particles");
    if ((int)d.types.size() > PIMC_MAX_TYPES) die("IOReadParams", "too many particle types");
    int atypes = 0, mtypes = 0;
    for (auto& s : d.types) (s.molecule ? mtypes : atypes)++;
    if (atypes > 1 || mtypes > 1)
        die("IOReadParams", "No more then one atom/molecule type: densities and potential energy");
    if (d.rotation) {
        if (rtypes != 1) die("IOReadParams", "Rotational degrees of freedom for one molecule's type only");
        bool found = false;
        for (auto& s : d.types)
            if (s.type == d.rot_type) {
                if (s.molecule == 0) die("IOReadParams", "Rotational degrees of freedom for molecules only");
                s.rtstep = d.rot_step;
                found = true;
                break;
            }
        if (!found) die("IOReadParams", "Can't find a particle type to sample rotational degrees of freedom");
    }
    if (d.minimage && d.rotation)  // the reference's linear-rotor move stops here (mc_piqmc.cc:2013-2016)
        for (auto& s : d.types)
            if (s.type == d.rot_type && s.molecule == 1) die("PotRotEnergy", "MIN IMAGE for orient pot");
    if (d.P <= 0 || d.passes <= 0 || d.blocks <= 0) die("IOReadParams", "NUMBEROFSLICES/PASSES/BLOCKS must be positive");
    if (d.skip_ratio <= 0 || d.skip_total <= 0 || d.skip_averg <= 0) die("IOReadParams", "MCSKIP_* must be positive");
    for (auto& s : d.types) s.mass = mass_of(s.type);
    return d;
}

// ------------------------------------------------------------------------------ table files
// Whitespace-separated numbers; '#' starts a comment running to the end of the line
// (read_datafile mc_poten.cc:757-788 skips comment lines).
std::vector<double> read_numbers(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f.good()) die("read_datafile", "cannot open " + path);
    std::string text((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    std::vector<double> out;
    const char* p = text.c_str();
    const char* end = p + text.size();
    while (p < end) {
        if (*p == '#') {
            while (p < end && *p != '\n') ++p;
            continue;
        }
        if (std::isspace((unsigned char)*p)) { ++p; continue; }
        char* q = nullptr;
        double v = std::strtod(p, &q);
        if (q == p) die("read_datafile", "bad number in " + path);
        out.push_back(v);
        p = q;
    }
    return out;
}

std::string fmt_default(double x) {  // ostream << double, precision 6 (file names, mc_poten.cc:443-448)
    std::ostringstream s;
    s << x;
    return s.str();
}

struct Context {
    pimc_ctx* ctx = nullptr;
    void check(int rc, const char* where) const {
        if (rc < 0) die(where, pimc_last_error(ctx));
    }
};

// Reads the tables of every species and fills the replica description.
pimc_replica_desc load_tables(const Deck& d, Context& c) {
    pimc_replica_desc r;
    std::memset(&r, 0, sizeof r);
    r.ntypes = (int)d.types.size();
    r.P = d.P;
    r.Prot = d.rotation ? d.Prot : 0;
    r.T = d.T;
    r.rotdens = -1;
    r.pot3d = -1;
    r.rot3d = -1;
    const std::string tag = "_T" + fmt_default(d.T) + "t" + std::to_string(d.Prot);
    bool has_atoms = false;
    for (auto& s : d.types) has_atoms |= s.molecule == 0;
    for (int t = 0; t < r.ntypes; ++t) {
        const Species& s = d.types[t];
        r.numb[t] = s.numb;
        r.molecule[t] = s.molecule;
        r.levels[t] = s.levels;
        r.mcstep[t] = s.mcstep;
        r.mass[t] = s.mass;
        r.pot1d[t] = r.pot2d[t] = -1;
        if (s.molecule == 0) {  // init_pot1D: two columns, r (Angstrom) and V (Kelvin)
            auto v = read_numbers(s.fpot + ".pot");
            if (v.size() < 4 || v.size() % 2) die("init_pot1D", s.fpot + ".pot: expected two columns");
            const int n = (int)v.size() / 2;
            std::vector<double> rr(n), vv(n);
            for (int i = 0; i < n; ++i) { rr[i] = v[2 * i]; vv[i] = v[2 * i + 1]; }
            r.pot1d[t] = pimc_add_pot1d(c.ctx, n, rr.data(), vv.data());
            c.check(r.pot1d[t], "init_pot1D");
        } else if (s.molecule == 1) {
            if (s.numb != 1) die("InitPotentials", "No more than one linear dopant molecule so far");
            if (has_atoms) {  // init_pot2D: rsize csize dr dc r-grid c-grid V[r][c]
                auto v = read_numbers(s.fpot + ".pot");
                if (v.size() < 4) die("init_pot2D", s.fpot + ".pot: short header");
                const int nr = (int)v[0], nc = (int)v[1];
                if (nr < 2 || nc < 2 || v.size() != 4 + (size_t)nr + nc + (size_t)nr * nc)
                    die("init_pot2D", s.fpot + ".pot: sizes do not match the data");
                const double* g = v.data() + 4;
                r.pot2d[t] = pimc_add_pot2d(c.ctx, nr, nc, v[2], v[3], g, g + nr, g + nr + nc);
                c.check(r.pot2d[t], "init_pot2D");
            }
            if (d.rotation) {  // init_rotdens: TYPE_T<T>t<Prot>.rot, four columns
                auto v = read_numbers(s.type + tag + ".rot");
                if (v.size() < 8 || v.size() % 4) die("init_rotdens", s.type + tag + ".rot: expected four columns");
                const int n = (int)v.size() / 4;
                std::vector<double> col[4];
                for (auto& x : col) x.resize(n);
                for (int i = 0; i < n; ++i)
                    for (int k = 0; k < 4; ++k) col[k][i] = v[4 * i + k];
                r.rotdens = pimc_add_rotdens(c.ctx, n, col[0].data(), col[1].data(), col[2].data(), col[3].data());
                c.check(r.rotdens, "init_rotdens");
            }
        } else {
            if (r.ntypes > 1) {  // init_pot3D: Rgrd THgrd CHgrd Rvmin Rvmax, then V
                auto v = read_numbers(s.fpot + ".pot");
                if (v.size() < 5) die("init_pot3D", s.fpot + ".pot: short header");
                const int nr = (int)v[0], nth = (int)v[1], nch = (int)v[2];
                if (v.size() != 5 + (size_t)nr * nth * nch) die("init_pot3D", s.fpot + ".pot: sizes do not match the data");
                r.pot3d = pimc_add_pot3d(c.ctx, nr, nth, nch, v[3], v[4], v.data() + 5);
                c.check(r.pot3d, "init_pot3D");
            }
            if (d.rotation) {  // init_rot3D: TYPE_T<T>t<Prot>.rho/.eng/.esq, 181*361*361 values each
                const size_t n = 181ull * 361 * 361;
                auto rho = read_numbers(s.type + tag + ".rho");
                auto eng = read_numbers(s.type + tag + ".eng");
                auto esq = read_numbers(s.type + tag + ".esq");
                if (rho.size() != n || eng.size() != n || esq.size() != n)
                    die("init_rot3D", s.type + tag + ": each of .rho/.eng/.esq must hold 181*361*361 values");
                r.rot3d = pimc_add_rot3d(c.ctx, rho.data(), eng.data(), esq.data());
                c.check(r.rot3d, "init_rot3D");
            }
        }
        if (s.type == d.rot_type) r.rtstep = s.rtstep;
    }
    return r;
}

// ------------------------------------------------------------------- initial configuration
// [3][natoms*P] coordinates and angles (phi, cos(theta), chi), slice fastest.
void initial_config(const Deck& d, std::vector<double>& coords, std::vector<double>& angles) {
    const int N = d.natoms(), P = d.P, n = N * P;
    coords.assign(3 * (size_t)n, 0.0);
    angles.assign(3 * (size_t)n, 0.0);
    for (int i = 0; i < n; ++i) angles[n + i] = 1.0;  // MCConfigInit: phi 0, cos 1, chi 0
    if (d.read_coords) {
        // xyz.init (initconf.f): a line with a count (and permutation indices), a comment line,
        // then one line per bead: label x phi y cos(theta) z chi
        std::ifstream f("xyz.init");
        if (!f.good()) die("initconf", "cannot open xyz.init");
        std::string line, label;
        std::getline(f, line);
        std::getline(f, line);
        for (int i = 0; i < n; ++i) {
            if (!std::getline(f, line)) die("initconf", "xyz.init holds fewer than natoms*P beads");
            std::istringstream s(line);
            s >> label;
            for (int j = 0; j < 3; ++j) s >> coords[(size_t)j * n + i] >> angles[(size_t)j * n + i];
            if (!s) die("initconf", "bad bead line in xyz.init");
        }
        return;
    }
    if (d.density <= 0.0) die("MCInit", "DENSITY must be positive to build the initial lattice");
    int natoms = 0, nmolcs = 0;
    for (auto& s : d.types) (s.molecule ? nmolcs : natoms) += s.numb;
    const double box = std::pow((natoms + nmolcs) / d.density, 1.0 / 3.0);
    int first = 0;
    double shift[3];
    // atoms: a simple cubic fill of the box, x fastest (initLattice_config)
    const double abox = natoms ? box / std::pow((double)natoms, 1.0 / 3.0) : 0.0;
    for (double& x : shift) x = 0.5 * abox;
    for (auto& s : d.types) {
        if (s.molecule == 0)
            for (int a = first; a < first + s.numb; ++a) {
                for (int id = 0; id < 3; ++id) coords[(size_t)id * n + a * P] = shift[id] - 0.5 * box;
                shift[0] += abox;
                for (int id = 0; id < 2; ++id)
                    if (shift[id] > box) { shift[id] = 0.5 * abox; shift[id + 1] += abox; }
            }
        first += s.numb;
    }
    // molecules: their own cubic fill, wrapped before placing; moved out of the atoms' cube
    // when the counts are equal
    const double amol = nmolcs ? box / std::pow((double)nmolcs, 1.0 / 3.0) : 0.0;
    for (double& x : shift) x = 0.5 * amol;
    first = 0;
    for (auto& s : d.types) {
        if (s.molecule)
            for (int a = first; a < first + s.numb; ++a) {
                for (int id = 0; id < 2; ++id)
                    if (shift[id] > box) { shift[id] = 0.5 * amol; shift[id + 1] += amol; }
                for (int id = 0; id < 3; ++id)
                    coords[(size_t)id * n + a * P] = shift[id] - 0.5 * box + (natoms == nmolcs ? box : 0.0);
                shift[0] += amol;
            }
        first += s.numb;
    }
    for (int id = 0; id < 3; ++id)  // replInitial_config: every slice starts at slice 0
        for (int a = 0; a < N; ++a)
            for (int it = 1; it < P; ++it) coords[(size_t)id * n + a * P + it] = coords[(size_t)id * n + a * P];
}

// -------------------------------------------------------------------------------- outputs
// One sample of MCGetAverage mc_main.cc:551-640: K, V, E_rot, E_rot^2, Cv, Cv_trans, Cv_rot, (E_rot term)^2.
struct Sums {
    double kin = 0, pot = 0, rot = 0, rotsq = 0, cv = 0, cvt = 0, cvr = 0, count = 0;
    void add(const double* s) {
        kin += s[0]; pot += s[1]; rot += s[2]; rotsq += s[3];
        cv += s[4]; cvt += s[5]; cvr += s[6]; count += 1.0;
    }
};

struct Chain {
    std::string stem;
    std::ofstream sum;
    Sums block, total;
    double sums_count = 0.0;
    std::vector<double> tot0, acc0;  // counters at the start of the block
    // bin counts and rotational correlation sums: this block, and all measured blocks
    std::vector<double> gr1, gr2, rcf, gr1_sum, gr2_sum, rcf_sum;
    // around a non-linear rotor: (r, theta, chi) bins of this block and of all measured blocks,
    // and the relative Euler angle bins of all measured blocks (never reset per block, as
    // _relthe_sum and its like are not)
    std::vector<double> gr3, gr3_sum, rel_sum;
};

constexpr int NBR = 300, NBT = 50, NBC = 100, NRCF = 10;  // MC_BINSR/T/C (mc_estim.cc:25-27), NUMB_RCF
constexpr double DR = 15.0 / NBR;                        // MAX_RADIUS / MC_BINSR, MIN_RADIUS = 0
const double DTH = M_PI / (NBT - 1), DCH = 2.0 * M_PI / (NBC - 1);
constexpr size_t NB3 = (size_t)NBR * NBT * NBC;

std::ofstream open_out(const std::string& name, const char* who) {
    std::ofstream f(name);
    if (!f.good()) die(who, "cannot open " + name);
    f << std::setprecision(6) << std::scientific;
    return f;
}

// SaveDensities2D: theta, r and the density per line, theta outer (mc_estim.cc:1666-1727)
void save_g2d(const std::vector<double>& g, const std::string& name, double acount, const Deck& d, int n_atoms) {
    auto f = open_out(name, "SaveDensities2D");
    const char* sp = "   ";
    const double norm3 = DR * DTH * d.P * acount;
    for (int it = 0; it < NBT; ++it) {
        for (int ir = 0; ir < NBR; ++ir) {
            if (n_atoms)
                f << std::setw(14) << (it + 0.5) * DTH * 180.0 / M_PI << sp << std::setw(14) << (ir + 0.5) * DR << sp
                  << std::setw(14) << g[ir * NBT + it] / norm3 << sp;
            f << "\n";
        }
    }
}

// SaveRCF: <n(tau).n(0)> and <P_l(n(tau).n(0))>, l = 1..9, over tau = k * tau_rot, k = 0..Prot
void save_rcf(const std::vector<double>& rcf, const std::string& name, double acount, const Deck& d) {
    auto f = open_out(name, "SaveRCF");
    const char* sp = "   ";
    const int Prot = d.Prot;
    const double norm = acount * Prot, tau = 1.0 / (d.T * Prot);
    for (int it = 0; it <= Prot; ++it)
        f << std::setw(14) << it * tau << sp << std::setw(14) << rcf[it % Prot] / norm << sp << "\n";
    f << "\n\n#\n";
    for (int it = 0; it <= Prot; ++it) {
        f << std::setw(14) << it * tau << sp;
        for (int l = 1; l < NRCF; ++l) f << std::setw(14) << rcf[(size_t)l * Prot + it % Prot] / norm << sp;
        f << "\n";
    }
}

// SaveDensities1D (.gra, and .gri/.grt around a linear rotor) and SaveDensities2D (.g2d) of
// one block; n_atoms is the number of structureless atoms (one atom type).
void save_block_densities(const Chain& c, const std::string& fname, double acount, const Deck& d, int n_atoms,
                          bool linear) {
    const char* sp = "   ";
    const double box = std::pow((double)d.natoms() / d.density, 1.0 / 3.0);
    {
        auto f = open_out(fname + ".gra", "SaveDensities1D");
        const double norm0 = 2.0 * M_PI * DR * acount * d.P / (box * box * box);
        for (int ir = 0; ir < NBR; ++ir) {
            const double r = (ir + 0.5) * DR;
            f << std::setw(14) << r << sp;
            if (n_atoms) f << std::setw(14) << c.gr1[ir] / (r * r * norm0 * n_atoms * (n_atoms - 1)) << sp;
            f << "\n";
        }
    }
    if (!linear) return;
    {
        auto f = open_out(fname + ".gri", "SaveDensities1D");
        const double norm1 = DR * d.P * acount * 4.0 * M_PI;
        for (int ir = 0; ir < NBR; ++ir) {
            const double r = (ir + 0.5) * DR;
            double dens = 0.0;
            for (int it = 0; it < NBT; ++it) dens += c.gr2[ir * NBT + it];
            f << std::setw(14) << r << sp;
            if (n_atoms) f << std::setw(14) << dens / (norm1 * r * r) << sp;
            f << "\n";
        }
    }
    {
        auto f = open_out(fname + ".grt", "SaveDensities1D");
        const double norm2 = d.P * acount * DTH;
        for (int it = 0; it < NBT; ++it) {
            double dens = 0.0;
            for (int ir = 0; ir < NBR; ++ir) dens += c.gr2[ir * NBT + it];
            f << std::setw(14) << (it + 0.5) * DTH * 180.0 / M_PI << sp;
            if (n_atoms) f << std::setw(14) << dens / (norm2 * n_atoms) << sp;
            f << "\n";
        }
    }
    save_g2d(c.gr2, fname + ".g2d", acount, d, n_atoms);
}


// Distributions around a non-linear rotor.  The reference keeps NUMB_DENS3D = atom types +
// molecule types columns, one per type in input order (densities_init mc_estim.cc:244-246);
// only the atom type's column has counts.  dens3d_columns gives, per column, the bins (or null)
// and the type's count.
struct Col3 {
    const std::vector<double>* bins;
    double numb;
};
std::vector<Col3> dens3d_columns(const Deck& d, const std::vector<double>& g3) {
    std::vector<Col3> cols;
    for (auto& s : d.types) cols.push_back({s.molecule == 0 ? &g3 : nullptr, (double)s.numb});
    return cols;
}

// SaveRho1D (mc_estim.cc:1445-1590): rho(r), rho(theta), rho(chi) of the 3D bins, and with
// rel (the accumulated files) the relative Euler angle distributions (:1591-1662).
void save_rho1d(const std::vector<double>& g3, const std::string& fname, double acount, const Deck& d,
                const std::vector<double>* rel) {
    const char* sp = "   ";
    const auto cols = dens3d_columns(d, g3);
    {
        auto f = open_out(fname + ".gri", "SaveRho1D");
        const double norm1 = DR * d.P * acount;
        for (int ir = 0; ir < NBR; ++ir) {
            f << std::setw(14) << (ir + 0.5) * DR << sp;
            for (auto& c : cols) {
                double dens = 0.0;
                if (c.bins)
                    for (size_t k = (size_t)ir * NBT * NBC; k < (size_t)(ir + 1) * NBT * NBC; ++k) dens += (*c.bins)[k];
                f << std::setw(14) << dens / norm1 << sp;
            }
            f << "\n";
        }
    }
    {
        auto f = open_out(fname + ".grt", "SaveRho1D");
        const double norm2 = d.P * acount * DTH * (180.0 / M_PI);
        for (int it = 0; it < NBT; ++it) {
            f << std::setw(14) << (it + 0.5) * DTH * 180.0 / M_PI << sp;
            for (auto& c : cols) {
                double dens = 0.0;
                if (c.bins)
                    for (int ir = 0; ir < NBR; ++ir)
                        for (int ic = 0; ic < NBC; ++ic) dens += (*c.bins)[((size_t)ir * NBT + it) * NBC + ic];
                f << std::setw(14) << dens / (norm2 * c.numb) << sp;
            }
            f << "\n";
        }
    }
    {
        auto f = open_out(fname + ".grc", "SaveRho1D");
        const double norm4 = d.P * acount * DCH * (180.0 / M_PI);
        for (int ic = 0; ic < NBC; ++ic) {
            f << std::setw(14) << (ic + 0.5) * DCH * 180.0 / M_PI << sp;
            for (auto& c : cols) {
                double dens = 0.0;
                if (c.bins)
                    for (int ir = 0; ir < NBR; ++ir)
                        for (int it = 0; it < NBT; ++it) dens += (*c.bins)[((size_t)ir * NBT + it) * NBC + ic];
                f << std::setw(14) << dens / (norm4 * c.numb) << sp;
            }
            f << "\n";
        }
    }
    if (!rel) return;
    const double nrot = d.rotation ? d.Prot : d.P;  // NumbRotTimes
    const double norm5 = nrot * acount * DCH * (180.0 / M_PI), norm6 = nrot * acount * DTH * (180.0 / M_PI);
    const std::vector<double>& q = *rel;
    {
        auto f = open_out(fname + ".eulphi", "SaveRho1D");
        for (int ip = 0; ip < NBC; ++ip)
            f << std::setw(14) << (ip + 0.5) * DCH * 180.0 / M_PI << sp << std::setw(14) << q[NBT + ip] / norm5 << "\n";
    }
    {
        auto f = open_out(fname + ".eulchi", "SaveRho1D");
        for (int ic = 0; ic < NBC; ++ic)
            f << std::setw(14) << (ic + 0.5) * DCH * 180.0 / M_PI << sp << std::setw(14) << q[NBT + NBC + ic] / norm5
              << "\n";
    }
    {
        auto f = open_out(fname + ".eulthe", "SaveRho1D");
        for (int it = 0; it < NBT; ++it)
            f << std::setw(14) << (it + 0.5) * DTH * 180.0 / M_PI << sp << std::setw(14) << q[it] / norm6 << "\n";
    }
}

// SaveRhoThetaChi (mc_estim.cc:1930-1995): rho(theta, chi) over all r, chi inner, a blank line
// after each theta
void save_gtc(const std::vector<double>& g3, const std::string& name, double acount, const Deck& d) {
    const char* sp = "   ";
    const auto cols = dens3d_columns(d, g3);
    auto f = open_out(name, "SaveRhoThetaChi");
    const double norm6 = DTH * DCH * d.P * acount * (180.0 * 180.0 / (M_PI * M_PI));
    for (int it = 0; it < NBT; ++it) {
        for (int ic = 0; ic < NBC; ++ic) {
            f << std::setw(14) << (it + 0.5) * DTH * 180.0 / M_PI << sp << std::setw(14)
              << (ic + 0.5) * DCH * 180.0 / M_PI << sp;
            for (auto& c : cols) {
                double dens = 0.0;
                if (c.bins)
                    for (int ir = 0; ir < NBR; ++ir) dens += (*c.bins)[((size_t)ir * NBT + it) * NBC + ic];
                f << std::setw(14) << dens / (norm6 * c.numb) << sp;
            }
            f << "\n";
        }
        f << "\n";
    }
}

// SaveDensities3D (mc_estim.cc:1728-1820): r, theta, chi and rho / (r^2 sin theta) per column,
// one line per bin, chi fastest.  Formatted into a buffer: the file has 1.5 million lines.  The
// r, theta and chi fields repeat, so each is formatted once; the densities go through
// sci_format::e14_6, which writes printf's "%14.6e" bytes.
void save_g3d(const std::vector<double>& g3, const std::string& name, double acount, const Deck& d) {
    const auto cols = dens3d_columns(d, g3);
    auto f = open_out(name, "SaveDensities3D");
    const double norm5 = DR * DTH * DCH * d.P * acount;
    auto field = [](double x) {
        char num[32];
        const int n = std::snprintf(num, sizeof num, "%14.6e   ", x);
        if (n < 0 || n >= (int)sizeof num) die("SaveDensities3D", "number format overflow");
        return std::string(num, (size_t)n);
    };
    std::vector<std::string> fr(NBR), ft(NBT), fc(NBC);
    for (int ir = 0; ir < NBR; ++ir) fr[ir] = field((ir + 0.5) * DR);
    for (int it = 0; it < NBT; ++it) ft[it] = field((it + 0.5) * DTH * 180.0 / M_PI);
    for (int ic = 0; ic < NBC; ++ic) fc[ic] = field((ic + 0.5) * DCH * 180.0 / M_PI);
    std::string buf;
    buf.reserve(1 << 22);
    char num[32];  // one density field (snprintf's longest is 15 characters)
    for (int ir = 0; ir < NBR; ++ir) {
        const double r = (ir + 0.5) * DR;
        for (int it = 0; it < NBT; ++it) {
            const double sth = std::sin((it + 0.5) * DTH);
            for (int ic = 0; ic < NBC; ++ic) {
                const size_t k = ((size_t)ir * NBT + it) * NBC + ic;
                for (auto& c : cols) {
                    const double v = c.bins ? (*c.bins)[k] / (norm5 * r * r * sth) : 0.0;
                    buf += fr[ir];
                    buf += ft[it];
                    buf += fc[ic];
                    buf.append(num, (size_t)sci_format::e14_6(v, num));
                    buf += "   ";
                }
                buf.push_back('\n');
            }
            if (buf.size() > (1u << 22) - 4096) {
                f.write(buf.data(), (std::streamsize)buf.size());
                buf.clear();
            }
        }
    }
    f.write(buf.data(), (std::streamsize)buf.size());
    if (!f.good()) die("SaveDensities3D", "cannot write " + name);
}

// Runs job(r) for r = 0..n-1 on up to 16 threads.
template <typename F>
void for_each_chain(int n, F job) {
    const int hw = (int)std::thread::hardware_concurrency();
    const int nt = std::max(1, std::min(std::min(n, 16), hw));
    if (nt == 1) {
        for (int r = 0; r < n; ++r) job(r);
        return;
    }
    std::atomic<int> next{0};
    std::vector<std::thread> pool;
    for (int t = 0; t < nt; ++t)
        pool.emplace_back([&] {
            for (int r = next++; r < n; r = next++) job(r);
        });
    for (auto& th : pool) th.join();
}

void set_sci(std::ostream& o) { o << std::setprecision(6) << std::scientific; }

void save_energy(const Chain& c, long block) {  // SaveEnergy
    std::ofstream f(c.stem + ".eng", std::ios::app);
    if (!f.good()) die("SaveEnergy", "cannot open " + c.stem + ".eng");
    set_sci(f);
    const Sums& b = c.block;
    const double n = b.count;
    const char* sp = "   ";
    f << std::setw(4) << block << sp << std::setw(14) << b.kin / n << sp << std::setw(14) << b.pot / n << sp
      << std::setw(14) << (b.kin + b.pot) / n << sp << std::setw(14) << b.rot / n << sp << std::setw(14) << b.rotsq / n
      << sp << std::setw(14) << (b.kin + b.pot + b.rot) / n << sp << std::setw(14) << b.cv / n << sp << std::setw(14)
      << b.cvt / n << sp << std::setw(14) << b.cvr / n << sp << "\n";
}

void save_sum_energy(Chain& c, const Deck& d, int natoms) {  // SaveSumEnergy
    const Sums& t = c.total;
    const double n = t.count, T = d.T, beta = 1.0 / d.T;
    const double e = (t.kin + t.pot + t.rot) / n;
    double cv = 0.5 * 3 * natoms * d.P * T - e;
    cv = -(cv * cv + t.cv / n) * beta / T;
    double cvt = 0.5 * 3 * natoms * d.P * T - t.kin / n;
    cvt = -(cvt * cvt + t.cvt / n) * beta / T;
    double cvr = -t.rot / n;
    cvr = -(cvr * cvr + t.cvr / n) * beta / T;
    const char* sp = "   ";
    c.sum << std::setw(4) << c.sums_count << sp << std::setw(14) << t.kin / n << sp << std::setw(14) << t.pot / n << sp
          << std::setw(14) << (t.kin + t.pot) / n << sp << std::setw(14) << t.rot / n << sp << std::setw(14)
          << t.rotsq / n << sp << std::setw(14) << e << sp << std::setw(14) << cv << sp << std::setw(14) << cvt << sp
          << std::setw(14) << cvr << sp << "\n";
    c.sum.flush();
}

void save_xyz(const Chain& c, const Deck& d, const std::vector<double>& coords, const std::vector<double>& cosines) {
    std::ofstream f(c.stem + ".xyz");  // IOxyz: label, then (coordinate, direction cosine) per axis
    if (!f.good()) die("IOxyz", "cannot open " + c.stem + ".xyz");
    set_sci(f);
    const int n = d.natoms() * d.P;
    const char* sp = "   ";
    f << n << "\n#" << sp << "xyz format:  [atom type]  x y z (Angstrom) \n";
    int i = 0;
    for (auto& s : d.types)
        for (int a = 1; a <= s.numb; ++a)
            for (int it = 0; it < d.P; ++it, ++i) {
                f << std::setw(5) << (s.type + std::to_string(a)) << sp;
                for (int id = 0; id < 3; ++id)
                    f << std::setw(14) << coords[(size_t)id * n + i] << sp << std::setw(14) << cosines[(size_t)id * n + i] << sp;
                f << "\n";
            }
}

void print_ratios(const Deck& d, Context& c, Chain& ch, int r, long block, long pass, long step) {
    const int nt = (int)d.types.size();
    std::vector<double> tot(3 * nt), acc(3 * nt);
    c.check(pimc_get_counters(c.ctx, r, tot.data(), acc.data()), "MCSaveAcceptRatio");
    std::ostringstream o;  // MCSaveAcceptRatio: default float format, width 8
    const char* sp = "   ";
    o << "BLOCK:" << std::setw(8) << block << sp << "PASS:" << std::setw(8) << pass << sp << "STEP:" << std::setw(8)
      << step << sp;
    auto ratio = [&](int t, int m) {
        const double T = tot[3 * t + m] - ch.tot0[3 * t + m], A = acc[3 * t + m] - ch.acc0[3 * t + m];
        return A / T;
    };
    int imtype = -1;
    for (int t = 0; t < nt; ++t) {
        if (d.types[t].type == d.rot_type) imtype = t;
        o << std::setw(8) << d.types[t].type << sp << std::setw(8) << ratio(t, PIMC_MCMOLEC) << sp << std::setw(8)
          << ratio(t, PIMC_MCMULTI) << sp;
    }
    if (d.rotation && imtype >= 0) o << sp << "Rot: " << std::setw(8) << ratio(imtype, PIMC_MCROTAT) << sp;
    if (r > 0) o << " [chain " << r << "]";
    std::cout << o.str() << std::endl;
}

// ------------------------------------------------------------------------------ checkpoint
const char* const F_STATUS = "yw001.stat";
const char* const F_CONFIG = "yw001.conf";
const char* const F_RANDOM = "yw001.rand";

bool file_exists(const char* name) { return std::ifstream(name).good(); }

void back_up(const char* name) {  // IOFileBackUp: the previous file as <name>.old
    if (!file_exists(name)) return;
    std::ifstream src(name, std::ios::binary);
    std::ofstream dst(std::string(name) + ".old", std::ios::binary | std::ios::trunc);
    dst << src.rdbuf();
}

struct Checkpoint {
    long block = 0;
    unsigned seed = 0;
    std::vector<unsigned> pass;
    std::vector<std::vector<double>> coords, cosines, angles;
};

void write_checkpoint(const Checkpoint& k, size_t n3) {
    back_up(F_STATUS);
    back_up(F_CONFIG);
    back_up(F_RANDOM);
    {
        std::ofstream f(F_STATUS, std::ios::trunc);
        f << "STARTBLOCK " << k.block << "\nREPLICAS " << k.pass.size() << "\n";
        if (!f.good()) die("StatusIO", std::string("cannot write ") + F_STATUS);
    }
    {
        std::ofstream f(F_CONFIG, std::ios::binary | std::ios::trunc);
        const long long bytes = (long long)(n3 * sizeof(double));
        for (size_t r = 0; r < k.pass.size(); ++r) {
            f.write(reinterpret_cast<const char*>(&bytes), sizeof bytes);
            f.write(reinterpret_cast<const char*>(k.coords[r].data()), bytes);
            f.write(reinterpret_cast<const char*>(k.cosines[r].data()), bytes);
            f.write(reinterpret_cast<const char*>(k.angles[r].data()), bytes);
        }
        if (!f.good()) die("ConfigIO", std::string("cannot write ") + F_CONFIG);
    }
    {
        std::ofstream f(F_RANDOM, std::ios::trunc);
        f << "PHILOX4X32-10 " << k.seed << "\n";
        for (size_t r = 0; r < k.pass.size(); ++r) f << r << " " << k.pass[r] << "\n";
        if (!f.good()) die("RandomIO", std::string("cannot write ") + F_RANDOM);
    }
}

Checkpoint read_checkpoint(int nrep, size_t n3) {
    Checkpoint k;
    {
        std::ifstream f(F_STATUS);
        if (!f.good()) die("StatusIO", std::string("RESTART: cannot open ") + F_STATUS);
        std::string key;
        long reps = -1;
        while (f >> key) {
            if (key == "STARTBLOCK") f >> k.block;
            else if (key == "REPLICAS") f >> reps;
            std::string rest;
            std::getline(f, rest);
        }
        if (reps != nrep) die("StatusIO", "RESTART: the checkpoint holds " + std::to_string(reps) + " chains, not " + std::to_string(nrep));
    }
    {
        std::ifstream f(F_RANDOM);
        std::string gen;
        if (!(f >> gen >> k.seed) || gen != "PHILOX4X32-10") die("RandomIO", std::string("RESTART: bad ") + F_RANDOM);
        k.pass.resize(nrep);
        for (int r = 0; r < nrep; ++r) {
            int idx;
            if (!(f >> idx >> k.pass[r]) || idx != r) die("RandomIO", std::string("RESTART: bad ") + F_RANDOM);
        }
    }
    {
        std::ifstream f(F_CONFIG, std::ios::binary);
        if (!f.good()) die("ConfigIO", std::string("RESTART: cannot open ") + F_CONFIG);
        k.coords.resize(nrep);
        k.cosines.resize(nrep);
        k.angles.resize(nrep);
        for (int r = 0; r < nrep; ++r) {
            long long bytes = 0;
            f.read(reinterpret_cast<char*>(&bytes), sizeof bytes);
            if (!f.good() || bytes != (long long)(n3 * sizeof(double)))
                die("ConfigIO", std::string("RESTART: ") + F_CONFIG + " does not match the deck's particles and slices");
            for (auto* v : {&k.coords[r], &k.cosines[r], &k.angles[r]}) {
                v->resize(n3);
                f.read(reinterpret_cast<char*>(v->data()), bytes);
            }
            if (!f.good()) die("ConfigIO", std::string("RESTART: ") + F_CONFIG + " is truncated");
        }
    }
    return k;
}

}  // namespace

int main(int argc, char** argv) {
    std::string input = "qmc.input";
    unsigned seed = 985456376u;  // mc_main.cc:151
    unsigned flags = 0;
    int nrep = 1;
    bool init_only = false, autotune = false;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        if (a == "--seed" && i + 1 < argc) seed = (unsigned)std::strtoul(argv[++i], nullptr, 10);
        else if (a == "--fp64") flags |= PIMC_FLAG_FP64;
        else if (a == "--ref-order") flags |= PIMC_FLAG_REF_ORDER;
        else if (a == "--replicas" && i + 1 < argc) nrep = std::atoi(argv[++i]);
        else if (a == "--init-only") init_only = true;
        else if (a == "--autotune") autotune = true;
        else if (a == "-h" || a == "--help") {
            std::cout << "usage: pimc [qmc.input] [--seed N] [--fp64] [--ref-order] [--replicas R] [--autotune] "
                         "[--init-only]\n";
            return 0;
        } else if (!a.empty() && a[0] != '-') input = a;
        else die("main", "unknown option " + a);
    }
    if (nrep < 1) die("main", "--replicas must be positive");
    Deck d = read_deck(input);
    const int N = d.natoms(), n = N * d.P;

    std::vector<double> coords, angles;
    initial_config(d, coords, angles);
    // direction cosines from (phi, cos(theta)) as MCConfigInit sets them; beads that carry no
    // orientation on the device keep these
    std::vector<double> cos0(3 * (size_t)n);
    for (int i = 0; i < n; ++i) {
        const double ph = angles[i], ct = angles[n + i], st = std::sqrt(std::max(0.0, 1.0 - ct * ct));
        cos0[i] = st * std::cos(ph);
        cos0[n + i] = st * std::sin(ph);
        cos0[2 * (size_t)n + i] = ct;
    }
    const std::string stem = d.outdir + d.prefix;
    if (init_only) {
        Chain ch;
        ch.stem = stem;
        save_xyz(ch, d, coords, cos0);
        return 0;
    }

    Checkpoint resume;
    if (d.restart) {
        resume = read_checkpoint(nrep, 3 * (size_t)n);
        seed = resume.seed;
        d.startblock = resume.block;
    } else if (file_exists(F_STATUS) || file_exists(F_CONFIG) || file_exists(F_RANDOM)) {
        // a new run does not overwrite a checkpoint (mc_main.cc:126-136)
        die("QMC", "a checkpoint (yw001.*) exists here: add RESTART to qmc.input to resume, or remove it");
    }

    Context c;
    if (pimc_create(nrep, flags, seed, &c.ctx) != 0)
        die("pimc_create", c.ctx ? pimc_last_error(c.ctx) : "no HIP device");
    pimc_replica_desc desc = load_tables(d, c);
    for (int r = 0; r < nrep; ++r) c.check(pimc_set_replica(c.ctx, r, &desc), "pimc_set_replica");
    if (d.minimage) {  // BoxSize of MCInit (mc_setup.cc:343-366): the volume per particle from DENSITY
        if (d.density <= 0.0) die("MCInit", "MINIMAGE needs a positive DENSITY for the box");
        const double L = std::pow((double)d.natoms() / d.density, 1.0 / 3.0), box[3] = {L, L, L};
        for (int r = 0; r < nrep; ++r) c.check(pimc_set_minimum_image(c.ctx, r, box), "pimc_set_minimum_image");
    }
    // MCSKIP_AVERG below NUMBEROFSLICES: samples inside a pass, after every P/k time steps in the
    // reference's order or every S/k phases in segment order; k the largest divisor of that grid
    // not above P / MCSKIP_AVERG
    int per_pass = 1;
    if (d.skip_averg < d.P) {
        int S = 1;
        for (auto& s : d.types) S = std::max(S, 1 << s.levels);
        const long grid = (flags & PIMC_FLAG_REF_ORDER) ? d.P : S;
        for (long k = std::min<long>(grid, d.P / d.skip_averg); k >= 1; --k)
            if (grid % k == 0) {
                per_pass = (int)k;
                break;
            }
        if (per_pass > 1) c.check(pimc_set_option(c.ctx, "samples_per_pass", per_pass), "pimc_set_option");
    }
    c.check(pimc_commit(c.ctx), "pimc_commit");
    for (int r = 0; r < nrep; ++r) c.check(pimc_set_state(c.ctx, r, coords.data(), angles.data()), "pimc_set_state");
    if (d.restart)
        for (int r = 0; r < nrep; ++r)
            c.check(pimc_restore_state(c.ctx, r, resume.coords[r].data(), resume.angles[r].data(),
                                       resume.cosines[r].data(), resume.pass[r]),
                    "pimc_restore_state");
    if (autotune) {
        long shapes[5];
        c.check(pimc_autotune(c.ctx, 2, shapes), "pimc_autotune");
        std::cerr << "launch shapes: bisection lanes " << shapes[0] << "/" << shapes[1] << "/" << shapes[2]
                  << ", rotor part LDS " << shapes[3] << " KB, centroid parts " << shapes[4] << "\n";
    }

    std::vector<Chain> chains(nrep);
    for (int r = 0; r < nrep; ++r) {
        Chain& ch = chains[r];
        ch.stem = r ? stem + "_r" + std::to_string(r) : stem;
        std::ofstream(ch.stem + ".eng", std::ios::trunc);
        ch.sum.open(ch.stem + "_sum.eng", std::ios::trunc);
        if (!ch.sum.good()) die("InitTotalAverage", "cannot open " + ch.stem + "_sum.eng");
        set_sci(ch.sum);
    }
    std::cout << "pimc: " << input << ": " << N << " particles, " << d.P << " slices, T = " << d.T << " K, "
              << nrep << " chain(s)" << std::endl;

    // Pass-granular schedule of the step counters (steps = passes * slices).
    const long P = d.P;
    const long every = std::max(1L, d.skip_averg / P);
    const long ratio_passes = std::max(1L, d.skip_ratio / P);
    // passes per library call: at most 1024 sampled passes, and no more samples than a replica's
    // buffer holds (4096, read and cleared after each call)
    const long chunk = every * std::max(1L, std::min(std::min(1024L, 4096L / per_pass), ratio_passes / every));
    auto crosses = [&](long q, long skip) { return (q * P) / skip > ((q - 1) * P) / skip; };

    const int nt = (int)d.types.size();
    int n_plain = 0;
    bool linear = false, nonlinear = false;
    for (auto& s : d.types) {
        if (s.molecule == 0) n_plain += s.numb;
        if (s.molecule == 1) linear = true;
        if (s.molecule == 2) nonlinear = true;
    }
    std::vector<double> samples;
    for (long block = d.startblock + 1; block <= d.startblock + d.blocks; ++block) {
        const bool measured = block > d.eqblocks;
        for (int r = 0; r < nrep; ++r) {
            Chain& ch = chains[r];
            ch.block = Sums();
            ch.tot0.assign(3 * nt, 0.0);
            ch.acc0.assign(3 * nt, 0.0);
            c.check(pimc_get_counters(c.ctx, r, ch.tot0.data(), ch.acc0.data()), "ResetMCCounts");
        }
        c.check(pimc_clear_samples(c.ctx), "pimc_clear_samples");
        for (auto& ch : chains) {
            ch.gr1.assign(NBR, 0.0);
            ch.gr2.assign(NBR * NBT, 0.0);
            ch.rcf.assign(d.rotation ? (size_t)NRCF * d.Prot : 0, 0.0);
            if (nonlinear) {
                ch.gr3.assign(NB3, 0.0);
                if (ch.rel_sum.empty()) ch.rel_sum.assign(NBT + 2 * NBC, 0.0);
            }
        }
        for (long done = 0; done < d.passes;) {
            const long m = std::min(chunk, d.passes - done);
            c.check(pimc_run(c.ctx, (int)m, measured ? (int)every : 0), "pimc_run");
            for (int r = 0; r < nrep; ++r) {
                Chain& ch = chains[r];
                int got = 0;
                if (measured) {
                    samples.resize(8 * (size_t)(m / every * per_pass + 1));
                    got = pimc_get_samples(c.ctx, r, samples.data(), (int)(m / every * per_pass + 1));
                    c.check(got, "pimc_get_samples");
                }
                int next = 0;
                bool ratio_line = false;
                for (long q = done + 1; q <= done + m; ++q) {
                    ratio_line |= crosses(q, d.skip_ratio);
                    if (measured && (q - done) % every == 0)
                        for (int i = 0; i < per_pass && next < got; ++i, ++next) {
                            ch.block.add(&samples[8 * (size_t)next]);
                            ch.total.add(&samples[8 * (size_t)next]);
                        }
                    if (measured && crosses(q, d.skip_total) && ch.block.count > 0) {
                        ch.sums_count += 1.0;
                        save_sum_energy(ch, d, N);
                    }
                }
                if (ratio_line) print_ratios(d, c, ch, r, block, done + m, (done + m) * P);
                if (measured && got > 0) {  // bins and correlation sums of this chunk's samples
                    std::vector<double> g1(NBR), g2(NBR * NBT);
                    c.check(pimc_get_histograms(c.ctx, r, g1.data(), g2.data()), "pimc_get_histograms");
                    for (int i = 0; i < NBR; ++i) ch.gr1[i] += g1[i];
                    for (int i = 0; i < NBR * NBT; ++i) ch.gr2[i] += g2[i];
                    if (nonlinear) {
                        std::vector<double> g3(NB3), rl(NBT + 2 * NBC);
                        const int has = pimc_get_histograms3d(c.ctx, r, g3.data(), rl.data());
                        c.check(has, "pimc_get_histograms3d");
                        if (has & 1)
                            for (size_t i = 0; i < NB3; ++i) ch.gr3[i] += g3[i];
                        if (has & 2)
                            for (int i = 0; i < NBT + 2 * NBC; ++i) ch.rel_sum[i] += rl[i];
                    }
                    if (d.rotation) {
                        std::vector<double> rc((size_t)NRCF * d.Prot);
                        c.check(pimc_get_rcf(c.ctx, r, rc.data(), (int)rc.size()), "pimc_get_rcf");
                        for (size_t i = 0; i < rc.size(); ++i) ch.rcf[i] += rc[i];
                    }
                }
            }
            if (measured) c.check(pimc_clear_samples(c.ctx), "pimc_clear_samples");
            done += m;
        }
        // every chain's configuration from the device first (device calls stay on this thread),
        // then each chain's files on a worker thread: the chains write disjoint files
        std::vector<std::vector<double>> cgs(nrep, std::vector<double>(3 * (size_t)n)), sgs = cgs;
        std::vector<double> ag(3 * (size_t)n);
        for (int r = 0; r < nrep; ++r) {
            std::vector<double>& sg = sgs[r];
            c.check(pimc_get_state(c.ctx, r, cgs[r].data(), ag.data(), sg.data()), "pimc_get_state");
            // beads without an orientation on the device keep their initial cosines
            int first = 0;
            for (auto& s : d.types) {
                if (!(d.rotation && s.type == d.rot_type))
                    for (int i = first * d.P; i < (first + s.numb) * d.P; ++i)
                        for (int id = 0; id < 3; ++id) sg[(size_t)id * n + i] = cos0[(size_t)id * n + i];
                first += s.numb;
            }
        }
        for_each_chain(nrep, [&](int r) {
            Chain& ch = chains[r];
            if (measured && ch.block.count > 0) {
                // MCSaveBlockAverages, then the accumulated files (mc_main.cc:446-462)
                std::ostringstream bn;
                bn << std::setw(3) << std::setfill('0') << block;
                const std::string fname = ch.stem + bn.str();
                // the block densities only with a molecule (MCSaveBlockAverages mc_main.cc:716-735)
                if (linear || nonlinear) save_block_densities(ch, fname, ch.block.count, d, n_plain, linear);
                if (nonlinear) {  // MCSaveBlockAverages mc_main.cc:722-735
                    save_rho1d(ch.gr3, fname, ch.block.count, d, nullptr);
                    save_gtc(ch.gr3, fname + ".gtc", ch.block.count, d);
                }
                if (d.rotation) save_rcf(ch.rcf, fname + ".rcf", ch.block.count, d);
                save_energy(ch, block);
                if (ch.gr1_sum.empty()) {
                    ch.gr1_sum.assign(NBR, 0.0);
                    ch.gr2_sum.assign(NBR * NBT, 0.0);
                    ch.rcf_sum.assign(ch.rcf.size(), 0.0);
                }
                for (int i = 0; i < NBR; ++i) ch.gr1_sum[i] += ch.gr1[i];
                for (int i = 0; i < NBR * NBT; ++i) ch.gr2_sum[i] += ch.gr2[i];
                for (size_t i = 0; i < ch.rcf.size(); ++i) ch.rcf_sum[i] += ch.rcf[i];
                {  // SaveGraSum
                    auto f = open_out(ch.stem + "_sum.gra", "SaveGraSum");
                    const double norma = DR * ch.total.count * d.P;
                    for (int ir = 0; ir < NBR; ++ir) {
                        f << std::setw(14) << (ir + 0.5) * DR << "   ";
                        if (n_plain)
                            f << std::setw(14) << ch.gr1_sum[ir] / (norma * n_plain * (n_plain - 1) / 2.0) << "   ";
                        f << "\n";
                    }
                }
                if (linear) save_g2d(ch.gr2_sum, ch.stem + "_sum.g2d", ch.total.count, d, n_plain);
                if (nonlinear) {  // mc_main.cc:456-461
                    if (ch.gr3_sum.empty()) ch.gr3_sum.assign(NB3, 0.0);
                    for (size_t i = 0; i < NB3; ++i) ch.gr3_sum[i] += ch.gr3[i];
                    save_g3d(ch.gr3_sum, ch.stem + "_sum.g3d", ch.total.count, d);
                    save_rho1d(ch.gr3_sum, ch.stem + "_sum", ch.total.count, d, &ch.rel_sum);
                }
                if (d.rotation) save_rcf(ch.rcf_sum, ch.stem + "_sum.rcf", ch.total.count, d);
            }
            save_xyz(ch, d, cgs[r], sgs[r]);
        });
        // checkpoint of every chain after the block (its exact device state)
        Checkpoint k;
        k.block = block;
        k.seed = seed;
        k.pass.resize(nrep);
        k.coords.assign(nrep, std::vector<double>(3 * (size_t)n));
        k.cosines = k.coords;
        k.angles = k.coords;
        for (int r = 0; r < nrep; ++r) {
            c.check(pimc_get_state(c.ctx, r, k.coords[r].data(), k.angles[r].data(), k.cosines[r].data()),
                    "pimc_get_state");
            c.check(pimc_get_pass(c.ctx, r, &k.pass[r]), "pimc_get_pass");
        }
        write_checkpoint(k, 3 * (size_t)n);
    }
    double pe = 0, bm = 0;
    c.check(pimc_get_work(c.ctx, &pe, &bm), "pimc_get_work");
    std::cout << "pimc: done, " << bm << " bead moves, " << pe << " pair evaluations" << std::endl;
    pimc_destroy(c.ctx);
    return 0;
}
