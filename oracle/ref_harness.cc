// ref_harness.cc -- drives the compiled upstream MoRiBS-PIMC objects for golden vectors.
//
// TEST INFRASTRUCTURE ONLY, built into oracle/_ref/libref.so by oracle/ref.mk from the
// reference sources where they lie.  It sets the reference's globals directly, calls the
// reference routines (mc_piqmc.cc, mc_poten.cc, mc_estim.cc, rotden.f, vcord.f,
// caleng_tip4p_gg.f) and replaces only the random-number sources: the SPRNG entry points
// that mc_randg.cc binds (init_rng/get_rn_dbl, sprng.h:55-70) and runif() of omprng.cc:34-39
// pop values from per-stream tapes, so a move can be replayed by the oracle on the same
// numbers.  Run with one OpenMP thread (the reference has data races, mc_piqmc.cc:1210).
#include "mc_confg.h"
#include "mc_const.h"
#include "mc_setup.h"
#include "mc_input.h"
#include "mc_poten.h"
#include "mc_piqmc.h"
#include "mc_estim.h"
#include "mc_utils.h"
#include "mc_randg.h"
#include "rngstream.h"
#include <omp.h>
#include <unistd.h>
#include <cstring>
#include <vector>
#include <cmath>

extern double** MCTotal;
extern double** MCAccep;
void PIMCPass(int, int);
void densities_init(void);
void densities_malloc(void);
void densities_reset(int);
extern double** _gr1D;
extern double*** _gr2D;
extern double** _gr3D;
extern double* _relthe_sum;
extern double* _relphi_sum;
extern double* _relchi_sum;
extern int NUMB_DENS3D;

extern "C" void rotden_(double*, double*, double*, double*, double*, double*, double*, double*, double*, int*);
extern "C" void vcord_(double*, double*, double*, double*, int*, int*, int*, double*, double*, double*, double*,
                       double*, double*, double*, double*, double*, double*, int*);
extern "C" void caleng_(double*, double*, double*, double*, double*);

namespace {
struct Tape {
    std::vector<double> v;
    size_t pos = 0;
};
Tape g_tape[16];
int g_short = 0;
double pop(int k) {
    Tape& t = g_tape[k];
    if (t.pos >= t.v.size()) { g_short = 1; return 0.5; }
    return t.v[t.pos++];
}
}  // namespace

// ---- RNG entry points bound by mc_randg.cc (C linkage) and omprng.cc ----
extern "C" int* init_rng(int gennum, int, int, int) { return reinterpret_cast<int*>(&g_tape[gennum & 15]); }
extern "C" double get_rn_dbl(int* id) {
    Tape* t = reinterpret_cast<Tape*>(id);
    return pop(int(t - g_tape));
}
extern "C" int free_rng(int*) { return 0; }
extern "C" int pack_rng(int*, char**) { return 0; }
extern "C" int* unpack_rng(char*) { return nullptr; }
double runif(RngStream*) { return pop(15); }
void randomseed() {}


extern "C" {

int h_setup(int ntypes, const char* names, const int* numb, const int* molecule, const int* levels,
            const double* mcstep, const double* mass, int P, int Prot, double T, double rtstep,
            const char* workdir, const char* potnames) {
    omp_set_num_threads(1);
    if (chdir(workdir) != 0) return -1;
    NDIM = 3;
    Temperature = T;
    Density = 0.02;
    NumbTimes = P;
    NumbRotTimes = Prot;
    ROTATION = false;
    IMPURITY = false;
    WORM = false;
    MINIMAGE = false;
    NUMB_ATOMS = NUMB_MOLCS = NUMB_ATOMTYPES = NUMB_MOLCTYPES = 0;
    int gatom = 0;
    const char* nm = names;
    const char* pn = potnames;
    for (int t = 0; t < ntypes; ++t) {
        std::strncpy(MCAtom[t].type, nm, MAX_STRING_LENGTH - 1);
        nm += std::strlen(nm) + 1;
        std::strncpy(MCAtom[t].fpot, pn, MAX_STRING_LENGTH - 1);
        pn += std::strlen(pn) + 1;
        MCAtom[t].numb = numb[t];
        MCAtom[t].molecule = molecule[t];
        MCAtom[t].levels = levels[t];
        MCAtom[t].mcstep = mcstep[t];
        MCAtom[t].mass = mass[t];
        MCAtom[t].brot = 0.0;
        MCAtom[t].stat = BOLT;
        MCAtom[t].pmod = PRIMITIVE;
        MCAtom[t].rtstep = rtstep;
        MCAtom[t].gatom = gatom;
        MCAtom[t].offset = gatom * P;
        gatom += numb[t];
        if (molecule[t] == 1 || molecule[t] == 2) {
            IMPURITY = true;
            NUMB_MOLCS += numb[t];
            NUMB_MOLCTYPES++;
            if (Prot > 0) ROTATION = true;
        } else {
            NUMB_ATOMS += numb[t];
            NUMB_ATOMTYPES++;
        }
    }
    NumbAtoms = gatom;
    NumbTypes = ntypes;
    MaxnTimes = gatom * P;
    MCAtom[NumbTypes].offset = MaxnTimes;
    MCSetUnits();
    MCMemAlloc();
    MemAllocMCCounts();
    MCInit();
    MCConfigInit();
    InitPotentials();
    // linear rotor: the .rot file in workdir (init_rotdens).  Non-linear rotor: the three
    // 181x361x361 tables come from h_set_rot3d instead of init_rot3D's text files.
    if (ROTATION && MCAtom[IMTYPE].molecule == 1) InitRotDensity();
    densities_init();
    densities_malloc();
    densities_reset(MC_TOTAL);
    densities_reset(MC_BLOCK);
    ResetMCCounts();
    RandomInit(0, 1);  // stream pointers of mc_randg.cc, bound to the tapes below
    return 0;
}

void h_set_state(const double* coords, const double* angles) {
    int n = NumbAtoms * NumbTimes;
    for (int id = 0; id < 3; ++id)
        for (int i = 0; i < n; ++i) {
            MCCoords[id][i] = coords[id * n + i];
            MCAngles[id][i] = angles[id * n + i];
        }
    for (int i = 0; i < n; ++i) {  // as mc_main.cc:186-197
        double phi = MCAngles[PHI][i], cost = MCAngles[CTH][i];
        double sint = sqrt(1.0 - cost * cost);
        MCCosine[AXIS_X][i] = sint * cos(phi);
        MCCosine[AXIS_Y][i] = sint * sin(phi);
        MCCosine[AXIS_Z][i] = cost;
    }
}

void h_set_newcoords(const double* c) {
    int n = NumbAtoms * NumbTimes;
    for (int id = 0; id < 3; ++id)
        for (int i = 0; i < n; ++i) newcoords[id][i] = c[id * n + i];
}

void h_get_state(double* coords, double* angles, double* cosines) {
    int n = NumbAtoms * NumbTimes;
    for (int id = 0; id < 3; ++id)
        for (int i = 0; i < n; ++i) {
            if (coords) coords[id * n + i] = MCCoords[id][i];
            if (angles) angles[id * n + i] = MCAngles[id][i];
            if (cosines) cosines[id * n + i] = MCCosine[id][i];
        }
}

void h_set_rot3d(double* rho, double* eng, double* esq) {
    rhoprp = rho;
    erotpr = eng;
    erotsq = esq;
}

// the globals init_pot3D (mc_poten.cc:254-303) fills from a file
void h_set_pot3d(int nr, int nth, int nch, double rmin, double rmax, double* v) {
    Rgrd = nr;
    THgrd = nth;
    CHgrd = nch;
    Rvmin = rmin;
    Rvmax = rmax;
    Rvstep = (Rvmax - Rvmin) / (double)(Rgrd - 1);
    vtable = v;
}

// minimum image (MINIMAGE, mc_input.cc:298-300) with the given BoxSize (mc_setup.cc:343-366)
void h_set_minimage(int on, const double* box) {
    MINIMAGE = on != 0;
    for (int id = 0; id < NDIM; ++id) BoxSize[id] = box[id];
}
double h_spot1d(double r, int type) { return SPot1D(r, type); }
double h_lpot2d(double r, double c, int type) { return LPot2D(r, c, type); }
double h_srotdens(double g, int type) { return SRotDens(g, type); }
double h_srotdens_deriv(double g, int type) { return SRotDensDeriv(g, type); }
double h_srotdens_esqrt(double g, int type) { return SRotDensEsqrt(g, type); }
double h_pot_energy_slice(int atom, int use_new, int it) { return PotEnergy(atom, use_new ? newcoords : MCCoords, it); }
double h_pot_energy_path(int atom, int use_new) { return PotEnergy(atom, use_new ? newcoords : MCCoords); }
double h_pot_rot_energy(int atom, int use_new, int it) { return PotRotEnergy(atom, use_new ? newcoords : MCCosine, it); }
double h_pot_rot_e3d(int atom, double* eul, int it) { return PotRotE3D(atom, eul, it); }

void h_rotden(double* e1, double* e2, double* erel, double* rho, double* erot, double* esq, int* istop) {
    rotden_(e1, e2, erel, rho, erot, esq, rhoprp, erotpr, erotsq, istop);
}

void h_vcord(double* eul, double* rcom, double* rpt, double* out /*vpot,rad,the,chi*/) {
    double hx[3], hy[3], hz[3];
    int iv = 0;
    vcord_(eul, rcom, rpt, vtable, &Rgrd, &THgrd, &CHgrd, &Rvmax, &Rvmin, &Rvstep, &out[0], &out[1], &out[2], &out[3],
           hx, hy, hz, &iv);
}

double h_caleng(double* c1, double* c2, double* e1, double* e2) {
    double e;
    caleng_(c1, c2, &e, e1, e2);
    return e;
}

double h_kin_energy() { return GetKinEnergy(); }
double h_pot_energy_densities(double* gr1d, double* gr2d) {
    densities_reset(MC_BLOCK);
    double v = GetPotEnergy_Densities();
    if (gr1d && NUMB_ATOMTYPES > 0)
        for (int ir = 0; ir < 300; ++ir) gr1d[ir] = _gr1D[0][ir];
    if (gr2d && IMPURITY && MCAtom[IMTYPE].molecule == 1)
        for (int ir = 0; ir < 300; ++ir)
            for (int it = 0; it < 50; ++it) gr2d[ir * 50 + it] = _gr2D[0][ir][it];
    return v;
}
// The bins of one estimator call around a non-linear rotor: GetPotEnergy_Densities fills
// _gr3D (bin_3Ddensity mc_estim.cc:1252-1273), GetRotE3D the relative Euler angle bins
// (mc_estim.cc:1046-1058).  gr3d takes NUMB_DENS3D x 300*50*100 values, rel the 50 theta,
// 100 phi and 100 chi bins; returns NUMB_DENS3D.
int h_nl_densities(double* gr3d, double* rel) {
    if (!(IMPURITY && MCAtom[IMTYPE].molecule == 2)) return 0;
    densities_reset(MC_BLOCK);
    densities_reset(MC_TOTAL);
    GetPotEnergy_Densities();
    GetRotE3D();
    const long nb = 300L * 50 * 100;
    for (int id = 0; id < NUMB_DENS3D; ++id)
        for (long i = 0; i < nb; ++i) gr3d[id * nb + i] = _gr3D[id][i];
    for (int i = 0; i < 50; ++i) rel[i] = _relthe_sum[i];
    for (int i = 0; i < 100; ++i) rel[50 + i] = _relphi_sum[i];
    for (int i = 0; i < 100; ++i) rel[150 + i] = _relchi_sum[i];
    return NUMB_DENS3D;
}
double h_rot_energy(double* esq, double* term) {
    double v = GetRotEnergy();
    *esq = ErotSQ;
    *term = Erot_termSQ;
    return v;
}
double h_rot_e3d(double* esq, double* term) {
    double v = GetRotE3D();
    *esq = ErotSQ;
    *term = Erot_termSQ;
    return v;
}

void h_set_tape(int stream, const double* v, long n) {
    g_tape[stream].v.assign(v, v + n);
    g_tape[stream].pos = 0;
    g_short = 0;
}
long h_tape_pos(int stream) { return long(g_tape[stream].pos); }
int h_tape_short() { return g_short; }

void h_molecular_move(int type) { MCMolecularMove(type); }
void h_bisection_move(int type, int time) { MCBisectionMove(type, time); }
void h_rotations_move(int type) { MCRotationsMove(type); }
void h_rotations_3d(int type) { MCRotations3D(type); }
void h_pimc_pass(int type, int time) { PIMCPass(type, time); }
void h_reset_counters() { ResetMCCounts(); }
void h_get_counters(double* total, double* accep) {
    for (int t = 0; t < NumbTypes; ++t)
        for (int m = 0; m < 3; ++m) {
            total[t * 3 + m] = MCTotal[t][m];
            accep[t * 3 + m] = MCAccep[t][m];
        }
}

}  // extern "C"
