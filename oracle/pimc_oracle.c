/*
 * pimc_oracle.c -- FP64 CPU restatement of MoRiBS-PIMC's sampling hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load this library;
 * the product never does.
 *
 * Citations are file:line in the upstream MoRiBS-PIMC tree.
 *
 * Parity pins (each is exercised by tests/test_oracle_*.py):
 *   or_spline/or_splint/or_init_spline  golden vectors from the compiled reference
 *                                        (tests/golden/ref_vectors.json) + knot identity
 *   or_spot1d, or_lpot2d, or_srotdens*  golden vectors + node identity + analytic branches
 *   or_pot_energy_slice/path, or_pot_rot_energy, or_pot_rot_e3d
 *                                        golden vectors + brute-force double loop
 *   or_deleul/or_rotden/or_rotpro        golden vectors + R1^T R2 round trip
 *   or_vcord/or_vcalc                    golden vectors + grid-node identity
 *   or_caleng                            golden vectors + hand-computed TIP4P dimer
 *   or_kin_energy, or_pot_energy_densities, or_rot_energy, or_rot_e3d
 *                                        golden vectors (estimators on fixed paths)
 *   or_nl_densities                      the compiled reference's own bins on the same
 *                                        paths + hand-placed partners and frames
 *   or_linden_rho                        reference-held .rot fixtures + closed form
 *   moves / schedule                     statistical parity against the compiled
 *                                        reference program and closed forms
 *                                        (free particle, free rotor)
 *   or_philox4x32_10                     Random123 known-answer vectors
 */
#include "pimc_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* constants, mc_const.h:9-15 and mc_confg.h:60 */
#define OR_HBAR 1.05457266
#define OR_AMU  1.6605402
#define OR_K_B  1.380658
#define OR_WNO2K 0.6950356
#define OR_RZERO 1.0e-10
#define OR_MC_BINSR 300
#define OR_MC_BINST 50
#define OR_MC_BINSC 100
#define OR_MAX_RADIUS 15.0

typedef struct {
    int n;
    double *x, *y, *y2;
    double alpha, unode, c6;
} pot1d_t;

typedef struct {
    int nr, nc;
    double dr, dc;
    double *rg, *cg, *v; /* v[ir*nc+ic] */
} pot2d_t;

typedef struct {
    int n;
    double *x, *rho, *erot, *esq, *rho2, *erot2, *esq2;
} rotdens_t;

typedef struct {
    int numb, molecule, levels, mlsegm, offset;
    double mcstep, mass, lambda;
} type_t;

struct or_sys {
    int ntypes, natoms, P, Prot, RotRatio, imtype, rotden_type;
    double T, beta, tau, rottau, rtstep;
    type_t type[OR_MAX_TYPES];
    int* mctype;
    double *coords, *cosine, *angles, *newc; /* each [3][natoms*P] contiguous */
    pot1d_t pot1d[OR_MAX_TYPES];
    pot2d_t pot2d[OR_MAX_TYPES];
    rotdens_t rotd[OR_MAX_TYPES];
    /* 3D tables (borrowed) */
    int nr3, nth3, nch3;
    double rvmin, rvmax, rvstep;
    const double* vtable;
    const double *rhoprp, *erotpr, *erotsq;
    int minimage;
    double box[3];
    double total[OR_MAX_TYPES][3], accep[OR_MAX_TYPES][3];
    uint32_t seed, replica;
    /* tape mode: draws come from per-stream tapes in the reference's call order
       (stream numbers of mc_randg.cc: 1 rnd1, 2 rnd2, 3 rnd3, 8/9 gauss r1/r2,
       15 runif of omprng.cc) instead of the counter-based generator */
    int tape_mode, tape_short;
    double* tape[16];
    long tape_n[16], tape_pos[16];
    char err[256];
};

#define C(s, id, i)  ((s)->coords[(size_t)(id) * (s)->natoms * (s)->P + (i)])
#define CS(s, id, i) ((s)->cosine[(size_t)(id) * (s)->natoms * (s)->P + (i)])
#define AN(s, id, i) ((s)->angles[(size_t)(id) * (s)->natoms * (s)->P + (i)])
#define NW(s, id, i) ((s)->newc[(size_t)(id) * (s)->natoms * (s)->P + (i)])
#define PHI 0
#define CTH 1
#define CHI 2

static int fail(or_sys* s, const char* msg) {
    snprintf(s->err, sizeof s->err, "%s", msg);
    return -1;
}

/* ======================= Philox4x32-10 ======================= */
/* Algorithm from Salmon et al., "Parallel random numbers: as easy as 1, 2, 3",
   SC'11; multipliers and Weyl constants as published there. */
void or_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; ++r) {
        if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* 53-bit uniform in [0,1) from two 32-bit words */
static double u53(uint32_t a, uint32_t b) {
    return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * (1.0 / 9007199254740992.0);
}

/* Counter map (identical in the HIP path):
     key  = {seed, replica}
     ctr0 = pass number
     ctr1 = kind << 28 | step        kind: 1 centroid, 2 bisection, 3 linear rotor, 4 3D rotor
     ctr2 = atom << 16 | slice       slice = segment start (bisection) or rotor slice
     ctr3 = level << 24 | point << 8 | id
   The same move (pass, atom, segment or slice, sweep) uses the same numbers in
   either schedule. */
#define K_CENTROID 1u
#define K_BISECT 2u
#define K_ROTLIN 3u
#define K_ROT3D 4u

static void draw(const or_sys* s, uint32_t pass, uint32_t kind, uint32_t step, uint32_t atom,
                 uint32_t slice, uint32_t c3, uint32_t out[4]) {
    uint32_t ctr[4] = {pass, (kind << 28) | (step & 0x0FFFFFFFu), (atom << 16) | (slice & 0xFFFFu), c3};
    uint32_t key[2] = {s->seed, s->replica};
    or_philox4x32_10(ctr, key, out);
}

double or_uniform(uint32_t seed, uint32_t replica, const uint32_t ctr[4], int which) {
    uint32_t key[2] = {seed, replica}, w[4];
    or_philox4x32_10(ctr, key, w);
    return which ? u53(w[2], w[3]) : u53(w[0], w[1]);
}

/* Gaussian convention of gauss() mc_randg.cc:138-150: density ~ exp(-alpha x^2),
   Box-Muller cosine branch. r1 in (0,1]. */
static double gauss_from(const uint32_t w[4], double alpha) {
    double r1 = 1.0 - u53(w[0], w[1]);
    double r2 = u53(w[2], w[3]);
    double x1 = sqrt(-log(r1)) * cos(2.0 * M_PI * r2);
    return x1 / sqrt(alpha);
}

double or_gauss(uint32_t seed, uint32_t replica, const uint32_t ctr[4], double alpha) {
    uint32_t key[2] = {seed, replica}, w[4];
    or_philox4x32_10(ctr, key, w);
    return gauss_from(w, alpha);
}

static double tape_pop(or_sys* s, int k) {
    if (s->tape_pos[k] >= s->tape_n[k]) { s->tape_short = 1; return 0.5; }
    return s->tape[k][s->tape_pos[k]++];
}

/* centroid move draws: idx 0..2 displacement (rnd1), idx 3 acceptance (rnd2) */
static double u_centroid(or_sys* s, uint32_t pass, int gatom, int idx) {
    if (s->tape_mode) return tape_pop(s, idx < 3 ? 1 : 2);
    uint32_t w[4];
    draw(s, pass, K_CENTROID, 0, (uint32_t)gatom, 0, (uint32_t)(idx >> 1), w);
    return (idx & 1) ? u53(w[2], w[3]) : u53(w[0], w[1]);
}

/* bisection Gaussian for coordinate id of midpoint j at level l (gauss, streams 8/9) */
static double g_bisect(or_sys* s, uint32_t pass, int gatom, int time, int level, int j, int id,
                       double alpha) {
    double r1, r2;
    if (s->tape_mode) {
        r1 = tape_pop(s, 8);
        r2 = tape_pop(s, 9);
    } else {
        uint32_t w[4];
        draw(s, pass, K_BISECT, 0, (uint32_t)gatom, (uint32_t)time,
             ((uint32_t)level << 24) | ((uint32_t)j << 8) | (uint32_t)id, w);
        r1 = 1.0 - u53(w[0], w[1]);
        r2 = u53(w[2], w[3]);
    }
    double x1 = sqrt(-log(r1)) * cos(2.0 * M_PI * r2);
    return x1 / sqrt(alpha);
}

/* bisection level acceptance (rnd3); only drawn when the level's deltav >= 0 */
static double u_bisect_accept(or_sys* s, uint32_t pass, int gatom, int time, int level) {
    if (s->tape_mode) return tape_pop(s, 3);
    uint32_t w[4];
    draw(s, pass, K_BISECT, 0, (uint32_t)gatom, (uint32_t)time, ((uint32_t)level << 24) | (0xFFFFu << 8), w);
    return u53(w[0], w[1]);
}

/* rotor move draws (runif), n = 3 (linear) or 4 (3D), drawn before the step */
static void u_rotor(or_sys* s, uint32_t kind, uint32_t pass, uint32_t sweep, int gatom, int k, int n,
                    double* r) {
    if (s->tape_mode) {
        for (int i = 0; i < n; ++i) r[i] = tape_pop(s, 15);
        return;
    }
    uint32_t w0[4], w1[4];
    draw(s, pass, kind, sweep, (uint32_t)gatom, (uint32_t)k, 0, w0);
    draw(s, pass, kind, sweep, (uint32_t)gatom, (uint32_t)k, 1, w1);
    r[0] = u53(w0[0], w0[1]);
    r[1] = u53(w0[2], w0[3]);
    r[2] = u53(w1[0], w1[1]);
    r[3] = u53(w1[2], w1[3]);
}

int or_set_tape(or_sys* s, int stream, const double* v, long n) {
    if (stream < 0 || stream > 15) return fail(s, "bad tape stream");
    free(s->tape[stream]);
    s->tape[stream] = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    if (n > 0) memcpy(s->tape[stream], v, sizeof(double) * (size_t)n);
    s->tape_n[stream] = n;
    s->tape_pos[stream] = 0;
    s->tape_mode = 1;
    s->tape_short = 0;
    return 0;
}

long or_tape_pos(or_sys* s, int stream) { return s->tape_pos[stream]; }
int  or_tape_short(or_sys* s) { return s->tape_short; }
void or_tape_off(or_sys* s) { s->tape_mode = 0; }

/* ======================= spline (mc_utils.cc:112-201) ======================= */
void or_spline(const double* x, const double* y, int n, double yp1, double ypn, double* y2) {
    double p, qn, sig, un;
    double* u = (double*)malloc(sizeof(double) * (size_t)n);
    if (yp1 > 0.99e30) y2[0] = u[0] = 0.0;
    else {
        y2[0] = -0.5;
        u[0] = (3. / (x[1] - x[0])) * ((y[1] - y[0]) / (x[1] - x[0]) - yp1);
    }
    for (int i = 1; i < n - 1; i++) {
        sig = (x[i] - x[i - 1]) / (x[i + 1] - x[i - 1]);
        p = sig * y2[i - 1] + 2.;
        y2[i] = (sig - 1.) / p;
        u[i] = (y[i + 1] - y[i]) / (x[i + 1] - x[i]) - (y[i] - y[i - 1]) / (x[i] - x[i - 1]);
        u[i] = (6. * u[i] / (x[i + 1] - x[i - 1]) - sig * u[i - 1]) / p;
    }
    if (ypn > 0.99e30) qn = un = 0.;
    else {
        qn = .5;
        un = (3. / (x[n - 1] - x[n - 2])) * (ypn - (y[n - 1] - y[n - 2]) / (x[n - 1] - x[n - 2]));
    }
    y2[n - 1] = (un - qn * u[n - 2]) / (qn * y2[n - 2] + 1.);
    for (int k = n - 2; k >= 0; k--) y2[k] = y2[k] * y2[k + 1] + u[k];
    free(u);
}

/* splint: bisection cell search, mc_utils.cc:156-186 */
double or_splint(const double* xa, const double* ya, const double* y2a, int n, double x) {
    int klo = 0, khi = n - 1, k;
    while (khi - klo > 1) {
        k = (khi + klo) >> 1;
        if (xa[k] > x) khi = k;
        else klo = k;
    }
    double h = xa[khi] - xa[klo];
    double a = (xa[khi] - x) / h;
    double b = (x - xa[klo]) / h;
    return a * ya[klo] + b * ya[khi] + ((a * a * a - a) * y2a[klo] + (b * b * b - b) * y2a[khi]) * (h * h) / 6.;
}

/* end slopes from first differences, mc_utils.cc:188-201 */
void or_init_spline(const double* grid, const double* data, double* sdata, int n) {
    double drl = grid[1] - grid[0];
    double dpl = (data[1] - data[0]) / drl;
    double drr = grid[n - 1] - grid[n - 2];
    double dpr = (data[n - 1] - data[n - 2]) / drr;
    or_spline(grid, data, n, dpl, dpr, sdata);
}

/* ======================= system ======================= */
or_sys* or_create(int ntypes, const int* numb, const int* molecule, const int* levels,
                  const double* mcstep, const double* mass, int P, int Prot, double T,
                  double rtstep, int rotden_type) {
    if (ntypes < 1 || ntypes > OR_MAX_TYPES || P < 2) return NULL;
    or_sys* s = (or_sys*)calloc(1, sizeof(or_sys));
    s->ntypes = ntypes;
    s->P = P;
    s->Prot = Prot > 0 ? Prot : P;
    s->T = T;
    s->rtstep = rtstep;
    s->rotden_type = rotden_type;
    s->imtype = -1;
    /* MCSetUnits mc_setup.cc:194-217: lambda = 0.5 * 100 hbar^2/(amu kB) / m */
    double lam = 100.0 * (OR_HBAR * OR_HBAR) / (OR_AMU * OR_K_B);
    int gatom = 0;
    for (int t = 0; t < ntypes; ++t) {
        type_t* ty = &s->type[t];
        ty->numb = numb[t];
        ty->molecule = molecule[t];
        ty->levels = levels[t];
        ty->mlsegm = (int)pow(2.0, levels[t]);
        ty->mcstep = mcstep[t];
        ty->mass = mass[t];
        ty->lambda = 0.5 * lam / mass[t];
        ty->offset = gatom * P; /* mc_input.cc:352-362 */
        gatom += numb[t];
        if (molecule[t] == 1 || molecule[t] == 2) s->imtype = t;
    }
    s->natoms = gatom;
    /* MCInit mc_setup.cc:374-398 */
    s->beta = 1.0 / T;
    s->tau = s->beta / (double)P;
    s->rottau = s->beta / (double)s->Prot;
    s->RotRatio = P / s->Prot;
    s->mctype = (int*)malloc(sizeof(int) * (size_t)(gatom > 0 ? gatom : 1));
    int na = 0;
    for (int t = 0; t < ntypes; ++t)
        for (int a = 0; a < numb[t]; ++a) s->mctype[na++] = t;
    size_t sz = (size_t)3 * gatom * P;
    s->coords = (double*)calloc(sz, sizeof(double));
    s->cosine = (double*)calloc(sz, sizeof(double));
    s->angles = (double*)calloc(sz, sizeof(double));
    s->newc = (double*)calloc(sz, sizeof(double));
    /* MCConfigInit mc_setup.cc:476-489: phi=0, cos(theta)=1, chi=0 */
    for (int i = 0; i < gatom * P; ++i) {
        AN(s, CTH, i) = 1.0;
        CS(s, 2, i) = 1.0;
    }
    return s;
}

static void free_tables(or_sys* s) {
    for (int t = 0; t < OR_MAX_TYPES; ++t) {
        free(s->pot1d[t].x); free(s->pot1d[t].y); free(s->pot1d[t].y2);
        free(s->pot2d[t].rg); free(s->pot2d[t].cg); free(s->pot2d[t].v);
        rotdens_t* r = &s->rotd[t];
        free(r->x); free(r->rho); free(r->erot); free(r->esq); free(r->rho2); free(r->erot2); free(r->esq2);
    }
}

void or_destroy(or_sys* s) {
    if (!s) return;
    free_tables(s);
    for (int k = 0; k < 16; ++k) free(s->tape[k]);
    free(s->mctype); free(s->coords); free(s->cosine); free(s->angles); free(s->newc);
    free(s);
}

const char* or_last_error(or_sys* s) { return s->err; }
int or_natoms(or_sys* s) { return s->natoms; }
int or_imtype(or_sys* s) { return s->imtype; }

static double* dup(const double* a, int n) {
    double* b = (double*)malloc(sizeof(double) * (size_t)n);
    memcpy(b, a, sizeof(double) * (size_t)n);
    return b;
}

/* init_pot1D mc_poten.cc:379-438 (units are K and Angstrom, so no scaling) */
int or_set_pot1d(or_sys* s, int type, int n, const double* x, const double* y) {
    if (n < 2) return fail(s, "Not enough data points for extrapolation");
    pot1d_t* p = &s->pot1d[type];
    p->n = n;
    p->x = dup(x, n);
    p->y = dup(y, n);
    p->y2 = (double*)malloc(sizeof(double) * (size_t)n);
    or_init_spline(p->x, p->y, p->y2, n);
    double fr = p->y[0] / p->y[1];
    double dr = p->x[1] - p->x[0];
    p->alpha = log(fr) / dr;
    p->unode = p->y[0] * exp(p->alpha * p->x[0]);
    double r0 = pow(p->x[n - 2], 6.0);
    double r1 = pow(p->x[n - 1], 6.0);
    fr = p->y[n - 1] - p->y[n - 2];
    p->c6 = fr / (1.0 / r0 - 1.0 / r1);
    return 0;
}

void or_get_pot1d_extrap(or_sys* s, int type, double* alpha, double* unode, double* c6, double* y2) {
    pot1d_t* p = &s->pot1d[type];
    *alpha = p->alpha; *unode = p->unode; *c6 = p->c6;
    if (y2) memcpy(y2, p->y2, sizeof(double) * (size_t)p->n);
}

/* init_pot2D mc_poten.cc:305-377 */
int or_set_pot2d(or_sys* s, int type, int nr, int nc, double dr, double dc, const double* rg,
                 const double* cg, const double* v) {
    pot2d_t* p = &s->pot2d[type];
    p->nr = nr; p->nc = nc; p->dr = dr; p->dc = dc;
    p->rg = dup(rg, nr); p->cg = dup(cg, nc); p->v = dup(v, nr * nc);
    return 0;
}

/* init_rotdens mc_poten.cc:508-546 */
int or_set_rotdens(or_sys* s, int type, int n, const double* x, const double* rho,
                   const double* erot, const double* esq) {
    rotdens_t* r = &s->rotd[type];
    r->n = n;
    r->x = dup(x, n); r->rho = dup(rho, n); r->erot = dup(erot, n); r->esq = dup(esq, n);
    r->rho2 = (double*)malloc(sizeof(double) * (size_t)n);
    r->erot2 = (double*)malloc(sizeof(double) * (size_t)n);
    r->esq2 = (double*)malloc(sizeof(double) * (size_t)n);
    or_init_spline(r->x, r->rho, r->rho2, n);
    or_init_spline(r->x, r->erot, r->erot2, n);
    or_init_spline(r->x, r->esq, r->esq2, n);
    return 0;
}

/* init_pot3D mc_poten.cc:254-303 */
int or_set_pot3d(or_sys* s, int nr, int nth, int nch, double rmin, double rmax, const double* v) {
    s->nr3 = nr; s->nth3 = nth; s->nch3 = nch;
    s->rvmin = rmin; s->rvmax = rmax;
    s->rvstep = (rmax - rmin) / (double)(nr - 1);
    s->vtable = v;
    return 0;
}

int or_set_rot3d(or_sys* s, const double* rho, const double* erot, const double* esq) {
    s->rhoprp = rho; s->erotpr = erot; s->erotsq = esq;
    return 0;
}

void or_set_minimage(or_sys* s, int on, const double box[3]) {
    s->minimage = on;
    if (box) for (int i = 0; i < 3; ++i) s->box[i] = box[i];
}

static void cosines_from_angles(or_sys* s) {
    int n = s->natoms * s->P;
    for (int i = 0; i < n; ++i) {
        double phi = AN(s, PHI, i), cost = AN(s, CTH, i);
        double sint = sqrt(1.0 - cost * cost);
        CS(s, 0, i) = sint * cos(phi);
        CS(s, 1, i) = sint * sin(phi);
        CS(s, 2, i) = cost;
    }
}

void or_set_state(or_sys* s, const double* coords, const double* angles) {
    size_t sz = (size_t)3 * s->natoms * s->P;
    if (coords) memcpy(s->coords, coords, sizeof(double) * sz);
    if (angles) {
        memcpy(s->angles, angles, sizeof(double) * sz);
        cosines_from_angles(s); /* as after xyz.init, mc_main.cc:186-197 */
    }
}

void or_get_state(or_sys* s, double* coords, double* angles, double* cosines) {
    size_t sz = (size_t)3 * s->natoms * s->P;
    if (coords) memcpy(coords, s->coords, sizeof(double) * sz);
    if (angles) memcpy(angles, s->angles, sizeof(double) * sz);
    if (cosines) memcpy(cosines, s->cosine, sizeof(double) * sz);
}

void or_get_counters(or_sys* s, double* total, double* accep) {
    for (int t = 0; t < s->ntypes; ++t)
        for (int m = 0; m < 3; ++m) {
            total[t * 3 + m] = s->total[t][m];
            accep[t * 3 + m] = s->accep[t][m];
        }
}

void or_reset_counters(or_sys* s) {
    memset(s->total, 0, sizeof s->total);
    memset(s->accep, 0, sizeof s->accep);
}

void or_set_rng(or_sys* s, uint32_t seed, uint32_t replica) {
    s->seed = seed;
    s->replica = replica;
}

/* ======================= interpolation ======================= */
/* SPot1D mc_poten.cc:624-639 */
double or_spot1d(or_sys* s, double r, int type) {
    pot1d_t* p = &s->pot1d[type];
    int n = p->n;
    if (r >= p->x[n - 1]) return (-p->c6 / pow(r, 6.0));
    if (r <= p->x[0]) return (p->unode * exp(-p->alpha * r));
    return or_splint(p->x, p->y, p->y2, n, r);
}

/* LPot2D mc_poten.cc:688-729: cells clamped, weights extrapolate */
double or_lpot2d(or_sys* s, double r, double cost, int type) {
    pot2d_t* p = &s->pot2d[type];
    double rmin = p->rg[0], cmin = p->cg[0];
    int ir = (int)floor((r - rmin) / p->dr);
    int ic = (int)floor((cost - cmin) / p->dc);
    if (ir < 0) ir = 0;
    else if (ir >= (p->nr - 1)) ir = p->nr - 2;
    if (ic < 0) ic = 0;
    else if (ic >= (p->nc - 1)) ic = p->nc - 2;
    int nc = p->nc;
    double y1 = p->v[ir * nc + ic];
    double y2 = p->v[(ir + 1) * nc + ic];
    double y3 = p->v[(ir + 1) * nc + ic + 1];
    double y4 = p->v[ir * nc + ic + 1];
    double r1 = p->rg[ir], r2 = p->rg[ir + 1];
    double c1 = p->cg[ic], c2 = p->cg[ic + 1];
    double dr = (r - r1) / (r2 - r1);
    double dc = (cost - c1) / (c2 - c1);
    return (1.0 - dr) * (1.0 - dc) * y1 + dr * (1.0 - dc) * y2 + dr * dc * y3 + (1.0 - dr) * dc * y4;
}

/* SRotDens mc_poten.cc:548-570 */
double or_srotdens(or_sys* s, double g, int type) {
    rotdens_t* r = &s->rotd[type];
    int n = r->n;
    if (g > r->x[n - 1]) return r->rho[n - 1];
    if (g < r->x[0]) {
        double rl = r->x[0], rr = r->x[1];
        double sa = (r->rho[1] - r->rho[0]) / (rr - rl);
        double sb = (r->rho[0] * rr - r->rho[1] * rl) / (rr - rl);
        return sa * g + sb;
    }
    return or_splint(r->x, r->rho, r->rho2, n, g);
}

/* SRotDensDeriv mc_poten.cc:572-596 */
double or_srotdens_deriv(or_sys* s, double g, int type) {
    rotdens_t* r = &s->rotd[type];
    int n = r->n;
    if (g > r->x[n - 1]) return 0.0;
    if (g < r->x[0]) {
        double rl = r->x[0], rr = r->x[1];
        double sa = (r->erot[1] - r->erot[0]) / (rr - rl);
        double sb = (r->erot[0] * rr - r->erot[1] * rl) / (rr - rl);
        return sa * g + sb;
    }
    return or_splint(r->x, r->erot, r->erot2, n, g);
}

/* SRotDensEsqrt mc_poten.cc:598-622 */
double or_srotdens_esqrt(or_sys* s, double g, int type) {
    rotdens_t* r = &s->rotd[type];
    int n = r->n;
    if (g > r->x[n - 1]) return 0.0;
    if (g < r->x[0]) {
        double rl = r->x[0], rr = r->x[1];
        double sa = (r->esq[1] - r->esq[0]) / (rr - rl);
        double sb = (r->esq[0] * rr - r->esq[1] * rl) / (rr - rl);
        return sa * g + sb;
    }
    return or_splint(r->x, r->esq, r->esq2, n, g);
}

/* ======================= rigid-body geometry (rotden.f, vcord.f) ======================= */
#define M3(m, i, j) (m)[((j) - 1) * 3 + ((i) - 1)] /* Fortran rotmat(i,j), column-major */

/* matpre rotden.f:136-163 */
void or_matpre(const double eul[3], double m[9]) {
    double phi = eul[0], theta = eul[1], chi = eul[2];
    double cp = cos(phi), sp = sin(phi), ct = cos(theta), st = sin(theta), ck = cos(chi), sk = sin(chi);
    M3(m, 1, 1) = cp * ct * ck - sp * sk;
    M3(m, 1, 2) = -cp * ct * sk - sp * ck;
    M3(m, 1, 3) = cp * st;
    M3(m, 2, 1) = sp * ct * ck + cp * sk;
    M3(m, 2, 2) = -sp * ct * sk + cp * ck;
    M3(m, 2, 3) = sp * st;
    M3(m, 3, 1) = -st * ck;
    M3(m, 3, 2) = st * sk;
    M3(m, 3, 3) = ct;
}

/* rottrn rotden.f:165-177: rsf = rcom + rotmat . rwf */
static void rottrn(const double m[9], const double rwf[3], double rsf[3], const double rcom[3]) {
    for (int i = 1; i <= 3; ++i) {
        double v = rcom[i - 1];
        for (int j = 1; j <= 3; ++j) v = v + M3(m, i, j) * rwf[j - 1];
        rsf[i - 1] = v;
    }
}

static double within(double v) { /* rotden.f:208-216 */
    if (v > 1.0) v = 1.0;
    if (v < -1.0) v = -1.0;
    return v;
}

/* deleul rotden.f:32-134 */
void or_deleul(const double e1[3], const double e2[3], double erel[3], int* istop) {
    const double small = 1.0e-08;
    double rotmat[9], rotma1[9], rotma2[9];
    *istop = 0;
    or_matpre(e2, rotmat);
    or_matpre(e1, rotma1);
    for (int i = 1; i <= 3; ++i)
        for (int j = 1; j <= 3; ++j) {
            double v = 0.0;
            for (int k = 1; k <= 3; ++k) v = v + M3(rotma1, k, i) * M3(rotmat, k, j);
            M3(rotma2, i, j) = v;
        }
    double cost = within(M3(rotma2, 3, 3));
    double theta2 = acos(cost);
    double sint = sin(theta2);
    double phi2, chi2, cchi, schi, cphi, sphi;
    if (fabs(1.0 - cost) < small) {
        phi2 = 0.0;
        cchi = within(M3(rotma2, 1, 1));
        schi = within(M3(rotma2, 2, 1));
        chi2 = (schi > 0.0) ? acos(cchi) : 2.0 * M_PI - acos(cchi);
    } else if (fabs(1.0 + cost) < small) {
        phi2 = 0.0;
        cchi = within(M3(rotma2, 2, 2));
        schi = within(M3(rotma2, 1, 2));
        chi2 = (schi > 0.0) ? acos(cchi) : 2.0 * M_PI - acos(cchi);
    } else {
        cphi = within(M3(rotma2, 1, 3) / sint);
        sphi = within(M3(rotma2, 2, 3) / sint);
        cchi = within(-M3(rotma2, 3, 1) / sint);
        schi = within(M3(rotma2, 3, 2) / sint);
        phi2 = (sphi > 0.0) ? acos(cphi) : 2.0 * M_PI - acos(cphi);
        chi2 = (schi > 0.0) ? acos(cchi) : 2.0 * M_PI - acos(cchi);
    }
    erel[0] = phi2;
    erel[1] = theta2;
    erel[2] = chi2;
}

/* rotpro rotpro_sub.f:1-64: corner value plus per-axis forward differences on the
   1-degree (theta, phi, chi) grid, index (itheta*361 + iphi)*361 + ichi */
void or_rotpro(double chi, double phi, double theta, double* rho, double* erot, double* esq,
               const double* rhoprp, const double* erotpr, const double* erotsq, int* jstop) {
    int ichi = (int)chi, iphi = (int)phi, itheta = (int)theta;
    if (ichi > 360 || ichi < 0) { ichi = 0; *jstop = 1; }
    if (iphi > 360 || iphi < 0) { iphi = 0; *jstop = 1; }
    if (itheta > 180 || itheta < 0) { itheta = 0; *jstop = 1; }
    long ind = ((long)itheta * 361 + iphi) * 361 + ichi;
    double rho0 = rhoprp[ind], erot0 = erotpr[ind], esq0 = erotsq[ind];
    double delchi = 0, delphi = 0, delthe = 0, delch2 = 0, delch3 = 0, delph2 = 0, delph3 = 0,
           delth2 = 0, delth3 = 0;
    if (ichi != 360) {
        ind = ((long)itheta * 361 + iphi) * 361 + ichi + 1;
        delchi = rhoprp[ind] - rho0; delch2 = erotpr[ind] - erot0; delch3 = erotsq[ind] - esq0;
    }
    if (iphi != 360) {
        ind = ((long)itheta * 361 + iphi + 1) * 361 + ichi;
        delphi = rhoprp[ind] - rho0; delph2 = erotpr[ind] - erot0; delph3 = erotsq[ind] - esq0;
    }
    if (itheta != 180) {
        ind = ((long)(itheta + 1) * 361 + iphi) * 361 + ichi;
        delthe = rhoprp[ind] - rho0; delth2 = erotpr[ind] - erot0; delth3 = erotsq[ind] - esq0;
    }
    *rho = rho0 + delchi * (chi - (double)ichi) + delphi * (phi - (double)iphi) + delthe * (theta - (double)itheta);
    *erot = erot0 + delch2 * (chi - (double)ichi) + delph2 * (phi - (double)iphi) + delth2 * (theta - (double)itheta);
    *esq = esq0 + delch3 * (chi - (double)ichi) + delph3 * (phi - (double)iphi) + delth3 * (theta - (double)itheta);
}

/* rotden rotden.f:1-31 */
void or_rotden(const double e1[3], const double e2[3], double erel[3], double* rho, double* erot,
               double* esq, const double* rhoprp, const double* erotpr, const double* erotsq, int* istop) {
    or_deleul(e1, e2, erel, istop);
    double phi = erel[0] * 180.0 / M_PI;
    double theta = erel[1] * 180.0 / M_PI;
    double chi = erel[2] * 180.0 / M_PI;
    int jstop = 0;
    or_rotpro(chi, phi, theta, rho, erot, esq, rhoprp, erotpr, erotsq, &jstop);
    if (jstop == 1) *istop = 1;
    *erot = *erot / OR_WNO2K;
    *esq = *esq / (OR_WNO2K * OR_WNO2K);
}

/* vcalc vcalc.f:1-65 (r in bohr, theta/chi in degrees) */
void or_vcalc(double r, double theta, double chi, double r0, double rmax, double rstep, int nrgrd,
              int nthgrd, int nchgrd, const double* vtable, double* vpes) {
    int maxrpt = nrgrd - 1, mxthpt = nthgrd - 1, mxchpt = nchgrd - 1;
    if (r < r0) r = r0;
    if (r > rmax) r = rmax;
    int ir = (int)((r - r0) / rstep);
    int ith = (int)theta;
    int ich = (int)chi;
    if (ich > mxchpt) ich = mxchpt;
    if (ich < 0) ich = 0;
    if (ith > mxthpt) ith = mxthpt;
    if (ith < 0) ith = 0;
    long ind = ((long)ir * nthgrd + ith) * nchgrd + ich;
    double v0 = vtable[ind];
    double gradr, delr, gradth, delth, gradch, delch;
    if (ir == maxrpt) { gradr = 0.0; delr = 0.0; }
    else {
        ind = ((long)(ir + 1) * nthgrd + ith) * nchgrd + ich;
        gradr = (vtable[ind] - v0) / rstep;
        delr = r - (r0 + ir * rstep);
    }
    if (ith == mxthpt) { gradth = 0.0; delth = 0.0; }
    else {
        ind = ((long)ir * nthgrd + (ith + 1)) * nchgrd + ich;
        gradth = vtable[ind] - v0;
        delth = theta - (double)ith;
    }
    if (ich == mxchpt) { gradch = 0.0; delch = 0.0; }
    else {
        ind = ((long)ir * nthgrd + ith) * nchgrd + ich + 1;
        gradch = vtable[ind] - v0;
        delch = chi - (double)ich;
    }
    *vpes = v0 + gradr * delr + gradth * delth + gradch * delch;
}

static double dotprd(const double a[3], const double b[3]) {
    double d = 0.0;
    for (int i = 0; i < 3; ++i) d = d + a[i] * b[i];
    return d;
}
static double dnorm(const double a[3]) { return sqrt(dotprd(a, a)); }

/* vcord vcord.f:1-98 with dotprd/dnorm/dotang vcord.f:155-188 */
void or_vcord(const double eulang[3], const double rcom[3], const double rpt[3], const double* vtable,
              int nrgrd, int nthgrd, int nchgrd, double rvmax, double rvmin, double rvstep,
              double* vpot, double* radret, double* theret, double* chiret,
              double hatx[3], double haty[3], double hatz[3], int ivcord) {
    const double small = 1.0e-08, bo2ang = 0.529177249;
    static const double unx[3] = {1, 0, 0}, uny[3] = {0, 1, 0}, unz[3] = {0, 0, 1}, origin[3] = {0, 0, 0};
    double rotmat[9], rh2com[3];
    or_matpre(eulang, rotmat);
    rottrn(rotmat, unx, hatx, origin);
    rottrn(rotmat, uny, haty, origin);
    rottrn(rotmat, unz, hatz, origin);
    for (int i = 0; i < 3; ++i) rh2com[i] = rpt[i] - rcom[i];
    if (ivcord == 1) return;
    double da = dotprd(hatz, rh2com) / (dnorm(hatz) * dnorm(rh2com));
    if (da > 1.0) da = 1.0;
    if (da < -1.0) da = -1.0;
    double thewff = acos(da);
    double chiwff;
    if (fabs(dotprd(rh2com, hatx)) < small) chiwff = M_PI / 2.0;
    else {
        double tanchi = dotprd(rh2com, haty) / dotprd(rh2com, hatx);
        chiwff = atan(fabs(tanchi));
    }
    double radwff = dnorm(rh2com);
    *radret = radwff;
    *theret = thewff;
    double rdx = dotprd(rh2com, hatx), rdy = dotprd(rh2com, haty);
    if (rdx >= 0.0 && rdy >= 0.0) *chiret = chiwff;
    else if (rdx < 0.0 && rdy >= 0.0) *chiret = M_PI - chiwff;
    else if (rdx < 0.0 && rdy < 0.0) *chiret = M_PI + chiwff;
    else *chiret = 2 * M_PI - chiwff;
    if (nchgrd == 91) {
    } else if (nchgrd == 181) {
        chiwff = *chiret;
        if (*chiret > M_PI) chiwff = 2 * M_PI - *chiret;
    } else if (nchgrd == 361) {
        chiwff = *chiret;
    }
    radwff = radwff / bo2ang;
    thewff = thewff * 180.0 / M_PI;
    chiwff = chiwff * 180.0 / M_PI;
    or_vcalc(radwff, thewff, chiwff, rvmin, rvmax, rvstep, nrgrd, nthgrd, nchgrd, vtable, vpot);
}

/* caleng caleng_tip4p_gg.f:2-186: rigid TIP4P dimer, body-frame sites in Angstrom */
void or_caleng(const double com1[3], const double com2[3], double* e, const double eul1[3],
               const double eul2[3]) {
    const double qm = -1.04, qh = 0.520;
    /* kcal2k is declared under implicit double precision(a-h,o-z) (caleng_tip4p_gg.f:22,34-38):
       its name starts with k, so the parameter is INTEGER and 503.218978939d0 truncates to 503 */
    const double br2ang = 0.52917721092, hr2k = 3.1577465e5, kcal2k = 503.0;
    static const double ROwf[3] = {0.0, 0.0, 0.06562}, RH1wf[3] = {0.7557, 0.0, -0.5223},
                        RH2wf[3] = {-0.7557, 0.0, -0.5223}, RMwf[3] = {0.0, 0.0, -0.08438};
    double m1[9], m2[9];
    double o1[3], mm1[3], h11[3], h21[3], o2[3], mm2[3], h12[3], h22[3];
    or_matpre(eul1, m1);
    rottrn(m1, ROwf, o1, com1);
    rottrn(m1, RMwf, mm1, com1);
    rottrn(m1, RH1wf, h11, com1);
    rottrn(m1, RH2wf, h21, com1);
    or_matpre(eul2, m2);
    rottrn(m2, ROwf, o2, com2);
    rottrn(m2, RMwf, mm2, com2);
    rottrn(m2, RH1wf, h12, com2);
    rottrn(m2, RH2wf, h22, com2);
    double roo = 0.0, rmm = 0.0;
    for (int i = 0; i < 3; ++i) {
        roo = roo + (o1[i] - o2[i]) * (o1[i] - o2[i]);
        rmm = rmm + (mm1[i] - mm2[i]) * (mm1[i] - mm2[i]);
    }
    rmm = sqrt(rmm);
    double roo4 = roo * roo, roo6 = roo4 * roo, roo12 = roo6 * roo6;
    roo = sqrt(roo);
    double A = 6.e5, B = 610.0;
    double v_o2lj = A / roo12 - B / roo6;
    double rhm[4] = {0, 0, 0, 0}, rhh[4] = {0, 0, 0, 0};
    for (int i = 0; i < 3; ++i) {
        rhm[0] = rhm[0] + (mm1[i] - h12[i]) * (mm1[i] - h12[i]);
        rhm[1] = rhm[1] + (mm1[i] - h22[i]) * (mm1[i] - h22[i]);
        rhm[2] = rhm[2] + (mm2[i] - h11[i]) * (mm2[i] - h11[i]);
        rhm[3] = rhm[3] + (mm2[i] - h21[i]) * (mm2[i] - h21[i]);
        rhh[0] = rhh[0] + (h11[i] - h12[i]) * (h11[i] - h12[i]);
        rhh[1] = rhh[1] + (h11[i] - h22[i]) * (h11[i] - h22[i]);
        rhh[2] = rhh[2] + (h21[i] - h12[i]) * (h21[i] - h12[i]);
        rhh[3] = rhh[3] + (h21[i] - h22[i]) * (h21[i] - h22[i]);
    }
    for (int k = 0; k < 4; ++k) { rhm[k] = sqrt(rhm[k]); rhh[k] = sqrt(rhh[k]); }
    double v_mh = qm * qh * (1.0 / rhm[0] + 1.0 / rhm[1] + 1.0 / rhm[2] + 1.0 / rhm[3]);
    double v_hh = qh * qh * (1.0 / rhh[0] + 1.0 / rhh[1] + 1.0 / rhh[2] + 1.0 / rhh[3]);
    double v_mm = qm * qm * (1.0 / rmm);
    *e = v_o2lj * kcal2k + (v_mh + v_mm + v_hh) * hr2k * br2ang;
}

/* ======================= linear-rotor density generator (linden.f) ======================= */
/* l cutoff linden.f:27-36: first l with exp(-tau B l(l+1)) < 1e-16 */
int or_linden_lmax(double tau, double bconst) {
    for (int l = 0; l <= 500; ++l)
        if (exp(-tau * bconst * l * (l + 1)) < 1.e-16) return l;
    return 500;
}

/* exarho linden.f:88-122 with LGND linden.f:180-193 */
double or_linden_rho(double cost, double tau, double bconst, int lmax, int iodevn, int nslice,
                     double* erot, double* erotsq) {
    const double pi = 3.14159265358979323846, boltz = 0.69503476;
    double pl[501];
    pl[0] = 1.0;
    pl[1] = cost;
    for (int L = 1; L <= lmax - 1; ++L) pl[L + 1] = ((2.0 * L + 1) * cost * pl[L] - L * pl[L - 1]) / (L + 1);
    double rho = 0.0, er = 0.0, es = 0.0;
    for (int l = 0; l <= lmax; ++l) {
        if ((l % 2) == iodevn || iodevn == -1) {
            double tmp = (double)(2 * l + 1) * pl[l];
            tmp = tmp * exp(-tau * bconst * (double)(l * (l + 1)));
            rho = rho + tmp;
            er = er + tmp * l * (l + 1) * bconst;
            es = es + tmp * l * (l + 1) * bconst * l * (l + 1) * bconst;
        }
    }
    er = er / (nslice * boltz);
    es = es / pow(nslice * boltz, 2.0);
    rho = rho / (4.0 * pi);
    er = er / (4.0 * pi);
    es = es / (4.0 * pi);
    if (erot) *erot = er;
    if (erotsq) *erotsq = es;
    return rho;
}

/* ======================= potential energies (mc_piqmc.cc) ======================= */
static double pair_dispatch(or_sys* s, int atom0, int atom1, const double* posbase, int it,
                            const double dr[3], double r) {
    int type0 = s->mctype[atom0], type1 = s->mctype[atom1];
    int offset0 = s->P * atom0, offset1 = s->P * atom1;
    int mol0 = s->type[type0].molecule, mol1 = s->type[type1].molecule;
    int NP = s->natoms * s->P;
    if (mol0 == 1 || mol1 == 1) { /* mc_piqmc.cc:1847-1870 */
        int sgn = 1, tm = offset1 + it / s->RotRatio, typep = type1;
        if (mol0 == 1) { sgn = -1; tm = offset0 + it / s->RotRatio; typep = type0; }
        double cost = 0.0;
        for (int id = 0; id < 3; ++id) cost += CS(s, id, tm) * dr[id];
        cost /= r;
        cost *= sgn;
        return or_lpot2d(s, r, cost, typep);
    } else if ((mol0 == 2 || mol1 == 2) && mol0 != mol1) { /* :1872-1911 */
        double rcom[3], rpt[3], eul[3], v, rr, th, ch, hx[3], hy[3], hz[3];
        int t0 = offset0 + it, t1 = offset1 + it, tm;
        if (mol0 == 2) {
            tm = offset0 + it / s->RotRatio;
            for (int id = 0; id < 3; ++id) { rcom[id] = posbase[id * NP + t0]; rpt[id] = C(s, id, t1); }
        } else {
            tm = offset1 + it / s->RotRatio;
            for (int id = 0; id < 3; ++id) { rpt[id] = posbase[id * NP + t0]; rcom[id] = C(s, id, t1); }
        }
        eul[PHI] = AN(s, PHI, tm);
        eul[CTH] = acos(AN(s, CTH, tm));
        eul[CHI] = AN(s, CHI, tm);
        or_vcord(eul, rcom, rpt, s->vtable, s->nr3, s->nth3, s->nch3, s->rvmax, s->rvmin, s->rvstep,
                 &v, &rr, &th, &ch, hx, hy, hz, 0);
        return v;
    } else if (mol0 == 2 && mol1 == 2 && s->type[s->imtype].numb > 1) { /* :1924-1955 */
        double c1[3], c2[3], e1[3], e2[3], e;
        int t0 = offset0 + it, t1 = offset1 + it;
        for (int id = 0; id < 3; ++id) { c1[id] = posbase[id * NP + t0]; c2[id] = C(s, id, t1); }
        int tm0 = offset0 + it / s->RotRatio, tm1 = offset1 + it / s->RotRatio;
        e1[PHI] = AN(s, PHI, tm0); e1[CTH] = acos(AN(s, CTH, tm0)); e1[CHI] = AN(s, CHI, tm0);
        e2[PHI] = AN(s, PHI, tm1); e2[CTH] = acos(AN(s, CTH, tm1)); e2[CHI] = AN(s, CHI, tm1);
        or_caleng(c1, c2, &e, e1, e2);
        return e;
    }
    return or_spot1d(s, r, type1); /* :1957-1958 */
}

/* PotEnergy(int atom0, double **pos, int it) mc_piqmc.cc:1796-1965 */
double or_pot_energy_slice(or_sys* s, int atom0, const double* pos, int it) {
    const double* pb = pos ? pos : s->coords;
    int NP = s->natoms * s->P;
    int offset0 = s->P * atom0;
    double spot = 0.0;
    for (int atom1 = 0; atom1 < s->natoms; atom1++) {
        if (atom1 == atom0) continue;
        int offset1 = s->P * atom1;
        int t0 = offset0 + it, t1 = offset1 + it;
        double dr[3], dr2 = 0.0;
        for (int id = 0; id < 3; id++) {
            dr[id] = pb[id * NP + t0] - C(s, id, t1);
            if (s->minimage) dr[id] -= (s->box[id] * rint(dr[id] / s->box[id]));
            dr2 += dr[id] * dr[id];
        }
        double r = sqrt(dr2);
        spot += pair_dispatch(s, atom0, atom1, pb, it, dr, r);
    }
    return spot;
}

/* PotEnergy(int atom0, double **pos) mc_piqmc.cc:1201-1383 (partner-major, slice sum) */
double or_pot_energy_path(or_sys* s, int atom0, const double* pos) {
    const double* pb = pos ? pos : s->coords;
    int NP = s->natoms * s->P;
    int offset0 = s->P * atom0;
    double spot = 0.0;
    for (int atom1 = 0; atom1 < s->natoms; atom1++) {
        if (atom1 == atom0) continue;
        int offset1 = s->P * atom1;
        double spot_pair = 0.0;
        for (int it = 0; it < s->P; it++) {
            int t0 = offset0 + it, t1 = offset1 + it;
            double dr[3], dr2 = 0.0;
            for (int id = 0; id < 3; id++) {
                dr[id] = pb[id * NP + t0] - C(s, id, t1);
                if (s->minimage) dr[id] -= (s->box[id] * rint(dr[id] / s->box[id]));
                dr2 += dr[id] * dr[id];
            }
            double r = sqrt(dr2);
            spot_pair += pair_dispatch(s, atom0, atom1, pb, it, dr, r);
        }
        spot += spot_pair;
    }
    return spot;
}

/* PotRotEnergy mc_piqmc.cc:1967-2044 */
double or_pot_rot_energy(or_sys* s, int atom0, const double* cosine, int it) {
    const double* cb = cosine ? cosine : s->cosine;
    int NP = s->natoms * s->P;
    int type0 = s->mctype[atom0];
    int offset0 = atom0 * s->P;
    double spot = 0.0;
    for (int atom1 = 0; atom1 < s->natoms; atom1++) {
        if (atom1 == atom0) continue;
        int offset1 = atom1 * s->P;
        int t0 = offset0 + it, t1 = offset1 + it;
        double dr[3], dr2 = 0.0;
        for (int id = 0; id < 3; id++) {
            dr[id] = C(s, id, t0) - C(s, id, t1);
            dr2 += dr[id] * dr[id];
        }
        double r = sqrt(dr2);
        int tm = offset0 + it / s->RotRatio;
        double cost = 0.0;
        for (int id = 0; id < 3; id++) cost += cb[id * NP + tm] * dr[id];
        cost /= r;
        cost *= -1;
        spot += or_lpot2d(s, r, cost, type0);
    }
    return spot;
}

/* PotRotE3D mc_piqmc.cc:2046-2151 */
double or_pot_rot_e3d(or_sys* s, int atom0, const double eul[3], int it) {
    int offset0 = atom0 * s->P;
    double spot = 0.0;
    for (int atom1 = 0; atom1 < s->natoms; atom1++) {
        if (atom1 == atom0) continue;
        int offset1 = atom1 * s->P;
        int type1 = s->mctype[atom1];
        int t0 = offset0 + it, t1 = offset1 + it;
        if (type1 != s->imtype) {
            double rcom[3], rpt[3], v, rr, th, ch, hx[3], hy[3], hz[3];
            for (int id = 0; id < 3; id++) { rcom[id] = C(s, id, t0); rpt[id] = C(s, id, t1); }
            or_vcord(eul, rcom, rpt, s->vtable, s->nr3, s->nth3, s->nch3, s->rvmax, s->rvmin,
                     s->rvstep, &v, &rr, &th, &ch, hx, hy, hz, 0);
            spot += v;
        } else {
            double c1[3], c2[3], e2[3], e;
            for (int id = 0; id < 3; id++) { c1[id] = C(s, id, t0); c2[id] = C(s, id, t1); }
            int tm1 = offset1 + it / s->RotRatio;
            e2[PHI] = AN(s, PHI, tm1); e2[CTH] = acos(AN(s, CTH, tm1)); e2[CHI] = AN(s, CHI, tm1);
            or_caleng(c1, c2, &e, eul, e2);
            spot += e;
        }
    }
    return spot;
}

/* ======================= moves ======================= */
/* MCMolecularMove mc_piqmc.cc:54-102.  Draws: ctr3=0 -> disp x,y; ctr3=1 -> disp z, accept */
int or_molecular_move(or_sys* s, int type, uint32_t pass) {
    type_t* ty = &s->type[type];
    int NP = s->natoms * s->P;
    for (int atom = 0; atom < ty->numb; atom++) {
        int offset = ty->offset + s->P * atom;
        int gatom = offset / s->P;
        double disp[3];
        for (int id = 0; id < 3; id++) disp[id] = ty->mcstep * (u_centroid(s, pass, gatom, id) - 0.5);
        for (int id = 0; id < 3; id++)
            for (int it = 0; it < s->P; it++) NW(s, id, offset + it) = C(s, id, offset + it) + disp[id];
        double deltav = or_pot_energy_path(s, gatom, s->newc) - or_pot_energy_path(s, gatom, NULL);
        int acc = 0;
        if (deltav < 0.0) acc = 1;
        else if (exp(-deltav * s->tau) > u_centroid(s, pass, gatom, 3)) acc = 1;
        s->total[type][OR_MCMOLEC] += 1.0;
        if (acc) {
            s->accep[type][OR_MCMOLEC] += 1.0;
            for (int id = 0; id < 3; id++)
                for (int it = 0; it < s->P; it++) C(s, id, offset + it) = NW(s, id, offset + it);
        }
    }
    (void)NP;
    return 0;
}

/* MCBisectionMove mc_piqmc.cc:194-294, for one atom.  Draws: Gaussian for midpoint j,
   coordinate id at level l: ctr3 = l<<24 | j<<8 | id; level acceptance: j = 0xFFFF. */
int or_bisection_atom(or_sys* s, int type, int atom, int time, uint32_t pass) {
    type_t* ty = &s->type[type];
    int P = s->P;
    double mclambda = ty->lambda;
    int mclevels = ty->levels, seg_size = ty->mlsegm;
    int offset = ty->offset + P * atom;
    int gatom = offset / P;
    int pit = (time + seg_size) % P;
    for (int id = 0; id < 3; id++) {
        NW(s, id, offset + time) = C(s, id, offset + time);
        NW(s, id, offset + pit) = C(s, id, offset + pit);
    }
    double bnorm = 1.0 / (mclambda * s->tau);
    int acc = 0;
    double pot0 = 0.0, pot1 = 0.0;
    for (int level = 0; level < mclevels; level++) {
        int lss = (int)pow(2.0, (mclevels - level));
        double bkin_norm = bnorm / (double)lss;
        double bpot_norm = s->tau * (double)(lss / 2);
        pot1 = pot0;
        pot0 = 0.0;
        int t0, t1, t2 = 0, j = 0;
        do {
            t0 = t2;
            t2 = t0 + lss;
            t1 = (t0 + t2) / 2;
            int pt0 = (time + t0) % P, pt1 = (time + t1) % P, pt2 = (time + t2) % P;
            for (int id = 0; id < 3; id++) {
                NW(s, id, offset + pt1) = 0.5 * (NW(s, id, offset + pt0) + NW(s, id, offset + pt2));
                NW(s, id, offset + pt1) += g_bisect(s, pass, gatom, time, level, j, id, bkin_norm);
            }
            pot0 += (or_pot_energy_slice(s, gatom, s->newc, pt1) - or_pot_energy_slice(s, gatom, NULL, pt1));
            if (t0 != 0)
                pot0 += (or_pot_energy_slice(s, gatom, s->newc, pt0) - or_pot_energy_slice(s, gatom, NULL, pt0));
            ++j;
        } while (t2 < seg_size);
        double deltav = (pot0 - 2.0 * pot1);
        deltav *= bpot_norm;
        acc = 0;
        if (deltav < 0.0) acc = 1;
        else if (exp(-deltav) > u_bisect_accept(s, pass, gatom, time, level)) acc = 1;
        if (!acc) break;
    }
    s->total[type][OR_MCMULTI] += 1.0;
    if (acc) {
        s->accep[type][OR_MCMULTI] += 1.0;
        for (int id = 0; id < 3; id++)
            for (int it = time; it <= time + seg_size; it++) {
                int p = it % P;
                C(s, id, offset + p) = NW(s, id, offset + p);
            }
    }
    return acc;
}

int or_bisection_move(or_sys* s, int type, int time, uint32_t pass) {
    for (int atom = 0; atom < s->type[type].numb; atom++) or_bisection_atom(s, type, atom, time, pass);
    return 0;
}

/* MCRotLinStep mc_piqmc.cc:781-936 (RotDenType 0) */
static int rot_lin_step(or_sys* s, int it1, int offset, int gatom, int type, double step,
                        double rand1, double rand2, double rand3) {
    int Prot = s->Prot;
    int NP = s->natoms * s->P;
    int it0 = it1 - 1, it2 = it1 + 1;
    if (it0 < 0) it0 += Prot;
    if (it2 >= Prot) it2 -= Prot;
    int t0 = offset + it0, t1 = offset + it1, t2 = offset + it2;
    double cost = AN(s, CTH, t1), phi = AN(s, PHI, t1);
    cost += (step * (rand1 - 0.5));
    phi += (step * (rand2 - 0.5));
    if (cost > 1.0) cost = 2.0 - cost;
    if (cost < -1.0) cost = -2.0 - cost;
    double sint = sqrt(1.0 - cost * cost);
    NW(s, 0, t1) = sint * cos(phi);
    NW(s, 1, t1) = sint * sin(phi);
    NW(s, 2, t1) = cost;
    double p0 = 0.0, p1 = 0.0;
    for (int id = 0; id < 3; id++) {
        p0 += (CS(s, id, t0) * CS(s, id, t1));
        p1 += (CS(s, id, t1) * CS(s, id, t2));
    }
    double dens_old = or_srotdens(s, p0, type) * or_srotdens(s, p1, type);
    if (fabs(dens_old) < OR_RZERO) dens_old = 0.0;
    if (dens_old < 0.0) { fail(s, "Rotational Moves: Negative rot density"); return -1; }
    double pot_old = 0.0;
    int itr0 = it1 * s->RotRatio, itr1 = itr0 + s->RotRatio;
    for (int it = itr0; it < itr1; it++) pot_old += or_pot_rot_energy(s, gatom, s->cosine, it);
    p0 = 0.0;
    p1 = 0.0;
    for (int id = 0; id < 3; id++) {
        p0 += (CS(s, id, t0) * NW(s, id, t1));
        p1 += (NW(s, id, t1) * CS(s, id, t2));
    }
    double dens_new = or_srotdens(s, p0, type) * or_srotdens(s, p1, type);
    if (fabs(dens_new) < OR_RZERO) dens_new = 0.0;
    if (dens_new < 0.0) { fail(s, "Rotational Moves: Negative rot density"); return -1; }
    double pot_new = 0.0;
    for (int it = itr0; it < itr1; it++) pot_new += or_pot_rot_energy(s, gatom, s->newc, it);
    double rd;
    if (dens_old > OR_RZERO) rd = dens_new / dens_old;
    else rd = 1.0;
    rd *= exp(-s->tau * (pot_new - pot_old));
    int acc = 0;
    if (rd > 1.0) acc = 1;
    else if (rd > rand3) acc = 1;
    s->total[type][OR_MCROTAT] += 1.0;
    if (acc) {
        s->accep[type][OR_MCROTAT] += 1.0;
        AN(s, CTH, t1) = cost;
        AN(s, PHI, t1) = phi;
        for (int id = 0; id < 3; id++) CS(s, id, t1) = NW(s, id, t1);
    }
    (void)NP;
    return acc;
}

/* MCRotationsMove mc_piqmc.cc:421-517: even rotor slices, then odd.
   Draws for slice k of sweep j: ctr3=0 -> rand1, rand2; ctr3=1 -> rand3 */
int or_rotations_move(or_sys* s, int type, uint32_t pass, uint32_t sweep) {
    int offset = s->type[type].offset;
    int gatom = offset / s->P;
    double step = s->rtstep;
    for (int parity = 0; parity < 2; ++parity)
        for (int itrot = parity; itrot < s->Prot; itrot += 2) {
            double r[4];
            u_rotor(s, K_ROTLIN, pass, sweep, gatom, itrot, 3, r);
            if (rot_lin_step(s, itrot, offset, gatom, type, step, r[0], r[1], r[2]) < 0) return -1;
        }
    return 0;
}

/* MCRot3Dstep mc_piqmc.cc:938-1199 (RotDenType 0) */
static int rot3d_step(or_sys* s, int it1, int offset, int gatom, int type, double step,
                      double rand1, double rand2, double rand3, double rand4) {
    int Prot = s->Prot;
    int it0 = it1 - 1, it2 = it1 + 1;
    if (it0 < 0) it0 += Prot;
    if (it2 >= Prot) it2 -= Prot;
    int t0 = offset + it0, t1 = offset + it1, t2 = offset + it2;
    double cost = AN(s, CTH, t1), phi = AN(s, PHI, t1), chi = AN(s, CHI, t1);
    cost += (step * (rand1 - 0.5));
    phi += 2.0 * M_PI * (step * (rand2 - 0.5));
    chi += 2.0 * M_PI * (step * (rand3 - 0.5));
    if (phi < 0.0) phi = 2.0 * M_PI + phi;
    if (chi < 0.0) chi = 2.0 * M_PI + chi;
    phi = fmod(phi, 2.0 * M_PI);
    chi = fmod(chi, 2.0 * M_PI);
    if (cost > 1.0) cost = 2.0 - cost;
    if (cost < -1.0) cost = -2.0 - cost;
    NW(s, PHI, t1) = phi;
    NW(s, CHI, t1) = chi;
    NW(s, CTH, t1) = cost;
    double rho, erot, esq, e1[3], e2[3], er[3];
    int istop = 0;
    e1[0] = AN(s, PHI, t0); e1[1] = acos(AN(s, CTH, t0)); e1[2] = AN(s, CHI, t0);
    e2[0] = AN(s, PHI, t1); e2[1] = acos(AN(s, CTH, t1)); e2[2] = AN(s, CHI, t1);
    or_rotden(e1, e2, er, &rho, &erot, &esq, s->rhoprp, s->erotpr, s->erotsq, &istop);
    if (istop == 1) { fail(s, "large matrix test error"); return -1; }
    double dens_old = rho;
    e1[0] = AN(s, PHI, t1); e1[1] = acos(AN(s, CTH, t1)); e1[2] = AN(s, CHI, t1);
    e2[0] = AN(s, PHI, t2); e2[1] = acos(AN(s, CTH, t2)); e2[2] = AN(s, CHI, t2);
    istop = 0;
    or_rotden(e1, e2, er, &rho, &erot, &esq, s->rhoprp, s->erotpr, s->erotsq, &istop);
    if (istop == 1) { fail(s, "large matrix test error"); return -1; }
    dens_old = dens_old * rho;
    if (fabs(dens_old) < OR_RZERO) dens_old = 0.0;
    if (dens_old < 0.0) dens_old = fabs(dens_old);
    double pot_old = 0.0;
    int itr0 = it1 * s->RotRatio, itr1 = itr0 + s->RotRatio;
    for (int it = itr0; it < itr1; it++) pot_old += or_pot_rot_e3d(s, gatom, e1, it);
    e1[0] = AN(s, PHI, t0); e1[1] = acos(AN(s, CTH, t0)); e1[2] = AN(s, CHI, t0);
    e2[0] = NW(s, PHI, t1); e2[1] = acos(NW(s, CTH, t1)); e2[2] = NW(s, CHI, t1);
    istop = 0;
    or_rotden(e1, e2, er, &rho, &erot, &esq, s->rhoprp, s->erotpr, s->erotsq, &istop);
    if (istop == 1) { fail(s, "large matrix test error"); return -1; }
    double dens_new = rho;
    e1[0] = NW(s, PHI, t1); e1[1] = acos(NW(s, CTH, t1)); e1[2] = NW(s, CHI, t1);
    e2[0] = AN(s, PHI, t2); e2[1] = acos(AN(s, CTH, t2)); e2[2] = AN(s, CHI, t2);
    istop = 0;
    or_rotden(e1, e2, er, &rho, &erot, &esq, s->rhoprp, s->erotpr, s->erotsq, &istop);
    if (istop == 1) { fail(s, "large matrix test error"); return -1; }
    dens_new = dens_new * rho;
    if (fabs(dens_new) < OR_RZERO) dens_new = 0.0;
    if (dens_new < 0.0) dens_new = fabs(dens_new);
    double pot_new = 0.0;
    for (int it = itr0; it < itr1; it++) pot_new += or_pot_rot_e3d(s, gatom, e1, it);
    double rd;
    if (dens_old > OR_RZERO) rd = dens_new / dens_old;
    else rd = 1.0;
    rd *= exp(-s->tau * (pot_new - pot_old));
    int acc = 0;
    if (rd > 1.0) acc = 1;
    else if (rd > rand4) acc = 1;
    s->total[type][OR_MCROTAT] += 1.0;
    if (acc) {
        s->accep[type][OR_MCROTAT] += 1.0;
        AN(s, CTH, t1) = cost;
        AN(s, PHI, t1) = phi;
        AN(s, CHI, t1) = chi;
        double sint = sqrt(1.0 - cost * cost);
        CS(s, 0, t1) = sint * cos(phi);
        CS(s, 1, t1) = sint * sin(phi);
        CS(s, 2, t1) = cost;
    }
    return acc;
}

/* MCRotations3D mc_piqmc.cc:692-779: even slices (all rotors), then odd.
   Draws for (rotor, slice k, sweep j): ctr3=0 -> rand1, rand2; ctr3=1 -> rand3, rand4 */
int or_rotations_3d(or_sys* s, int type, uint32_t pass, uint32_t sweep) {
    double step = s->rtstep;
    for (int parity = 0; parity < 2; ++parity)
        for (int itrot = parity; itrot < s->Prot; itrot += 2)
            for (int atom0 = 0; atom0 < s->type[type].numb; atom0++) {
                int offset = s->type[type].offset + s->P * atom0;
                int gatom = offset / s->P;
                double r[4];
                u_rotor(s, K_ROT3D, pass, sweep, gatom, itrot, 4, r);
                if (rot3d_step(s, itrot, offset, gatom, type, step, r[0], r[1], r[2], r[3]) < 0) return -1;
            }
    return 0;
}

/* ======================= schedule ======================= */
static int rot_sweep(or_sys* s, int type, uint32_t pass, uint32_t sweep) {
    if (type != s->imtype) return 0;
    if (s->type[type].molecule == 1) return or_rotations_move(s, type, pass, sweep);
    if (s->type[type].molecule == 2) return or_rotations_3d(s, type, pass, sweep);
    return 0;
}

/* One pass.
   OR_ORDER_REFERENCE: mc_main.cc:348-381 with PIMCPass mc_main.cc:510-522:
     for t in 0..P-1, for each type: [t==0: centroid], bisection at t, [rotor: sweep t].
   OR_ORDER_SEGMENTS: the same set of moves with the same random numbers, grouped so
     that segments [s, s+S] (S = max segment size, P % S == 0) with equal s mod S move
     together: for u in 0..S-1, for each type: [u==0: centroid], bisection of every
     atom on every segment s = u + kS; then sweeps j = u + kS.  Segments of one
     offset share only fixed end points, so they are independent. */
int or_pass(or_sys* s, uint32_t pass, int order) {
    int P = s->P;
    if (order == OR_ORDER_REFERENCE) {
        for (int t = 0; t < P; t++)
            for (int type = 0; type < s->ntypes; type++) {
                if (t == 0) or_molecular_move(s, type, pass);
                or_bisection_move(s, type, t, pass);
                if (rot_sweep(s, type, pass, (uint32_t)t) < 0) return -1;
            }
        return 0;
    }
    int S = 1;
    for (int type = 0; type < s->ntypes; ++type)
        if (s->type[type].mlsegm > S) S = s->type[type].mlsegm;
    if (P % S != 0) return fail(s, "segment order needs P divisible by the segment size");
    for (int u = 0; u < S; ++u) {
        for (int type = 0; type < s->ntypes; type++) {
            if (u == 0) or_molecular_move(s, type, pass);
            for (int atom = 0; atom < s->type[type].numb; ++atom)
                for (int k = 0; k < P / S; ++k) or_bisection_atom(s, type, atom, u + k * S, pass);
        }
        for (int type = 0; type < s->ntypes; type++)
            for (int k = 0; k < P / S; ++k)
                if (rot_sweep(s, type, pass, (uint32_t)(u + k * S)) < 0) return -1;
    }
    return 0;
}

/* ======================= estimators (mc_estim.cc) ======================= */
/* GetKinEnergy mc_estim.cc:876-937 (Boltzmann: PIndex identity) */
double or_kin_energy(or_sys* s) {
    double r2avr = 0.0;
    int numb = 0;
    for (int atom = 0; atom < s->natoms; atom++) {
        numb++;
        int type = s->mctype[atom];
        int offset0 = s->P * atom;
        double sum = 0.0;
        for (int it = 0; it < s->P; it++) {
            int t0 = offset0 + it, t1 = offset0 + (it + 1) % s->P;
            for (int dim = 0; dim < 3; dim++) {
                double dr = C(s, dim, t0) - C(s, dim, t1);
                if (s->minimage) dr -= (s->box[dim] * rint(dr / s->box[dim]));
                sum += (dr * dr);
            }
        }
        r2avr += (sum / (4.0 * s->beta * s->type[type].lambda));
    }
    return (double)s->P * s->T * (0.5 * (double)(3 * numb) - r2avr);
}

/* GetPotEnergy_Densities mc_estim.cc:500-687 with bin_1Ddensity/bin_2Ddensity :1233-1286.
   gr1d[300] (atom-atom), gr2d[300*50] (atom around linear rotor) for one atom type. */
double or_pot_energy_densities(or_sys* s, double* gr1d, double* gr2d) {
    const double dradius = OR_MAX_RADIUS / (double)OR_MC_BINSR;
    const double dtheta = M_PI / (double)(OR_MC_BINST - 1);
    double spot = 0.0;
    for (int atom0 = 0; atom0 < s->natoms - 1; atom0++)
        for (int atom1 = atom0 + 1; atom1 < s->natoms; atom1++) {
            int type0 = s->mctype[atom0], type1 = s->mctype[atom1];
            int mol0 = s->type[type0].molecule, mol1 = s->type[type1].molecule;
            int offset0 = s->P * atom0, offset1 = s->P * atom1;
            double spot_pair = 0.0;
            for (int it = 0; it < s->P; it++) {
                int t0 = offset0 + it, t1 = offset1 + it;
                double dr[3], dr2 = 0.0;
                for (int id = 0; id < 3; id++) {
                    dr[id] = C(s, id, t0) - C(s, id, t1);
                    if (s->minimage) dr[id] -= (s->box[id] * rint(dr[id] / s->box[id]));
                    dr2 += dr[id] * dr[id];
                }
                double r = sqrt(dr2);
                if (mol0 == 1 || mol1 == 1) {
                    int sgn = 1, tm = offset1 + it / s->RotRatio, typep = type1;
                    if (mol0 == 1) { sgn = -1; tm = offset0 + it / s->RotRatio; typep = type0; }
                    double cost = 0.0;
                    for (int id = 0; id < 3; id++) cost += CS(s, id, tm) * dr[id];
                    cost /= r;
                    cost *= sgn;
                    if (gr2d) {
                        int br = (int)floor(r / dradius);
                        if (br < OR_MC_BINSR && br >= 0) {
                            int bt = (int)floor(acos(cost) / dtheta);
                            if (bt < OR_MC_BINST && bt >= 0) gr2d[br * OR_MC_BINST + bt] += 1.0;
                        }
                    }
                    spot_pair += or_lpot2d(s, r, cost, typep);
                } else if ((mol0 == 2 || mol1 == 2) && mol0 != mol1) {
                    double rcom[3], rpt[3], eul[3], v, rr, th, ch, hx[3], hy[3], hz[3];
                    int tm;
                    if (mol0 == 2) {
                        tm = offset0 + it / s->RotRatio;
                        for (int id = 0; id < 3; id++) { rcom[id] = C(s, id, t0); rpt[id] = C(s, id, t1); }
                    } else {
                        tm = offset1 + it / s->RotRatio;
                        for (int id = 0; id < 3; id++) { rpt[id] = C(s, id, t0); rcom[id] = C(s, id, t1); }
                    }
                    eul[PHI] = AN(s, PHI, tm); eul[CTH] = acos(AN(s, CTH, tm)); eul[CHI] = AN(s, CHI, tm);
                    or_vcord(eul, rcom, rpt, s->vtable, s->nr3, s->nth3, s->nch3, s->rvmax, s->rvmin,
                             s->rvstep, &v, &rr, &th, &ch, hx, hy, hz, 0);
                    spot_pair += v;
                } else if (mol0 == 2 && mol1 == 2 && s->type[s->imtype].numb > 1) {
                    double c1[3], c2[3], e1[3], e2[3], e;
                    for (int id = 0; id < 3; id++) { c1[id] = C(s, id, t0); c2[id] = C(s, id, t1); }
                    int tm0 = offset0 + it / s->RotRatio, tm1 = offset1 + it / s->RotRatio;
                    e1[PHI] = AN(s, PHI, tm0); e1[CTH] = acos(AN(s, CTH, tm0)); e1[CHI] = AN(s, CHI, tm0);
                    e2[PHI] = AN(s, PHI, tm1); e2[CTH] = acos(AN(s, CTH, tm1)); e2[CHI] = AN(s, CHI, tm1);
                    or_caleng(c1, c2, &e, e1, e2);
                    spot_pair += e;
                } else if (type0 == type1 && mol0 == 0) {
                    if (gr1d) {
                        int br = (int)floor(r / dradius);
                        if (br < OR_MC_BINSR && br >= 0) gr1d[br] += 1.0;
                    }
                    spot_pair += or_spot1d(s, r, type1);
                }
            }
            spot += spot_pair;
        }
    return spot / (double)s->P;
}

/* GetRotEnergy mc_estim.cc:939-987 (RotDenType 0) */
double or_rot_energy(or_sys* s, double* erotsq, double* erot_termsq) {
    int type = s->imtype;
    int offset = s->type[type].offset;
    double srot = 0.0, esq = 0.0, term = 0.0;
    for (int it0 = 0; it0 < s->Prot; it0++) {
        int t0 = offset + it0, t1 = offset + (it0 + 1) % s->Prot;
        double p0 = 0.0;
        for (int id = 0; id < 3; id++) p0 += (CS(s, id, t0) * CS(s, id, t1));
        double rdens = or_srotdens(s, p0, type);
        double d = or_srotdens_deriv(s, p0, type);
        if (fabs(rdens) > OR_RZERO) srot += (d / rdens);
        term += (d / rdens) * (d / rdens);
        esq += or_srotdens_esqrt(s, p0, type) / rdens;
    }
    if (erotsq) *erotsq = esq;
    if (erot_termsq) *erot_termsq = term;
    return srot;
}

/* GetRotE3D mc_estim.cc:989-1096 (RotDenType 0, RNskip 1).  The reference advances
   offset cumulatively over rotors (offset += NumbTimes*atom), which is right for up to
   two rotors; more are rejected here. */
double or_rot_e3d(or_sys* s, double* erotsq, double* erot_termsq) {
    int type = s->imtype;
    int offset = s->type[type].offset;
    double e3d = 0.0, esq_all = 0.0, term_all = 0.0;
    if (s->type[type].numb > 2) { fail(s, "GetRotE3D restatement supports at most two rotors"); return NAN; }
    for (int atom = 0; atom < s->type[type].numb; atom++) {
        offset += (s->P * atom);
        double srot = 0.0, sesq = 0.0, se_termsq = 0.0;
        for (int it0 = 0; it0 < s->Prot; it0++) {
            int t0 = offset + it0, t1 = offset + (it0 + 1) % s->Prot;
            double e1[3], e2[3], er[3], rho, erot, esq;
            int istop = 0;
            e1[0] = AN(s, PHI, t0); e1[1] = acos(AN(s, CTH, t0)); e1[2] = AN(s, CHI, t0);
            e2[0] = AN(s, PHI, t1); e2[1] = acos(AN(s, CTH, t1)); e2[2] = AN(s, CHI, t1);
            or_rotden(e1, e2, er, &rho, &erot, &esq, s->rhoprp, s->erotpr, s->erotsq, &istop);
            srot += erot;
            sesq += esq;
            se_termsq += erot * erot;
        }
        double n = (double)s->Prot;
        srot = srot / n;
        sesq = sesq / (n * n);
        se_termsq = se_termsq / (n * n);
        e3d += srot;
        esq_all += sesq;
        term_all += se_termsq;
    }
    if (erotsq) *erotsq = esq_all;
    if (erot_termsq) *erot_termsq = term_all;
    return e3d;
}

/* The bins around a non-linear rotor of one estimator call.  gr3d[(ir*50 + it)*100 + ic]:
   bin_3Ddensity (mc_estim.cc:1252-1273) of vcord's (radret, theret, chiret) for every
   atom / rotor pair and slice, as GetPotEnergy_Densities makes them (mc_estim.cc:582-641);
   the atom partners are of the one atom type.  rel[0..49], rel[50..149], rel[150..249]:
   the relative Euler angles theta, phi, chi of neighbouring rotor slices as GetRotE3D bins
   them (mc_estim.cc:1046-1058, RotDenType 0 so each adds 1), with its rotor offsets.
   Either array may be NULL; bins are added to what the arrays hold.  Returns -1 for more
   than two rotors, as or_rot_e3d does. */
int or_nl_densities(or_sys* s, double* gr3d, double* rel) {
    const double dradius = OR_MAX_RADIUS / (double)OR_MC_BINSR;
    const double dtheta = M_PI / (double)(OR_MC_BINST - 1);
    const double dchi = 2.0 * M_PI / (double)(OR_MC_BINSC - 1);
    if (s->imtype < 0 || s->type[s->imtype].molecule != 2) return 0;
    if (s->type[s->imtype].numb > 2) { fail(s, "GetRotE3D restatement supports at most two rotors"); return -1; }
    if (gr3d)
        for (int atom0 = 0; atom0 < s->natoms - 1; atom0++)
            for (int atom1 = atom0 + 1; atom1 < s->natoms; atom1++) {
                int type0 = s->mctype[atom0], type1 = s->mctype[atom1];
                int mol0 = s->type[type0].molecule, mol1 = s->type[type1].molecule;
                if (!((mol0 == 2 || mol1 == 2) && mol0 != mol1)) continue;
                int offset0 = s->P * atom0, offset1 = s->P * atom1;
                for (int it = 0; it < s->P; it++) {
                    int t0 = offset0 + it, t1 = offset1 + it, tm;
                    double rcom[3], rpt[3], eul[3], v, rr, th, ch, hx[3], hy[3], hz[3];
                    if (mol0 == 2) {
                        tm = offset0 + it / s->RotRatio;
                        for (int id = 0; id < 3; id++) { rcom[id] = C(s, id, t0); rpt[id] = C(s, id, t1); }
                    } else {
                        tm = offset1 + it / s->RotRatio;
                        for (int id = 0; id < 3; id++) { rpt[id] = C(s, id, t0); rcom[id] = C(s, id, t1); }
                    }
                    eul[PHI] = AN(s, PHI, tm); eul[CTH] = acos(AN(s, CTH, tm)); eul[CHI] = AN(s, CHI, tm);
                    or_vcord(eul, rcom, rpt, s->vtable, s->nr3, s->nth3, s->nch3, s->rvmax, s->rvmin,
                             s->rvstep, &v, &rr, &th, &ch, hx, hy, hz, 0);
                    int br = (int)floor(rr / dradius);
                    if (br < OR_MC_BINSR && br >= 0) {
                        int bt = (int)floor(th / dtheta);
                        if (bt < OR_MC_BINST && bt >= 0) {
                            int bc = (int)floor(ch / dchi);
                            if (bc < OR_MC_BINSC && bc >= 0) gr3d[(br * OR_MC_BINST + bt) * OR_MC_BINSC + bc] += 1.0;
                        }
                    }
                }
            }
    if (rel) {
        int type = s->imtype;
        int offset = s->type[type].offset;
        for (int atom = 0; atom < s->type[type].numb; atom++) {
            offset += (s->P * atom);
            for (int it0 = 0; it0 < s->Prot; it0++) {
                int t0 = offset + it0, t1 = offset + (it0 + 1) % s->Prot;
                double e1[3], e2[3], er[3], rho, erot, esq;
                int istop = 0;
                e1[0] = AN(s, PHI, t0); e1[1] = acos(AN(s, CTH, t0)); e1[2] = AN(s, CHI, t0);
                e2[0] = AN(s, PHI, t1); e2[1] = acos(AN(s, CTH, t1)); e2[2] = AN(s, CHI, t1);
                or_rotden(e1, e2, er, &rho, &erot, &esq, s->rhoprp, s->erotpr, s->erotsq, &istop);
                int bt = (int)floor(er[1] / dtheta);
                if (bt < OR_MC_BINST && bt >= 0) rel[bt] += 1.0;
                int bp = (int)floor(er[0] / dchi);
                if (bp < OR_MC_BINSC && bp >= 0) rel[OR_MC_BINST + bp] += 1.0;
                int bc = (int)floor(er[2] / dchi);
                if (bc < OR_MC_BINSC && bc >= 0) rel[OR_MC_BINST + OR_MC_BINSC + bc] += 1.0;
            }
        }
    }
    return 0;
}

/* MCGetAverage mc_main.cc:551-616 (energy and Cv terms of one sample) */
void or_sample(or_sys* s, double out[8]) {
    double skin = or_kin_energy(s);
    double spot = or_pot_energy_densities(s, NULL, NULL);
    double srot = 0.0, rsq = 0.0, rterm = 0.0;
    if (s->imtype >= 0 && s->type[s->imtype].molecule == 1 && s->rotd[s->imtype].n > 0)
        srot = or_rot_energy(s, &rsq, &rterm);
    else if (s->imtype >= 0 && s->type[s->imtype].molecule == 2 && s->rhoprp)
        srot = or_rot_e3d(s, &rsq, &rterm);
    double n3 = (double)(3 * s->natoms);
    double a = n3 * 0.5 / s->tau - skin - spot - srot;
    double sCv = -0.5 * n3 / (s->beta * s->tau) - a * a + (2.0 / s->beta) * (0.5 * n3 / s->tau - skin) + rterm - rsq;
    double b = n3 * 0.5 / s->tau - skin;
    double sCv_trans = -0.5 * n3 / (s->beta * s->tau) - b * b + (2.0 / s->beta) * (0.5 * n3 / s->tau - skin);
    double sCv_rot = -srot * srot + rterm - rsq;
    out[0] = skin; out[1] = spot; out[2] = srot; out[3] = rsq;
    out[4] = sCv; out[5] = sCv_trans; out[6] = sCv_rot; out[7] = rterm;
}

/* One pass sampled per_pass times inside it, as the reference samples when its count of time
   steps is a multiple of MCSKIP_AVERG (mc_main.cc:348-386): in reference order after time
   step t when (t + 1) % (P / per_pass) == 0; in segment order after phase u (the bisections
   starting at u, u + S, ... and the rotor sweeps u, u + S, ...) when (u + 1) % (S / per_pass)
   == 0.  Returns the number of samples written to out (8 values each), -1 on error. */
int or_pass_sampled(or_sys* s, uint32_t pass, int order, int per_pass, double* out) {
    int P = s->P, ns = 0;
    if (per_pass < 1) return fail(s, "per_pass must be at least 1");
    if (order == OR_ORDER_REFERENCE) {
        if (P % per_pass) return fail(s, "samples per pass must divide P");
        for (int t = 0; t < P; t++) {
            for (int type = 0; type < s->ntypes; type++) {
                if (t == 0) or_molecular_move(s, type, pass);
                or_bisection_move(s, type, t, pass);
                if (rot_sweep(s, type, pass, (uint32_t)t) < 0) return -1;
            }
            if ((t + 1) % (P / per_pass) == 0) or_sample(s, out + 8 * (ns++));
        }
        return ns;
    }
    int S = 1;
    for (int type = 0; type < s->ntypes; ++type)
        if (s->type[type].mlsegm > S) S = s->type[type].mlsegm;
    if (P % S != 0) return fail(s, "segment order needs P divisible by the segment size");
    if (S % per_pass) return fail(s, "samples per pass must divide the segment size");
    for (int u = 0; u < S; ++u) {
        for (int type = 0; type < s->ntypes; type++) {
            if (u == 0) or_molecular_move(s, type, pass);
            for (int atom = 0; atom < s->type[type].numb; ++atom)
                for (int k = 0; k < P / S; ++k) or_bisection_atom(s, type, atom, u + k * S, pass);
        }
        for (int type = 0; type < s->ntypes; type++)
            for (int k = 0; k < P / S; ++k)
                if (rot_sweep(s, type, pass, (uint32_t)(u + k * S)) < 0) return -1;
        if ((u + 1) % (S / per_pass) == 0) or_sample(s, out + 8 * (ns++));
    }
    return ns;
}

int or_run(or_sys* s, uint32_t pass0, int npasses, int order, int every, double* out) {
    int ns = 0;
    for (int n = 0; n < npasses; ++n) {
        if (or_pass(s, pass0 + (uint32_t)n, order) < 0) return -1;
        if (out && every > 0 && ((n + 1) % every) == 0) or_sample(s, out + 8 * (ns++));
    }
    return ns;
}
