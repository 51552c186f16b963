/*
 * moribs_pimc.h -- C ABI of the MI355X (gfx950) path-integral Monte Carlo sampler.
 *
 * The library replaces MoRiBS-PIMC's sampling hot path: the translational centroid and
 * multilevel-bisection moves, the rotor moves on tabulated free-rotor density matrices,
 * the pair / rotor-particle potential-energy differences they need, and the energy
 * estimators evaluated on the sampled paths.  Plain C types only; every buffer passed in
 * is caller-owned host memory and is copied.  All device work of one context is ordered
 * on one HIP stream.  Errors are reported by a negative return value and a message from
 * pimc_last_error(); nothing calls exit() (the reference exits from nrerror(),
 * mc_utils.cc:99-108, and _io_error(), mc_input.h:35-42).
 *
 * Reference interfaces replaced (file:line in upstream MoRiBS-PIMC):
 *   moves        MCMolecularMove mc_piqmc.cc:54, MCBisectionMove :194, MCRotationsMove :421,
 *                MCRotations3D :692 (declared mc_piqmc.h:4-23), scheduled by PIMCPass
 *                mc_main.cc:510-522 inside the pass loop mc_main.cc:348-381
 *   energies     PotEnergy(int,double**,int) mc_piqmc.cc:1796, PotEnergy(int,double**) :1201,
 *                PotRotEnergy :1967, PotRotE3D :2046
 *   tables       InitPotentials/InitRotDensity mc_poten.cc:85-167 reading init_pot1D :379,
 *                init_pot2D :305, init_pot3D :254, init_rotdens :508, init_rot3D :440
 *   state        MCCoords/MCAngles/MCCosine mc_setup.h:176-190, laid out [3][NumbAtoms*NumbTimes]
 *   counters     MCTotal/MCAccep[type][MCMOLEC|MCMULTI|MCROTAT] mc_piqmc.h:25-26, mc_setup.h:161-164
 *   estimators   MCGetAverage mc_main.cc:551-616 (GetKinEnergy mc_estim.cc:876,
 *                GetPotEnergy_Densities :500, GetRotEnergy :939, GetRotE3D :989)
 */
#ifndef MORIBS_PIMC_H
#define MORIBS_PIMC_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pimc_ctx pimc_ctx;

#define PIMC_MAX_TYPES 4

/* context flags */
#define PIMC_FLAG_FP64        0x1u  /* pair terms in FP64 instead of FP32 (parity runs) */
#define PIMC_FLAG_REF_ORDER   0x2u  /* move order of mc_main.cc:348-381 (one segment at a time) */
#define PIMC_FLAG_PROFILE     0x4u  /* time every kernel launch with HIP events */

/* move kinds, the second index of MCTotal/MCAccep (mc_setup.h:161-164) */
#define PIMC_MCMOLEC 0
#define PIMC_MCMULTI 1
#define PIMC_MCROTAT 2

/* One replica (an independent Markov chain): the system of qmc.input.
   Types follow the reference: atoms first, then at most one molecule type
   (mc_input.cc:194-224, 386-387). */
typedef struct {
    int    ntypes;                   /* NumbTypes */
    int    numb[PIMC_MAX_TYPES];     /* MCAtom[t].numb */
    int    molecule[PIMC_MAX_TYPES]; /* 0 atom, 1 linear rotor, 2 non-linear rotor (MCAtom[t].molecule) */
    int    levels[PIMC_MAX_TYPES];   /* bisection levels (MCAtom[t].levels) */
    double mcstep[PIMC_MAX_TYPES];   /* centroid step, Angstrom (MCAtom[t].mcstep) */
    double mass[PIMC_MAX_TYPES];     /* amu (MCInitParams mc_setup.cc:244-319) */
    int    pot1d[PIMC_MAX_TYPES];    /* 1D table used when the partner is of this type, or -1 */
    int    pot2d[PIMC_MAX_TYPES];    /* 2D table of the linear-rotor type, or -1 */
    int    rotdens;                  /* .rot table of the linear rotor, or -1 */
    int    P;                        /* NUMBEROFSLICES */
    int    Prot;                     /* rotational slices (ROTATION line), 0 = no rotation */
    double T;                        /* TEMPERATURE, Kelvin */
    double rtstep;                   /* rotational step (ROTATION line) */
    int    pot3d;                    /* 3D table of the atom / non-linear rotor pairs, or -1 */
    int    rot3d;                    /* rho/E/E^2 tables of the non-linear rotor, or -1 */
} pimc_replica_desc;

/* ---- lifetime ---- */
/* seed and replica index form the key of the counter-based generator (Philox4x32-10)
   that replaces SPRNG streams (mc_randg.cc) and RngStream (omprng.cc). */
int  pimc_create(int n_replicas, unsigned flags, unsigned seed, pimc_ctx** out);
void pimc_destroy(pimc_ctx* ctx);
const char* pimc_last_error(const pimc_ctx* ctx);
int  pimc_device_name(char* buf, int len);

/* ---- tables (host arrays in Kelvin / Angstrom / reference file order) ---- */
/* init_pot1D mc_poten.cc:379-438: clamped spline + U0 exp(-alpha r) and -C6/r^6 tails */
int  pimc_add_pot1d(pimc_ctx* ctx, int n, const double* r, const double* v);
/* init_pot2D mc_poten.cc:305-377: header sizes/steps, grids, V[r][cos] */
int  pimc_add_pot2d(pimc_ctx* ctx, int nr, int nc, double dr, double dc, const double* rgrid,
                    const double* cgrid, const double* v);
/* init_rotdens mc_poten.cc:508-546: the four .rot columns */
int  pimc_add_rotdens(pimc_ctx* ctx, int n, const double* cosg, const double* rho,
                      const double* erot, const double* esq);
/* init_pot3D mc_poten.cc:254-303: header Rgrd THgrd CHgrd Rvmin Rvmax (R in bohr), then
   V[(ir*THgrd + ith)*CHgrd + ich] in the file's order */
int  pimc_add_pot3d(pimc_ctx* ctx, int nr, int nth, int nch, double rmin, double rmax, const double* v);
/* init_rot3D mc_poten.cc:440-506: the .rho, .eng and .esq tables, 181*361*361 values each in
   rotpro's order (rotpro_sub.f:7-26); copied to the device at once (566 MB) */
int  pimc_add_rot3d(pimc_ctx* ctx, const double* rho, const double* erot, const double* esq);
/* Free-rotor density matrices computed on the GPU, in the reference generators' units and
   order; ctx needs no replicas or commit (any created context).
   linden (linear_prop/linden.f:88-122): the four .rot columns cos(gamma), rho, E and E^2
   numerators on npt points of [-1, 1], for rotational constant bconst (cm^-1), T (K) and
   nslice rotor slices; iodevn -1 all l, 0 even, 1 odd; lmax <= 0 takes linden's cut-off (first
   l with weight below 1e-16).  Returns lmax used.
   asymrho (nmv_prop/asymrho.f:34-731): rho, E (cm^-1) and E^2 (cm^-2) on the 181 x 361 x 361
   whole-degree (theta, phi, chi) grid that pimc_add_rot3d reads, for rotational constants
   A, B, C (cm^-1), J up to maxj, iodevn -1/0/1 (both, even, odd k).  Fails like asymrho when
   exp(-tau E_max) > 1e-8 ("raise maxj"). */
int  pimc_linden(pimc_ctx* ctx, double T, int nslice, double bconst, int npt, int iodevn, int lmax,
                 double* cost, double* rho, double* erot, double* esq);
int  pimc_asymrho(pimc_ctx* ctx, double T, int nslice, double A, double B, double C, int maxj, int iodevn,
                  double* rho, double* eng, double* esq);

/* ---- replicas and state ---- */
int  pimc_set_replica(pimc_ctx* ctx, int replica, const pimc_replica_desc* desc);
/* The checks pimc_set_replica applies, without a context or a device: the reference's input
   checks on the sampled path (MCInit mc_setup.cc:384-398, IOReadParams mc_input.cc:210-330, with
   their messages; the reference exits where this returns -1), the limits of this library, and an
   even rotor slice count for the checkerboard sweeps (mc_piqmc.cc:490-512, :728-760).  Returns 0
   if accepted, else -1 with the reason in msg (len bytes, may be NULL). */
int  pimc_check_replica(const pimc_replica_desc* desc, char* msg, int len);
/* Minimum image (the MINIMAGE keyword, mc_input.cc:298-300) of a set replica in a periodic box
   of edges box[0..2] (BoxSize, mc_setup.cc:343-366), before pimc_commit: the separation of every
   atom pair in the moves and the potential estimator, and every kinetic link, is taken to its
   nearest image, dr - BoxSize rint(dr / BoxSize) (mc_piqmc.cc:1239-1240,1832-1833;
   mc_estim.cc:536-537,916-917).  Pairs with a non-linear rotor take the positions themselves, as
   vcord and caleng do there.  box == NULL turns it off.  Refused for a rotating linear rotor: the
   reference's rotor move stops with "MIN IMAGE for orient pot" (mc_piqmc.cc:2013-2016). */
int  pimc_set_minimum_image(pimc_ctx* ctx, int replica, const double* box);
int  pimc_commit(pimc_ctx* ctx);   /* allocates device state after all replicas are set */
/* How pimc_commit splits linear-rotor chains over workgroups, without a device (no counterpart
   in the reference, whose sweep is one OpenMP loop per chain, mc_piqmc.cc:490-512).  Chain i has
   prot[i] rotor slices (even), rot_ratio[i] translational slices per rotor slice and natoms[i]
   atoms.  Parts hold an even number of slices, sized so their LDS stays within lds_kb (at most
   part_slices_max slices when > 0); a chain's parts wait on one another on one XCD, so a chain
   gets at most cus_per_xcd * min(blocks_per_cu, lds_per_cu / lds) parts, fewer and larger parts
   when the LDS budget alone would give more.  Writes parts_per_chain[i], slices_per_part[i] and
   the launch's LDS bytes; returns -1 with a message when no plan fits (a part above 160 KB, or
   parts that cannot be resident together). */
int  pimc_plan_rotor_parts(int nchains, const int* prot, const int* rot_ratio, const int* natoms, int fp64,
                           int lds_kb, int part_slices_max, int blocks_per_cu, int lds_per_cu, int cus_per_xcd,
                           int* parts_per_chain, int* slices_per_part, long* lds_bytes, char* msg, int len);
/* How pimc_commit splits non-linear rotor chains (MCRotations3D's checkerboard, mc_piqmc.cc:728-
   760) over workgroups, without a device; no counterpart in the reference.  One count W for all
   nchains chains of up to protmax slices with nrot rotors: the largest power of two up to 64 whose
   parts hold at least 16 slices and fit LDS, with all parts resident at once (ncu * occupancy)
   and the W parts of one chain, which wait on one another on one XCD, within cus_per_xcd *
   occupancy; occupancy = min(blocks_per_cu, lds_per_cu / part LDS).  Returns W (1: whole
   chains), -1 for arguments out of range. */
int  pimc_plan_rotor3d_split(int protmax, int nchains, int nrot, int blocks_per_cu, int lds_per_cu, int ncu,
                             int cus_per_xcd);
/* Launch-shape options, set before pimc_commit (no counterpart in the reference, whose
   OpenMP loops have no tunable shape):
     "rotor_part_slices"  most rotor slices per workgroup of the rotor sweep (even, >= 2);
                          0 (default) sizes parts from the LDS budget.
     "rotor_lds_kb"       LDS budget of one part of the linear-rotor sweep, 8 to 150 KB
                          (default 52: three parts per CU).
     "centroid_parts"     about how many workgroups the centroid move's replicas are split into
                          in all (0, the default: min(1024, half of what the device holds)).  Results do not
                          depend on it: every draw is addressed by (sweep, slice).
     "bisect_lanes_small", "bisect_lanes_medium", "bisect_lanes_large"
                          lanes (64, 128 or 256) per bisection segment for replicas of at most
                          24, at most 64 and more atoms; defaults 64, 128, 256.  Results do not
                          depend on them.
     "delayed_acceptance" 1: the non-linear rotor move (MCRot3Dstep mc_piqmc.cc:938-1199) runs
                          its test as three Metropolis factors: the density, the surrogate
                          network, then the exact correction (see the surrogate section below);
                          needs atoms with a 3D potential (rotor-atom network) or two rotors
                          (rotor-rotor network), and a surrogate loaded before pimc_run.
     "samples_per_pass"   k >= 1 (default 1): a sampled pass takes k estimator samples inside it,
                          after every P/k time steps in the reference's order, or every S/k
                          phases in segment order (S = 2^levels; a phase moves P/S time steps'
                          worth of slices), as the reference samples every MCSKIP_AVERG time
                          steps (mc_main.cc:383-386).  Needs the same P and S in every replica
                          and k dividing P (reference order) or S (segment order).
   One option may be set at any time, also after pimc_commit:
     "profile"            1: time every launch with HIP events (pimc_kernel_time), issuing passes
                          launch by launch; 0: replay the captured pass graph (the default). */
int  pimc_set_option(pimc_ctx* ctx, const char* name, long value);
/* coords/angles [3][natoms*P] like MCCoords/MCAngles; cosines follow from angles
   (mc_main.cc:186-197).  angles may be NULL (phi=0, cos(theta)=1, chi=0, mc_setup.cc:476-489).
   Only rotor entries of angles are kept; pimc_get_state returns those defaults for atoms. */
int  pimc_set_state(pimc_ctx* ctx, int replica, const double* coords, const double* angles);
int  pimc_get_state(pimc_ctx* ctx, int replica, double* coords, double* angles, double* cosines);
/* restart (ConfigIO mc_input.cc:560-609 and the generator streams of RandomIO mc_randg.cc:36-88):
   the state as pimc_get_state returned it, cosines included, and the replica's pass counter
   (the generator's whole state: every draw is addressed by pass).  Rotation matrices of
   non-linear rotors are rebuilt on the device with the moves' own code, so a resumed run
   continues exactly as the uninterrupted one. */
int  pimc_restore_state(pimc_ctx* ctx, int replica, const double* coords, const double* angles,
                        const double* cosines, unsigned pass);
int  pimc_get_pass(pimc_ctx* ctx, int replica, unsigned* pass);

/* ---- sampling ---- */
/* n_passes passes of PIMCPass over all replicas; one MCGetAverage sample per replica
   every sample_every passes (0: none).  Pass numbers continue across calls. */
int  pimc_run(pimc_ctx* ctx, int n_passes, int sample_every);
/* Autotuner of launch shapes (no counterpart in the reference, whose OpenMP loops have no
   tunable shape).  After pimc_commit and the initial state: times each bisection size class
   at 64, 128 and 256 lanes per segment, the linear-rotor sweep at part LDS budgets of 40,
   52, 64 and 78 KB and the centroid move split into about 256, 512 or 1024 parts, each
   candidate over `passes` passes' launches from the same saved state, and keeps the fastest.  The chains' coordinates, orientations, counters and pass numbers are
   restored, so results are the same as without tuning.  chosen (may be NULL, else 5 entries)
   receives the lanes of the small, medium and large classes, the rotor budget in KB and the
   centroid part target. */
int  pimc_autotune(pimc_ctx* ctx, int passes, long* chosen);
/* samples of one replica, 8 doubles each: K, V, E_rot, E_rot^2 term (ErotSQ), Cv, Cv_trans,
   Cv_rot, Erot_termSQ (mc_main.cc:551-616).  Returns the number copied.  A replica keeps up to
   4096 samples between pimc_clear_samples calls; if it took more, the later ones were not
   stored and this call fails instead of returning a truncated series. */
int  pimc_get_samples(pimc_ctx* ctx, int replica, double* out, int max_samples);
int  pimc_clear_samples(pimc_ctx* ctx);
/* MCTotal/MCAccep of one replica, [ntypes][3] */
int  pimc_get_counters(pimc_ctx* ctx, int replica, double* total, double* accep);
/* accumulated g(r) (300 bins) and g(r,theta) (300x50) counts of one replica
   (bin_1Ddensity/bin_2Ddensity mc_estim.cc:1233-1286) */
int  pimc_get_histograms(pimc_ctx* ctx, int replica, double* gr1d, double* gr2d);
/* accumulated bins around a non-linear rotor of one replica since pimc_clear_samples:
   gr3d[(ir*50 + it)*100 + ic] counts the atoms at (r, theta, chi) of the rotor's body frame
   (300 radial bins to 15 A, theta by pi/49, chi by 2 pi/99; bin_3Ddensity mc_estim.cc:1252-1273
   on vcord's radret/theret/chiret, GetPotEnergy_Densities mc_estim.cc:582-641); rel[0..49],
   rel[50..149], rel[150..249] count the relative Euler angles theta, phi, chi of neighbouring
   rotor slices (GetRotE3D mc_estim.cc:1046-1058).  Either pointer may be NULL.  Returns a
   bit mask of what the replica has: 1 the 3D bins (non-linear rotor and atoms), 2 the
   relative angle bins (rotating non-linear rotor with a density table); arrays it lacks are
   left untouched. */
int  pimc_get_histograms3d(pimc_ctx* ctx, int replica, double* gr3d, double* rel);
/* rotational correlation sums of one replica since pimc_clear_samples, out[10][Prot]:
   row 0 sum of n(t0).n(t0+t), rows l = 1..9 sum of P_l of it, over time origins and samples
   (GetRCF mc_estim.cc:1098-1139; the reference adds 1.0 in rows 1..9).  Returns Prot, or 0
   when the replica does not rotate. */
int  pimc_get_rcf(pimc_ctx* ctx, int replica, double* out, int max_values);
/* pair evaluations performed and bead moves proposed since creation (all replicas) */
int  pimc_get_work(pimc_ctx* ctx, double* pair_evals, double* bead_moves);

/* ---- energies on the current state (batched) ---- */
/* PotEnergy(atom, pos, it) mc_piqmc.cc:1796: out[q] for queries (replica, atom, it) with
   the atom's bead at slice it replaced by pos[3q..3q+2] */
int  pimc_pot_energy_slice(pimc_ctx* ctx, int nq, const int* replica, const int* atom,
                           const int* it, const double* pos, double* out);
/* pimc_pot_energy_slice without any device allocation.  No counterpart in the reference: it is
   for callers that cannot allocate device memory (code under HIP graph capture, serving loops,
   harnesses that forbid hipMalloc).  Queries, positions and results are staged in a device
   array of the library (704 KiB, PIMC_NOALLOC_MAX_QUERIES queries) that every GPU holds for the
   rest of the process; more queries per call are refused with an error.  Calls that overlap on
   one GPU (other threads or contexts) share that array and race; the caller serialises them. */
#define PIMC_NOALLOC_MAX_QUERIES 16384
int  pimc_pot_energy_slice_noalloc(pimc_ctx* ctx, int nq, const int* replica, const int* atom,
                                   const int* it, const double* pos, double* out);
/* one MCGetAverage sample of every replica on the current state, out[8*replica] */
int  pimc_estimate(pimc_ctx* ctx, double* out);

/* ---- measurement ---- */
/* accumulated device time (ms) and launch count of a kernel family while profiling is on
   (PIMC_FLAG_PROFILE at creation, or the option "profile" set to 1 at any time; passes are then
   issued launch by launch between HIP events instead of replaying the captured pass graph):
   "bisect", "centroid", "rotor", "estimate" */
int  pimc_kernel_time(pimc_ctx* ctx, const char* family, double* ms, long* launches);
int  pimc_synchronize(pimc_ctx* ctx);
/* pair evaluations performed by one kernel family since creation */
int  pimc_family_evals(pimc_ctx* ctx, const char* family, double* evals);

/* ---- surrogate network and delayed acceptance ----
   No counterpart in the reference: its rotor moves always evaluate PotRotE3D
   (mc_piqmc.cc:2046-2151) in full.  The network (4 -> 128 -> 128 -> 128 -> 1, SiLU) maps the
   body-frame geometry of an atom-rotor pair (R, ux, uy, uz) to the pair energy in Kelvin.
   theta holds 33793 floats: W1[128][4] b1[128] W2[128][128] b2[128] W3[128][128] b3[128]
   w4[128] b4; norm holds xm[4] xs[4] ym ys (input (x - xm)/xs, energy ys*net + ym). */
/* training rows from one replica's current configuration: every (rotor slice, translational
   slice, atom partner) pair, x[n][4] features and v[n] exact energies; returns n */
int  pimc_surrogate_collect(pimc_ctx* ctx, int replica, float* x, float* v, int max_rows);
/* normalises the rows, initialises the weights from seed and trains with Adam on minibatches
   of `batch` rows for `epochs` epochs (learning rate decaying tenfold); loss[epochs] receives
   the mean normalised loss of each epoch; the result becomes the sampler's surrogate */
int  pimc_surrogate_train(pimc_ctx* ctx, const float* x, const float* v, int n, int epochs, int batch,
                          double lr, unsigned seed, double* loss);
/* loss (1/2n) sum (net(x) - y)^2 and its gradient (theta layout) of the raw network at theta */
int  pimc_surrogate_grad(pimc_ctx* ctx, const float* theta, const float* x, const float* y, int n,
                         float* grad, double* loss);
int  pimc_surrogate_set(pimc_ctx* ctx, const float* theta, const float* norm);
int  pimc_surrogate_get(pimc_ctx* ctx, float* theta, float* norm);
/* the sampler's evaluation (BF16 matrix cores) on raw features: out[n] energies in Kelvin */
int  pimc_surrogate_eval(pimc_ctx* ctx, const float* x, int n, float* out);
/* out[4]: delayed-acceptance proposals, proposals that passed stages 0 and 1 (charged the exact
   network-pair sums), accepted moves, network evaluations */
int  pimc_get_da_counters(pimc_ctx* ctx, int replica, double* out);

#ifdef __cplusplus
}
#endif
#endif
