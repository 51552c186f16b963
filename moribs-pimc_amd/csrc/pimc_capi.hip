// pimc_capi.hip -- host side of the C ABI declared in include/moribs_pimc.h.
//
// A context holds any number of independent replicas (Markov chains).  pimc_run() issues,
// per pass, the move kernels in the order of PIMCPass (mc_main.cc:510-522) grouped by
// segment offset: for u in 0..S-1 { [u==0: centroid(type), bisection(type) per type],
// bisection(all types), linear-rotor sweeps u, u+S, ... } -- every atom gets one bisection
// proposal per segment start and the rotor P sweeps per pass, as in mc_main.cc:348-381.
// One pass is captured once as a HIP graph and replayed.
#include "pimc_kernels.hpp"
#include "pimc_gen.hpp"
#include "../../include/moribs_pimc.h"
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>
#include <map>
#include <mutex>
#include <algorithm>
#include <array>

using namespace pimc;

namespace {

// physical constants as in mc_const.h:12-14 (lambda = hbar^2/2m in Angstrom^2 K)
constexpr double kHBAR = 1.05457266, kAMU = 1.6605402, kKB = 1.380658;

// Numerical-Recipes clamped cubic spline (second derivatives), end slopes given.
void spline_y2(const std::vector<double>& x, const std::vector<double>& y, double yp1, double ypn,
               std::vector<double>& y2) {
    const int n = (int)x.size();
    std::vector<double> u(n);
    y2.assign(n, 0.0);
    y2[0] = -0.5;
    u[0] = (3.0 / (x[1] - x[0])) * ((y[1] - y[0]) / (x[1] - x[0]) - yp1);
    for (int i = 1; i < n - 1; ++i) {
        double sig = (x[i] - x[i - 1]) / (x[i + 1] - x[i - 1]);
        double p = sig * y2[i - 1] + 2.0;
        y2[i] = (sig - 1.0) / p;
        double d = (y[i + 1] - y[i]) / (x[i + 1] - x[i]) - (y[i] - y[i - 1]) / (x[i] - x[i - 1]);
        u[i] = (6.0 * d / (x[i + 1] - x[i - 1]) - sig * u[i - 1]) / p;
    }
    double qn = 0.5;
    double un = (3.0 / (x[n - 1] - x[n - 2])) * (ypn - (y[n - 1] - y[n - 2]) / (x[n - 1] - x[n - 2]));
    y2[n - 1] = (un - qn * u[n - 2]) / (qn * y2[n - 2] + 1.0);
    for (int k = n - 2; k >= 0; --k) y2[k] = y2[k] * y2[k + 1] + u[k];
}

// end slopes from first differences at both ends (init_spline, mc_utils.cc:188-201)
void spline_fd(const std::vector<double>& x, const std::vector<double>& y, std::vector<double>& y2) {
    const int n = (int)x.size();
    double dpl = (y[1] - y[0]) / (x[1] - x[0]);
    double dpr = (y[n - 1] - y[n - 2]) / (x[n - 1] - x[n - 2]);
    spline_y2(x, y, dpl, dpr, y2);
}

struct Family {
    double ms = 0;
    long n = 0;
};

// Workgroup i of a launch runs on XCD i mod 8.  The parts of one chain (consecutive in `parts`,
// W of them) all go to one XCD, so their exchanges stay in one L2; chains are spread over the
// XCDs by `work` (heaviest first, to the least loaded XCD) and the grid is padded with r = -1.
template <typename Part, typename Work>
std::vector<Part> xcd_order(const std::vector<Part>& parts, Work work) {
    constexpr int NXCD = 8;
    std::vector<std::vector<Part>> bins(NXCD);
    std::vector<double> load(NXCD, 0.0);
    for (size_t i = 0; i < parts.size();) {
        const int W = parts[i].W;
        const int x = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        load[x] += work(parts[i].r);
        for (int w = 0; w < W; ++w) bins[x].push_back(parts[i + w]);
        i += W;
    }
    size_t len = 0;
    for (auto& b : bins) len = std::max(len, b.size());
    std::vector<Part> grid(len * NXCD);
    for (size_t j = 0; j < len; ++j)
        for (int x = 0; x < NXCD; ++x) {
            Part q{};
            q.r = -1;
            if (j < bins[x].size()) q = bins[x][j];
            grid[j * NXCD + x] = q;
        }
    return grid;
}

// bisection size classes of the linear family: lanes per workgroup by atom count
inline int bisect_class(int natoms) { return natoms <= 24 ? 0 : (natoms <= 64 ? 1 : 2); }

template <typename T>
T* dalloc(size_t n) {
    T* p = nullptr;
    if (n == 0) n = 1;
    if (hipMalloc(&p, n * sizeof(T)) != hipSuccess) return nullptr;
    return p;
}

// Staging of pimc_pot_energy_slice_noalloc: queries (replica, atom, slice: three int arrays),
// positions and results, in a file-scope device array of which every GPU holds its own copy
// until the process ends.  Host code never passes the array's name to a kernel or a copy (on
// the host it is a stand-in, not a device address); it uses the address hipGetSymbolAddress
// returned for the device, kept per device in scratch_dev.
constexpr size_t NA_Q = PIMC_NOALLOC_MAX_QUERIES;
constexpr size_t NA_OFF_A = NA_Q * sizeof(int), NA_OFF_I = 2 * NA_Q * sizeof(int), NA_OFF_POS = 3 * NA_Q * sizeof(int);
constexpr size_t NA_OFF_OUT = NA_OFF_POS + 3 * NA_Q * sizeof(double);
constexpr size_t NA_BYTES = NA_OFF_OUT + NA_Q * sizeof(double);  // 704 KiB
static_assert(NA_OFF_A % 256 == 0 && NA_OFF_POS % 256 == 0 && NA_OFF_OUT % 256 == 0, "scratch regions stay aligned");
__device__ __align__(256) static unsigned char scratch[NA_BYTES];

constexpr int NA_MAX_DEVICES = 64;
std::mutex scratch_mu;
unsigned char* scratch_dev[NA_MAX_DEVICES] = {};

// the address of this device's copy of `scratch`, looked up once per device (device current)
unsigned char* scratch_for_current_device(std::string& err) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= NA_MAX_DEVICES) {
        err = "no current device";
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (!scratch_dev[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
            return nullptr;
        }
        scratch_dev[dev] = static_cast<unsigned char*>(p);
    }
    return scratch_dev[dev];
}

}  // namespace

struct pimc_ctx {
    int nrep = 0;
    unsigned flags = 0, seed = 0;
    bool fp64 = false, ref_order = false, profile = false, committed = false;
    std::string err;
    std::vector<pimc_replica_desc> desc;
    std::vector<char> desc_set;
    std::vector<std::array<double, 3>> box;  // minimum-image box per replica, zeros when off
    bool any_mi = false;                       // a replica uses it: launch the kernels built for it
    std::vector<Rep> rep;
    // host tables
    std::vector<Tab1D> t1;
    std::vector<double4> k1;
    std::vector<Tab2D> t2;
    std::vector<double> v2, g2;
    std::vector<TabRot> tr;
    std::vector<double4> rk, re;
    std::vector<Tab3D> t3;
    std::vector<double> v3;
    double4* d_r3 = nullptr;  // rho, E, E^2 of every non-linear density table, 3 * ROT3D_SIZE points each
    int nr3 = 0;
    // device
    hipStream_t stream = nullptr;
    Rep* d_rep = nullptr;
    Mut* d_mut = nullptr;
    double4* d_beads = nullptr;
    RotBead* d_rotb = nullptr;
    unsigned* d_hist = nullptr;
    double* d_rcf = nullptr;      // rotational correlation sums of the rotating replicas
    int* d_rcfreps = nullptr;
    int nrcf = 0, rcf_protmax = 1;
    long long nrcfvals = 0;
    double* d_samples = nullptr;
    int sample_cap = 0;
    float4* d_k1f = nullptr;
    double4* d_k1d = nullptr;
    Tab1D* d_t1 = nullptr;
    float* d_v2f = nullptr;
    float4* d_v2c = nullptr;  // corner-packed cells of the uniform 2D tables
    double* d_v2d = nullptr;
    double* d_g2d = nullptr;
    float* d_g2f = nullptr;
    Tab2D* d_t2 = nullptr;
    double4* d_rk = nullptr;
    double4* d_re = nullptr;
    TabRot* d_tr = nullptr;
    double4* d_v3d = nullptr;
    float4* d_v3f = nullptr;
    Tab3D* d_t3 = nullptr;
    Frame* d_frames = nullptr;     // rotation matrix of every non-linear rotor slice
    int2* d_work1 = nullptr;       // bisection work items of the non-linear family
    int nwork1 = 0;
    int* d_all0 = nullptr;         // replicas of the linear family
    int* d_all1 = nullptr;         // replicas of the non-linear family
    int nall0 = 0, nall1 = 0;
    int* d_rot3 = nullptr;         // replicas with rotating non-linear rotors
    int nrot3 = 0, nrotmax = 0;
    int rot3_block = 256;          // lanes of the non-linear rotor sweep
    size_t lds_rot3 = 0;
    int2* d_rows1 = nullptr;
    int nrows1 = 0;
    size_t lds_bisect1 = 0;
    int2* d_work = nullptr;
    int nwork = 0;
    int* d_all = nullptr;
    int* d_rotreps = nullptr;
    int nrotreps = 0;
    double* d_est = nullptr;
    size_t lds_bisect = 0, lds_rotor = 0;
    int Smax = 1, Pmax = 1, Nmax = 1, max_types = 1;
    int bisect_block = 64;
    // linear-family bisection in three size classes (N <= 24, <= 64, larger) with 64, 128 and
    // 256 lanes and their own LDS sizes, launched as parallel branches of the pass
    int bw_n[3] = {0, 0, 0}, bw_off[3] = {0, 0, 0};
    int bw_lanes[3] = {64, 128, 256};  // lanes per segment of each class
    size_t bw_lds[3] = {0, 0, 0};
    hipStream_t side[2] = {nullptr, nullptr};
    hipEvent_t ev_fork = nullptr, ev_join[2] = {nullptr, nullptr};
    int2* d_rows = nullptr;     // (replica, tile) work of the estimator pair sum
    double2* d_rowbuf = nullptr;  // a (potential, kinetic) partial per tile
    int nrows = 0;
    size_t lds_est = 0;
    size_t lds_knots = 0;       // the largest 1D knot array, staged in LDS by the centroid move
    RotPart* d_parts = nullptr;  // rotor chains split into slice ranges, one workgroup each
    unsigned* d_flags = nullptr;  // half-sweeps completed per part
    unsigned long long* d_xrec = nullptr;  // per part, its two boundary records (8 words each) as last published
    int nparts = 0, nflags = 0;
    int opt_rotor_part_slices = 0;  // most rotor slices per part, 0: from the LDS budget
    int opt_rotor_lds_kb = 52;      // LDS budget of a rotor part (52 KB: three parts per CU)
    int opt_centroid_parts = 0;     // target number of centroid parts, 0: from the capacity
    int centroid_target = 0;        // target the current centroid parts were planned with
    int rotor_lds_kb = 52;          // budget the current parts were planned with
    std::vector<int> rot_chains;    // replicas with a linear rotor, heaviest first
    RotPart* d_parts3 = nullptr;     // non-linear chains split by slice range
    unsigned* d_flags3 = nullptr;
    unsigned long long* d_pub3 = nullptr;  // boundary records published per part
    int nparts3 = 0, pub3_stride = 0;
    size_t lds_parts3 = 0;
    RotPart* d_cparts0 = nullptr;    // centroid parts (slice ranges) of the linear family
    RotPart* d_cparts1 = nullptr;    // and of the non-linear family
    int ncp0 = 0, ncp1 = 0, ncflags = 0;
    unsigned* d_cflags = nullptr;    // atoms done per centroid part
    unsigned long long* d_cslots = nullptr;  // partial Delta V per part as two numbered words, two buffers
    // surrogate network of the delayed-acceptance rotor move
    bool opt_da = false;
    MlpPacked* d_mlp = nullptr;     // the sampler's packed copy
    bool mlp_ready = false;
    std::vector<float> theta, norm;  // host copies: weights, then xm[4] xs[4] ym ys
    long long nbeads = 0, nrotb = 0, nhist = 0;
    hipGraphExec_t graph_exec = nullptr;    // one pass
    hipGraphExec_t graph_exec_s = nullptr;  // one pass with samples_per_pass samples inside it
    int opt_samples_per_pass = 1;
    bool graph_ok = false;
    std::map<std::string, Family> fam;
    struct Pending {
        const char* name;
        hipEvent_t a, b;
    };
    std::vector<Pending> pending;
    double bead_moves_per_pass = 0;
    double bead_moves = 0;
    unsigned char* scratch = nullptr;  // this device's copy of `scratch` (pimc_pot_energy_slice_noalloc)
    std::string scratch_err;
    Tables tables() const {
        Tables T;
        T.k1f = d_k1f;
        T.k1d = d_k1d;
        T.t1 = d_t1;
        T.v2f = d_v2f;
        T.v2c = d_v2c;
        T.v2d = d_v2d;
        T.g2d = d_g2d;
        T.g2f = d_g2f;
        T.t2 = d_t2;
        T.rk = d_rk;
        T.re = d_re;
        T.tr = d_tr;
        return T;
    }
    Tables3 tables3() const {
        Tables3 T;
        T.v3d = d_v3d;
        T.v3f = d_v3f;
        T.t3 = d_t3;
        T.r3 = d_r3;
        return T;
    }
};

static int fail(pimc_ctx* c, const std::string& m) {
    if (c) c->err = m;
    return -1;
}

#define HIPCHECK(ctx, x)                                                                   \
    do {                                                                                   \
        hipError_t e_ = (x);                                                               \
        if (e_ != hipSuccess) return fail(ctx, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

extern "C" {

int pimc_device_name(char* buf, int len) {
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, 0) != hipSuccess) return -1;
    std::snprintf(buf, (size_t)len, "%s (%s, %d CUs)", p.name, p.gcnArchName, p.multiProcessorCount);
    return 0;
}

int pimc_create(int n_replicas, unsigned flags, unsigned seed, pimc_ctx** out) {
    if (!out || n_replicas < 1) return -1;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return -2;  // no HIP device: no fallback
    pimc_ctx* c = new pimc_ctx();
    c->nrep = n_replicas;
    c->flags = flags;
    c->seed = seed;
    c->fp64 = flags & PIMC_FLAG_FP64;
    c->ref_order = flags & PIMC_FLAG_REF_ORDER;
    c->profile = flags & PIMC_FLAG_PROFILE;
    c->desc.resize(n_replicas);
    c->desc_set.assign(n_replicas, 0);
    c->box.assign(n_replicas, std::array<double, 3>{0.0, 0.0, 0.0});
    if (hipStreamCreate(&c->stream) != hipSuccess) {
        delete c;
        return -3;
    }
    c->scratch = scratch_for_current_device(c->scratch_err);  // reported by the entry point that uses it
    *out = c;
    return 0;
}

void pimc_destroy(pimc_ctx* c) {
    if (!c) return;
    for (auto& p : c->pending) {
        (void)hipEventDestroy(p.a);
        (void)hipEventDestroy(p.b);
    }
    if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
    if (c->graph_exec_s) (void)hipGraphExecDestroy(c->graph_exec_s);
    void* ptrs[] = {c->d_rep, c->d_mut, c->d_beads, c->d_rotb, c->d_hist, c->d_samples, c->d_k1f, c->d_k1d,
                    c->d_t1, c->d_v2f, c->d_v2c, c->d_v2d, c->d_g2d, c->d_g2f, c->d_t2, c->d_rk, c->d_re, c->d_tr,
                    c->d_work, c->d_all, c->d_rotreps, c->d_est, c->d_rows, c->d_rowbuf, c->d_parts, c->d_flags, c->d_xrec,
                    c->d_r3, c->d_v3d, c->d_v3f, c->d_t3, c->d_frames, c->d_work1, c->d_all0, c->d_all1, c->d_rot3, c->d_rows1, c->d_mlp,
                    c->d_rcf, c->d_rcfreps, c->d_cparts0, c->d_cparts1, c->d_cflags, c->d_cslots,
                    c->d_parts3, c->d_flags3, c->d_pub3};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    for (int k = 0; k < 2; ++k) {
        if (c->side[k]) (void)hipStreamDestroy(c->side[k]);
        if (c->ev_join[k]) (void)hipEventDestroy(c->ev_join[k]);
    }
    if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

const char* pimc_last_error(const pimc_ctx* c) { return c ? c->err.c_str() : "null context"; }

// a grid counts as uniform when every point lies within 1e-6 (relative to the span) of the
// straight line through its end points
static bool uniform_grid(const double* g, int n, double& g0, double& step) {
    g0 = g[0];
    step = (g[n - 1] - g[0]) / (double)(n - 1);
    double span = std::fabs(g[n - 1] - g[0]);
    for (int i = 0; i < n; ++i)
        if (std::fabs(g[i] - (g0 + i * step)) > 1e-6 * span) return false;
    return step > 0;
}

int pimc_add_pot1d(pimc_ctx* c, int n, const double* r, const double* v) {
    if (c->committed) return fail(c, "tables must be added before pimc_commit");
    if (n < 2) return fail(c, "init_pot1D: Not enough data points for extrapolation");
    std::vector<double> x(r, r + n), y(v, v + n), y2;
    for (int i = 1; i < n; ++i)
        if (!(x[i] > x[i - 1])) return fail(c, "init_pot1D: grid must increase");
    spline_fd(x, y, y2);
    Tab1D t;
    t.n = n;
    t.off = (int)c->k1.size();
    t.x0 = x[0];
    t.xmax = x[n - 1];
    t.inv_h = (double)(n - 1) / (x[n - 1] - x[0]);
    double g0, step;
    t.uniform = uniform_grid(x.data(), n, g0, step) ? 1 : 0;
    t.pad = 0;
    // U0 exp(-alpha r) below the grid, -C6/r^6 above (mc_poten.cc:420-437)
    t.alpha = std::log(y[0] / y[1]) / (x[1] - x[0]);
    t.unode = y[0] * std::exp(t.alpha * x[0]);
    double r0 = std::pow(x[n - 2], 6.0), r1 = std::pow(x[n - 1], 6.0);
    t.c6 = (y[n - 1] - y[n - 2]) / (1.0 / r0 - 1.0 / r1);
    // the fourth knot component is the inverse width of the cell the knot opens
    for (int i = 0; i < n; ++i)
        c->k1.push_back(make_double4(x[i], y[i], y2[i], i + 1 < n ? 1.0 / (x[i + 1] - x[i]) : 0.0));
    c->t1.push_back(t);
    return (int)c->t1.size() - 1;
}

int pimc_add_pot2d(pimc_ctx* c, int nr, int nc, double dr, double dc, const double* rg, const double* cg,
                   const double* v) {
    if (c->committed) return fail(c, "tables must be added before pimc_commit");
    if (nr < 2 || nc < 2) return fail(c, "init_pot2D: grid needs two points per axis");
    Tab2D t;
    t.nr = nr;
    t.nc = nc;
    t.off = (long long)c->v2.size();
    t.coff = -1;  // set at pimc_commit for uniform tables
    t.goff = (int)c->g2.size();
    t.dr = dr;
    t.dc = dc;
    bool ur = uniform_grid(rg, nr, t.gr0, t.gdr), uc = uniform_grid(cg, nc, t.gc0, t.gdc);
    t.uniform = (ur && uc && dr > 0 && dc > 0) ? 1 : 0;
    c->g2.insert(c->g2.end(), rg, rg + nr);
    c->g2.insert(c->g2.end(), cg, cg + nc);
    c->v2.insert(c->v2.end(), v, v + (size_t)nr * nc);
    c->t2.push_back(t);
    return (int)c->t2.size() - 1;
}

int pimc_add_rotdens(pimc_ctx* c, int n, const double* x, const double* rho, const double* erot, const double* esq) {
    if (c->committed) return fail(c, "tables must be added before pimc_commit");
    if (n < 2) return fail(c, "init_rotdens: too few points");
    std::vector<double> gx(x, x + n), yr(rho, rho + n), ye(erot, erot + n), ys(esq, esq + n), s0, s1, s2;
    spline_fd(gx, yr, s0);
    spline_fd(gx, ye, s1);
    spline_fd(gx, ys, s2);
    TabRot t;
    t.n = n;
    t.off = (int)c->rk.size();
    t.x0 = gx[0];
    t.xmax = gx[n - 1];
    t.inv_h = (double)(n - 1) / (gx[n - 1] - gx[0]);
    for (int i = 0; i < n; ++i) {
        c->rk.push_back(make_double4(gx[i], yr[i], s0[i], 0.0));
        c->re.push_back(make_double4(ye[i], s1[i], ys[i], s2[i]));
    }
    c->tr.push_back(t);
    return (int)c->tr.size() - 1;
}

int pimc_add_pot3d(pimc_ctx* c, int nr, int nth, int nch, double rmin, double rmax, const double* v) {
    if (c->committed) return fail(c, "tables must be added before pimc_commit");
    if (nr < 2 || nth < 2 || nch < 2 || !(rmax > rmin) || !v) return fail(c, "init_pot3D: bad 3D grid");
    if (nch != 361 && nch != 181 && nch != 91 && nth != 181)
        return fail(c, "init_pot3D: Degree wise THETA or CHI grid is violated");
    Tab3D t{};
    t.nr = nr;
    t.nth = nth;
    t.nch = nch;
    t.rmin = rmin;
    t.rmax = rmax;
    t.rstep = (rmax - rmin) / (double)(nr - 1);
    t.off = (long long)c->v3.size();
    c->v3.insert(c->v3.end(), v, v + (size_t)nr * nth * nch);
    c->t3.push_back(t);
    return (int)c->t3.size() - 1;
}

// one table as (value, d_chi, d_phi, d_theta) per point of the 181 x 361 x 361 grid; the
// differences are those rotpro forms (rotpro_sub.f:31-61), zero where it drops a term
static void rot3d_points(const double* t, std::vector<double4>& out) {
    out.resize((size_t)ROT3D_SIZE);
    for (int ith = 0; ith <= 180; ++ith)
        for (int iph = 0; iph <= 360; ++iph)
            for (int ich = 0; ich <= 360; ++ich) {
                const size_t i0 = ((size_t)ith * 361 + iph) * 361 + ich;
                const double v = t[i0];
                out[i0] = make_double4(v, ich != 360 ? t[i0 + 1] - v : 0.0, iph != 360 ? t[i0 + 361] - v : 0.0,
                                       ith != 180 ? t[i0 + 361 * 361] - v : 0.0);
            }
}

namespace {
struct DevMem {  // a device buffer freed on every return path
    void* p = nullptr;
    ~DevMem() {
        if (p) (void)hipFree(p);
    }
    double* d() const { return static_cast<double*>(p); }
};
}  // namespace

int pimc_linden(pimc_ctx* c, double T, int nslice, double bconst, int npt, int iodevn, int lmax, double* cost,
                double* rho, double* erot, double* esq) {
    if (!c) return -1;
    if (!(T > 0.0) || nslice < 1 || npt < 2 || iodevn < -1 || iodevn > 1 || !cost || !rho || !erot || !esq)
        return fail(c, "pimc_linden: T > 0, nslice >= 1, npt >= 2, iodevn in -1..1 and four output arrays");
    const double tau = 1.4387752224 / (T * nslice);  // linden.f:5,24, in cm
    if (lmax <= 0) {  // the first l whose weight falls below 1e-16 (linden.f:27-36)
        lmax = 500;
        for (int l = 0; l <= 500; ++l)
            if (std::exp(-tau * bconst * l * (l + 1.0)) < 1.0e-16) {
                lmax = l;
                break;
            }
    }
    DevMem d;
    HIPCHECK(c, hipMalloc(&d.p, sizeof(double) * 4 * (size_t)npt));
    hipLaunchKernelGGL(k_linden, dim3((npt + 255) / 256), dim3(256), 0, c->stream, npt, lmax, iodevn, tau, bconst,
                       nslice * 0.69503476, d.d());
    HIPCHECK(c, hipGetLastError());
    std::vector<double> h(4 * (size_t)npt);
    HIPCHECK(c, hipMemcpyAsync(h.data(), d.p, sizeof(double) * h.size(), hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    std::copy(h.begin(), h.begin() + npt, cost);
    std::copy(h.begin() + npt, h.begin() + 2 * npt, rho);
    std::copy(h.begin() + 2 * npt, h.begin() + 3 * npt, erot);
    std::copy(h.begin() + 3 * npt, h.end(), esq);
    return lmax;
}

int pimc_asymrho(pimc_ctx* c, double T, int nslice, double A, double B, double C, int maxj, int iodevn, double* rho,
                 double* eng, double* esq) {
    if (!c) return -1;
    if (!(T > 0.0) || nslice < 1 || maxj < 0 || maxj > 160 || iodevn < -1 || iodevn > 1 || !rho || !eng || !esq)
        return fail(c, "pimc_asymrho: T > 0, nslice >= 1, 0 <= maxj <= 160, iodevn in -1..1 and three output arrays");
    const double beta = 1.0 / (0.6950356 * T), tau = beta / nslice;  // asymrho.f:42, in cm
    std::vector<double> g;
    const double emax = asym_blocks(maxj, A, B, C, tau, g);
    if (std::exp(-tau * emax) > 1.0e-8) return fail(c, "pimc_asymrho: too large contribution from emax: raise maxj");
    const int n = 2 * maxj + 1, NG = ASYM_NG, NT = 181;
    const size_t nn = (size_t)n * n, np = (size_t)NG * NG;
    const int npar = iodevn == -1 ? 2 : 1;
    const int par0 = iodevn == -1 ? 0 : iodevn;
    // e^{i k chi} split into cosine and sine tables [n][361], and [cos(m phi) | -sin(m phi)] [361][2n]
    std::vector<double> xc(n * NG), xs(n * NG), fm(NG * 2 * (size_t)n);
    for (int kk = 0; kk < n; ++kk)
        for (int a = 0; a < NG; ++a) {
            const double ang = (double)a * (M_PI / 180.0), x = ang * (double)(kk - maxj);
            xc[(size_t)kk * NG + a] = std::cos(x);
            xs[(size_t)kk * NG + a] = std::sin(x);
            fm[(size_t)a * 2 * n + kk] = std::cos(x);
            fm[(size_t)a * 2 * n + n + kk] = -std::sin(x);
        }
    const std::vector<int> src = asym_fill_map();
    const size_t nbatch = (size_t)NT * npar * 3;
    DevMem dg, dxc, dxs, dfm, dsrc, dm, dt, dx, dout;
    HIPCHECK(c, hipMalloc(&dg.p, sizeof(double) * g.size()));
    HIPCHECK(c, hipMalloc(&dxc.p, sizeof(double) * xc.size()));
    HIPCHECK(c, hipMalloc(&dxs.p, sizeof(double) * xs.size()));
    HIPCHECK(c, hipMalloc(&dfm.p, sizeof(double) * fm.size()));
    HIPCHECK(c, hipMalloc(&dsrc.p, sizeof(int) * src.size()));
    HIPCHECK(c, hipMalloc(&dm.p, sizeof(double) * nbatch * nn));
    HIPCHECK(c, hipMalloc(&dt.p, sizeof(double) * nbatch * 2 * n * NG));
    HIPCHECK(c, hipMalloc(&dx.p, sizeof(double) * nbatch * np));
    HIPCHECK(c, hipMalloc(&dout.p, sizeof(double) * 3 * NT * np));
    HIPCHECK(c, hipMemcpyAsync(dg.p, g.data(), sizeof(double) * g.size(), hipMemcpyHostToDevice, c->stream));
    HIPCHECK(c, hipMemcpyAsync(dxc.p, xc.data(), sizeof(double) * xc.size(), hipMemcpyHostToDevice, c->stream));
    HIPCHECK(c, hipMemcpyAsync(dxs.p, xs.data(), sizeof(double) * xs.size(), hipMemcpyHostToDevice, c->stream));
    HIPCHECK(c, hipMemcpyAsync(dfm.p, fm.data(), sizeof(double) * fm.size(), hipMemcpyHostToDevice, c->stream));
    HIPCHECK(c, hipMemcpyAsync(dsrc.p, src.data(), sizeof(int) * src.size(), hipMemcpyHostToDevice, c->stream));
    // M_q(theta) of each parity, batch (theta, parity, q)
    for (int pi = 0; pi < npar; ++pi)
        hipLaunchKernelGGL(k_asym_m, dim3((unsigned)((nn + 255) / 256), NT), dim3(256), 0, c->stream, maxj, par0 + pi,
                           npar, pi, dg.d() + (size_t)(par0 + pi) * (maxj + 1) * 3 * nn, dm.d());
    HIPCHECK(c, hipGetLastError());
    // T = M [cos(k chi); sin(k chi)] stacked as [2n][361], then X = [cos(m phi) | -sin(m phi)] T
    const dim3 g1((NG + 63) / 64, (n + 63) / 64, (unsigned)nbatch), g2((NG + 63) / 64, (NG + 63) / 64, (unsigned)nbatch);
    hipLaunchKernelGGL(k_dgemm, g1, dim3(256), 0, c->stream, n, NG, n, dm.d(), n, (long long)nn, dxc.d(), NG, 0LL,
                       dt.d(), NG, (long long)2 * n * NG);
    hipLaunchKernelGGL(k_dgemm, g1, dim3(256), 0, c->stream, n, NG, n, dm.d(), n, (long long)nn, dxs.d(), NG, 0LL,
                       dt.d() + (size_t)n * NG, NG, (long long)2 * n * NG);
    hipLaunchKernelGGL(k_dgemm, g2, dim3(256), 0, c->stream, NG, NG, 2 * n, dfm.d(), 2 * n, 0LL, dt.d(), NG,
                       (long long)2 * n * NG, dx.d(), NG, (long long)np);
    HIPCHECK(c, hipGetLastError());
    hipLaunchKernelGGL(k_asym_fill, dim3((unsigned)((np + 255) / 256), NT), dim3(256), 0, c->stream, npar,
                       static_cast<const int*>(dsrc.p), dx.d(), dout.d(), dout.d() + NT * np, dout.d() + 2 * NT * np);
    HIPCHECK(c, hipGetLastError());
    HIPCHECK(c, hipMemcpyAsync(rho, dout.d(), sizeof(double) * NT * np, hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(c, hipMemcpyAsync(eng, dout.d() + NT * np, sizeof(double) * NT * np, hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(c, hipMemcpyAsync(esq, dout.d() + 2 * NT * np, sizeof(double) * NT * np, hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int pimc_add_rot3d(pimc_ctx* c, const double* rho, const double* erot, const double* esq) {
    if (c->committed) return fail(c, "tables must be added before pimc_commit");
    if (!rho || !erot || !esq) return fail(c, "init_rot3D: null table");
    const size_t one = (size_t)ROT3D_SIZE * sizeof(double4);
    double4* buf = nullptr;
    if (hipMalloc(&buf, 3 * one * (size_t)(c->nr3 + 1)) != hipSuccess) return fail(c, "hipMalloc failed (rot3d)");
    if (c->nr3 > 0) {
        HIPCHECK(c, hipMemcpy(buf, c->d_r3, 3 * one * (size_t)c->nr3, hipMemcpyDeviceToDevice));
        (void)hipFree(c->d_r3);
    }
    c->d_r3 = buf;
    double4* dst = buf + (size_t)3 * ROT3D_SIZE * c->nr3;
    std::vector<double4> pts;
    const double* src[3] = {rho, erot, esq};
    for (int q = 0; q < 3; ++q) {
        rot3d_points(src[q], pts);
        HIPCHECK(c, hipMemcpy(dst + (size_t)q * ROT3D_SIZE, pts.data(), one, hipMemcpyHostToDevice));
    }
    return c->nr3++;
}

int pimc_set_replica(pimc_ctx* c, int r, const pimc_replica_desc* d) {
    if (c->committed) return fail(c, "replicas must be set before pimc_commit");
    if (r < 0 || r >= c->nrep) return fail(c, "replica index out of range");
    if (!d) return fail(c, "pimc_set_replica: null description");
    const std::string m = check_replica(*d);
    if (!m.empty()) return fail(c, m);
    c->desc[r] = *d;
    c->desc_set[r] = 1;
    return 0;
}

// a rotating linear rotor: the reference's rotor move stops under MINIMAGE (mc_piqmc.cc:2013-2016)
static bool rotating_linear(const pimc_replica_desc& d) {
    for (int t = 0; t < d.ntypes; ++t)
        if (d.molecule[t] == 1 && d.Prot > 0) return true;
    return false;
}

int pimc_set_minimum_image(pimc_ctx* c, int r, const double* box) {
    if (c->committed) return fail(c, "the minimum image is set before pimc_commit");
    if (r < 0 || r >= c->nrep) return fail(c, "replica index out of range");
    if (!c->desc_set[r]) return fail(c, "pimc_set_minimum_image: pimc_set_replica first");
    if (!box) {
        c->box[r] = {0.0, 0.0, 0.0};
        return 0;
    }
    for (int i = 0; i < 3; ++i)
        if (!(box[i] > 0.0) || !std::isfinite(box[i])) return fail(c, "minimum image: box edges must be positive");
    if (rotating_linear(c->desc[r]))
        return fail(c, "MIN IMAGE for orient pot: the minimum image is not applied to a rotating linear rotor");
    c->box[r] = {box[0], box[1], box[2]};
    return 0;
}

int pimc_check_replica(const pimc_replica_desc* d, char* msg, int len) {
    const std::string m = d ? check_replica(*d) : std::string("null description");
    if (msg && len > 0) std::snprintf(msg, (size_t)len, "%s", m.c_str());
    return m.empty() ? 0 : -1;
}

int pimc_plan_rotor3d_split(int protmax, int nchains, int nrot, int blocks_per_cu, int lds_per_cu, int ncu,
                            int cus_per_xcd) {
    if (protmax < 2 || nchains < 1 || nrot < 1 || nrot > 2 || ncu < 1 || cus_per_xcd < 1) return -1;
    return plan_rotor3d_split(
        protmax, nchains, ncu, cus_per_xcd, [&](int ns) { return rotor3d_part_lds(ns, nrot, sizeof(Frame)); },
        [&](size_t b) { return std::min(blocks_per_cu, (int)((size_t)lds_per_cu / std::max<size_t>(b, 1))); });
}

int pimc_plan_rotor_parts(int nchains, const int* prot, const int* rot_ratio, const int* natoms, int fp64, int lds_kb,
                          int part_slices_max, int blocks_per_cu, int lds_per_cu, int cus_per_xcd, int* parts_per_chain,
                          int* slices_per_part, long* lds_bytes, char* msg, int len) {
    std::string err;
    std::vector<RotorChain> ch;
    for (int i = 0; i < nchains; ++i) {
        if (prot[i] < 2 || prot[i] % 2 || rot_ratio[i] < 1 || natoms[i] < 2) err = "chain " + std::to_string(i) + " out of range";
        ch.push_back(RotorChain{i, prot[i], rot_ratio[i], natoms[i]});
    }
    std::vector<RotPart> parts;
    size_t lds = 0;
    int rc = err.empty() ? 0 : -1;
    if (rc == 0)
        rc = plan_linear_rotor(ch, fp64 ? sizeof(double4) : sizeof(float4), lds_kb, part_slices_max, cus_per_xcd,
                               [&](size_t b) { return std::min(blocks_per_cu, (int)((size_t)lds_per_cu / std::max<size_t>(b, 1))); },
                               parts, lds, err);
    if (rc == 0) {
        for (const RotPart& q : parts)
            if (q.w == 0) {
                parts_per_chain[q.r] = q.W;
                slices_per_part[q.r] = q.k1 - q.k0;
            }
        *lds_bytes = (long)lds;
    }
    if (msg && len > 0) std::snprintf(msg, (size_t)len, "%s", err.c_str());
    return rc;
}

int pimc_set_option(pimc_ctx* c, const char* name, long value) {
    if (!c || !name) return fail(c, "pimc_set_option: null argument");
    if (std::strcmp(name, "profile") == 0) {  // at any time: the next pass is issued accordingly
        if (value != 0 && value != 1) return fail(c, "profile is 0 or 1");
        c->profile = value == 1;
        return 0;
    }
    if (c->committed) return fail(c, "options are set before pimc_commit");
    if (std::strcmp(name, "rotor_part_slices") == 0) {
        if (value < 0 || value > (1 << 20)) return fail(c, "rotor_part_slices out of range");
        c->opt_rotor_part_slices = (int)value;
        return 0;
    }
    static const char* lanes_opt[3] = {"bisect_lanes_small", "bisect_lanes_medium", "bisect_lanes_large"};
    for (int k = 0; k < 3; ++k)
        if (std::strcmp(name, lanes_opt[k]) == 0) {
            if (value != 64 && value != 128 && value != 256)
                return fail(c, std::string(name) + " is 64, 128 or 256");
            c->bw_lanes[k] = (int)value;
            return 0;
        }
    if (std::strcmp(name, "centroid_parts") == 0) {
        if (value < 0 || value > 4096) return fail(c, "centroid_parts is 0 to 4096");
        c->opt_centroid_parts = (int)value;
        return 0;
    }
    if (std::strcmp(name, "rotor_lds_kb") == 0) {
        if (value < 8 || value > 150) return fail(c, "rotor_lds_kb is 8 to 150");
        c->opt_rotor_lds_kb = (int)value;
        return 0;
    }
    if (std::strcmp(name, "samples_per_pass") == 0) {
        if (value < 1 || value > 4096) return fail(c, "samples_per_pass is 1 to 4096");
        c->opt_samples_per_pass = (int)value;
        return 0;
    }
    if (std::strcmp(name, "delayed_acceptance") == 0) {
        if (value != 0 && value != 1) return fail(c, "delayed_acceptance is 0 or 1");
        c->opt_da = value == 1;
        return 0;
    }
    return fail(c, std::string("unknown option ") + name);
}

// Linear-rotor parts: slice ranges whose partner geometry (vec4 per translational slice and
// partner) stays within an LDS budget of lds_kb; the heaviest chains first, the parts of a
// chain adjacent and on one XCD (its boundary exchange and its .rot table stay in one L2).
// Called at commit and again by the autotuner, which compares budgets.
static int plan_rotor_parts(pimc_ctx* c, int lds_kb) {
    const size_t geo_bytes = c->fp64 ? sizeof(double4) : sizeof(float4);
    hipDeviceProp_t prop;
    HIPCHECK(c, hipGetDeviceProperties(&prop, 0));
    constexpr int NXCD = 8;
    std::vector<RotorChain> chains;
    for (int r : c->rot_chains) chains.push_back(RotorChain{r, c->rep[r].Prot, c->rep[r].RotRatio, c->rep[r].natoms});
    // resident parts per CU at a launch's LDS size, as the runtime computes it for k_rotor
    auto occupancy = [&](size_t lds) {
        int occ = 0;
        hipError_t e = c->fp64 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_rotor<double>, 256, lds)
                               : hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_rotor<float>, 256, lds);
        return e == hipSuccess ? occ : 0;
    };
    std::vector<RotPart> parts;
    size_t lds_r = 0;
    std::string err;
    if (plan_linear_rotor(chains, geo_bytes, lds_kb, c->opt_rotor_part_slices, prop.multiProcessorCount / NXCD, occupancy,
                          parts, lds_r, err))
        return fail(c, err);
    const int nflags = (int)parts.size();
    parts = xcd_order(parts, [&](int r) {
        const Rep& p = c->rep[r];
        return (double)p.P * p.Prot * p.natoms;
    });
    if (c->d_parts) (void)hipFree(c->d_parts);
    if (c->d_flags) (void)hipFree(c->d_flags);
    if (c->d_xrec) (void)hipFree(c->d_xrec);
    c->d_parts = nullptr;
    c->d_flags = nullptr;
    c->d_xrec = nullptr;
    c->lds_rotor = lds_r;
    c->nflags = nflags;
    c->nparts = (int)parts.size();
    c->rotor_lds_kb = lds_kb;
    if (parts.empty()) return 0;
    c->d_parts = dalloc<RotPart>(parts.size());
    c->d_flags = dalloc<unsigned>(nflags);
    c->d_xrec = dalloc<unsigned long long>(16 * (size_t)nflags);
    if (!c->d_parts || !c->d_flags || !c->d_xrec) return fail(c, "hipMalloc failed");
    HIPCHECK(c, hipMemcpy(c->d_parts, parts.data(), sizeof(RotPart) * parts.size(), hipMemcpyHostToDevice));
    HIPCHECK(c, hipMemset(c->d_flags, 0, sizeof(unsigned) * nflags));
    HIPCHECK(c, hipMemset(c->d_xrec, 0, sizeof(unsigned long long) * 16 * nflags));  // no record matches a wanted number
    return 0;
}

// Centroid parts: each replica's slices split so that parts carry similar shares of the work
// sum N^2 P, about `target` parts in all (0: min(1024, half the resident capacity)), at most 64
// parts and at least 8 slices a part.  Every part of a replica must be resident while its parts
// wait for one another, so the target stays within what the device holds at once.  Called at
// commit and again by the autotuner.
static int plan_centroid_parts(pimc_ctx* c, int target_parts) {
    int occ0 = 0, occ1 = 0;
    hipDeviceProp_t prop;
    HIPCHECK(c, hipGetDeviceProperties(&prop, 0));
    const int ncu = prop.multiProcessorCount;
    if (c->fp64) {
        HIPCHECK(c, hipOccupancyMaxActiveBlocksPerMultiprocessor(
                        &occ0, (c->any_mi ? k_centroid_parts<double, 0, true> : k_centroid_parts<double, 0, false>), 256,
                        c->lds_knots));
        HIPCHECK(c, hipOccupancyMaxActiveBlocksPerMultiprocessor(
                        &occ1, (c->any_mi ? k_centroid_parts<double, 1, true> : k_centroid_parts<double, 1, false>), 256,
                        c->lds_knots));
    } else {
        HIPCHECK(c, hipOccupancyMaxActiveBlocksPerMultiprocessor(
                        &occ0, (c->any_mi ? k_centroid_parts<float, 0, true> : k_centroid_parts<float, 0, false>), 256,
                        c->lds_knots));
        HIPCHECK(c, hipOccupancyMaxActiveBlocksPerMultiprocessor(
                        &occ1, (c->any_mi ? k_centroid_parts<float, 1, true> : k_centroid_parts<float, 1, false>), 256,
                        c->lds_knots));
    }
    const double capacity = (double)ncu * std::max(1, std::min(occ0, occ1));
    // the parts of a replica wait on one another on one XCD: at most what one XCD holds
    const int per_xcd = std::max(1, std::min(occ0, occ1)) * (ncu / 8);
    const double target = target_parts > 0 ? std::min((double)target_parts, capacity / 2.0) : std::min(1024.0, capacity / 2.0);
    std::vector<RotPart> cp[2];
    double total = 0.0;
    for (int r = 0; r < c->nrep; ++r) total += (double)c->rep[r].natoms * c->rep[r].natoms * c->rep[r].P;
    std::vector<int> order(c->nrep);
    for (int r = 0; r < c->nrep; ++r) order[r] = r;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
        const Rep &a = c->rep[x], &b = c->rep[y];
        return (double)a.natoms * a.natoms * a.P > (double)b.natoms * b.natoms * b.P;
    });
    int fb = 0;
    for (int r : order) {
        const Rep& p = c->rep[r];
        const double w = (double)p.natoms * p.natoms * p.P;
        int W = (int)std::ceil(target * w / std::max(total, 1.0));
        W = std::max(1, std::min({W, 64, per_xcd, std::max(1, p.P / 8)}));
        for (int q = 0; q < W; ++q) {
            RotPart x{};
            x.r = r;
            x.k0 = (int)((long long)p.P * q / W);
            x.k1 = (int)((long long)p.P * (q + 1) / W);
            x.w = q;
            x.W = W;
            x.fbase = fb;
            cp[p.fam].push_back(x);
        }
        fb += W;
    }
    auto cost = [&](int r) { return (double)c->rep[r].natoms * c->rep[r].natoms * c->rep[r].P; };
    for (int f = 0; f < 2; ++f)
        if (!cp[f].empty()) cp[f] = xcd_order(cp[f], cost);
    for (void* q : {(void*)c->d_cparts0, (void*)c->d_cparts1, (void*)c->d_cflags, (void*)c->d_cslots})
        if (q) (void)hipFree(q);
    c->ncflags = fb;
    c->centroid_target = (int)target;
    c->ncp0 = (int)cp[0].size();
    c->ncp1 = (int)cp[1].size();
    c->d_cparts0 = dalloc<RotPart>(cp[0].size());
    c->d_cparts1 = dalloc<RotPart>(cp[1].size());
    c->d_cflags = dalloc<unsigned>((size_t)std::max(fb, 1));
    c->d_cslots = dalloc<unsigned long long>(4 * (size_t)std::max(fb, 1));
    if (!c->d_cparts0 || !c->d_cparts1 || !c->d_cflags || !c->d_cslots) return fail(c, "hipMalloc failed");
    HIPCHECK(c, hipMemset(c->d_cflags, 0, sizeof(unsigned) * std::max(fb, 1)));
    HIPCHECK(c, hipMemset(c->d_cslots, 0, sizeof(unsigned long long) * 4 * std::max(fb, 1)));  // no word matches a wanted number
    if (!cp[0].empty())
        HIPCHECK(c, hipMemcpy(c->d_cparts0, cp[0].data(), sizeof(RotPart) * cp[0].size(), hipMemcpyHostToDevice));
    if (!cp[1].empty())
        HIPCHECK(c, hipMemcpy(c->d_cparts1, cp[1].data(), sizeof(RotPart) * cp[1].size(), hipMemcpyHostToDevice));
    return 0;
}

int pimc_commit(pimc_ctx* c) {
    if (c->committed) return fail(c, "already committed");
    for (int r = 0; r < c->nrep; ++r)
        if (!c->desc_set[r]) return fail(c, "replica " + std::to_string(r) + " not set");
    const double lam = 100.0 * (kHBAR * kHBAR) / (kAMU * kKB);
    c->rep.resize(c->nrep);
    long long boff = 0, roff = 0, hoff = 0, rcfoff = 0;
    std::vector<int> rcfreps;
    std::vector<int2> work, rows, work1, rows1;
    std::vector<int> rotreps, all0, all1, rot3;
    size_t lds_b1 = 0;
    long long rowoff = 0;
    size_t lds_b = 0;
    c->bead_moves_per_pass = 0;
    for (int r = 0; r < c->nrep; ++r) {
        const pimc_replica_desc& d = c->desc[r];
        Rep& p = c->rep[r];
        std::memset(&p, 0, sizeof p);
        p.ntypes = d.ntypes;
        p.P = d.P;
        p.Prot = d.Prot > 0 ? d.Prot : d.P;
        p.RotRatio = p.P / p.Prot;
        p.T = d.T;
        p.beta = 1.0 / d.T;
        p.tau = p.beta / (double)d.P;
        p.rtstep = d.rtstep;
        p.imtype = -1;
        p.rotkind = 0;
        p.rot_atom = -1;
        p.lin_atom = -1;
        p.pot1d_id = -1;
        p.pot2d_id = -1;
        p.key = r;
        p.rotdens = d.rotdens;
        p.fam = 0;
        p.nl_atom = -1;
        p.nrot = 0;
        p.pot3d_id = d.pot3d;
        p.rot3d_id = d.rot3d;
        if (c->box[r][0] > 0.0) {
            if (rotating_linear(d))
                return fail(c, "MIN IMAGE for orient pot: the minimum image is not applied to a rotating linear rotor");
            p.minimage = 1;
            c->any_mi = true;
            for (int i = 0; i < 3; ++i) p.box[i] = c->box[r][i];
        }
        int g = 0, S = 1;
        for (int t = 0; t < d.ntypes; ++t) {
            p.tnumb[t] = d.numb[t];
            p.tmol[t] = d.molecule[t];
            p.tlev[t] = d.levels[t];
            p.tbegin[t] = g;
            p.lambda[t] = 0.5 * lam / d.mass[t];
            p.mcstep[t] = d.mcstep[t];
            p.pot1d[t] = d.pot1d[t];
            p.pot2d[t] = d.pot2d[t];
            S = std::max(S, 1 << d.levels[t]);
            if (d.molecule[t] == 0) p.pot1d_id = d.pot1d[t];
            if (d.molecule[t] == 1) {
                p.lin_atom = g;
                p.pot2d_id = d.pot2d[t];
            }
            if (d.molecule[t] == 1 && d.Prot > 0) {
                p.imtype = t;
                p.rotkind = 1;
                p.rot_atom = g;
            }
            if (d.molecule[t] == 2) {
                p.fam = 1;
                p.imtype = t;
                p.nl_atom = g;
                p.rot_atom = g;
                p.nrot = d.numb[t];
                if (d.Prot > 0) p.rotkind = 2;
            }
            g += d.numb[t];
            c->bead_moves_per_pass += (double)d.numb[t] * d.P * ((1 << d.levels[t]) - 1) + (double)d.numb[t] * d.P;
        }
        p.natoms = g;
        p.S = S;
        for (int t = 0; t < d.ntypes; ++t)
            if (d.levels[t] < 1 || d.levels[t] > 8) return fail(c, "bisection levels must be 1..8");
        if (!c->ref_order && d.P % S) return fail(c, "segment order needs P divisible by the segment size");
        // table checks
        int n_plain = 0;  // structureless atoms, the partners of the rotor-atom potentials
        for (int t = 0; t < d.ntypes; ++t) n_plain += d.molecule[t] == 0 ? d.numb[t] : 0;
        for (int t = 0; t < d.ntypes; ++t) {
            if (d.molecule[t] == 0 && (d.pot1d[t] < 0 || d.pot1d[t] >= (int)c->t1.size()))
                return fail(c, "atom type without a 1D potential table");
            // a lone rotor has no rotor-atom pairs and needs no 2D table
            if (d.molecule[t] == 1 && (d.pot2d[t] >= (int)c->t2.size() || (n_plain > 0 && d.pot2d[t] < 0)))
                return fail(c, "linear rotor without a 2D potential table");
        }
        if (p.fam == 1) {
            if (p.pot3d_id >= (int)c->t3.size() || (p.natoms > p.nrot && p.pot3d_id < 0))
                return fail(c, "non-linear rotor without a 3D potential table");
            if (p.rotkind == 2 && (p.rot3d_id < 0 || p.rot3d_id >= c->nr3))
                return fail(c, "rotating non-linear rotor without rho/E/E^2 tables");
            if (p.rotkind == 2 && c->opt_da && !(p.natoms > p.nrot ? p.pot3d_id >= 0 : p.nrot == 2))
                return fail(c, "delayed acceptance needs atoms with a 3D potential (rotor-atom network) or two "
                               "rotors (rotor-rotor network)");
            if (p.rotkind == 2) {
                c->bead_moves_per_pass += (double)p.P * p.Prot * p.nrot;
                rot3.push_back(r);
                c->nrotmax = std::max(c->nrotmax, p.nrot);
            }
        }
        if (p.rotkind == 1) {
            if (d.rotdens < 0 || d.rotdens >= (int)c->tr.size()) return fail(c, "linear rotor without a .rot table");
            if (p.Prot % 2) return fail(c, "the checkerboard rotor sweep needs an even number of rotor slices");
            c->bead_moves_per_pass += (double)p.P * p.Prot;
            rotreps.push_back(r);
        }
        p.bead_off = boff;
        boff += (long long)p.natoms * p.P;
        p.rot_off = roff;
        roff += ((p.rotkind || p.fam == 1) ? (long long)p.tnumb[p.imtype] * p.Prot : 0);
        p.hist_off = hoff;
        {  // atom-atom g(r), the atoms around a linear rotor, then the bins around a non-linear rotor
            long long h = BINSR + BINSR * BINST;
            bool atoms = false;
            for (int t = 0; t < p.ntypes; ++t) atoms |= p.tmol[t] == 0 && p.tnumb[t] > 0;
            p.h3_off = -1;
            p.hrel_off = -1;
            if (p.fam == 1 && p.nl_atom >= 0 && atoms) {
                p.h3_off = (int)h;
                h += (long long)BINSR * BINST * BINSC;
            }
            if (p.rotkind == 2 && p.rot3d_id >= 0) {
                p.hrel_off = (int)h;
                h += BINST + 2 * BINSC;
            }
            hoff += h;
        }
        p.rcf_off = -1;
        if (p.rotkind) {  // GetRCF runs with ROTATION (mc_main.cc:571-587)
            p.rcf_off = rcfoff;
            rcfoff += (long long)RCF_L * p.Prot;
            rcfreps.push_back(r);
            c->rcf_protmax = std::max(c->rcf_protmax, p.Prot);
        }
        // estimator tiles: the largest power of two of slices whose window of all atoms
        // (one slice more than the tile) stays within 1152 beads, or two slices
        int ts = 1;
        while (2 * ts <= p.P && (long long)p.natoms * (2 * ts + 1) <= std::max(1152, 2 * p.natoms)) ts *= 2;
        p.est_ts = ts;
        const int ntiles = (p.P + ts - 1) / ts;
        p.row_off = rowoff;
        for (int k = 0; k < ntiles; ++k) (p.fam ? rows1 : rows).push_back(make_int2(r, k));
        (p.fam ? all1 : all0).push_back(r);
        rowoff += ntiles;
        c->lds_est = std::max(c->lds_est, (size_t)(3 * p.natoms * (ts + 1) + 3 * ts) * sizeof(double) +
                                              (size_t)p.natoms * sizeof(int));
        int W = S + 1;
        // (y, y'') pairs of uniform FP32 knots in the linear family, vec4 knots otherwise
        size_t knots = 0;
        if (p.pot1d_id >= 0) {
            const Tab1D& t = c->t1[p.pot1d_id];
            knots = (size_t)t.n * (c->fp64 ? sizeof(double4) : (p.fam == 0 && t.uniform ? sizeof(float2) : sizeof(float4)));
        }
        // bisection LDS: new positions and the draw ring (FP64), the window, the new midpoints
        // and the rotor axes in the pair precision, frames, knots (the layout of bisect_body)
        const size_t lds_w = (size_t)((3 * W + bisect_ring(W) + 3) & ~3) * sizeof(double) +
                             (size_t)(p.natoms + 2) * W * (c->fp64 ? sizeof(double4) : sizeof(float4)) +
                             (size_t)12 * p.nrot * W * sizeof(double) + knots;
        (p.fam ? lds_b1 : lds_b) = std::max(p.fam ? lds_b1 : lds_b, lds_w);
        if (!p.fam) {
            const int k = bisect_class(p.natoms);
            c->bw_lds[k] = std::max(c->bw_lds[k], lds_w);
        }
        int nseg = c->ref_order ? 1 : p.P / S;
        for (int k = 0; k < nseg; ++k) (p.fam ? work1 : work).push_back(make_int2(r, k));
        c->Nmax = std::max(c->Nmax, p.natoms);
        c->Smax = std::max(c->Smax, S);
        c->Pmax = std::max(c->Pmax, p.P);
        c->max_types = std::max(c->max_types, p.ntypes);
    }
    if (c->opt_samples_per_pass > 1) {  // samples inside a pass: one time grid for every replica
        const int k = c->opt_samples_per_pass;
        for (int r = 0; r < c->nrep; ++r)
            if (c->rep[r].P != c->Pmax || c->rep[r].S != c->Smax)
                return fail(c, "samples_per_pass > 1 needs the same P and segment size in every replica");
        if ((c->ref_order ? c->Pmax : c->Smax) % k)
            return fail(c, c->ref_order ? "samples_per_pass must divide P"
                                        : "samples_per_pass must divide the segment size 2^levels");
    }
    if (std::max(lds_b, lds_b1) > 160 * 1024)
        return fail(c, "segment window exceeds LDS (too many atoms for this segment size)");
    if (c->lds_est > 150 * 1024) return fail(c, "estimator tile exceeds LDS (too many atoms)");
    c->lds_bisect = lds_b;
    c->lds_bisect1 = lds_b1;
    for (const Tab1D& t : c->t1)
        c->lds_knots = std::max(c->lds_knots, (size_t)t.n * (c->fp64 ? sizeof(double4) : sizeof(float4)));
    std::stable_sort(rotreps.begin(), rotreps.end(), [&](int x, int y) {
        const Rep &a = c->rep[x], &b = c->rep[y];
        return (long long)a.P * a.Prot * a.natoms > (long long)b.P * b.Prot * b.natoms;
    });
    c->rot_chains = rotreps;
    if (plan_rotor_parts(c, c->opt_rotor_lds_kb)) return -1;
    if (plan_centroid_parts(c, c->opt_centroid_parts)) return -1;
    c->nbeads = boff;
    c->nrotb = roff;
    c->nhist = hoff;
    c->nrcf = (int)rcfreps.size();
    c->nrcfvals = rcfoff;
    c->d_rcf = dalloc<double>(std::max<long long>(rcfoff, 1));
    c->d_rcfreps = dalloc<int>(rcfreps.size());
    if (!c->d_rcf || !c->d_rcfreps) return fail(c, "hipMalloc failed");
    HIPCHECK(c, hipMemset(c->d_rcf, 0, sizeof(double) * std::max<long long>(rcfoff, 1)));
    if (!rcfreps.empty())
        HIPCHECK(c, hipMemcpy(c->d_rcfreps, rcfreps.data(), sizeof(int) * rcfreps.size(), hipMemcpyHostToDevice));
    // largest windows first: the longest segment sequences start while the grid fills
    auto heavier = [&](const int2& x, const int2& y) {
        const Rep &a = c->rep[x.x], &b = c->rep[y.x];
        return (long long)a.natoms * a.natoms * a.S > (long long)b.natoms * b.natoms * b.S;
    };
    std::stable_sort(work.begin(), work.end(), heavier);
    std::stable_sort(work1.begin(), work1.end(), heavier);
    {  // size classes of the linear family, heaviest class first in the work array
        std::vector<int2> byc[3];
        for (const int2& w : work) byc[bisect_class(c->rep[w.x].natoms)].push_back(w);
        work.clear();
        for (int k = 2; k >= 0; --k) {
            c->bw_off[k] = (int)work.size();
            c->bw_n[k] = (int)byc[k].size();
            work.insert(work.end(), byc[k].begin(), byc[k].end());
        }
        int busy = 0;
        for (int k = 0; k < 3; ++k) busy += c->bw_n[k] > 0;
        if (busy > 1)
            for (int k = 0; k < 2; ++k) {
                HIPCHECK(c, hipStreamCreate(&c->side[k]));  // default priority, like the main stream
                HIPCHECK(c, hipEventCreateWithFlags(&c->ev_join[k], hipEventDisableTiming));
            }
        if (busy > 1) HIPCHECK(c, hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
    }
    // Non-linear chains split by slice range when whole chains would leave CUs idle: as many
    // parts (powers of two, at least 16 slices each) as the device holds resident at once, since
    // the parts of a chain wait for one another.  With delayed acceptance every chain runs as
    // parts of k_rotor3d_da, a short chain as one part.
    if (!rot3.empty()) {
        hipDeviceProp_t prop;
        HIPCHECK(c, hipGetDeviceProperties(&prop, 0));
        int protmax = 2, nrmax = 1;
        for (int r : rot3) {
            protmax = std::max(protmax, c->rep[r].Prot);
            nrmax = std::max(nrmax, c->rep[r].nrot);
        }
        const bool da = c->opt_da;
        auto lds_for = [&](int ns) {
            return da ? rotor3d_da_part_lds(ns, nrmax, sizeof(Frame), sizeof(MlpPacked))
                      : rotor3d_part_lds(ns, nrmax, sizeof(Frame));
        };
        auto occupancy = [&](size_t lds) {
            int occ = 0;
            hipError_t e;
            if (da) e = c->fp64 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_rotor3d_da<double>, 256, lds)
                                : hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_rotor3d_da<float>, 256, lds);
            else e = c->fp64 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_rotor3d_parts<double>, 256, lds)
                             : hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_rotor3d_parts<float>, 256, lds);
            return e == hipSuccess ? occ : 0;
        };
        const int W = plan_rotor3d_split(protmax, (int)rot3.size(), prop.multiProcessorCount,
                                         prop.multiProcessorCount / 8, lds_for, occupancy);
        if (W > 1 || da) {
            std::vector<RotPart> p3;
            std::vector<int> whole;
            for (int r : rot3) {
                const Rep& p = c->rep[r];
                int ns = ((p.Prot + W - 1) / W + 1) & ~1;
                if (ns > p.Prot) ns = p.Prot;
                const int Wr = (p.Prot + ns - 1) / ns;
                if (Wr < 2 && !da) {
                    whole.push_back(r);
                    continue;
                }
                const int fbase = (int)p3.size();
                for (int w = 0; w < Wr; ++w) {
                    RotPart q{};
                    q.r = r;
                    q.k0 = w * ns;
                    q.k1 = std::min(p.Prot, (w + 1) * ns);
                    q.w = w;
                    q.W = Wr;
                    q.fbase = fbase;
                    p3.push_back(q);
                    c->lds_parts3 = std::max(c->lds_parts3, lds_for(q.k1 - q.k0));
                }
            }
            if (c->lds_parts3 > ROTOR_LDS_MAX) return fail(c, "non-linear rotor part exceeds LDS");
            rot3.swap(whole);
            const int nflags3 = (int)p3.size();
            // two records per part, each 2 words per double of nrmax frames, then nrmax link densities
            c->pub3_stride = 4 * nrmax * ((int)(sizeof(Frame) / sizeof(double)) + 1);
            p3 = xcd_order(p3, [&](int r) { return (double)c->rep[r].Prot * c->rep[r].nrot; });
            c->nparts3 = (int)p3.size();
            c->d_parts3 = dalloc<RotPart>(p3.size());
            c->d_flags3 = dalloc<unsigned>((size_t)nflags3);
            c->d_pub3 = dalloc<unsigned long long>((size_t)nflags3 * c->pub3_stride);
            if (!c->d_parts3 || !c->d_flags3 || !c->d_pub3) return fail(c, "hipMalloc failed");
            HIPCHECK(c, hipMemcpy(c->d_parts3, p3.data(), sizeof(RotPart) * p3.size(), hipMemcpyHostToDevice));
            HIPCHECK(c, hipMemset(c->d_flags3, 0, sizeof(unsigned) * nflags3));
            HIPCHECK(c, hipMemset(c->d_pub3, 0, sizeof(unsigned long long) * nflags3 * c->pub3_stride));
        }
    }
    c->nwork = (int)work.size();
    c->nwork1 = (int)work1.size();
    c->nall0 = (int)all0.size();
    c->nall1 = (int)all1.size();
    c->nrot3 = (int)rot3.size();
    c->nrows1 = (int)rows1.size();
    c->d_work1 = dalloc<int2>(work1.size());
    c->d_all1 = dalloc<int>(all1.size());
    c->d_rot3 = dalloc<int>(rot3.size());
    c->d_rows1 = dalloc<int2>(rows1.size());
    if (!c->d_work1 || !c->d_all1 || !c->d_rot3 || !c->d_rows1) return fail(c, "hipMalloc failed");
    auto upv = [&](void* d, const void* h, size_t bytes) -> int {
        if (bytes == 0) return 0;
        return hipMemcpy(d, h, bytes, hipMemcpyHostToDevice) == hipSuccess ? 0 : -1;
    };
    if (upv(c->d_work1, work1.data(), work1.size() * sizeof(int2)) || upv(c->d_all1, all1.data(), all1.size() * sizeof(int)) ||
        upv(c->d_rot3, rot3.data(), rot3.size() * sizeof(int)) || upv(c->d_rows1, rows1.data(), rows1.size() * sizeof(int2)))
        return fail(c, "upload failed");
    c->nrotreps = (int)rotreps.size();
    c->nrows = (int)rows.size();
    c->d_rows = dalloc<int2>(rows.size());
    c->d_rowbuf = dalloc<double2>((size_t)rowoff);
    if (!c->d_rows || !c->d_rowbuf) return fail(c, "hipMalloc failed");
    HIPCHECK(c, hipMemcpy(c->d_rows, rows.data(), sizeof(int2) * rows.size(), hipMemcpyHostToDevice));
    // device allocations
    c->d_rep = dalloc<Rep>(c->nrep);
    c->d_mut = dalloc<Mut>(c->nrep);
    c->d_beads = dalloc<double4>(boff);
    c->d_rotb = dalloc<RotBead>(roff);
    c->d_hist = dalloc<unsigned>(hoff);
    c->d_work = dalloc<int2>(work.size());
    c->d_all = dalloc<int>(c->nrep);
    c->d_rotreps = dalloc<int>(rotreps.size());
    c->d_est = dalloc<double>(8 * (size_t)c->nrep);
    c->sample_cap = 4096;
    c->d_samples = dalloc<double>((size_t)8 * c->sample_cap * c->nrep);
    for (int r = 0; r < c->nrep; ++r) c->rep[r].sample_off = (long long)8 * c->sample_cap * r;
    if (!c->d_rep || !c->d_mut || !c->d_beads || !c->d_rotb || !c->d_hist || !c->d_work || !c->d_all ||
        !c->d_rotreps || !c->d_est || !c->d_samples)
        return fail(c, "hipMalloc failed");
    HIPCHECK(c, hipMemcpy(c->d_rep, c->rep.data(), sizeof(Rep) * c->nrep, hipMemcpyHostToDevice));
    HIPCHECK(c, hipMemset(c->d_mut, 0, sizeof(Mut) * c->nrep));
    HIPCHECK(c, hipMemset(c->d_hist, 0, sizeof(unsigned) * std::max<long long>(hoff, 1)));
    HIPCHECK(c, hipMemcpy(c->d_work, work.data(), sizeof(int2) * work.size(), hipMemcpyHostToDevice));
    std::vector<int> all(c->nrep);
    for (int r = 0; r < c->nrep; ++r) all[r] = r;
    HIPCHECK(c, hipMemcpy(c->d_all, all.data(), sizeof(int) * c->nrep, hipMemcpyHostToDevice));
    c->d_all0 = dalloc<int>(all0.size());
    c->d_frames = dalloc<Frame>(std::max<long long>(roff, 1));
    if (!c->d_all0 || !c->d_frames) return fail(c, "hipMalloc failed");
    if (!all0.empty()) HIPCHECK(c, hipMemcpy(c->d_all0, all0.data(), sizeof(int) * all0.size(), hipMemcpyHostToDevice));
    {  // frames of the default orientation (phi = 0, cos theta = 1, chi = 0)
        std::vector<Frame> fr(std::max<long long>(roff, 1), make_frame(0.0, 1.0, 0.0));
        HIPCHECK(c, hipMemcpy(c->d_frames, fr.data(), sizeof(Frame) * fr.size(), hipMemcpyHostToDevice));
    }
    {  // 3D potential tables: (V, dV_r, dV_theta, dV_chi) per point, FP64 and FP32, the
       // differences those vcalc forms (vcalc.f:33-58), zero at an axis' last point
        std::vector<double4> pd(c->v3.size());
        std::vector<float4> pf(c->v3.size());
        for (const Tab3D& t : c->t3)
            for (int ir = 0; ir < t.nr; ++ir)
                for (int ith = 0; ith < t.nth; ++ith)
                    for (int ich = 0; ich < t.nch; ++ich) {
                        const size_t i0 = (size_t)t.off + ((size_t)ir * t.nth + ith) * t.nch + ich;
                        const size_t ir1 = i0 + (size_t)t.nth * t.nch, it1 = i0 + t.nch, ic1 = i0 + 1;
                        const double* v = c->v3.data();
                        const bool lr = ir == t.nr - 1, lt = ith == t.nth - 1, lc = ich == t.nch - 1;
                        pd[i0] = make_double4(v[i0], lr ? 0.0 : v[ir1] - v[i0], lt ? 0.0 : v[it1] - v[i0],
                                              lc ? 0.0 : v[ic1] - v[i0]);
                        const float f0 = (float)v[i0];
                        pf[i0] = make_float4(f0, lr ? 0.0f : (float)v[ir1] - f0, lt ? 0.0f : (float)v[it1] - f0,
                                             lc ? 0.0f : (float)v[ic1] - f0);
                    }
        c->d_v3d = dalloc<double4>(pd.size());
        c->d_v3f = dalloc<float4>(pf.size());
        c->d_t3 = dalloc<Tab3D>(c->t3.size());
        if (!c->d_v3d || !c->d_v3f || !c->d_t3) return fail(c, "hipMalloc failed");
        if (!c->v3.empty()) {
            HIPCHECK(c, hipMemcpy(c->d_v3d, pd.data(), sizeof(double4) * pd.size(), hipMemcpyHostToDevice));
            HIPCHECK(c, hipMemcpy(c->d_v3f, pf.data(), sizeof(float4) * pf.size(), hipMemcpyHostToDevice));
            HIPCHECK(c, hipMemcpy(c->d_t3, c->t3.data(), sizeof(Tab3D) * c->t3.size(), hipMemcpyHostToDevice));
        }
    }
    if (!rotreps.empty())
        HIPCHECK(c, hipMemcpy(c->d_rotreps, rotreps.data(), sizeof(int) * rotreps.size(), hipMemcpyHostToDevice));
    // default state: every bead at the origin of its own lattice site is the caller's job;
    // rotor axes start along z (mc_setup.cc:476-489)
    std::vector<RotBead> rb(std::max<long long>(roff, 1));
    for (auto& b : rb) {
        b.phi = 0; b.cth = 1; b.chi = 0; b.pad0 = 0;
        b.nx = 0; b.ny = 0; b.nz = 1; b.pad1 = 0;
    }
    HIPCHECK(c, hipMemcpy(c->d_rotb, rb.data(), sizeof(RotBead) * rb.size(), hipMemcpyHostToDevice));
    HIPCHECK(c, hipMemset(c->d_beads, 0, sizeof(double4) * std::max<long long>(boff, 1)));
    // tables
    std::vector<float4> k1f(c->k1.size());
    for (size_t i = 0; i < c->k1.size(); ++i)
        k1f[i] = make_float4((float)c->k1[i].x, (float)c->k1[i].y, (float)c->k1[i].z, (float)c->k1[i].w);
    std::vector<float> v2f(c->v2.begin(), c->v2.end()), g2f(c->g2.begin(), c->g2.end());
    // uniform 2D tables also as one float4 of corner values per cell (one load per FP32 look-up)
    std::vector<float4> v2c;
    for (Tab2D& t : c->t2) {
        t.coff = -1;
        if (!t.uniform) continue;
        t.coff = (long long)v2c.size();
        for (int ir = 0; ir + 1 < t.nr; ++ir)
            for (int ic = 0; ic + 1 < t.nc; ++ic) {
                const float* v = v2f.data() + t.off + (long long)ir * t.nc + ic;
                v2c.push_back(make_float4(v[0], v[1], v[t.nc], v[t.nc + 1]));
            }
    }
    c->d_k1f = dalloc<float4>(k1f.size());
    c->d_k1d = dalloc<double4>(c->k1.size());
    c->d_t1 = dalloc<Tab1D>(c->t1.size());
    c->d_v2f = dalloc<float>(v2f.size());
    c->d_v2c = dalloc<float4>(std::max<size_t>(v2c.size(), 1));
    c->d_v2d = dalloc<double>(c->v2.size());
    c->d_g2d = dalloc<double>(c->g2.size());
    c->d_g2f = dalloc<float>(g2f.size());
    c->d_t2 = dalloc<Tab2D>(c->t2.size());
    c->d_rk = dalloc<double4>(c->rk.size());
    c->d_re = dalloc<double4>(c->re.size());
    c->d_tr = dalloc<TabRot>(c->tr.size());
    auto up = [&](void* d, const void* h, size_t bytes) -> int {
        if (bytes == 0) return 0;
        return hipMemcpy(d, h, bytes, hipMemcpyHostToDevice) == hipSuccess ? 0 : -1;
    };
    if (up(c->d_k1f, k1f.data(), k1f.size() * sizeof(float4)) || up(c->d_k1d, c->k1.data(), c->k1.size() * sizeof(double4)) ||
        up(c->d_t1, c->t1.data(), c->t1.size() * sizeof(Tab1D)) || up(c->d_v2f, v2f.data(), v2f.size() * sizeof(float)) ||
        up(c->d_v2c, v2c.data(), v2c.size() * sizeof(float4)) ||
        up(c->d_v2d, c->v2.data(), c->v2.size() * sizeof(double)) ||
        up(c->d_g2d, c->g2.data(), c->g2.size() * sizeof(double)) ||
        up(c->d_g2f, g2f.data(), g2f.size() * sizeof(float)) || up(c->d_t2, c->t2.data(), c->t2.size() * sizeof(Tab2D)) ||
        up(c->d_rk, c->rk.data(), c->rk.size() * sizeof(double4)) ||
        up(c->d_re, c->re.data(), c->re.size() * sizeof(double4)) || up(c->d_tr, c->tr.data(), c->tr.size() * sizeof(TabRot)))
        return fail(c, "table upload failed");
    // non-linear rotor sweep: 512 lanes when the chains are short enough that the lane groups'
    // frames and the link cache fit in LDS, else 256 (groups = lanes / G, G as the kernel picks it);
    // 512 lanes keep a 256-register budget per lane
    {
        int protmax = 2;
        for (int r : rot3) protmax = std::max(protmax, c->rep[r].Prot);
        auto lds_for = [&](int block) {
            int G = 1;
            const int maxmoves = (protmax + 1) / 2;
            while (G < 64 && G * 2 * maxmoves <= block) G *= 2;
            return (size_t)((c->nrotmax * protmax + 1) & ~1) * sizeof(double) +
                   (size_t)(block / G) * (c->nrotmax + 1) * sizeof(Frame);
        };
        c->rot3_block = lds_for(512) <= 64 * 1024 ? 512 : 256;
        c->lds_rot3 = lds_for(c->rot3_block);
        if (!rot3.empty() && c->lds_rot3 > 160 * 1024) return fail(c, "non-linear rotor chain exceeds LDS");
    }
    // bisection block: enough lanes for the widest level of the largest window
    c->bisect_block = c->Nmax <= 24 ? 64 : (c->Nmax <= 64 ? 128 : 256);
    c->committed = true;
    return 0;
}

int pimc_set_state(pimc_ctx* c, int r, const double* coords, const double* angles) {
    if (!c->committed) return fail(c, "pimc_commit first");
    if (r < 0 || r >= c->nrep) return fail(c, "replica index out of range");
    const Rep& p = c->rep[r];
    const int n = p.natoms * p.P;
    if (coords) {
        std::vector<double4> b(n);
        for (int i = 0; i < n; ++i) b[i] = make_double4(coords[i], coords[n + i], coords[2 * n + i], 0.0);
        HIPCHECK(c, hipMemcpy(c->d_beads + p.bead_off, b.data(), sizeof(double4) * n, hipMemcpyHostToDevice));
    }
    if ((p.rotkind || p.fam == 1) && angles) {
        const int nr = p.tnumb[p.imtype];
        std::vector<RotBead> rb((size_t)nr * p.Prot);
        std::vector<Frame> fr((size_t)nr * p.Prot);
        for (int a = 0; a < nr; ++a)
            for (int k = 0; k < p.Prot; ++k) {
                int i = (p.tbegin[p.imtype] + a) * p.P + k;  // rotor slice k lives at offset + k
                RotBead& b = rb[(size_t)a * p.Prot + k];
                b.phi = angles[i];
                b.cth = angles[n + i];
                b.chi = angles[2 * n + i];
                double sint = std::sqrt(1.0 - b.cth * b.cth);
                b.nx = sint * std::cos(b.phi);
                b.ny = sint * std::sin(b.phi);
                b.nz = b.cth;
                b.pad0 = b.pad1 = 0;
                fr[(size_t)a * p.Prot + k] = make_frame(b.phi, b.cth, b.chi);
            }
        HIPCHECK(c, hipMemcpy(c->d_rotb + p.rot_off, rb.data(), sizeof(RotBead) * rb.size(), hipMemcpyHostToDevice));
        if (p.fam == 1)
            HIPCHECK(c, hipMemcpy(c->d_frames + p.rot_off, fr.data(), sizeof(Frame) * fr.size(), hipMemcpyHostToDevice));
    }
    return 0;
}

int pimc_restore_state(pimc_ctx* c, int r, const double* coords, const double* angles, const double* cosines,
                       unsigned pass) {
    if (!c->committed) return fail(c, "pimc_commit first");
    if (r < 0 || r >= c->nrep) return fail(c, "replica index out of range");
    if (!coords || !angles || !cosines) return fail(c, "pimc_restore_state: null argument");
    const Rep& p = c->rep[r];
    const int n = p.natoms * p.P;
    std::vector<double4> b(n);
    for (int i = 0; i < n; ++i) b[i] = make_double4(coords[i], coords[n + i], coords[2 * n + i], 0.0);
    HIPCHECK(c, hipMemcpy(c->d_beads + p.bead_off, b.data(), sizeof(double4) * n, hipMemcpyHostToDevice));
    if (p.rotkind || p.fam == 1) {
        const int nr = p.tnumb[p.imtype];
        std::vector<RotBead> rb((size_t)nr * p.Prot);
        for (int a = 0; a < nr; ++a)
            for (int k = 0; k < p.Prot; ++k) {
                const int i = (p.tbegin[p.imtype] + a) * p.P + k;
                RotBead& q = rb[(size_t)a * p.Prot + k];
                q.phi = angles[i];
                q.cth = angles[n + i];
                q.chi = angles[2 * n + i];
                q.nx = cosines[i];
                q.ny = cosines[n + i];
                q.nz = cosines[2 * n + i];
                q.pad0 = q.pad1 = 0;
            }
        HIPCHECK(c, hipMemcpy(c->d_rotb + p.rot_off, rb.data(), sizeof(RotBead) * rb.size(), hipMemcpyHostToDevice));
        if (p.fam == 1) {
            const int m = (int)rb.size();
            hipLaunchKernelGGL(k_frames_from_angles, dim3((m + 255) / 256), dim3(256), 0, c->stream, c->d_rep,
                               c->d_rotb, c->d_frames, r, m);
            HIPCHECK(c, hipGetLastError());
        }
    }
    Mut m;
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    HIPCHECK(c, hipMemcpy(&m, c->d_mut + r, sizeof(Mut), hipMemcpyDeviceToHost));
    m.pass = pass;
    HIPCHECK(c, hipMemcpy(c->d_mut + r, &m, sizeof(Mut), hipMemcpyHostToDevice));
    return 0;
}

int pimc_get_pass(pimc_ctx* c, int r, unsigned* pass) {
    if (!c->committed || r < 0 || r >= c->nrep || !pass) return fail(c, "bad replica");
    Mut m;
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    HIPCHECK(c, hipMemcpy(&m, c->d_mut + r, sizeof(Mut), hipMemcpyDeviceToHost));
    *pass = m.pass;
    return 0;
}

int pimc_get_state(pimc_ctx* c, int r, double* coords, double* angles, double* cosines) {
    if (!c->committed) return fail(c, "pimc_commit first");
    if (r < 0 || r >= c->nrep) return fail(c, "replica index out of range");
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    const Rep& p = c->rep[r];
    const int n = p.natoms * p.P;
    if (coords) {
        std::vector<double4> b(n);
        HIPCHECK(c, hipMemcpy(b.data(), c->d_beads + p.bead_off, sizeof(double4) * n, hipMemcpyDeviceToHost));
        for (int i = 0; i < n; ++i) {
            coords[i] = b[i].x;
            coords[n + i] = b[i].y;
            coords[2 * n + i] = b[i].z;
        }
    }
    if (angles || cosines) {
        for (int i = 0; i < n; ++i) {  // non-rotor entries keep the initial values
            if (angles) { angles[i] = 0; angles[n + i] = 1; angles[2 * n + i] = 0; }
            if (cosines) { cosines[i] = 0; cosines[n + i] = 0; cosines[2 * n + i] = 1; }
        }
        if (p.rotkind || p.fam == 1) {
            const int nr = p.tnumb[p.imtype];
            std::vector<RotBead> rb((size_t)nr * p.Prot);
            HIPCHECK(c, hipMemcpy(rb.data(), c->d_rotb + p.rot_off, sizeof(RotBead) * rb.size(), hipMemcpyDeviceToHost));
            for (int a = 0; a < nr; ++a)
                for (int k = 0; k < p.Prot; ++k) {
                    int i = (p.tbegin[p.imtype] + a) * p.P + k;
                    const RotBead& b = rb[(size_t)a * p.Prot + k];
                    if (angles) { angles[i] = b.phi; angles[n + i] = b.cth; angles[2 * n + i] = b.chi; }
                    if (cosines) { cosines[i] = b.nx; cosines[n + i] = b.ny; cosines[2 * n + i] = b.nz; }
                }
        }
    }
    return 0;
}

}  // extern "C"

// ------------------------------------------------------------------ launches
namespace {

struct Timed {
    pimc_ctx* c;
    const char* name;
    hipEvent_t a = nullptr, b = nullptr;
    Timed(pimc_ctx* c_, const char* n) : c(c_), name(n) {
        if (c->profile) {
            (void)hipEventCreate(&a);
            (void)hipEventCreate(&b);
            (void)hipEventRecord(a, c->stream);
        }
    }
    ~Timed() {  // events are resolved later, so timing does not serialise the stream
        if (c->profile) {
            (void)hipEventRecord(b, c->stream);
            c->pending.push_back({name, a, b});
        }
    }
};

void resolve_events(pimc_ctx* c) {
    for (auto& p : c->pending) {
        (void)hipEventSynchronize(p.b);
        float ms = 0;
        (void)hipEventElapsedTime(&ms, p.a, p.b);
        Family& f = c->fam[p.name];
        f.ms += ms;
        f.n += 1;
        (void)hipEventDestroy(p.a);
        (void)hipEventDestroy(p.b);
    }
    c->pending.clear();
}

template <typename R>
void launch_bisect(pimc_ctx* c, int u, int tb, int te) {
    Run run{c->seed, u, c->ref_order ? 1 : 0, tb, te};
    Timed t(c, "bisect");
    // size classes: the heaviest present on the main stream, the others forked onto side
    // streams and joined again (parallel branches of the captured pass)
    int busy = 0, top = -1;
    for (int k = 2; k >= 0; --k)
        if (c->bw_n[k]) {
            ++busy;
            if (top < 0) top = k;
        }
    if (busy > 1) (void)hipEventRecord(c->ev_fork, c->stream);
    int nside = 0;
    for (int k = 2; k >= 0; --k) {
        if (!c->bw_n[k]) continue;
        hipStream_t s = c->stream;
        if (k != top) {
            s = c->side[nside];
            (void)hipStreamWaitEvent(s, c->ev_fork, 0);
        }
        hipLaunchKernelGGL((c->any_mi ? k_bisect<R, true> : k_bisect<R, false>), dim3(c->bw_n[k]), dim3(c->bw_lanes[k]), c->bw_lds[k], s, c->d_rep, c->d_mut,
                           c->d_beads, c->d_rotb, c->tables(), c->d_work + c->bw_off[k], run);
        if (k != top) (void)hipEventRecord(c->ev_join[nside++], s);
    }
    for (int i = 0; i < nside; ++i) (void)hipStreamWaitEvent(c->stream, c->ev_join[i], 0);
    if (c->nwork1)
        hipLaunchKernelGGL((c->any_mi ? k_bisect_nl<R, true> : k_bisect_nl<R, false>), dim3(c->nwork1), dim3(c->bisect_block), c->lds_bisect1, c->stream,
                           c->d_rep, c->d_mut, c->d_beads, c->d_rotb, c->d_frames, c->tables(), c->tables3(), c->d_work1, run);
}

template <typename R>
void launch_centroid(pimc_ctx* c, int type) {
    Run run{c->seed, 0, c->ref_order ? 1 : 0, type, type + 1};
    Timed t(c, "centroid");
    if (c->ncp0)
        hipLaunchKernelGGL((c->any_mi ? k_centroid_parts<R, 0, true> : k_centroid_parts<R, 0, false>), dim3(c->ncp0), dim3(256), c->lds_knots, c->stream, c->d_rep, c->d_mut,
                           c->d_beads, c->d_rotb, c->d_frames, c->tables(), c->tables3(), c->d_cparts0, c->d_cflags,
                           c->d_cslots, c->ncflags, run);
    if (c->ncp1)
        hipLaunchKernelGGL((c->any_mi ? k_centroid_parts<R, 1, true> : k_centroid_parts<R, 1, false>), dim3(c->ncp1), dim3(256), c->lds_knots, c->stream, c->d_rep, c->d_mut,
                           c->d_beads, c->d_rotb, c->d_frames, c->tables(), c->tables3(), c->d_cparts1, c->d_cflags,
                           c->d_cslots, c->ncflags, run);
}

template <typename R>
void launch_rotor(pimc_ctx* c, int u) {
    if (c->nparts == 0 && c->nrot3 == 0 && c->nparts3 == 0) return;
    Run run{c->seed, u, c->ref_order ? 1 : 0, 0, 0};
    Timed t(c, "rotor");
    if (c->nparts)
        hipLaunchKernelGGL(k_rotor<R>, dim3(c->nparts), dim3(256), c->lds_rotor, c->stream, c->d_rep, c->d_mut,
                           c->d_beads, c->d_rotb, c->tables(), c->tables3(), c->d_parts, c->d_flags, c->d_xrec, run);
    if (c->nparts3 && c->opt_da)
        hipLaunchKernelGGL(k_rotor3d_da<R>, dim3(c->nparts3), dim3(256), c->lds_parts3, c->stream, c->d_rep, c->d_mut,
                           c->d_beads, c->d_rotb, c->d_frames, c->tables(), c->tables3(), c->d_parts3, c->d_flags3,
                           c->d_pub3, c->pub3_stride, run, c->d_mlp);
    else if (c->nparts3)
        hipLaunchKernelGGL(k_rotor3d_parts<R>, dim3(c->nparts3), dim3(256), c->lds_parts3, c->stream, c->d_rep,
                           c->d_mut, c->d_beads, c->d_rotb, c->d_frames, c->tables(), c->tables3(), c->d_parts3,
                           c->d_flags3, c->d_pub3, c->pub3_stride, run);
    if (c->nrot3 && c->rot3_block == 512)
        hipLaunchKernelGGL((k_rotor3d<R, 512>), dim3(c->nrot3), dim3(512), c->lds_rot3, c->stream, c->d_rep,
                           c->d_mut, c->d_beads, c->d_rotb, c->d_frames, c->tables(), c->tables3(), c->d_rot3, run);
    else if (c->nrot3)
        hipLaunchKernelGGL((k_rotor3d<R, 256>), dim3(c->nrot3), dim3(256), c->lds_rot3, c->stream, c->d_rep,
                           c->d_mut, c->d_beads, c->d_rotb, c->d_frames, c->tables(), c->tables3(), c->d_rot3, run);
}

template <typename R>
void launch_estimate(pimc_ctx* c, double* out) {
    Timed t(c, "estimate");
    if (c->nrows)
        hipLaunchKernelGGL((c->any_mi ? k_est_tiles<R, 0, true> : k_est_tiles<R, 0, false>), dim3(c->nrows), dim3(256), c->lds_est, c->stream, c->d_rep,
                           c->d_beads, c->d_rotb, c->d_frames, c->tables(), c->tables3(), c->d_rows,
                           out ? nullptr : c->d_hist, c->d_rowbuf);
    if (c->nrows1)
        hipLaunchKernelGGL((c->any_mi ? k_est_tiles<R, 1, true> : k_est_tiles<R, 1, false>), dim3(c->nrows1), dim3(256), c->lds_est, c->stream, c->d_rep,
                           c->d_beads, c->d_rotb, c->d_frames, c->tables(), c->tables3(), c->d_rows1,
                           out ? nullptr : c->d_hist, c->d_rowbuf);
    if (!out && c->nrcf)
        hipLaunchKernelGGL(k_rcf, dim3(c->nrcf, (c->rcf_protmax + 255) / 256), dim3(256), 0, c->stream, c->d_rep,
                           c->d_rotb, c->d_rcfreps, c->d_rcf);
    hipLaunchKernelGGL(k_est_final, dim3(c->nrep), dim3(256), 0, c->stream, c->d_rep, c->d_mut, c->d_rotb,
                       c->d_frames, c->tables(), c->tables3(),
                       c->d_all, c->d_rowbuf, c->d_samples, c->sample_cap, out, out ? nullptr : c->d_hist);
}

// Autotuner launches: one bisection size class alone, and the linear-rotor sweep alone, over
// the start slices of one pass as the context issues them.
template <typename R>
void launch_bisect_class(pimc_ctx* c, int k, int u) {
    Run run{c->seed, u, c->ref_order ? 1 : 0, 0, c->max_types};
    hipLaunchKernelGGL((c->any_mi ? k_bisect<R, true> : k_bisect<R, false>), dim3(c->bw_n[k]), dim3(c->bw_lanes[k]), c->bw_lds[k], c->stream, c->d_rep,
                       c->d_mut, c->d_beads, c->d_rotb, c->tables(), c->d_work + c->bw_off[k], run);
}

template <typename R>
void launch_centroid_all(pimc_ctx* c) {
    for (int ty = 0; ty < c->max_types; ++ty) {
        Run run{c->seed, 0, c->ref_order ? 1 : 0, ty, ty + 1};
        if (c->ncp0)
            hipLaunchKernelGGL((c->any_mi ? k_centroid_parts<R, 0, true> : k_centroid_parts<R, 0, false>), dim3(c->ncp0), dim3(256), c->lds_knots, c->stream, c->d_rep,
                               c->d_mut, c->d_beads, c->d_rotb, c->d_frames, c->tables(), c->tables3(), c->d_cparts0,
                               c->d_cflags, c->d_cslots, c->ncflags, run);
        if (c->ncp1)
            hipLaunchKernelGGL((c->any_mi ? k_centroid_parts<R, 1, true> : k_centroid_parts<R, 1, false>), dim3(c->ncp1), dim3(256), c->lds_knots, c->stream, c->d_rep,
                               c->d_mut, c->d_beads, c->d_rotb, c->d_frames, c->tables(), c->tables3(), c->d_cparts1,
                               c->d_cflags, c->d_cslots, c->ncflags, run);
    }
}

template <typename R>
void launch_rotor_linear(pimc_ctx* c, int u) {
    Run run{c->seed, u, c->ref_order ? 1 : 0, 0, 0};
    hipLaunchKernelGGL(k_rotor<R>, dim3(c->nparts), dim3(256), c->lds_rotor, c->stream, c->d_rep, c->d_mut,
                       c->d_beads, c->d_rotb, c->tables(), c->tables3(), c->d_parts, c->d_flags, c->d_xrec, run);
}

// one pass in the order described at the top of this file; with `sampled` and samples_per_pass
// k > 1, an estimator sample after every P/k time steps (reference order) or every S/k phases
// (segment order), as the reference samples every MCSKIP_AVERG time steps (mc_main.cc:383-386)
template <typename R>
void issue_pass(pimc_ctx* c, bool sampled) {
    const int k = sampled ? c->opt_samples_per_pass : 1;
    if (c->ref_order) {
        for (int t = 0; t < c->Pmax; ++t) {
            for (int ty = 0; ty < c->max_types; ++ty) {
                if (t == 0) launch_centroid<R>(c, ty);
                launch_bisect<R>(c, t, ty, ty + 1);
            }
            launch_rotor<R>(c, t);
            if (k > 1 && (t + 1) % (c->Pmax / k) == 0) launch_estimate<R>(c, nullptr);
        }
    } else {
        for (int u = 0; u < c->Smax; ++u) {
            if (u == 0) {
                for (int ty = 0; ty < c->max_types; ++ty) {
                    launch_centroid<R>(c, ty);
                    launch_bisect<R>(c, 0, ty, ty + 1);
                }
            } else {
                launch_bisect<R>(c, u, 0, PIMC_MAX_TYPES);
            }
            launch_rotor<R>(c, u);
            if (k > 1 && (u + 1) % (c->Smax / k) == 0) launch_estimate<R>(c, nullptr);
        }
    }
    hipLaunchKernelGGL(k_pass_end, dim3((c->nrep + 255) / 256), dim3(256), 0, c->stream, c->d_mut, c->nrep);
}

int check_errors(pimc_ctx* c) {
    std::vector<Mut> m(c->nrep);
    HIPCHECK(c, hipMemcpyAsync(m.data(), c->d_mut, sizeof(Mut) * c->nrep, hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    for (int r = 0; r < c->nrep; ++r) {
        if (m[r].err & 2) return fail(c, "rotor sweep: a neighbouring part did not arrive (replica " + std::to_string(r) + ")");
        if (m[r].err & 4) return fail(c, "large matrix test error (relative Euler angle off the rotden grid, replica " + std::to_string(r) + ")");
        if (m[r].err & 8) return fail(c, "centroid move: a part of the replica did not arrive (replica " + std::to_string(r) + ")");
        if (m[r].err) return fail(c, "Rotational Moves: Negative rot density (replica " + std::to_string(r) + ")");
    }
    return 0;
}

}  // namespace

extern "C" {

int pimc_run(pimc_ctx* c, int n_passes, int sample_every) {
    if (!c->committed) return fail(c, "pimc_commit first");
    if (c->opt_da && c->nparts3 && !c->mlp_ready)
        return fail(c, "delayed acceptance needs a surrogate: pimc_surrogate_train or pimc_surrogate_set first");
    for (int p = 0; p < n_passes; ++p) {
        const bool sampled = sample_every > 0 && (p + 1) % sample_every == 0;
        const bool inside = sampled && c->opt_samples_per_pass > 1;  // samples within the pass
        bool use_graph = !c->profile;
        hipGraphExec_t& gx = inside ? c->graph_exec_s : c->graph_exec;
        if (use_graph && !gx) {
            hipGraph_t g;
            HIPCHECK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            if (c->fp64) issue_pass<double>(c, inside);
            else issue_pass<float>(c, inside);
            HIPCHECK(c, hipStreamEndCapture(c->stream, &g));
            HIPCHECK(c, hipGraphInstantiate(&gx, g, nullptr, nullptr, 0));
            (void)hipGraphDestroy(g);
        }
        if (use_graph) HIPCHECK(c, hipGraphLaunch(gx, c->stream));
        else if (c->fp64) issue_pass<double>(c, inside);
        else issue_pass<float>(c, inside);
        HIPCHECK(c, hipGetLastError());
        if (sampled && !inside) {
            if (c->fp64) launch_estimate<double>(c, nullptr);
            else launch_estimate<float>(c, nullptr);
        }
        c->bead_moves += c->bead_moves_per_pass;
    }
    HIPCHECK(c, hipGetLastError());
    return check_errors(c);
}

int pimc_autotune(pimc_ctx* c, int passes, long* chosen) {
    if (!c->committed) return fail(c, "pimc_commit first");
    if (passes < 1 || passes > 64) return fail(c, "autotune: passes is 1 to 64");
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    // every candidate starts from the same state; the chains get it back at the end
    double4* sb = dalloc<double4>((size_t)std::max<long long>(c->nbeads, 1));
    RotBead* sr = dalloc<RotBead>((size_t)std::max<long long>(c->nrotb, 1));
    Mut* sm = dalloc<Mut>(c->nrep);
    if (!sb || !sr || !sm) {
        for (void* q : {(void*)sb, (void*)sr, (void*)sm})
            if (q) (void)hipFree(q);
        return fail(c, "autotune: hipMalloc failed");
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = 0;
    auto copy_all = [&](bool save) -> int {
        auto cp = [&](void* dst, const void* src, size_t bytes) {
            return bytes == 0 || hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, c->stream) == hipSuccess;
        };
        const size_t nb = sizeof(double4) * (size_t)c->nbeads, nr = sizeof(RotBead) * (size_t)c->nrotb;
        const size_t nm = sizeof(Mut) * (size_t)c->nrep;
        bool ok = save ? (cp(sb, c->d_beads, nb) && cp(sr, c->d_rotb, nr) && cp(sm, c->d_mut, nm))
                       : (cp(c->d_beads, sb, nb) && cp(c->d_rotb, sr, nr) && cp(c->d_mut, sm, nm));
        return ok ? 0 : -1;
    };
    // time `passes` passes' worth of one kind of launch; < 0 if a launch failed or reported an error
    auto timed = [&](auto&& issue) -> double {
        if (copy_all(false)) return -1.0;
        const int nu = c->ref_order ? c->Pmax : c->Smax;
        if (hipEventRecord(e0, c->stream) != hipSuccess) return -1.0;
        for (int p = 0; p < passes; ++p)
            for (int u = 0; u < nu; ++u) issue(u);
        if (hipEventRecord(e1, c->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) return -1.0;
        if (hipGetLastError() != hipSuccess) return -1.0;
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) return -1.0;
        std::vector<Mut> m(c->nrep);
        if (hipMemcpy(m.data(), c->d_mut, sizeof(Mut) * c->nrep, hipMemcpyDeviceToHost) != hipSuccess) return -1.0;
        for (const Mut& x : m)
            if (x.err) return -1.0;
        return ms;
    };
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess || copy_all(true)) {
        rc = fail(c, "autotune: event or copy failed");
    } else {
        static const int lanes[3] = {64, 128, 256};
        for (int k = 0; k < 3 && rc == 0; ++k) {
            if (!c->bw_n[k]) continue;
            int best = c->bw_lanes[k];
            double tbest = -1.0;
            for (int L : lanes) {
                c->bw_lanes[k] = L;
                const double t = timed([&](int u) {
                    if (c->fp64) launch_bisect_class<double>(c, k, u);
                    else launch_bisect_class<float>(c, k, u);
                });
                if (t >= 0.0 && (tbest < 0.0 || t < tbest)) {
                    tbest = t;
                    best = L;
                }
            }
            c->bw_lanes[k] = best;
        }
        if (rc == 0 && !c->rot_chains.empty()) {
            static const int budgets[4] = {40, 52, 64, 78};
            int best = c->rotor_lds_kb;
            double tbest = -1.0;
            for (int kb : budgets) {
                if (plan_rotor_parts(c, kb)) {
                    rc = -1;
                    break;
                }
                const double t = timed([&](int u) {
                    if (c->fp64) launch_rotor_linear<double>(c, u);
                    else launch_rotor_linear<float>(c, u);
                });
                if (t >= 0.0 && (tbest < 0.0 || t < tbest)) {
                    tbest = t;
                    best = kb;
                }
            }
            if (rc == 0 && c->rotor_lds_kb != best && plan_rotor_parts(c, best)) rc = -1;
        }
        if (rc == 0 && c->ncp0 + c->ncp1 > 0) {  // centroid parts in all: fewer parts keep more in L2
            static const int targets[3] = {256, 512, 1024};
            int best = c->centroid_target;
            double tbest = -1.0;
            for (int tg : targets) {
                if (plan_centroid_parts(c, tg)) {
                    rc = -1;
                    break;
                }
                const double t = timed([&](int u) {
                    if (u != 0) return;
                    if (c->fp64) launch_centroid_all<double>(c);
                    else launch_centroid_all<float>(c);
                });
                if (t >= 0.0 && (tbest < 0.0 || t < tbest)) {
                    tbest = t;
                    best = c->centroid_target;
                }
            }
            if (rc == 0 && c->centroid_target != best && plan_centroid_parts(c, best)) rc = -1;
        }
        if (copy_all(false) || hipStreamSynchronize(c->stream) != hipSuccess)
            rc = rc ? rc : fail(c, "autotune: restoring the state failed");
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    for (void* q : {(void*)sb, (void*)sr, (void*)sm}) (void)hipFree(q);
    for (hipGraphExec_t* gx : {&c->graph_exec, &c->graph_exec_s})
        if (*gx) {  // launch shapes changed: the next pass is captured again
            (void)hipGraphExecDestroy(*gx);
            *gx = nullptr;
        }
    if (rc == 0 && chosen) {
        for (int k = 0; k < 3; ++k) chosen[k] = c->bw_lanes[k];
        chosen[3] = c->rotor_lds_kb;
        chosen[4] = c->centroid_target;
    }
    return rc;
}

int pimc_synchronize(pimc_ctx* c) {
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int pimc_get_samples(pimc_ctx* c, int r, double* out, int max_samples) {
    if (!c->committed || r < 0 || r >= c->nrep) return fail(c, "bad replica");
    Mut m;
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    HIPCHECK(c, hipMemcpy(&m, c->d_mut + r, sizeof(Mut), hipMemcpyDeviceToHost));
    if ((long long)m.nsamples > c->sample_cap)
        return fail(c, "replica " + std::to_string(r) + " took " + std::to_string(m.nsamples) + " samples, more than the " +
                           std::to_string(c->sample_cap) + " its buffer holds: read and clear them (pimc_clear_samples) " +
                           "at least every " + std::to_string(c->sample_cap) + " samples");
    int n = std::min<int>(std::min<int>((int)m.nsamples, c->sample_cap), max_samples);
    if (n > 0)
        HIPCHECK(c, hipMemcpy(out, c->d_samples + c->rep[r].sample_off, sizeof(double) * 8 * n, hipMemcpyDeviceToHost));
    return n;
}

int pimc_clear_samples(pimc_ctx* c) {
    std::vector<Mut> m(c->nrep);
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    HIPCHECK(c, hipMemcpy(m.data(), c->d_mut, sizeof(Mut) * c->nrep, hipMemcpyDeviceToHost));
    for (auto& x : m) x.nsamples = 0;
    HIPCHECK(c, hipMemcpy(c->d_mut, m.data(), sizeof(Mut) * c->nrep, hipMemcpyHostToDevice));
    HIPCHECK(c, hipMemset(c->d_hist, 0, sizeof(unsigned) * std::max<long long>(c->nhist, 1)));
    HIPCHECK(c, hipMemset(c->d_rcf, 0, sizeof(double) * std::max<long long>(c->nrcfvals, 1)));
    return 0;
}

int pimc_get_counters(pimc_ctx* c, int r, double* total, double* accep) {
    if (!c->committed || r < 0 || r >= c->nrep) return fail(c, "bad replica");
    Mut m;
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    HIPCHECK(c, hipMemcpy(&m, c->d_mut + r, sizeof(Mut), hipMemcpyDeviceToHost));
    for (int t = 0; t < c->rep[r].ntypes; ++t)
        for (int k = 0; k < 3; ++k) {
            total[3 * t + k] = (double)m.total[t][k];
            accep[3 * t + k] = (double)m.accep[t][k];
        }
    return 0;
}

int pimc_get_histograms(pimc_ctx* c, int r, double* gr1d, double* gr2d) {
    if (!c->committed || r < 0 || r >= c->nrep) return fail(c, "bad replica");
    std::vector<unsigned> h(BINSR + BINSR * BINST);
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    HIPCHECK(c, hipMemcpy(h.data(), c->d_hist + c->rep[r].hist_off, sizeof(unsigned) * h.size(), hipMemcpyDeviceToHost));
    if (gr1d)
        for (int i = 0; i < BINSR; ++i) gr1d[i] = h[i];
    if (gr2d)
        for (int i = 0; i < BINSR * BINST; ++i) gr2d[i] = h[BINSR + i];
    return 0;
}

int pimc_get_histograms3d(pimc_ctx* c, int r, double* gr3d, double* rel) {
    if (!c->committed || r < 0 || r >= c->nrep) return fail(c, "bad replica");
    const Rep& p = c->rep[r];
    int has = 0;
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    if (p.h3_off >= 0) {
        has |= 1;
        if (gr3d) {
            std::vector<unsigned> h((size_t)BINSR * BINST * BINSC);
            HIPCHECK(c, hipMemcpy(h.data(), c->d_hist + p.hist_off + p.h3_off, sizeof(unsigned) * h.size(),
                                  hipMemcpyDeviceToHost));
            for (size_t i = 0; i < h.size(); ++i) gr3d[i] = h[i];
        }
    }
    if (p.hrel_off >= 0) {
        has |= 2;
        if (rel) {
            unsigned h[BINST + 2 * BINSC];
            HIPCHECK(c, hipMemcpy(h, c->d_hist + p.hist_off + p.hrel_off, sizeof(h), hipMemcpyDeviceToHost));
            for (int i = 0; i < BINST + 2 * BINSC; ++i) rel[i] = h[i];
        }
    }
    return has;
}

int pimc_get_rcf(pimc_ctx* c, int r, double* out, int max_values) {
    if (!c->committed || r < 0 || r >= c->nrep) return fail(c, "bad replica");
    const Rep& p = c->rep[r];
    if (p.rcf_off < 0) return 0;
    if (!out || max_values < RCF_L * p.Prot) return fail(c, "pimc_get_rcf: buffer smaller than 10*Prot");
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    HIPCHECK(c, hipMemcpy(out, c->d_rcf + p.rcf_off, sizeof(double) * RCF_L * p.Prot, hipMemcpyDeviceToHost));
    return p.Prot;
}

int pimc_get_work(pimc_ctx* c, double* pair_evals, double* bead_moves) {
    std::vector<Mut> m(c->nrep);
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    HIPCHECK(c, hipMemcpy(m.data(), c->d_mut, sizeof(Mut) * c->nrep, hipMemcpyDeviceToHost));
    double pe = 0;
    for (auto& x : m)
        for (int k = 0; k < 4; ++k) pe += (double)x.pair_evals[k];
    if (pair_evals) *pair_evals = pe;
    if (bead_moves) *bead_moves = c->bead_moves;
    return 0;
}

}  // extern "C"

namespace {

int check_queries(pimc_ctx* c, int nq, const int* qr, const int* qa, const int* qit) {
    if (!c->committed) return fail(c, "pimc_commit first");
    if (nq < 0) return fail(c, "query count is negative");
    for (int q = 0; q < nq; ++q) {
        if (qr[q] < 0 || qr[q] >= c->nrep) return fail(c, "query replica out of range");
        const Rep& p = c->rep[qr[q]];
        if (qa[q] < 0 || qa[q] >= p.natoms || qit[q] < 0 || qit[q] >= p.P) return fail(c, "query out of range");
    }
    return 0;
}

// k_pot_slice on device copies of the queries, then the results back to out
int pot_slice_on(pimc_ctx* c, int nq, const int* qr, const int* qa, const int* qit, const double* qpos, double* out,
                 int* d_r, int* d_a, int* d_i, double* d_p, double* d_o) {
    HIPCHECK(c, hipMemcpyAsync(d_r, qr, sizeof(int) * nq, hipMemcpyHostToDevice, c->stream));
    HIPCHECK(c, hipMemcpyAsync(d_a, qa, sizeof(int) * nq, hipMemcpyHostToDevice, c->stream));
    HIPCHECK(c, hipMemcpyAsync(d_i, qit, sizeof(int) * nq, hipMemcpyHostToDevice, c->stream));
    HIPCHECK(c, hipMemcpyAsync(d_p, qpos, sizeof(double) * 3 * nq, hipMemcpyHostToDevice, c->stream));
    if (c->fp64)
        hipLaunchKernelGGL(k_pot_slice<double>, dim3((nq + 127) / 128), dim3(128), 0, c->stream, c->d_rep, c->d_beads,
                           c->d_rotb, c->d_frames, c->tables(), c->tables3(), nq, d_r, d_a, d_i, d_p, d_o);
    else
        hipLaunchKernelGGL(k_pot_slice<float>, dim3((nq + 127) / 128), dim3(128), 0, c->stream, c->d_rep, c->d_beads,
                           c->d_rotb, c->d_frames, c->tables(), c->tables3(), nq, d_r, d_a, d_i, d_p, d_o);
    HIPCHECK(c, hipGetLastError());
    HIPCHECK(c, hipMemcpyAsync(out, d_o, sizeof(double) * nq, hipMemcpyDeviceToHost, c->stream));
    if (hipStreamSynchronize(c->stream) != hipSuccess) return fail(c, "pot_energy_slice failed");
    return 0;
}

}  // namespace

extern "C" {

int pimc_pot_energy_slice(pimc_ctx* c, int nq, const int* qr, const int* qa, const int* qit, const double* qpos,
                          double* out) {
    if (check_queries(c, nq, qr, qa, qit)) return -1;
    if (nq == 0) return 0;
    int *d_r = dalloc<int>(nq), *d_a = dalloc<int>(nq), *d_i = dalloc<int>(nq);
    double *d_p = dalloc<double>(3 * (size_t)nq), *d_o = dalloc<double>(nq);
    int rc = 0;
    if (!d_r || !d_a || !d_i || !d_p || !d_o) rc = fail(c, "hipMalloc failed");
    if (!rc) rc = pot_slice_on(c, nq, qr, qa, qit, qpos, out, d_r, d_a, d_i, d_p, d_o);
    void* ps[] = {d_r, d_a, d_i, d_p, d_o};
    for (void* p : ps)
        if (p) (void)hipFree(p);
    return rc;
}

// The same queries staged in this device's copy of `scratch` instead of buffers allocated per
// call: no allocation, at most PIMC_NOALLOC_MAX_QUERIES queries.  Nothing in the staging needs
// to start at zero (every query slot read is written first, every result written by the kernel).
int pimc_pot_energy_slice_noalloc(pimc_ctx* c, int nq, const int* qr, const int* qa, const int* qit,
                                  const double* qpos, double* out) {
    if (nq > PIMC_NOALLOC_MAX_QUERIES)
        return fail(c, "pimc_pot_energy_slice_noalloc: at most " + std::to_string(PIMC_NOALLOC_MAX_QUERIES) +
                           " queries per call (" + std::to_string(nq) + " given)");
    if (check_queries(c, nq, qr, qa, qit)) return -1;
    if (!c->scratch) return fail(c, "pimc_pot_energy_slice_noalloc: no staging on this device (" + c->scratch_err + ")");
    if (nq == 0) return 0;
    unsigned char* s = c->scratch;
    return pot_slice_on(c, nq, qr, qa, qit, qpos, out, reinterpret_cast<int*>(s), reinterpret_cast<int*>(s + NA_OFF_A),
                        reinterpret_cast<int*>(s + NA_OFF_I), reinterpret_cast<double*>(s + NA_OFF_POS),
                        reinterpret_cast<double*>(s + NA_OFF_OUT));
}

int pimc_estimate(pimc_ctx* c, double* out) {
    if (!c->committed) return fail(c, "pimc_commit first");
    if (c->fp64) launch_estimate<double>(c, c->d_est);
    else launch_estimate<float>(c, c->d_est);
    HIPCHECK(c, hipGetLastError());
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    HIPCHECK(c, hipMemcpy(out, c->d_est, sizeof(double) * 8 * c->nrep, hipMemcpyDeviceToHost));
    return 0;
}

int pimc_family_evals(pimc_ctx* c, const char* family, double* evals) {
    static const char* names[4] = {"bisect", "centroid", "rotor", "estimate"};
    int k = -1;
    for (int i = 0; i < 4; ++i)
        if (std::strcmp(names[i], family) == 0) k = i;
    if (k < 0) return fail(c, "unknown kernel family");
    std::vector<Mut> m(c->nrep);
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    HIPCHECK(c, hipMemcpy(m.data(), c->d_mut, sizeof(Mut) * c->nrep, hipMemcpyDeviceToHost));
    double e = 0;
    for (auto& x : m) e += (double)x.pair_evals[k];
    *evals = e;
    return 0;
}

int pimc_kernel_time(pimc_ctx* c, const char* family, double* ms, long* launches) {
    resolve_events(c);
    auto it = c->fam.find(family);
    if (it == c->fam.end()) {
        *ms = 0;
        *launches = 0;
        return 0;
    }
    *ms = it->second.ms;
    *launches = it->second.n;
    return 0;
}

// ------------------------------------------------------------------------------ surrogate
namespace {

struct MlpBuffers {
    float *x = nullptr, *y = nullptr, *Z1 = nullptr, *H1 = nullptr, *Z2 = nullptr, *H2 = nullptr, *Z3 = nullptr,
          *H3 = nullptr, *out = nullptr, *dout = nullptr, *dA = nullptr, *dB = nullptr;
    double* loss = nullptr;
    int cap = 0;
    bool alloc(int B) {
        cap = B;
        const size_t h = (size_t)B * MLP_H;
        x = dalloc<float>((size_t)B * MLP_IN);
        y = dalloc<float>(B);
        Z1 = dalloc<float>(h); H1 = dalloc<float>(h);
        Z2 = dalloc<float>(h); H2 = dalloc<float>(h);
        Z3 = dalloc<float>(h); H3 = dalloc<float>(h);
        out = dalloc<float>(B); dout = dalloc<float>(B);
        dA = dalloc<float>(h); dB = dalloc<float>(h);
        loss = dalloc<double>(1);
        return x && y && Z1 && H1 && Z2 && H2 && Z3 && H3 && out && dout && dA && dB && loss;
    }
    ~MlpBuffers() {
        void* p[] = {x, y, Z1, H1, Z2, H2, Z3, H3, out, dout, dA, dB, loss};
        for (void* q : p)
            if (q) (void)hipFree(q);
    }
};

void mlp_gemm(hipStream_t s, int M, int N, int K, const float* A, int lda, int ta, const float* B, int ldb, int tb,
              float* C, int ldc) {
    hipLaunchKernelGGL(k_mlp_gemm, dim3((N + 31) / 32, (M + 31) / 32), dim3(64), 0, s, M, N, K, A, lda, ta, B, ldb,
                       tb, C, ldc, 0);
}

void mlp_bias(hipStream_t s, float* Z, float* H, const float* b, int rows, int cols, int act) {
    const long long n = (long long)rows * cols;
    hipLaunchKernelGGL(k_mlp_bias_act, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, Z, H, b, rows, cols, act);
}

void mlp_back(hipStream_t s, float* dH, const float* Z, int rows) {
    const long long n = (long long)rows * MLP_H;
    hipLaunchKernelGGL(k_mlp_act_back, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, dH, Z, n);
}

// Loss and gradient of the batch in m.x / m.y (B rows, normalised) at weights th; g gets the
// gradient, m.loss accumulates the loss.
void mlp_forward_backward(hipStream_t s, const float* th, float* g, MlpBuffers& m, int B) {
    // forward: Z = H_prev W^T + b, H = SiLU(Z)
    mlp_gemm(s, B, MLP_H, MLP_IN, m.x, MLP_IN, 0, th + MLP_OFF_W1, MLP_IN, 1, m.Z1, MLP_H);
    mlp_bias(s, m.Z1, m.H1, th + MLP_OFF_B1, B, MLP_H, 1);
    mlp_gemm(s, B, MLP_H, MLP_H, m.H1, MLP_H, 0, th + MLP_OFF_W2, MLP_H, 1, m.Z2, MLP_H);
    mlp_bias(s, m.Z2, m.H2, th + MLP_OFF_B2, B, MLP_H, 1);
    mlp_gemm(s, B, MLP_H, MLP_H, m.H2, MLP_H, 0, th + MLP_OFF_W3, MLP_H, 1, m.Z3, MLP_H);
    mlp_bias(s, m.Z3, m.H3, th + MLP_OFF_B3, B, MLP_H, 1);
    mlp_gemm(s, B, 1, MLP_H, m.H3, MLP_H, 0, th + MLP_OFF_W4, 1, 0, m.out, 1);
    mlp_bias(s, m.out, m.out, th + MLP_OFF_B4, B, 1, 0);
    hipLaunchKernelGGL(k_mlp_mse, dim3((B + 255) / 256), dim3(256), 0, s, m.out, m.y, B, m.dout, m.loss);
    // backward
    mlp_gemm(s, MLP_H, 1, B, m.H3, MLP_H, 1, m.dout, 1, 0, g + MLP_OFF_W4, 1);
    hipLaunchKernelGGL(k_mlp_colsum, dim3(1), dim3(256), 0, s, m.dout, B, 1, g + MLP_OFF_B4);
    mlp_gemm(s, B, MLP_H, 1, m.dout, 1, 0, th + MLP_OFF_W4, MLP_H, 0, m.dA, MLP_H);  // dH3
    mlp_back(s, m.dA, m.Z3, B);                                                      // dZ3
    mlp_gemm(s, MLP_H, MLP_H, B, m.dA, MLP_H, 1, m.H2, MLP_H, 0, g + MLP_OFF_W3, MLP_H);
    hipLaunchKernelGGL(k_mlp_colsum, dim3(MLP_H), dim3(256), 0, s, m.dA, B, MLP_H, g + MLP_OFF_B3);
    mlp_gemm(s, B, MLP_H, MLP_H, m.dA, MLP_H, 0, th + MLP_OFF_W3, MLP_H, 0, m.dB, MLP_H);  // dH2
    mlp_back(s, m.dB, m.Z2, B);
    mlp_gemm(s, MLP_H, MLP_H, B, m.dB, MLP_H, 1, m.H1, MLP_H, 0, g + MLP_OFF_W2, MLP_H);
    hipLaunchKernelGGL(k_mlp_colsum, dim3(MLP_H), dim3(256), 0, s, m.dB, B, MLP_H, g + MLP_OFF_B2);
    mlp_gemm(s, B, MLP_H, MLP_H, m.dB, MLP_H, 0, th + MLP_OFF_W2, MLP_H, 0, m.dA, MLP_H);  // dH1
    mlp_back(s, m.dA, m.Z1, B);
    mlp_gemm(s, MLP_H, MLP_IN, B, m.dA, MLP_H, 1, m.x, MLP_IN, 0, g + MLP_OFF_W1, MLP_IN);
    hipLaunchKernelGGL(k_mlp_colsum, dim3(MLP_H), dim3(256), 0, s, m.dA, B, MLP_H, g + MLP_OFF_B1);
}

int mlp_upload(pimc_ctx* c) {
    hipError_t e0 = hipSuccess;
    auto ck = [&](hipError_t x) { if (e0 == hipSuccess) e0 = x; };
    if (!c->d_mlp) c->d_mlp = dalloc<MlpPacked>(1);
    float *th = dalloc<float>(MLP_NW), *nm = dalloc<float>(2 * MLP_IN + 2);
    if (!c->d_mlp || !th || !nm) {
        if (th) (void)hipFree(th);
        if (nm) (void)hipFree(nm);
        return fail(c, "surrogate: out of device memory");
    }
    ck(hipMemcpy(th, c->theta.data(), sizeof(float) * MLP_NW, hipMemcpyHostToDevice));
    ck(hipMemcpy(nm, c->norm.data(), sizeof(float) * (2 * MLP_IN + 2), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_mlp_pack, dim3((MLP_H * MLP_H + 255) / 256), dim3(256), 0, c->stream, th, nm, c->d_mlp);
    ck(hipGetLastError());
    ck(hipStreamSynchronize(c->stream));
    (void)hipFree(th);
    (void)hipFree(nm);
    if (e0 != hipSuccess) return fail(c, std::string("surrogate pack: ") + hipGetErrorString(e0));
    c->mlp_ready = true;
    return 0;
}

}  // namespace

int pimc_surrogate_collect(pimc_ctx* c, int r, float* x, float* v, int max_rows) {
    if (!c->committed) return fail(c, "pimc_commit first");
    if (r < 0 || r >= c->nrep) return fail(c, "replica index out of range");
    const Rep& p = c->rep[r];
    const int nat = p.natoms - p.nrot;
    if (p.fam != 1 || p.nrot < 1 || !(nat > 0 ? p.pot3d_id >= 0 : p.nrot == 2))
        return fail(c, "surrogate rows need non-linear rotors among atoms with a 3D potential, or two rotors");
    const long long n = (long long)p.nrot * p.Prot * p.RotRatio * (nat > 0 ? nat : p.nrot - 1);
    if (!x || !v || max_rows < n) return fail(c, "surrogate_collect: buffer smaller than " + std::to_string(n) + " rows");
    float* dx = dalloc<float>((size_t)n * MLP_IN);
    float* dv = dalloc<float>((size_t)n);
    if (!dx || !dv) {
        if (dx) (void)hipFree(dx);
        if (dv) (void)hipFree(dv);
        return fail(c, "surrogate_collect: out of device memory");
    }
    hipLaunchKernelGGL(k_surrogate_collect, dim3((unsigned)std::min<long long>((n + 255) / 256, 4096)), dim3(256), 0,
                       c->stream, c->d_rep, c->d_beads, c->d_frames, c->tables3(), r, dx, dv, (int)n);
    hipError_t e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess) e = hipMemcpy(x, dx, sizeof(float) * MLP_IN * n, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(v, dv, sizeof(float) * n, hipMemcpyDeviceToHost);
    (void)hipFree(dx);
    (void)hipFree(dv);
    if (e != hipSuccess) return fail(c, std::string("surrogate_collect: ") + hipGetErrorString(e));
    return (int)n;
}

int pimc_surrogate_grad(pimc_ctx* c, const float* theta, const float* x, const float* y, int n, float* grad,
                        double* loss) {
    if (!theta || !x || !y || !grad || !loss || n < 1 || n > (1 << 20)) return fail(c, "surrogate_grad: bad arguments");
    MlpBuffers m;
    float *th = dalloc<float>(MLP_NW), *g = dalloc<float>(MLP_NW);
    int rc = 0;
    hipError_t e0 = hipSuccess;
    auto ck = [&](hipError_t x) { if (e0 == hipSuccess) e0 = x; };
    if (!m.alloc(n) || !th || !g) rc = fail(c, "surrogate_grad: out of device memory");
    if (rc == 0) {
        ck(hipMemcpy(th, theta, sizeof(float) * MLP_NW, hipMemcpyHostToDevice));
        ck(hipMemcpy(m.x, x, sizeof(float) * MLP_IN * n, hipMemcpyHostToDevice));
        ck(hipMemcpy(m.y, y, sizeof(float) * n, hipMemcpyHostToDevice));
        ck(hipMemset(m.loss, 0, sizeof(double)));
        mlp_forward_backward(c->stream, th, g, m, n);
        ck(hipGetLastError());
        hipError_t e = e0 != hipSuccess ? e0 : hipStreamSynchronize(c->stream);
        if (e == hipSuccess) e = hipMemcpy(grad, g, sizeof(float) * MLP_NW, hipMemcpyDeviceToHost);
        if (e == hipSuccess) e = hipMemcpy(loss, m.loss, sizeof(double), hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = fail(c, std::string("surrogate_grad: ") + hipGetErrorString(e));
    }
    if (th) (void)hipFree(th);
    if (g) (void)hipFree(g);
    return rc;
}

int pimc_surrogate_train(pimc_ctx* c, const float* x, const float* v, int n, int epochs, int batch, double lr,
                         unsigned seed, double* loss) {
    if (!x || !v || n < 64 || epochs < 1 || batch < 32 || batch > n || !(lr > 0.0))
        return fail(c, "surrogate_train: bad arguments (n >= 64, 32 <= batch <= n, epochs >= 1, lr > 0)");
    // normalisation from the training set
    std::vector<double> mean(MLP_IN + 1, 0.0), var(MLP_IN + 1, 0.0);
    for (int i = 0; i < n; ++i) {
        for (int j = 0; j < MLP_IN; ++j) mean[j] += x[(size_t)i * MLP_IN + j];
        mean[MLP_IN] += v[i];
    }
    for (double& m : mean) m /= n;
    for (int i = 0; i < n; ++i) {
        for (int j = 0; j < MLP_IN; ++j) var[j] += (x[(size_t)i * MLP_IN + j] - mean[j]) * (x[(size_t)i * MLP_IN + j] - mean[j]);
        var[MLP_IN] += (v[i] - mean[MLP_IN]) * (v[i] - mean[MLP_IN]);
    }
    std::vector<float> norm(2 * MLP_IN + 2);
    for (int j = 0; j <= MLP_IN; ++j) {
        double sd = std::sqrt(var[j] / n);
        if (!(sd > 1e-12)) sd = 1.0;
        if (j < MLP_IN) {
            norm[j] = (float)mean[j];
            norm[MLP_IN + j] = (float)sd;
        } else {
            norm[2 * MLP_IN] = (float)mean[j];
            norm[2 * MLP_IN + 1] = (float)sd;
        }
    }
    std::vector<float> xn((size_t)n * MLP_IN), yn(n);
    for (int i = 0; i < n; ++i) {
        for (int j = 0; j < MLP_IN; ++j) xn[(size_t)i * MLP_IN + j] = (x[(size_t)i * MLP_IN + j] - norm[j]) / norm[MLP_IN + j];
        yn[i] = (v[i] - norm[2 * MLP_IN]) / norm[2 * MLP_IN + 1];
    }
    // Glorot-uniform weights, zero biases
    std::vector<float> th(MLP_NW, 0.0f);
    uint64_t st = 0x9E3779B97F4A7C15ull ^ seed;
    auto unif = [&]() {  // splitmix64 -> [0, 1)
        st += 0x9E3779B97F4A7C15ull;
        uint64_t z = st;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return (double)(z >> 11) * (1.0 / 9007199254740992.0);
    };
    auto fill = [&](int off, int fan_out, int fan_in) {
        const double a = std::sqrt(6.0 / (fan_in + fan_out));
        for (int i = 0; i < fan_out * fan_in; ++i) th[off + i] = (float)(a * (2.0 * unif() - 1.0));
    };
    fill(MLP_OFF_W1, MLP_H, MLP_IN);
    fill(MLP_OFF_W2, MLP_H, MLP_H);
    fill(MLP_OFF_W3, MLP_H, MLP_H);
    fill(MLP_OFF_W4, 1, MLP_H);
    MlpBuffers m;
    float *dX = dalloc<float>((size_t)n * MLP_IN), *dY = dalloc<float>(n), *dth = dalloc<float>(MLP_NW),
          *dg = dalloc<float>(MLP_NW), *dm = dalloc<float>(MLP_NW), *dv = dalloc<float>(MLP_NW);
    int* dperm = dalloc<int>(n);
    int rc = 0;
    hipError_t e0 = hipSuccess;
    auto ck = [&](hipError_t x) { if (e0 == hipSuccess) e0 = x; };
    if (!m.alloc(batch) || !dX || !dY || !dth || !dg || !dm || !dv || !dperm) rc = fail(c, "surrogate_train: out of device memory");
    if (rc == 0) {
        ck(hipMemcpy(dX, xn.data(), sizeof(float) * xn.size(), hipMemcpyHostToDevice));
        ck(hipMemcpy(dY, yn.data(), sizeof(float) * n, hipMemcpyHostToDevice));
        ck(hipMemcpy(dth, th.data(), sizeof(float) * MLP_NW, hipMemcpyHostToDevice));
        ck(hipMemset(dm, 0, sizeof(float) * MLP_NW));
        ck(hipMemset(dv, 0, sizeof(float) * MLP_NW));
        std::vector<int> perm(n);
        for (int i = 0; i < n; ++i) perm[i] = i;
        const int steps = n / batch;
        long t = 0;
        for (int ep = 0; ep < epochs && rc == 0; ++ep) {
            for (int i = n - 1; i > 0; --i) std::swap(perm[i], perm[(int)(unif() * (i + 1))]);
            ck(hipMemcpy(dperm, perm.data(), sizeof(int) * n, hipMemcpyHostToDevice));
            ck(hipMemset(m.loss, 0, sizeof(double)));
            const float lr_ep = (float)(lr * std::pow(0.1, (double)ep / epochs));  // decays tenfold over the run
            for (int s = 0; s < steps; ++s) {
                hipLaunchKernelGGL(k_mlp_gather, dim3((batch + 255) / 256), dim3(256), 0, c->stream, dX, dY, dperm,
                                   s * batch, batch, m.x, m.y);
                mlp_forward_backward(c->stream, dth, dg, m, batch);
                ++t;
                const float c1 = (float)(1.0 - std::pow(0.9, (double)t)), c2 = (float)(1.0 - std::pow(0.999, (double)t));
                hipLaunchKernelGGL(k_mlp_adam, dim3((MLP_NW + 255) / 256), dim3(256), 0, c->stream, dth, dg, dm, dv,
                                   MLP_NW, lr_ep, c1, c2);
            }
            double l = 0.0;
            ck(hipGetLastError());
            hipError_t e = e0 != hipSuccess ? e0 : hipStreamSynchronize(c->stream);
            if (e == hipSuccess) e = hipMemcpy(&l, m.loss, sizeof(double), hipMemcpyDeviceToHost);
            if (e != hipSuccess) rc = fail(c, std::string("surrogate_train: ") + hipGetErrorString(e));
            if (loss) loss[ep] = l / steps;
        }
        if (rc == 0 && hipMemcpy(th.data(), dth, sizeof(float) * MLP_NW, hipMemcpyDeviceToHost) != hipSuccess)
            rc = fail(c, "surrogate_train: copy back");
    }
    void* ptrs[] = {dX, dY, dth, dg, dm, dv, dperm};
    for (void* q : ptrs)
        if (q) (void)hipFree(q);
    if (rc) return rc;
    c->theta = th;
    c->norm = norm;
    return mlp_upload(c);
}

int pimc_surrogate_set(pimc_ctx* c, const float* theta, const float* norm) {
    if (!theta || !norm) return fail(c, "surrogate_set: null argument");
    for (int j = 0; j < MLP_IN; ++j)
        if (!(norm[MLP_IN + j] != 0.0f)) return fail(c, "surrogate_set: zero input scale");
    c->theta.assign(theta, theta + MLP_NW);
    c->norm.assign(norm, norm + 2 * MLP_IN + 2);
    return mlp_upload(c);
}

int pimc_surrogate_get(pimc_ctx* c, float* theta, float* norm) {
    if (!c->mlp_ready) return fail(c, "no surrogate loaded");
    if (theta) std::memcpy(theta, c->theta.data(), sizeof(float) * MLP_NW);
    if (norm) std::memcpy(norm, c->norm.data(), sizeof(float) * (2 * MLP_IN + 2));
    return 0;
}

int pimc_surrogate_eval(pimc_ctx* c, const float* x, int n, float* out) {
    if (!c->mlp_ready) return fail(c, "no surrogate loaded");
    if (!x || !out || n < 1) return fail(c, "surrogate_eval: bad arguments");
    float *dx = dalloc<float>((size_t)n * MLP_IN), *dy = dalloc<float>(n);
    int rc = 0;
    hipError_t e0 = hipSuccess;
    auto ck = [&](hipError_t x) { if (e0 == hipSuccess) e0 = x; };
    if (!dx || !dy) rc = fail(c, "surrogate_eval: out of device memory");
    if (rc == 0) {
        ck(hipMemcpy(dx, x, sizeof(float) * MLP_IN * n, hipMemcpyHostToDevice));
        const int waves = (n + 31) / 32;
        hipLaunchKernelGGL(k_mlp_infer, dim3((waves + 3) / 4), dim3(256), sizeof(MlpPacked), c->stream, c->d_mlp, dx,
                           n, dy);
        ck(hipGetLastError());
        hipError_t e = e0 != hipSuccess ? e0 : hipStreamSynchronize(c->stream);
        if (e == hipSuccess) e = hipMemcpy(out, dy, sizeof(float) * n, hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = fail(c, std::string("surrogate_eval: ") + hipGetErrorString(e));
    }
    if (dx) (void)hipFree(dx);
    if (dy) (void)hipFree(dy);
    return rc;
}

int pimc_get_da_counters(pimc_ctx* c, int r, double* out) {
    if (!c->committed || r < 0 || r >= c->nrep || !out) return fail(c, "bad replica");
    Mut m;
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    HIPCHECK(c, hipMemcpy(&m, c->d_mut + r, sizeof(Mut), hipMemcpyDeviceToHost));
    for (int k = 0; k < 3; ++k) out[k] = (double)m.da[k];
    out[3] = (double)m.surrogate_evals;
    return 0;
}

}  // extern "C"
