// pimc_device.hpp -- device-side data model and arithmetic of the gfx950 PIMC sampler.
//
// Layout in HBM (one context, all replicas):
//   beads   double4 [sum_r natoms_r * P_r]      bead (x, y, z, -) of atom a, slice it at
//                                               rep.bead_off + a*P + it   (FP64 source of truth)
//   rotb    RotBead [sum_r nrot_r * Prot_r]     (phi, cos theta, chi) + unit axis, per rotor slice
//   tables  1D knots float4/double4 (x, y, y2, -), 2D values float/double [nr*nc] + FP64 grids,
//           .rot knots double4 (x, rho, rho'', -) and (erot, erot'', esq, esq'')
// Pair terms are evaluated in `Real` (float in production, double in parity runs); every
// sum over pairs is accumulated in double.  Displacements are formed in double before
// rounding so absolute coordinates never lose bits to FP32.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <type_traits>

namespace pimc {

constexpr int MAXT = 4;
constexpr double RZERO = 1.0e-10;  // zero of the rotational density (mc_confg.h:60)
constexpr int BINSR = 300;         // radial bins (mc_estim.cc:25)
constexpr int BINST = 50;          // polar bins (mc_estim.cc:26)
constexpr int BINSC = 100;         // chi bins around a non-linear rotor (mc_estim.cc:27)
constexpr double MAX_RADIUS = 15.0;

enum : uint32_t { K_CENTROID = 1u, K_BISECT = 2u, K_ROTLIN = 3u, K_ROT3D = 4u };

struct Rep {
    int natoms, P, Prot, RotRatio, ntypes, imtype, rotkind, S;
    int tnumb[MAXT], tmol[MAXT], tlev[MAXT], tbegin[MAXT];
    int pot1d[MAXT], pot2d[MAXT], rotdens;
    int rot_atom;  // global index of the (first) rotor atom, -1 if none
    int lin_atom;  // global index of the linear molecule, -1 if none
    int pot1d_id;  // 1D table of the atom type, -1 if none
    int pot2d_id;  // 2D table of the linear-molecule type, -1 if none
    int fam;       // 0: atoms and at most one linear molecule; 1: atoms and non-linear rotors
    int nl_atom;   // global index of the first non-linear rotor (the rotor type is last), -1 if none
    int nrot;      // number of non-linear rotors
    int pot3d_id;  // 3D table of the atom / non-linear rotor pairs, -1 if none
    int rot3d_id;  // rho/E/E^2 tables of the non-linear rotor, -1 if none
    int key;       // replica index, second word of the generator key
    int est_ts;    // slices per estimator tile (a power of two)
    int h3_off;    // (r, theta, chi) bins around a non-linear rotor, from hist_off; -1 if none
    int hrel_off;  // relative Euler angle bins (theta, phi, chi) of rotor links, from hist_off; -1 if none
    int minimage;  // 1: separations of atom pairs and kinetic links take the nearest periodic image
    double lambda[MAXT], mcstep[MAXT];
    double T, beta, tau, rtstep;
    double box[3];  // BoxSize (mc_setup.cc:343-366), used when minimage is set
    long long bead_off, rot_off, hist_off, sample_off;
    long long row_off;  // first estimator tile partial of the replica
    long long rcf_off;  // [RCF_L][Prot] accumulators of the rotational correlation functions, -1 if none
};
constexpr int RCF_L = 10;  // NUMB_RCF mc_confg.h:58

struct Mut {
    unsigned pass;
    int err;
    unsigned nsamples, pad;
    unsigned long long total[MAXT][3];
    unsigned long long accep[MAXT][3];
    unsigned long long pair_evals[4];  // bisect, centroid, rotor, estimate
    unsigned long long da[3];          // delayed acceptance: proposals, stage-1 passes, accepted
    unsigned long long surrogate_evals;
};

struct RotBead {
    double phi, cth, chi, pad0;
    double nx, ny, nz, pad1;
};

struct Tab1D {
    int n, off;
    int uniform, pad;  // uniform: every knot within 1e-6 of the span of the straight line
    double x0, xmax, inv_h, alpha, unode, c6;
};
struct Tab2D {
    int nr, nc;
    long long off;   // into values
    long long coff;  // into the corner-packed FP32 cells of a uniform table, -1 if none
    int goff;       // into grids: r grid at goff, cos grid at goff + nr
    int uniform;    // both grids uniform to 1e-6 relative: FP32 lookups use grid arithmetic
    double dr, dc;  // steps of the file header (cell index)
    double gr0, gdr, gc0, gdc;  // first point and spacing of the actual grids
};
struct TabRot {
    int n, off;
    double x0, xmax, inv_h;
};

// 3D potential table of vcalc.f:1-65: V[(ir*nth + ith)*nch + ich], R in bohr on a uniform grid
struct Tab3D {
    int nr, nth, nch, pad;
    double rmin, rmax, rstep;
    long long off;
};

struct Tables {
    const float4* k1f;
    const double4* k1d;
    const Tab1D* t1;
    const float* v2f;
    const float4* v2c;  // per cell of a uniform table: V[ir][ic], V[ir][ic+1], V[ir+1][ic], V[ir+1][ic+1]
    const double* v2d;
    const double* g2d;
    const float* g2f;
    const Tab2D* t2;
    const double4* rk;  // x, rho, rho'', -
    const double4* re;  // erot, erot'', esq, esq''
    const TabRot* tr;
};

// tables of the non-linear rotors, a kernel argument of its own so kernels of the linear
// family pay no registers for them
// A grid point of a 3D table holds its value and the forward differences to the next point
// along each axis (zero at the last point of an axis), in the order of its look-up
// (vcalc: r, theta, chi; rotpro: chi, phi, theta), so one look-up is one aligned load.
struct Tables3 {
    const double4* v3d;  // 3D potential (vtable order): V, dV_r, dV_theta, dV_chi, FP64
    const float4* v3f;   // the same in FP32 (differences taken in FP32)
    const Tab3D* t3;
    const double4* r3;   // per density table: rho, then E, then E^2 blocks of ROT3D_SIZE points,
                         // each point (value, d_chi, d_phi, d_theta)
};

// rotation matrix of the Euler angles of one rotor slice (matpre, rotden.f:136-163), Fortran
// column order m[(j-1)*3 + (i-1)] = rotmat(i, j), so the body axes are its three columns;
// zn = |z axis| as vcord's dotang takes it (vcord.f:176-188).  80 bytes, 16-byte aligned: five
// 16-byte loads per frame, and the split water sweep's LDS copies leave room for two parts per CU
struct alignas(16) Frame {
    double m[9];
    double zn;
};

constexpr long long ROT3D_SIZE = 181LL * 361 * 361;  // rotpro_sub.f:7-26, mc_setup.h:116

// ---------------------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).  Counter map:
//   key = {seed, replica}; ctr = {pass, kind<<28 | step, atom<<16 | slice, level<<24 | point<<8 | id}
__device__ __forceinline__ uint4 philox(uint4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        uint32_t lo0 = 0xD2511F53u * c.x, hi0 = __umulhi(0xD2511F53u, c.x);
        uint32_t lo1 = 0xCD9E8D57u * c.z, hi1 = __umulhi(0xCD9E8D57u, c.z);
        c = make_uint4(hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0);
    }
    return c;
}

__device__ __forceinline__ double u53(uint32_t a, uint32_t b) {
    return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * (1.0 / 9007199254740992.0);
}

__device__ __forceinline__ uint4 draw(uint32_t seed, uint32_t key, uint32_t pass, uint32_t kind, uint32_t step,
                                      uint32_t atom, uint32_t slice, uint32_t c3) {
    return philox(make_uint4(pass, (kind << 28) | (step & 0x0FFFFFFFu), (atom << 16) | (slice & 0xFFFFu), c3), seed,
                  key);
}

// Gaussian of density ~ exp(-alpha x^2), Box-Muller cosine branch (gauss() convention).
__device__ __forceinline__ double gauss_from(uint4 w, double alpha) {
    double r1 = 1.0 - u53(w.x, w.y);
    double r2 = u53(w.z, w.w);
    return sqrt(-log(r1)) * cos(6.283185307179586476925286766559 * r2) / sqrt(alpha);
}

// ---------------------------------------------------------------------------------------
// Interpolation.  Table descriptors are resolved once per kernel, so a lookup costs one
// round trip to the table.  Cell search: a guess from the mean spacing, then local
// correction to the cell a bisection search would return (largest k with x[k] <= r,
// capped at n-2).
template <typename R>
using vec4_t = typename std::conditional<std::is_same<R, float>::value, float4, double4>::type;

template <typename R>
struct Lut1D {
    const vec4_t<R>* k;  // knots (x, y, y'', 1/(x[i+1]-x[i]))
    int n, uniform;
    R x0, xmax, inv_h, alpha, unode, c6;
};

template <typename R>
struct Lut2D {
    const R* v;  // V[r][cos]
    const R* g;  // r grid, then cos grid
    const float4* vc;  // corner-packed cells (FP32 uniform tables)
    int nr, nc, uniform;
    R rmin, cmin, dr, dc, inv_dr, inv_dc;  // grid origins and file-header steps
    R gdr, gdc, inv_gdr, inv_gdc;          // actual grid spacings (uniform grids)
};

__device__ __forceinline__ const float4* knots_ptr(const Tables& T, float) { return T.k1f; }
__device__ __forceinline__ const double4* knots_ptr(const Tables& T, double) { return T.k1d; }
__device__ __forceinline__ const float* values_ptr(const Tables& T, float) { return T.v2f; }
__device__ __forceinline__ const double* values_ptr(const Tables& T, double) { return T.v2d; }
__device__ __forceinline__ const float* grid_ptr(const Tables& T, float) { return T.g2f; }
__device__ __forceinline__ const double* grid_ptr(const Tables& T, double) { return T.g2d; }

template <typename R>
__device__ __forceinline__ Lut1D<R> lut1d(const Tables& T, int tid) {
    Lut1D<R> L;
    if (tid < 0) {
        L.k = nullptr;
        L.n = 0;
        L.uniform = 0;
        L.x0 = L.xmax = L.inv_h = L.alpha = L.unode = L.c6 = (R)0;
        return L;
    }
    const Tab1D t = T.t1[tid];
    L.k = knots_ptr(T, R()) + t.off;
    L.n = t.n;
    L.uniform = t.uniform;
    L.x0 = (R)t.x0;
    L.xmax = (R)t.xmax;
    L.inv_h = (R)t.inv_h;
    L.alpha = (R)t.alpha;
    L.unode = (R)t.unode;
    L.c6 = (R)t.c6;
    return L;
}

template <typename R>
__device__ __forceinline__ Lut2D<R> lut2d(const Tables& T, int tid) {
    Lut2D<R> L;
    if (tid < 0) {
        L.v = nullptr;
        L.g = nullptr;
        L.vc = nullptr;
        L.nr = L.nc = 2;
        L.uniform = 0;
        L.rmin = L.cmin = L.dr = L.dc = L.inv_dr = L.inv_dc = (R)1;
        L.gdr = L.gdc = L.inv_gdr = L.inv_gdc = (R)1;
        return L;
    }
    const Tab2D t = T.t2[tid];
    L.v = values_ptr(T, R()) + t.off;
    L.g = grid_ptr(T, R()) + t.goff;
    L.vc = t.coff >= 0 ? T.v2c + t.coff : nullptr;
    L.nr = t.nr;
    L.nc = t.nc;
    L.uniform = t.uniform;
    L.rmin = L.g[0];
    L.cmin = L.g[t.nr];
    L.dr = (R)t.dr;
    L.dc = (R)t.dc;
    L.inv_dr = (R)(1.0 / t.dr);
    L.inv_dc = (R)(1.0 / t.dc);
    L.gdr = (R)t.gdr;
    L.gdc = (R)t.gdc;
    L.inv_gdr = (R)(1.0 / t.gdr);
    L.inv_gdc = (R)(1.0 / t.gdc);
    return L;
}

template <typename R>
constexpr bool is_fp32() {
    return std::is_same<R, float>::value;
}

// FP32 square root and reciprocal of the move kernels: the hardware instructions (1 ulp)
// instead of the correctly rounded sequences; FP64 keeps IEEE arithmetic.
__device__ __forceinline__ float fast_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ double fast_sqrt(double x) { return sqrt(x); }
__device__ __forceinline__ float fast_div(float x, float y) { return x * __builtin_amdgcn_rcpf(y); }
__device__ __forceinline__ double fast_div(double x, double y) { return x / y; }

// SPot1D: spline inside the grid, U0 exp(-alpha r) below it, -C6/r^6 above it.
// `k` is the knot array (LDS or global copy of L.k).  FP64 keeps the reference's
// arithmetic; FP32 multiplies by the stored inverse cell width instead of dividing, and on
// uniform knots takes the cell from the grid arithmetic alone (at a cell boundary the two
// cells' cubics agree to rounding) and selects between the spline and the tails without
// branches, so that independent look-ups interleave.
// The tails of the uniform FP32 spline (mc_poten.cc:629-634): U0 e^{-alpha r} at or below the
// first knot, -C6/r^6 at or beyond the last, in branches that a wave skips unless one of its
// lanes needs them.  The tests are wave-uniform because a masked region this short would still
// be issued, and the empty asm keeps the compiler from computing the tails ahead of the test.
__device__ __forceinline__ float spot1d_tails(const Lut1D<float>& L, float r, float v) {
    if (__any(r <= L.x0)) {
        float rw = r;
        asm volatile("" : "+v"(rw));
        if (rw <= L.x0) v = L.unode * __expf(-L.alpha * rw);
    }
    if (__any(r >= L.xmax)) {
        float rw = r;
        asm volatile("" : "+v"(rw));
        if (rw >= L.xmax) {
            const float r2 = rw * rw;
            v = -fast_div(L.c6, r2 * r2 * r2);
        }
    }
    return v;
}

template <typename R>
__device__ __forceinline__ R spot1d(const Lut1D<R>& L, const vec4_t<R>* __restrict__ k, R r) {
    if constexpr (is_fp32<R>()) {
        if (L.uniform) {  // the table without branches, the tails behind wave-uniform tests
            float f = floor((r - L.x0) * L.inv_h);
            f = fmin(fmax(f, 0.f), (float)(L.n - 2));
            const int i = (int)f;
            const float4 a = k[i], b = k[i + 1];
            const float A = (b.x - r) * a.w, B = (r - a.x) * a.w, h = b.x - a.x;
            const float v = A * a.y + B * b.y + ((A * A * A - A) * a.z + (B * B * B - B) * b.z) * (h * h) * (1.0f / 6.0f);
            return spot1d_tails(L, r, v);
        }
    }
    if (r >= L.xmax) {
        R r2 = r * r;
        return -fast_div(L.c6, r2 * r2 * r2);
    }
    if constexpr (is_fp32<R>()) {
        if (r <= L.x0) return L.unode * __expf(-L.alpha * r);
    } else {
        if (r <= L.x0) return L.unode * exp(-L.alpha * r);
    }
    R f = floor((r - L.x0) * L.inv_h);
    f = fmin(fmax(f, (R)0), (R)(L.n - 2));
    int i = (int)f;
    vec4_t<R> a = k[i], b = k[i + 1];
    while (i > 0 && a.x > r) {
        --i;
        b = a;
        a = k[i];
    }
    while (i < L.n - 2 && b.x <= r) {
        ++i;
        a = b;
        b = k[i + 1];
    }
    R h = b.x - a.x;
    if constexpr (is_fp32<R>()) {
        R A = (b.x - r) * a.w, B = (r - a.x) * a.w;
        return A * a.y + B * b.y + ((A * A * A - A) * a.z + (B * B * B - B) * b.z) * (h * h) * (R)(1.0 / 6.0);
    } else {
        R A = (b.x - r) / h, B = (r - a.x) / h;
        return A * a.y + B * b.y + ((A * A * A - A) * a.z + (B * B * B - B) * b.z) * (h * h) / (R)6;
    }
}

// SPot1D on uniform FP32 knots from (y, y'') pairs alone (x and the cell width follow from the
// grid), for kernels that stage the knots in LDS at half the size; the same branchless form
__device__ __forceinline__ float spot1d_u2(const Lut1D<float>& L, const float2* __restrict__ k2, float r) {
    const float t = (r - L.x0) * L.inv_h;
    float f = floor(t);
    f = fmin(fmax(f, 0.f), (float)(L.n - 2));
    const int i = (int)f;
    const float2 a = k2[i], b = k2[i + 1];
    const float h = fast_div(1.0f, L.inv_h);
    const float B = t - f, A = 1.0f - B;
    const float v = A * a.x + B * b.x + ((A * A * A - A) * a.y + (B * B * B - B) * b.y) * (h * h) * (1.0f / 6.0f);
    return spot1d_tails(L, r, v);
}

// LPot2D on a uniform FP32 grid in two steps, so that the loads of many look-ups can be in
// flight together: the cell and the weights, then the lerps of its four corner values, which
// the corner-packed copy holds in one 16-byte load.
__device__ __forceinline__ void lpot2d_u_cell(const Lut2D<float>& L, float r, float c, int& cell, float& wr,
                                              float& wc) {
    float fr = floor((r - L.rmin) * L.inv_dr), fc = floor((c - L.cmin) * L.inv_dc);
    fr = fmin(fmax(fr, 0.f), (float)(L.nr - 2));
    fc = fmin(fmax(fc, 0.f), (float)(L.nc - 2));
    cell = (int)fr * (L.nc - 1) + (int)fc;
    wr = (r - (L.rmin + fr * L.gdr)) * L.inv_gdr;
    wc = (c - (L.cmin + fc * L.gdc)) * L.inv_gdc;
}

__device__ __forceinline__ float lpot2d_u_lerp(float y1, float y4, float y2, float y3, float wr, float wc) {
    const float lo = y1 + wc * (y4 - y1), hi = y2 + wc * (y3 - y2);
    return lo + wr * (hi - lo);
}

// LPot2D: cells clamped to the grid, weights extrapolate linearly off-grid.
template <typename R>
__device__ __forceinline__ R lpot2d(const Lut2D<R>& L, R r, R c) {
    if constexpr (is_fp32<R>()) {
        if (L.uniform) {  // grid positions by arithmetic, reciprocal multiplies, lerp form
            int cell;
            float wr, wc;
            lpot2d_u_cell(L, r, c, cell, wr, wc);
            const float4 q = L.vc[cell];
            return lpot2d_u_lerp(q.x, q.y, q.z, q.w, wr, wc);
        }
    }
    R fr = floor((r - L.rmin) / L.dr), fc = floor((c - L.cmin) / L.dc);
    fr = fmin(fmax(fr, (R)0), (R)(L.nr - 2));
    fc = fmin(fmax(fc, (R)0), (R)(L.nc - 2));
    int ir = (int)fr, ic = (int)fc;
    const R* v = L.v + (long long)ir * L.nc + ic;
    R y1 = v[0], y4 = v[1];
    R y2 = v[L.nc], y3 = v[L.nc + 1];
    R r1 = L.g[ir], r2 = L.g[ir + 1];
    R c1 = L.g[L.nr + ic], c2 = L.g[L.nr + ic + 1];
    R dr = (r - r1) / (r2 - r1);
    R dc = (c - c1) / (c2 - c1);
    return ((R)1 - dr) * ((R)1 - dc) * y1 + dr * ((R)1 - dc) * y2 + dr * dc * y3 + ((R)1 - dr) * dc * y4;
}

// SRotDens (mc_poten.cc:548-570) with the table descriptor already resolved: the last value
// above the grid, the line through the first two knots below it, the spline inside.
__device__ __forceinline__ double rot_rho(const Tables& T, const TabRot& t, double g) {
    const double4* k = T.rk + t.off;
    if (g > t.xmax) return k[t.n - 1].y;
    if (g < t.x0) {
        double4 k0 = k[0], k1 = k[1];
        double sa = (k1.y - k0.y) / (k1.x - k0.x);
        double sb = (k0.y * k1.x - k1.y * k0.x) / (k1.x - k0.x);
        return sa * g + sb;
    }
    double f = floor((g - t.x0) * t.inv_h);
    f = fmin(fmax(f, 0.0), (double)(t.n - 2));
    int i = (int)f;
    double4 a = k[i], b = k[i + 1];
    while (i > 0 && a.x > g) {
        --i;
        b = a;
        a = k[i];
    }
    while (i < t.n - 2 && b.x <= g) {
        ++i;
        a = b;
        b = k[i + 1];
    }
    double h = b.x - a.x;
    double A = (b.x - g) / h, B = (g - a.x) / h;
    return A * a.y + B * b.y + ((A * A * A - A) * a.z + (B * B * B - B) * b.z) * (h * h) / 6.;
}

// rot_rho in two steps, so that its knot loads can be in flight with other loads: the knots of
// the cell the grid arithmetic guesses, then the value (with the cell corrected if needed).
struct RhoCell {
    double4 a, b;
    int i;
};

__device__ __forceinline__ RhoCell rot_rho_load(const Tables& T, const TabRot& t, double g) {
    const double4* k = T.rk + t.off;
    double f = floor((g - t.x0) * t.inv_h);
    f = fmin(fmax(f, 0.0), (double)(t.n - 2));
    RhoCell c;
    c.i = (int)f;
    c.a = k[c.i];
    c.b = k[c.i + 1];
    return c;
}

__device__ __forceinline__ double rot_rho_eval(const Tables& T, const TabRot& t, double g, RhoCell c) {
    const double4* k = T.rk + t.off;
    if (g > t.xmax) return k[t.n - 1].y;
    if (g < t.x0) {
        double4 k0 = k[0], k1 = k[1];
        double sa = (k1.y - k0.y) / (k1.x - k0.x);
        double sb = (k0.y * k1.x - k1.y * k0.x) / (k1.x - k0.x);
        return sa * g + sb;
    }
    int i = c.i;
    double4 a = c.a, b = c.b;
    while (i > 0 && a.x > g) {
        --i;
        b = a;
        a = k[i];
    }
    while (i < t.n - 2 && b.x <= g) {
        ++i;
        a = b;
        b = k[i + 1];
    }
    double h = b.x - a.x;
    double A = (b.x - g) / h, B = (g - a.x) / h;
    return A * a.y + B * b.y + ((A * A * A - A) * a.z + (B * B * B - B) * b.z) * (h * h) / 6.;
}

// .rot spline in FP64 (x, f, f'') at column `col`: 0 rho, 1 erot, 2 esq.
__device__ __forceinline__ double rot_spline(const Tables& T, int tid, double g, int col) {
    const TabRot t = T.tr[tid];
    if (g > t.xmax) {
        if (col == 0) return T.rk[t.off + t.n - 1].y;
        return 0.0;
    }
    if (g < t.x0) {
        double4 k0 = T.rk[t.off], k1 = T.rk[t.off + 1];
        double y0, y1;
        if (col == 0) { y0 = k0.y; y1 = k1.y; }
        else if (col == 1) { y0 = T.re[t.off].x; y1 = T.re[t.off + 1].x; }
        else { y0 = T.re[t.off].z; y1 = T.re[t.off + 1].z; }
        double rl = k0.x, rr = k1.x;
        double sa = (y1 - y0) / (rr - rl);
        double sb = (y0 * rr - y1 * rl) / (rr - rl);
        return sa * g + sb;
    }
    double f = floor((g - t.x0) * t.inv_h);
    f = fmin(fmax(f, 0.0), (double)(t.n - 2));
    int k = (int)f;
    double xa = T.rk[t.off + k].x, xb = T.rk[t.off + k + 1].x;
    while (k > 0 && xa > g) {
        --k;
        xb = xa;
        xa = T.rk[t.off + k].x;
    }
    while (k < t.n - 2 && xb <= g) {
        ++k;
        xa = xb;
        xb = T.rk[t.off + k + 1].x;
    }
    double ya, yb, sa, sb;
    if (col == 0) {
        double4 ka = T.rk[t.off + k], kb = T.rk[t.off + k + 1];
        ya = ka.y; yb = kb.y; sa = ka.z; sb = kb.z;
    } else {
        double4 ea = T.re[t.off + k], eb = T.re[t.off + k + 1];
        if (col == 1) { ya = ea.x; yb = eb.x; sa = ea.y; sb = eb.y; }
        else { ya = ea.z; yb = eb.z; sa = ea.w; sb = eb.w; }
    }
    double h = xb - xa;
    double A = (xb - g) / h, B = (g - xa) / h;
    return A * ya + B * yb + ((A * A * A - A) * sa + (B * B * B - B) * sb) * (h * h) / 6.;
}

// ---------------------------------------------------------------------------------------
// Pair potential of a mover and a partner at one slice (dispatch of PotEnergy(int,double**,int),
// mc_piqmc.cc:1847-1958, for one atom type and one linear-molecule type per system,
// mc_input.cc:386-387).  rot_a / rot_j: the mover / partner is the linear molecule; (nx,ny,nz)
// is its axis at that slice.
// ... on the separation (dx, dy, dz) = mover - partner.  HALF: k1 points to (y, y'') pairs of
// uniform FP32 knots (spot1d_u2), otherwise to vec4 knots.
template <typename R, bool HALF = false>
__device__ __forceinline__ R pair_d(const Lut1D<R>& l1, const void* __restrict__ k1, const Lut2D<R>& l2,
                                    bool rot_a, bool rot_j, R dx, R dy, R dz, double nx, double ny, double nz) {
    const R r = fast_sqrt(dx * dx + dy * dy + dz * dz);
    if (rot_a || rot_j) {
        R cost = fast_div((R)nx * dx + (R)ny * dy + (R)nz * dz, r);
        if (rot_a) cost = -cost;
        return lpot2d<R>(l2, r, cost);
    }
    if constexpr (HALF && is_fp32<R>()) return spot1d_u2(l1, static_cast<const float2*>(k1), r);
    else return spot1d<R>(l1, static_cast<const vec4_t<R>*>(k1), r);
}

// MINIMAGE (mc_piqmc.cc:1239-1240,1832-1833; mc_estim.cc:536-537,916-917): component id of a
// separation moved to its nearest periodic image, d - BoxSize * rint(d / BoxSize); box == nullptr
// (no minimum image) leaves it as it is
template <typename T>
__device__ __forceinline__ T nearest_image(T d, const double* box, int id) {
    if (!box) return d;
    const T L = (T)box[id];
    return d - L * rint(d / L);
}

// positions in FP64: the separation is formed in FP64 and rounded once
template <typename R>
__device__ __forceinline__ R pair_v(const Lut1D<R>& l1, const vec4_t<R>* __restrict__ k1, const Lut2D<R>& l2,
                                    bool rot_a, bool rot_j, double xa, double ya, double za, double xj, double yj,
                                    double zj, double nx, double ny, double nz, const double* box = nullptr) {
    return pair_d<R>(l1, k1, l2, rot_a, rot_j, (R)nearest_image(xa - xj, box, 0), (R)nearest_image(ya - yj, box, 1),
                     (R)nearest_image(za - zj, box, 2), nx, ny, nz);
}

// positions already in the pair precision (the bisection window)
template <typename R, bool HALF = false>
__device__ __forceinline__ R pair_vr(const Lut1D<R>& l1, const void* __restrict__ k1, const Lut2D<R>& l2,
                                     bool rot_a, bool rot_j, R xa, R ya, R za, R xj, R yj, R zj, double nx, double ny,
                                     double nz, const double* box = nullptr) {
    return pair_d<R, HALF>(l1, k1, l2, rot_a, rot_j, nearest_image(xa - xj, box, 0), nearest_image(ya - yj, box, 1),
                           nearest_image(za - zj, box, 2), nx, ny, nz);
}

// One DPP lane move within rows of 16 lanes (ctrl: quad_perm 0x00-0xff, row_mirror 0x140,
// row_half_mirror 0x141) of a 32- or 64-bit value
template <int CTRL, typename T>
__device__ __forceinline__ T dpp_f(T v) {
    static_assert(sizeof(T) == 4 || sizeof(T) == 8, "32- or 64-bit values");
    T r;
    if constexpr (sizeof(T) == 4) {
        int b;
        __builtin_memcpy(&b, &v, 4);
        b = __builtin_amdgcn_update_dpp(0, b, CTRL, 0xf, 0xf, true);
        __builtin_memcpy(&r, &b, 4);
    } else {
        long long b;
        __builtin_memcpy(&b, &v, 8);
        const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffff), CTRL, 0xf, 0xf, true);
        const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xf, 0xf, true);
        const long long c = ((long long)hi << 32) | (long long)(unsigned)lo;
        __builtin_memcpy(&r, &c, 8);
    }
    return r;
}

// Lane l's copy of a 32- or 64-bit value, wave-uniform (v_readlane)
template <typename T>
__device__ __forceinline__ T read_lane(T v, int l) {
    T r;
    if constexpr (sizeof(T) == 4) {
        int b;
        __builtin_memcpy(&b, &v, 4);
        b = __builtin_amdgcn_readlane(b, l);
        __builtin_memcpy(&r, &b, 4);
    } else {
        long long b;
        __builtin_memcpy(&b, &v, 8);
        const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffff), l);
        const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
        const long long c = ((long long)hi << 32) | (long long)(unsigned)lo;
        __builtin_memcpy(&r, &c, 8);
    }
    return r;
}

// Sum over aligned groups of G lanes (G a power of two up to 64, all lanes of a group active);
// every lane of a group gets the same sum.  Within a row of 16 the butterfly runs on DPP moves
// (xor 1, xor 2 by quad permutes; then the half-row and row mirrors, which pair the quads and the
// half-rows once each holds one value), across rows on lane shuffles, or for the whole wave as
// the four row sums read into scalar registers and added as the shuffles would add them.
template <typename T>
__device__ __forceinline__ T group_sum(T v, int G) {
    if (G >= 2) v += dpp_f<0xb1>(v);   // quad_perm [1,0,3,2]
    if (G >= 4) v += dpp_f<0x4e>(v);   // quad_perm [2,3,0,1]
    if (G >= 8) v += dpp_f<0x141>(v);  // row_half_mirror
    if (G >= 16) v += dpp_f<0x140>(v); // row_mirror
    if (G == 64) return (read_lane(v, 0) + read_lane(v, 16)) + (read_lane(v, 32) + read_lane(v, 48));
    if (G >= 32) v += __shfl_xor(v, 16, 64);
    return v;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
    return group_sum(v, 64);
}

}  // namespace pimc
