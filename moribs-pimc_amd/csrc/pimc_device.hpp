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

namespace pimc {

constexpr int MAXT = 4;
constexpr double RZERO = 1.0e-10;  // zero of the rotational density (mc_confg.h:60)
constexpr int BINSR = 300;         // radial bins (mc_estim.cc:25)
constexpr int BINST = 50;          // polar bins (mc_estim.cc:26)
constexpr double MAX_RADIUS = 15.0;

enum : uint32_t { K_CENTROID = 1u, K_BISECT = 2u, K_ROTLIN = 3u, K_ROT3D = 4u };

struct Rep {
    int natoms, P, Prot, RotRatio, ntypes, imtype, rotkind, S;
    int tnumb[MAXT], tmol[MAXT], tlev[MAXT], tbegin[MAXT];
    int pot1d[MAXT], pot2d[MAXT], rotdens;
    int rot_atom;  // global index of the (first) rotor atom, -1 if none
    int key;       // replica index, second word of the generator key
    double lambda[MAXT], mcstep[MAXT];
    double T, beta, tau, rtstep;
    long long bead_off, rot_off, hist_off, sample_off;
};

struct Mut {
    unsigned pass;
    int err;
    unsigned nsamples, pad;
    unsigned long long total[MAXT][3];
    unsigned long long accep[MAXT][3];
    unsigned long long pair_evals;
};

struct RotBead {
    double phi, cth, chi, pad0;
    double nx, ny, nz, pad1;
};

struct Tab1D {
    int n, off;
    double x0, xmax, inv_h, alpha, unode, c6;
};
struct Tab2D {
    int nr, nc;
    long long off;  // into values
    int goff;       // into grids: r grid at goff, cos grid at goff + nr
    int pad;
    double dr, dc;
};
struct TabRot {
    int n, off;
    double x0, xmax, inv_h;
};

struct Tables {
    const float4* k1f;
    const double4* k1d;
    const Tab1D* t1;
    const float* v2f;
    const double* v2d;
    const double* g2d;
    const float* g2f;
    const Tab2D* t2;
    const double4* rk;  // x, rho, rho'', -
    const double4* re;  // erot, erot'', esq, esq''
    const TabRot* tr;
};

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
// Interpolation.  Cell search: a guess from the mean spacing, then local correction to the
// cell a bisection search would return (largest k with x[k] <= r, capped at n-2).
__device__ __forceinline__ float4 knot1(const Tables& T, int i, float) { return T.k1f[i]; }
__device__ __forceinline__ double4 knot1(const Tables& T, int i, double) { return T.k1d[i]; }

template <typename R>
__device__ __forceinline__ R spot1d(const Tables& T, int tid, R r) {
    const Tab1D t = T.t1[tid];
    if (r >= (R)t.xmax) {
        R r2 = r * r;
        return -(R)t.c6 / (r2 * r2 * r2);
    }
    if (r <= (R)t.x0) return (R)t.unode * exp(-(R)t.alpha * r);
    R f = floor((r - (R)t.x0) * (R)t.inv_h);
    f = fmin(fmax(f, (R)0), (R)(t.n - 2));
    int k = (int)f;
    auto a = knot1(T, t.off + k, R());
    auto b = knot1(T, t.off + k + 1, R());
    while (k > 0 && a.x > r) {
        --k;
        b = a;
        a = knot1(T, t.off + k, R());
    }
    while (k < t.n - 2 && b.x <= r) {
        ++k;
        a = b;
        b = knot1(T, t.off + k + 1, R());
    }
    R h = b.x - a.x;
    R A = (b.x - r) / h, B = (r - a.x) / h;
    return A * a.y + B * b.y + ((A * A * A - A) * a.z + (B * B * B - B) * b.z) * (h * h) / (R)6;
}

__device__ __forceinline__ float v2(const Tables& T, long long i, float) { return T.v2f[i]; }
__device__ __forceinline__ double v2(const Tables& T, long long i, double) { return T.v2d[i]; }
__device__ __forceinline__ float g2(const Tables& T, int i, float) { return T.g2f[i]; }
__device__ __forceinline__ double g2(const Tables& T, int i, double) { return T.g2d[i]; }

// LPot2D: cells clamped to the grid, weights extrapolate linearly off-grid.
template <typename R>
__device__ __forceinline__ R lpot2d(const Tables& T, int tid, R r, R c) {
    const Tab2D t = T.t2[tid];
    R rmin = g2(T, t.goff, R()), cmin = g2(T, t.goff + t.nr, R());
    R fr = floor((r - rmin) / (R)t.dr), fc = floor((c - cmin) / (R)t.dc);
    fr = fmin(fmax(fr, (R)0), (R)(t.nr - 2));
    fc = fmin(fmax(fc, (R)0), (R)(t.nc - 2));
    int ir = (int)fr, ic = (int)fc;
    long long base = t.off + (long long)ir * t.nc + ic;
    R y1 = v2(T, base, R()), y4 = v2(T, base + 1, R());
    R y2 = v2(T, base + t.nc, R()), y3 = v2(T, base + t.nc + 1, R());
    R r1 = g2(T, t.goff + ir, R()), r2 = g2(T, t.goff + ir + 1, R());
    R c1 = g2(T, t.goff + t.nr + ic, R()), c2 = g2(T, t.goff + t.nr + ic + 1, R());
    R dr = (r - r1) / (r2 - r1);
    R dc = (c - c1) / (c2 - c1);
    return ((R)1 - dr) * ((R)1 - dc) * y1 + dr * ((R)1 - dc) * y2 + dr * dc * y3 + ((R)1 - dr) * dc * y4;
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
// Pair potential of a mover (type ta) and a partner (type tj) at one slice; `nh` is the
// linear rotor's axis at that slice.  Dispatch of PotEnergy(int,double**,int).
template <typename R>
__device__ __forceinline__ R pair_v(const Tables& T, const Rep& rp, int ta, int tj, double xa, double ya,
                                    double za, double xj, double yj, double zj, double nx, double ny, double nz) {
    R dx = (R)(xa - xj), dy = (R)(ya - yj), dz = (R)(za - zj);
    R r = sqrt(dx * dx + dy * dy + dz * dz);
    int ma = rp.tmol[ta], mj = rp.tmol[tj];
    if (ma == 1 || mj == 1) {
        int typep = (ma == 1) ? ta : tj;
        R cost = ((R)nx * dx + (R)ny * dy + (R)nz * dz) / r;
        if (ma == 1) cost = -cost;
        return lpot2d<R>(T, rp.pot2d[typep], r, cost);
    }
    return spot1d<R>(T, rp.pot1d[tj], r);
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

}  // namespace pimc
