// pimc_nl.hpp -- non-linear rotors: Euler frames, the density-matrix look-up of rotden/rotpro,
// the 3D rotor-particle potential of vcord/vcalc and the TIP4P rotor-rotor energy of caleng.
//
// A rotor slice's rotation matrix is computed once when its angles change and cached as a
// Frame; rotden's relative Euler angles are then R1^T R2 of two cached frames, and vcord's body
// axes are the frame's columns.  Density factors and caleng are FP64; vcord/vcalc run in the
// pair precision R (float in production, double in parity runs).
#pragma once
#include "pimc_device.hpp"

namespace pimc {

constexpr double PI_D = 3.14159265358979323846;

// matpre (rotden.f:136-163) of (phi, theta = acos(cos theta), chi), as MCRot3Dstep passes the
// angles (mc_piqmc.cc:1019-1024).  Host and device build frames with the same expressions.
__host__ __device__ inline Frame make_frame(double phi, double cth, double chi) {
    const double theta = acos(cth);
    const double cp = cos(phi), sp = sin(phi), ct = cos(theta), st = sin(theta), ck = cos(chi), sk = sin(chi);
    Frame f;
    f.m[0] = cp * ct * ck - sp * sk;   // (1,1)
    f.m[1] = sp * ct * ck + cp * sk;   // (2,1)
    f.m[2] = -st * ck;                 // (3,1)
    f.m[3] = -cp * ct * sk - sp * ck;  // (1,2)
    f.m[4] = -sp * ct * sk + cp * ck;  // (2,2)
    f.m[5] = st * sk;                  // (3,2)
    f.m[6] = cp * st;                  // (1,3)
    f.m[7] = sp * st;                  // (2,3)
    f.m[8] = ct;                       // (3,3)
    double d = 0.0;
    for (int i = 6; i < 9; ++i) d = d + f.m[i] * f.m[i];
    f.zn = sqrt(d);
    return f;
}

__device__ __forceinline__ double fm(const Frame& f, int i, int j) { return f.m[(j - 1) * 3 + (i - 1)]; }

__device__ __forceinline__ double within1(double v) {  // rotden.f:208-216
    if (v > 1.0) v = 1.0;
    if (v < -1.0) v = -1.0;
    return v;
}

// deleul (rotden.f:32-134): Euler angles of R(e1)^T R(e2), radians
__device__ inline void deleul_frames(const Frame& f1, const Frame& f2, double& phi2, double& theta2, double& chi2) {
    const double small = 1.0e-08;
    double r2[9];
    for (int i = 1; i <= 3; ++i)
        for (int j = 1; j <= 3; ++j) {
            double v = 0.0;
            for (int k = 1; k <= 3; ++k) v = v + fm(f1, k, i) * fm(f2, k, j);
            r2[(j - 1) * 3 + (i - 1)] = v;
        }
#define R2(i, j) r2[((j) - 1) * 3 + ((i) - 1)]
    const double cost = within1(R2(3, 3));
    theta2 = acos(cost);
    const double sint = sin(theta2);
    if (fabs(1.0 - cost) < small) {
        phi2 = 0.0;
        const double cchi = within1(R2(1, 1)), schi = within1(R2(2, 1));
        chi2 = (schi > 0.0) ? acos(cchi) : 2.0 * PI_D - acos(cchi);
    } else if (fabs(1.0 + cost) < small) {
        phi2 = 0.0;
        const double cchi = within1(R2(2, 2)), schi = within1(R2(1, 2));
        chi2 = (schi > 0.0) ? acos(cchi) : 2.0 * PI_D - acos(cchi);
    } else {
        const double cphi = within1(R2(1, 3) / sint), sphi = within1(R2(2, 3) / sint);
        const double cchi = within1(-R2(3, 1) / sint), schi = within1(R2(3, 2) / sint);
        phi2 = (sphi > 0.0) ? acos(cphi) : 2.0 * PI_D - acos(cphi);
        chi2 = (schi > 0.0) ? acos(cchi) : 2.0 * PI_D - acos(cchi);
    }
#undef R2
}

// rotden (rotden.f:1-31) with rotpro (rotpro_sub.f:1-64): corner value plus one forward
// difference per axis on the 1-degree grid, index (itheta*361 + iphi)*361 + ichi; E and E^2
// converted from cm^-1 to Kelvin.  Returns 1 for an out-of-range angle (the reference's istop).
// rho alone, for the moves: one 32-byte gather instead of three
__device__ inline int rotden_rho(const Frame& f1, const Frame& f2, const double4* __restrict__ tab, double& rho) {
    double p2, t2, c2;
    deleul_frames(f1, f2, p2, t2, c2);
    const double phi = p2 * 180.0 / PI_D, theta = t2 * 180.0 / PI_D, chi = c2 * 180.0 / PI_D;
    int stop = 0;
    int ichi = (int)chi, iphi = (int)phi, ith = (int)theta;
    if (ichi > 360 || ichi < 0) { ichi = 0; stop = 1; }
    if (iphi > 360 || iphi < 0) { iphi = 0; stop = 1; }
    if (ith > 180 || ith < 0) { ith = 0; stop = 1; }
    const long long i0 = ((long long)ith * 361 + iphi) * 361 + ichi;
    const double xc = chi - (double)ichi, xp = phi - (double)iphi, xt = theta - (double)ith;
    const double4 r = tab[i0];
    rho = r.x + r.y * xc + r.z * xp + r.w * xt;
    return stop;
}

// rel (if not null) receives the relative Euler angles (phi, theta, chi) in radians
__device__ inline int rotden_frames(const Frame& f1, const Frame& f2, const double4* __restrict__ tab, double& rho,
                                    double& erot, double& esq, double* rel = nullptr) {
    const double wno2k = 0.6950356;
    double p2, t2, c2;
    deleul_frames(f1, f2, p2, t2, c2);
    if (rel) {
        rel[0] = p2;
        rel[1] = t2;
        rel[2] = c2;
    }
    const double phi = p2 * 180.0 / PI_D, theta = t2 * 180.0 / PI_D, chi = c2 * 180.0 / PI_D;
    int stop = 0;
    int ichi = (int)chi, iphi = (int)phi, ith = (int)theta;
    if (ichi > 360 || ichi < 0) { ichi = 0; stop = 1; }
    if (iphi > 360 || iphi < 0) { iphi = 0; stop = 1; }
    if (ith > 180 || ith < 0) { ith = 0; stop = 1; }
    const long long i0 = ((long long)ith * 361 + iphi) * 361 + ichi;
    const double xc = chi - (double)ichi, xp = phi - (double)iphi, xt = theta - (double)ith;
    const double4 r = tab[i0];
    rho = r.x + r.y * xc + r.z * xp + r.w * xt;
    const double4 e = tab[ROT3D_SIZE + i0];
    erot = (e.x + e.y * xc + e.z * xp + e.w * xt) / wno2k;
    const double4 q = tab[2 * ROT3D_SIZE + i0];
    esq = (q.x + q.y * xc + q.z * xp + q.w * xt) / (wno2k * wno2k);
    return stop;
}

template <typename R>
struct Lut3D {
    const vec4_t<R>* v;
    int nr, nth, nch;
    R rmin, rmax, rstep, inv_rstep;
};

__device__ __forceinline__ const float4* v3_ptr(const Tables3& T, float) { return T.v3f; }
__device__ __forceinline__ const double4* v3_ptr(const Tables3& T, double) { return T.v3d; }

template <typename R>
__device__ __forceinline__ Lut3D<R> lut3d(const Tables3& T, int tid) {
    Lut3D<R> L;
    if (tid < 0) {
        L.v = nullptr;
        L.nr = L.nth = L.nch = 2;
        L.rmin = L.rmax = L.rstep = L.inv_rstep = (R)1;
        return L;
    }
    const Tab3D t = T.t3[tid];
    L.v = v3_ptr(T, R()) + t.off;
    L.nr = t.nr;
    L.nth = t.nth;
    L.nch = t.nch;
    L.rmin = (R)t.rmin;
    L.rmax = (R)t.rmax;
    L.rstep = (R)t.rstep;
    L.inv_rstep = (R)(1.0 / t.rstep);
    return L;
}

// vcalc (vcalc.f:1-65): r in bohr clamped to the grid, theta and chi in degrees; corner value
// plus one forward difference per axis
template <typename R>
__device__ __forceinline__ R vcalc(const Lut3D<R>& L, R r, R theta, R chi) {
    if (r < L.rmin) r = L.rmin;
    if (r > L.rmax) r = L.rmax;
    int ir;
    if constexpr (std::is_same<R, float>::value) ir = (int)((r - L.rmin) * L.inv_rstep);
    else ir = (int)((r - L.rmin) / L.rstep);
    int ith = (int)theta, ich = (int)chi;
    if (ich > L.nch - 1) ich = L.nch - 1;
    if (ich < 0) ich = 0;
    if (ith > L.nth - 1) ith = L.nth - 1;
    if (ith < 0) ith = 0;
    if (ir > L.nr - 1) ir = L.nr - 1;  // r == rmax with rounding up
    const vec4_t<R> q = L.v[((long long)ir * L.nth + ith) * L.nch + ich];
    // the differences are zero at an axis' last point, where vcalc drops that term
    return q.x + (q.y / L.rstep) * (r - (L.rmin + (R)ir * L.rstep)) + q.z * (theta - (R)ith) + q.w * (chi - (R)ich);
}

// vcord (vcord.f:1-98): partner at rpt around the rotor centre rcom with body frame f
template <typename R>
__device__ __forceinline__ R vcord_v(const Lut3D<R>& L, const Frame& f, double cx, double cy, double cz, double px,
                                     double py, double pz) {
    const R pi = (R)PI_D;
    const R rx = (R)(px - cx), ry = (R)(py - cy), rz = (R)(pz - cz);
    const R hx0 = (R)f.m[0], hx1 = (R)f.m[1], hx2 = (R)f.m[2];
    const R hy0 = (R)f.m[3], hy1 = (R)f.m[4], hy2 = (R)f.m[5];
    const R hz0 = (R)f.m[6], hz1 = (R)f.m[7], hz2 = (R)f.m[8];
    R dz = (R)0;
    dz = dz + hz0 * rx;
    dz = dz + hz1 * ry;
    dz = dz + hz2 * rz;
    R rr = (R)0;
    rr = rr + rx * rx;
    rr = rr + ry * ry;
    rr = rr + rz * rz;
    const R rn = sqrt(rr);
    R da = dz / ((R)f.zn * rn);
    if (da > (R)1) da = (R)1;
    if (da < (R)-1) da = (R)-1;
    R the = acos(da);
    R rdx = (R)0, rdy = (R)0;
    rdx = rdx + rx * hx0;
    rdx = rdx + ry * hx1;
    rdx = rdx + rz * hx2;
    rdy = rdy + rx * hy0;
    rdy = rdy + ry * hy1;
    rdy = rdy + rz * hy2;
    R chiw;
    if (fabs(rdx) < (R)1.0e-08) chiw = pi / (R)2;
    else chiw = atan(fabs(rdy / rdx));
    R chir;
    if (rdx >= (R)0 && rdy >= (R)0) chir = chiw;
    else if (rdx < (R)0 && rdy >= (R)0) chir = pi - chiw;
    else if (rdx < (R)0 && rdy < (R)0) chir = pi + chiw;
    else chir = (R)2 * pi - chiw;
    if (L.nch == 181) {
        chiw = chir;
        if (chir > pi) chiw = (R)2 * pi - chir;
    } else if (L.nch == 361) {
        chiw = chir;
    }
    const R rad = rn / (R)0.529177249;
    return vcalc<R>(L, rad, the * (R)180 / pi, chiw * (R)180 / pi);
}

// vcord's geometry for the density bins (vcord.f:40-75), FP64: radret = |rpt - rcom| in
// Angstrom, theret = angle to the body z axis, chiret = azimuth about it in [0, 2 pi)
__device__ inline void vcord_geometry(const Frame& f, double cx, double cy, double cz, double px, double py, double pz,
                                      double& rad, double& the, double& chi) {
    const double rx = px - cx, ry = py - cy, rz = pz - cz;
    double dz = 0.0;
    dz = dz + f.m[6] * rx;
    dz = dz + f.m[7] * ry;
    dz = dz + f.m[8] * rz;
    double rr = 0.0;
    rr = rr + rx * rx;
    rr = rr + ry * ry;
    rr = rr + rz * rz;
    rad = sqrt(rr);
    the = acos(within1(dz / (f.zn * rad)));
    double rdx = 0.0, rdy = 0.0;
    rdx = rdx + rx * f.m[0];
    rdx = rdx + ry * f.m[1];
    rdx = rdx + rz * f.m[2];
    rdy = rdy + rx * f.m[3];
    rdy = rdy + ry * f.m[4];
    rdy = rdy + rz * f.m[5];
    const double chiw = fabs(rdx) < 1.0e-08 ? PI_D / 2.0 : atan(fabs(rdy / rdx));
    if (rdx >= 0.0 && rdy >= 0.0) chi = chiw;
    else if (rdx < 0.0 && rdy >= 0.0) chi = PI_D - chiw;
    else if (rdx < 0.0 && rdy < 0.0) chi = PI_D + chiw;
    else chi = 2 * PI_D - chiw;
}

// caleng (caleng_tip4p_gg.f:2-186): rigid TIP4P dimer in Kelvin, always FP64 (the Coulomb
// terms of distant pairs cancel to a small fraction of their size).  kcal2k is an INTEGER
// parameter there (implicit typing of names starting with k), so the LJ term is scaled by 503.
__device__ inline double caleng_v(double c1x, double c1y, double c1z, double c2x, double c2y, double c2z,
                                  const Frame& f1, const Frame& f2) {
    const double qm = -1.04, qh = 0.520, br2ang = 0.52917721092, hr2k = 3.1577465e5, kcal2k = 503.0;
    const double site[4][3] = {{0.0, 0.0, 0.06562}, {0.0, 0.0, -0.08438}, {0.7557, 0.0, -0.5223},
                               {-0.7557, 0.0, -0.5223}};  // O, M, H1, H2
    double s1[4][3], s2[4][3];
    const double c1[3] = {c1x, c1y, c1z}, c2[3] = {c2x, c2y, c2z};
    for (int a = 0; a < 4; ++a)
        for (int i = 1; i <= 3; ++i) {
            double v = c1[i - 1], w = c2[i - 1];
            for (int j = 1; j <= 3; ++j) {
                v = v + fm(f1, i, j) * site[a][j - 1];
                w = w + fm(f2, i, j) * site[a][j - 1];
            }
            s1[a][i - 1] = v;
            s2[a][i - 1] = w;
        }
    double roo = 0.0, rmm = 0.0;
    for (int i = 0; i < 3; ++i) {
        roo = roo + (s1[0][i] - s2[0][i]) * (s1[0][i] - s2[0][i]);
        rmm = rmm + (s1[1][i] - s2[1][i]) * (s1[1][i] - s2[1][i]);
    }
    rmm = sqrt(rmm);
    const double roo4 = roo * roo, roo6 = roo4 * roo, roo12 = roo6 * roo6;
    const double vlj = 6.e5 / roo12 - 610.0 / roo6;
    double rhm[4] = {0, 0, 0, 0}, rhh[4] = {0, 0, 0, 0};
    for (int i = 0; i < 3; ++i) {
        rhm[0] = rhm[0] + (s1[1][i] - s2[2][i]) * (s1[1][i] - s2[2][i]);
        rhm[1] = rhm[1] + (s1[1][i] - s2[3][i]) * (s1[1][i] - s2[3][i]);
        rhm[2] = rhm[2] + (s2[1][i] - s1[2][i]) * (s2[1][i] - s1[2][i]);
        rhm[3] = rhm[3] + (s2[1][i] - s1[3][i]) * (s2[1][i] - s1[3][i]);
        rhh[0] = rhh[0] + (s1[2][i] - s2[2][i]) * (s1[2][i] - s2[2][i]);
        rhh[1] = rhh[1] + (s1[2][i] - s2[3][i]) * (s1[2][i] - s2[3][i]);
        rhh[2] = rhh[2] + (s1[3][i] - s2[2][i]) * (s1[3][i] - s2[2][i]);
        rhh[3] = rhh[3] + (s1[3][i] - s2[3][i]) * (s1[3][i] - s2[3][i]);
    }
    for (int k = 0; k < 4; ++k) {
        rhm[k] = sqrt(rhm[k]);
        rhh[k] = sqrt(rhh[k]);
    }
    const double vmh = qm * qh * (1.0 / rhm[0] + 1.0 / rhm[1] + 1.0 / rhm[2] + 1.0 / rhm[3]);
    const double vhh = qh * qh * (1.0 / rhh[0] + 1.0 / rhh[1] + 1.0 / rhh[2] + 1.0 / rhh[3]);
    const double vmm = qm * qm * (1.0 / rmm);
    return vlj * kcal2k + (vmh + vmm + vhh) * hr2k * br2ang;
}

}  // namespace pimc
