// pimc_gen.hpp -- free-rotor density matrices on the GPU: the linear rotor of linden
// (linear_prop/linden.f:88-122) and the asymmetric top of asymrho (nmv_prop/asymrho.f:34-731;
// J. Chem. Phys. 134, 054117 (2011) eqs 15-16).
//
// Asymmetric top.  On the whole-degree grid of relative Euler angles,
//   X_q(theta, phi, chi) = Re sum_{m,k} e^{i m phi} M_q(theta)[m,k] e^{i k chi},
//   M_q(theta)[m,k]      = sum_J (2J+1)/(8 pi^2) G_{J,q}[m,k] d^J_mk(theta),
//   G_{J,q}[m,k]         = sum_s exp(-tau E_Js) E_Js^q c_s(m) c_s(k),
// over the eigenstates s of each rigid-rotor block (J, parity of k), states with
// weight (2J+1)/(8 pi^2) e^{-tau E} below 1e-16 dropped.  rho = X_0, E = X_1 / X_0 and
// E^2 = X_2 / X_0 per parity, the parities combined as asymrho does.
//   host:      block Hamiltonians and their eigenvectors (Jacobi), G, the grid's fill map
//   k_asym_m:  M_q(theta) per (theta, parity): one lane per (m, k), d^J_mk by the three-term
//              recursion in J from its closed form at J = max(|m|, |k|)
//   k_dgemm:   e^{ik chi} and e^{im phi} applied as two FP64 products per (theta, parity, q)
//   k_asym_fill: parity combination and the symmetry fill of asymrho.f:665-709, one lane per point
#pragma once
#include <hip/hip_runtime.h>
#include <cmath>
#include <vector>
#include <numeric>

namespace pimc {

// ------------------------------------------------------------------------- linden
// One lane per cos(gamma) point: rho, E and E^2 numerators summed over l with the Legendre
// recursion, then the normalisations of linden.f:108-121.  out = [4][npt]: cos, rho, E, E^2.
__global__ __launch_bounds__(256) void k_linden(int npt, int lmax, int iodevn, double tau, double bconst,
                                                double nslice_boltz, double* __restrict__ out) {
#pragma clang fp contract(off)  // every product rounded before its sum, as linden.f's arithmetic is
    const int ic = blockIdx.x * blockDim.x + threadIdx.x;
    if (ic >= npt) return;
    const double cstep = 2.0 / (double)(npt - 1);
    const double x = (double)ic * cstep - 1.0;
    double p0 = 1.0, p1 = x;  // P_0, P_1
    double rho = 0.0, er = 0.0, es = 0.0;
    for (int l = 0; l <= lmax; ++l) {
        const double pl = l == 0 ? p0 : p1;
        if (l % 2 == iodevn || iodevn == -1) {
            const double ll = (double)l * (double)(l + 1);
            const double tmp = (double)(2 * l + 1) * pl * exp(-tau * bconst * ll);
            rho += tmp;
            er += tmp * ll * bconst;
            es += tmp * ll * bconst * ll * bconst;
        }
        if (l >= 1) {  // P_{l+1} = ((2l+1) x P_l - l P_{l-1}) / (l+1)
            const double pn = ((double)(2 * l + 1) * x * p1 - (double)l * p0) / (double)(l + 1);
            p0 = p1;
            p1 = pn;
        }
    }
    er = er / nslice_boltz;
    es = es / (nslice_boltz * nslice_boltz);
    const double q = 4.0 * 3.14159265358979323846;
    out[ic] = x;
    out[npt + ic] = rho / q;
    out[2 * npt + ic] = er / q;
    out[3 * npt + ic] = es / q;
}

// ------------------------------------------------------------------------- asymrho
constexpr int ASYM_NG = 361;  // phi and chi: whole degrees 0..360

__device__ __forceinline__ double log_fact(int n) { return lgamma((double)n + 1.0); }

// d^J_mk(beta) = <J m| exp(-i beta J_y) |J k> at J = max(|m|, |k|), where the factorial sum
// (Zare eq. 3.57) has a single term
__device__ inline double wigner_d_seed(int j, int m, int k, double c, double s) {
    const int lo = max(0, k - m), hi = min(j + k, j - m);
    double d = 0.0;
    for (int v = lo; v <= hi; ++v) {
        const double lw = 0.5 * (log_fact(j + k) + log_fact(j - k) + log_fact(j + m) + log_fact(j - m)) -
                          log_fact(j - m - v) - log_fact(j + k - v) - log_fact(v + m - k) - log_fact(v);
        const double t = exp(lw) * pow(c, (double)(2 * j + k - m - 2 * v)) * pow(s, (double)(m - k + 2 * v));
        d += ((v + m - k) & 1) ? -t : t;
    }
    return d;
}

// M_q(theta)[m, k] of parity p: grid (n*n / 256, 181), lane (m, k).  g = [maxj+1][3][n][n] of
// this parity (the (2J+1)/(8 pi^2) factor folded in); mq = [181][npar][3][n][n], this parity at
// index pi.
__global__ __launch_bounds__(256) void k_asym_m(int maxj, int p, int npar, int pi, const double* __restrict__ g,
                                                double* __restrict__ mq) {
    const int n = 2 * maxj + 1;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    const int ith = blockIdx.y;
    if (e >= n * n) return;
    const int m = e / n - maxj, k = e % n - maxj;
    const size_t nn = (size_t)n * n;
    double* out = mq + (size_t)(ith * npar + pi) * 3 * nn + e;
    if ((((m % 2) + 2) % 2) != p || (((k % 2) + 2) % 2) != p) {
        out[0] = out[nn] = out[2 * nn] = 0.0;
        return;
    }
    const double beta = (double)ith * (3.14159265358979323846 / 180.0);
    const double x = cos(beta), c = cos(0.5 * beta), s = sin(0.5 * beta);
    const int j0 = max(abs(m), abs(k));
    double dprev = 0.0, d = wigner_d_seed(j0, m, k, c, s);
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    const double mk = (double)m * (double)k, m2 = (double)m * m, k2 = (double)k * k;
    for (int J = j0; J <= maxj; ++J) {
        const double* gj = g + (size_t)J * 3 * nn + e;
        a0 += gj[0] * d;
        a1 += gj[nn] * d;
        a2 += gj[2 * nn] * d;
        if (J == maxj) break;
        // J sqrt(((J+1)^2-m^2)((J+1)^2-k^2)) d^{J+1}
        //   = (2J+1)(J(J+1) cos beta - m k) d^J - (J+1) sqrt((J^2-m^2)(J^2-k^2)) d^{J-1}
        double dn;
        if (J == 0) {
            dn = x * d;  // m = k = 0: P_1
        } else {
            const double Jd = (double)J, J1 = Jd + 1.0;
            const double num = (2.0 * Jd + 1.0) * (Jd * J1 * x - mk) * d - J1 * sqrt((Jd * Jd - m2) * (Jd * Jd - k2)) * dprev;
            dn = num / (Jd * sqrt((J1 * J1 - m2) * (J1 * J1 - k2)));
        }
        dprev = d;
        d = dn;
    }
    out[0] = a0;
    out[nn] = a1;
    out[2 * nn] = a2;
}

// Batched row-major C = A B (FP64), blockIdx.z = batch; 64 x 64 tile per 256 lanes, each lane a
// 4 x 4 block of C, K in steps of 16 through LDS.
__global__ __launch_bounds__(256) void k_dgemm(int M, int N, int K, const double* __restrict__ A, int lda, long long sA,
                                               const double* __restrict__ B, int ldb, long long sB, double* __restrict__ C,
                                               int ldc, long long sC) {
    __shared__ double As[16][64 + 1];
    __shared__ double Bs[16][64 + 1];
    const int bz = blockIdx.z;
    A += (size_t)bz * sA;
    B += (size_t)bz * sB;
    C += (size_t)bz * sC;
    const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
    const int row0 = blockIdx.y * 64, col0 = blockIdx.x * 64;
    double acc[4][4] = {};
    for (int k0 = 0; k0 < K; k0 += 16) {
        for (int i = threadIdx.x; i < 16 * 64; i += 256) {
            const int kk = i % 16, r = i / 16;  // A tile: 64 rows x 16 k
            const int gr = row0 + r, gk = k0 + kk;
            As[kk][r] = (gr < M && gk < K) ? A[(size_t)gr * lda + gk] : 0.0;
            const int cc = i % 64, kb = i / 64;  // B tile: 16 k x 64 columns
            const int gc = col0 + cc, gkb = k0 + kb;
            Bs[kb][cc] = (gc < N && gkb < K) ? B[(size_t)gkb * ldb + gc] : 0.0;
        }
        __syncthreads();
#pragma unroll 4
        for (int kk = 0; kk < 16; ++kk) {
            double a[4], b[4];
            for (int i = 0; i < 4; ++i) a[i] = As[kk][ty + 16 * i];
            for (int j = 0; j < 4; ++j) b[j] = Bs[kk][tx + 16 * j];
            for (int i = 0; i < 4; ++i)
                for (int j = 0; j < 4; ++j) acc[i][j] += a[i] * b[j];
        }
        __syncthreads();
    }
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            const int r = row0 + ty + 16 * i, cc = col0 + tx + 16 * j;
            if (r < M && cc < N) C[(size_t)r * ldc + cc] = acc[i][j];
        }
}

// rho, E, E^2 from the per-parity X_q (x = [181][npar][3][361*361]), read at the fill map's
// source point; grid (361*361 / 256, 181).  iodevn -1 divides each parity's E by its own rho first.
__global__ __launch_bounds__(256) void k_asym_fill(int npar, const int* __restrict__ src, const double* __restrict__ x,
                                                   double* __restrict__ rho, double* __restrict__ eng,
                                                   double* __restrict__ esq) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int np = ASYM_NG * ASYM_NG;
    if (i >= np) return;
    const size_t th = blockIdx.y;
    x += th * npar * 3 * (size_t)np;
    rho += th * np;
    eng += th * np;
    esq += th * np;
    const int s = src[i];
    const double eps = 1.0e-16;
    double r, e, q;
    if (npar == 2) {
        const double r0 = x[s], e0 = x[np + s], q0 = x[2 * np + s];
        const double r1 = x[3 * np + s], e1 = x[4 * np + s], q1 = x[5 * np + s];
        r = r0 + r1;
        const double ee = fabs(r0) > eps ? e0 / r0 : 0.0, eo = fabs(r1) > eps ? e1 / r1 : 0.0;
        const double se = fabs(r0) > eps ? q0 / r0 : 0.0, so = fabs(r1) > eps ? q1 / r1 : 0.0;
        e = (eo * r1 + ee * r0) / r;
        q = (so * r1 + se * r0) / r;
        if (fabs(r) < eps) e = q = 0.0;
    } else {
        r = x[s];
        e = fabs(r) > eps ? x[np + s] / r : 0.0;
        q = fabs(r) > eps ? x[2 * np + s] / r : 0.0;
    }
    rho[i] = r;
    eng[i] = e;
    esq[i] = q;
}

// ---- host preparation

// Cyclic Jacobi eigen-decomposition of a symmetric n x n matrix (row-major, overwritten):
// w = eigenvalues, v[i*n + s] = component i of eigenvector s.
inline void jacobi_eigen(int n, std::vector<double>& a, std::vector<double>& w, std::vector<double>& v) {
    v.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) v[(size_t)i * n + i] = 1.0;
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0, diag = 0.0;
        for (int p = 0; p < n; ++p) {
            diag += a[(size_t)p * n + p] * a[(size_t)p * n + p];
            for (int q = p + 1; q < n; ++q) off += a[(size_t)p * n + q] * a[(size_t)p * n + q];
        }
        if (off <= 1e-32 * diag || off == 0.0) break;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q) {
                const double apq = a[(size_t)p * n + q];
                if (apq == 0.0) continue;
                const double app = a[(size_t)p * n + p], aqq = a[(size_t)q * n + q];
                const double th = (aqq - app) / (2.0 * apq);
                const double t = (th >= 0.0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for (int r = 0; r < n; ++r) {  // columns p, q
                    const double arp = a[(size_t)r * n + p], arq = a[(size_t)r * n + q];
                    a[(size_t)r * n + p] = c * arp - s * arq;
                    a[(size_t)r * n + q] = s * arp + c * arq;
                }
                for (int r = 0; r < n; ++r) {  // rows p, q
                    const double apr = a[(size_t)p * n + r], aqr = a[(size_t)q * n + r];
                    a[(size_t)p * n + r] = c * apr - s * aqr;
                    a[(size_t)q * n + r] = s * apr + c * aqr;
                }
                for (int r = 0; r < n; ++r) {
                    const double vrp = v[(size_t)r * n + p], vrq = v[(size_t)r * n + q];
                    v[(size_t)r * n + p] = c * vrp - s * vrq;
                    v[(size_t)r * n + q] = s * vrp + c * vrq;
                }
            }
    }
    w.resize(n);
    for (int i = 0; i < n; ++i) w[i] = a[(size_t)i * n + i];
}

// The (phi, chi) point whose value each grid point takes after the symmetry fill of
// asymrho.f:665-709 (its four copy loops applied in order to the point indices).
inline std::vector<int> asym_fill_map() {
    const int G = ASYM_NG;
    std::vector<int> x(G * G);
    std::iota(x.begin(), x.end(), 0);
    auto at = [&](int ip, int ic) -> int& { return x[ip * G + ic]; };
    for (int ip = 90; ip <= 180; ++ip)
        for (int ic = 180 - ip; ic <= ip; ++ic) at(ip, ic) = at(180 - ic, 180 - ip);
    for (int ip = 180; ip <= 270; ++ip)
        for (int ic = ip - 180; ic <= 360 - ip; ++ic) at(ip, ic) = at(180 + ic, ip - 180);
    for (int ip = 180; ip <= 360; ++ip)
        for (int ic = 360 - ip; ic <= ip; ++ic) at(ip, ic) = at(360 - ic, 360 - ip);
    for (int ip = 0; ip <= 360; ++ip)
        for (int ic = ip; ic <= 360; ++ic) at(ip, ic) = at(ic, ip);
    return x;
}

// G_{J,q} of both parities, [2][maxj+1][3][n][n] with (2J+1)/(8 pi^2) folded in, from the rigid-
// rotor blocks of asymrho's rotmat (asymrho.f:792-813): diagonal (A+C)/2 J(J+1) + (B-(A+C)/2) k^2,
// coupling of k and k-2 by (A-C)/4 c+(J,k-2) c+(J,k-1).  Returns the largest block energy.
inline double asym_blocks(int maxj, double A, double B, double C, double tau, std::vector<double>& g) {
    const int n = 2 * maxj + 1;
    const size_t nn = (size_t)n * n;
    g.assign(2 * (size_t)(maxj + 1) * 3 * nn, 0.0);
    double emax = -1e300;
    auto cplus = [](int j, int k) { return (k >= j || k < -j) ? 0.0 : std::sqrt(j * (j + 1.0) - k * (k + 1.0)); };
    for (int J = 0; J <= maxj; ++J) {
        const double pre = (2.0 * J + 1.0) / (8.0 * M_PI * M_PI);
        for (int par = 0; par < 2; ++par) {
            std::vector<int> ks;
            for (int k = -J; k <= J; ++k)
                if (((k % 2) + 2) % 2 == par) ks.push_back(k);
            const int nb = (int)ks.size();
            if (!nb) continue;
            std::vector<double> h((size_t)nb * nb, 0.0), w, v;
            for (int i = 0; i < nb; ++i) {
                const int k = ks[i];
                h[(size_t)i * nb + i] = 0.5 * (A + C) * J * (J + 1) + (B - 0.5 * (A + C)) * k * k;
                if (i > 0) h[(size_t)i * nb + i - 1] = h[(size_t)(i - 1) * nb + i] = 0.25 * (A - C) * cplus(J, k - 2) * cplus(J, k - 1);
            }
            jacobi_eigen(nb, h, w, v);
            for (int s = 0; s < nb; ++s) {
                emax = std::max(emax, w[s]);
                const double ws = std::exp(-tau * w[s]);
                if (ws * pre < 1.0e-16) continue;
                double* gb = g.data() + ((size_t)par * (maxj + 1) + J) * 3 * nn;
                for (int a = 0; a < nb; ++a)
                    for (int b = 0; b < nb; ++b) {
                        const size_t e = (size_t)(ks[a] + maxj) * n + (ks[b] + maxj);
                        const double cc = v[(size_t)a * nb + s] * v[(size_t)b * nb + s];
                        gb[e] += pre * ws * cc;
                        gb[nn + e] += pre * ws * w[s] * cc;
                        gb[2 * nn + e] += pre * ws * w[s] * w[s] * cc;
                    }
            }
        }
    }
    return emax;
}

}  // namespace pimc
