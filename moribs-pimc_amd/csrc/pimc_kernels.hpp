// pimc_kernels.hpp -- gfx950 kernels of the sampling hot path.
//
//   k_bisect     one workgroup per (replica, imaginary-time segment).  The segment window
//                (all atoms' beads on the S+1 slices of the segment, plus the linear rotor's
//                axis there) is staged in LDS once and every atom of the selected types is
//                moved in turn, as MCBisectionMove does (mc_piqmc.cc:194-294).  Segments of
//                one offset share only fixed end points, so they run concurrently.
//   k_centroid_parts  a replica's slices split over workgroups; atoms in turn, Delta V over the
//                part's slices x partners, partials exchanged as numbered words in memory
//                (MCMolecularMove mc_piqmc.cc:54-102).
//   k_rotor      one workgroup per part of a linear rotor's chain (a range of slices): the part's
//                orientations and its rotor-partner geometry live in LDS for all sweeps of a
//                launch; each half of the even/odd checkerboard (mc_piqmc.cc:490-512) is a
//                parallel step over slices, a lane group per slice sums Delta V over partners
//                (MCRotLinStep :781-936); parts trade one boundary slice per half-sweep as a
//                record of numbered words (each 64-bit word: the publishing half's number over
//                32 bits of a double), so a poll of the record returns the data with its check.
//   k_est_tiles  one workgroup per (replica, tile of slices): the tile's part of the pair sum of
//                GetPotEnergy_Densities (mc_estim.cc:500-687, every pair, with the g(r),
//                g(r,theta) and (r,theta,chi) bins) and of GetKinEnergy (mc_estim.cc:876-937),
//                beads read once.
//   k_est_final  one workgroup per replica: tiles summed in order (deterministic), GetRotEnergy
//                (mc_estim.cc:939-987) or GetRotE3D with its relative Euler angle bins
//                (:989-1096), and the Cv terms of MCGetAverage (mc_main.cc:551-616).
//   k_pot_slice  batched PotEnergy(atom, pos, it) (mc_piqmc.cc:1796-1965).
//   k_rotor3d_da the non-linear rotor sweep with a surrogate first stage (delayed acceptance).
#pragma once
#include <type_traits>
#include "pimc_device.hpp"
#include "pimc_nl.hpp"
#include "pimc_mlp.hpp"
#include "pimc_plan.hpp"

namespace pimc {

struct Run {
    uint32_t seed;
    int u;           // offset within the segment stride (segment order) or time slice (reference order)
    int ref_order;   // 1: one segment/sweep per replica per launch (mc_main.cc order)
    int type_begin, type_end;
};

__device__ __forceinline__ double4 bead(const double4* beads, const Rep& rp, int a, int it) {
    return beads[rp.bead_off + (long long)a * rp.P + it];
}

__device__ __forceinline__ int type_of(const Rep& rp, int a) {
    int t = 0;
    while (t + 1 < rp.ntypes && a >= rp.tbegin[t + 1]) ++t;
    return t;
}

// sum over the block; `red` holds one value per wave (16 for 1024 lanes)
template <typename T>
__device__ __forceinline__ T block_sum(T v, T* red) {
    v = wave_sum(v);
    const int wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[wid] = v;
    __syncthreads();
    T s = 0;
    for (int i = 0; i < nw; ++i) s += red[i];
    return s;
}

__device__ __forceinline__ void rotor_axis(const RotBead* rotb, const Rep& rp, int it, double& nx, double& ny,
                                           double& nz) {
    if (rp.rotkind == 1 && rp.rot_atom >= 0) {
        int ar = rp.rot_atom - rp.tbegin[rp.imtype];
        RotBead rb = rotb[rp.rot_off + (long long)ar * rp.Prot + it / rp.RotRatio];
        nx = rb.nx; ny = rb.ny; nz = rb.nz;
    } else {
        nx = 0; ny = 0; nz = 1;
    }
}

// PotEnergy's dispatch (mc_piqmc.cc:1847-1958) for the non-linear family: pairs of the atom
// type by SPot1D, atom / non-linear rotor pairs by vcord with the rotor's frame at that slice
// (rcom = the rotor, rpt = the atom), rotor / rotor pairs by caleng.  vcord and caleng take the
// positions themselves, so the minimum image (box) reaches the atom pairs only, as there.
template <typename R>
__device__ __forceinline__ double pair_nl(const Lut1D<R>& l1, const vec4_t<R>* __restrict__ k1, const Lut3D<R>& l3,
                                          bool rot_a, bool rot_j, double xa, double ya, double za, double xj,
                                          double yj, double zj, const Frame& fa, const Frame& fj,
                                          const double* box = nullptr) {
    if (rot_a && rot_j) return caleng_v(xa, ya, za, xj, yj, zj, fa, fj);
    if (rot_a) return (double)vcord_v<R>(l3, fa, xa, ya, za, xj, yj, zj);
    if (rot_j) return (double)vcord_v<R>(l3, fj, xj, yj, zj, xa, ya, za);
    const R dx = (R)nearest_image(xa - xj, box, 0), dy = (R)nearest_image(ya - yj, box, 1),
            dz = (R)nearest_image(za - zj, box, 2);
    return (double)spot1d<R>(l1, k1, sqrt(dx * dx + dy * dy + dz * dz));
}

__device__ __forceinline__ const Frame& frame_at(const Frame* __restrict__ frames, const Rep& rp, int a, int it) {
    return frames[rp.rot_off + (long long)(a - rp.nl_atom) * rp.Prot + it / rp.RotRatio];
}

// ------------------------------------------------------------------------- bisection
constexpr int BIS_RING = 128;  // doubles of draws staged for a batch of atoms

// doubles of the draw ring for a window of W slots: a batch of atoms, or one atom's draws
__host__ __device__ constexpr int bisect_ring(int W) { return BIS_RING > 3 * W + 16 ? BIS_RING : 3 * W + 16; }

// The window of all atoms' positions on the segment's S+1 slots is held in the pair precision
// (float4 in FP32 runs, double4 in FP64 runs); new positions are built in FP64 from FP64 end
// points, and accepted ones go to HBM in FP64 at once.  blockDim a multiple of 64.
// FAM 1 (non-linear rotors) adds the rotors' frames at the window's slots: fr [nrot*(S+1)].
// MI: built for contexts that hold a replica under the minimum image (its pairs then take their
// nearest image); without it the kernels carry no trace of it
template <typename R, int FAM, bool MI>
__device__ __forceinline__ void bisect_body(const Rep* __restrict__ reps, Mut* __restrict__ mut,
                                            double4* __restrict__ beads, const RotBead* __restrict__ rotb,
                                            const Frame* __restrict__ frames, Tables T, Tables3 T3,
                                            const int2* __restrict__ work, Run run) {
    extern __shared__ double smem[];
    __shared__ double red[16];
    const int2 wk = work[blockIdx.x];
    const Rep& rp = reps[wk.x];
    const int tid = threadIdx.x, nth = blockDim.x;
    const int S = rp.S, P = rp.P, N = rp.natoms;
    int s0;
    if (run.ref_order) {
        if (run.u >= P) return;
        s0 = run.u;
    } else {
        if (run.u >= S) return;
        s0 = run.u + wk.y * S;
    }
    const int W = S + 1;
    using P4 = vec4_t<R>;
    // LDS: the moving atom's new positions and its draws (FP64), then in the pair precision the
    // window of all atoms' positions (x, y, z, -), the new midpoints again (one read per item)
    // and the rotor axes at the window's slices, then rotor frames (FAM 1) and the 1D spline
    // knots: (y, y'') pairs of uniform FP32 knots (linear family), vec4 knots otherwise
    double* nw = smem;
    double* gw = nw + 3 * W;  // draws of a batch of atoms: per atom 3 per midpoint, then one per level
    P4* win = reinterpret_cast<P4*>(smem + ((3 * W + bisect_ring(W) + 3) & ~3));
    P4* nwf = win + N * W;
    P4* nhf = nwf + W;
    Frame* fr = reinterpret_cast<Frame*>(nhf + W);
    const int nfr = FAM == 1 ? rp.nrot * W : 0;
    P4* kn = reinterpret_cast<P4*>(fr + nfr);
    const Lut1D<R> l1 = lut1d<R>(T, rp.pot1d_id);
    const Lut2D<R> l2 = lut2d<R>(T, rp.pot2d_id);
    const Lut3D<R> l3 = lut3d<R>(T3, FAM == 1 ? rp.pot3d_id : -1);
    const bool half_knots = FAM == 0 && is_fp32<R>() && l1.uniform;
    const double* box = MI && rp.minimage ? rp.box : nullptr;  // MINIMAGE (not with a rotating linear rotor)
    if (half_knots) {
        float2* k2 = reinterpret_cast<float2*>(kn);
        for (int i = tid; i < l1.n; i += nth) k2[i] = make_float2((float)l1.k[i].y, (float)l1.k[i].z);
    } else {
        for (int i = tid; i < l1.n; i += nth) kn[i] = l1.k[i];
    }
    if constexpr (FAM == 1)
        for (int i = tid; i < nfr; i += nth) {
            const int q = i / W, w = i - q * W;
            fr[i] = frame_at(frames, rp, rp.nl_atom + q, (s0 + w) % P);
        }
    for (int i = tid; i < N * W; i += nth) {
        const int a = i / W, w = i - a * W;
        const double4 b = bead(beads, rp, a, (s0 + w) % P);
        P4 q;
        q.x = (R)b.x;
        q.y = (R)b.y;
        q.z = (R)b.z;
        q.w = (R)0;
        win[i] = q;
    }
    for (int w = tid; w < W; w += nth) {
        double x, y, z;
        rotor_axis(rotb, rp, (s0 + w) % P, x, y, z);
        P4 q;
        q.x = (R)x;  // the pair functions take the axis in the pair precision
        q.y = (R)y;
        q.z = (R)z;
        q.w = (R)0;
        nhf[w] = q;
    }
    __syncthreads();

    const uint32_t pass = mut[wk.x].pass;
    unsigned long long evals = 0;
    unsigned long long tot[MAXT] = {0, 0, 0, 0}, acc_n[MAXT] = {0, 0, 0, 0};
    const int tend = run.type_end < rp.ntypes ? run.type_end : rp.ntypes;

    for (int ty = run.type_begin; ty < tend; ++ty) {
        const int L = rp.tlev[ty];
        const int St = 1 << L;
        const double bnorm = 1.0 / (rp.lambda[ty] * rp.tau);
        const int nm = St - 1;
        const int per = 3 * nm + L;  // draws of one atom's move
        const int K = per <= BIS_RING ? (nth < BIS_RING ? nth : BIS_RING) / per : 1;  // atoms per draw batch
        // end points of an atom's segment (slots 0 and St) in FP64 from HBM (no move of this launch
        // changes them), loaded one atom ahead on the three coordinate lanes
        double e0 = 0.0, e1 = 0.0;
        if (tid < 3 && rp.tnumb[ty] > 0) {
            const int a = rp.tbegin[ty];
            const double4 b0 = bead(beads, rp, a, s0 % P), b1 = bead(beads, rp, a, (s0 + St) % P);
            e0 = tid == 0 ? b0.x : (tid == 1 ? b0.y : b0.z);
            e1 = tid == 0 ? b1.x : (tid == 1 ? b1.y : b1.z);
        }
        for (int ia = 0; ia < rp.tnumb[ty]; ++ia) {
            const int a = rp.tbegin[ty] + ia;
            const int kk = ia % K;
            if (kk == 0) {
                // every draw of the next K moves at once: Gaussians of midpoint m (level-major
                // order) and the level acceptance uniforms u, kept as -log(u): the level test
                // exp(-dV) > u (mc_piqmc.cc:265-273) is then dV < -log(u), a comparison
                const int nk = rp.tnumb[ty] - ia < K ? rp.tnumb[ty] - ia : K;
                for (int i = tid; i < nk * per; i += nth) {
                    const int k = i / per, d = i - k * per;
                    const uint32_t ab = (uint32_t)(a + k);
                    if (d < 3 * nm) {
                        int m = d / 3, id = d - 3 * m;
                        int level = 31 - __clz(m + 1);
                        int q = m + 1 - (1 << level);
                        uint4 w4 = draw(run.seed, rp.key, pass, K_BISECT, 0, ab, (uint32_t)s0,
                                        ((uint32_t)level << 24) | ((uint32_t)q << 8) | (uint32_t)id);
                        gw[i] = gauss_from(w4, bnorm / (double)(1 << (L - level)));
                    } else {
                        int level = d - 3 * nm;
                        uint4 w4 = draw(run.seed, rp.key, pass, K_BISECT, 0, ab, (uint32_t)s0,
                                        ((uint32_t)level << 24) | (0xFFFFu << 8));
                        gw[i] = -log(u53(w4.x, w4.y));
                    }
                }
            }
            const double* g = gw + kk * per;  // this move's draws
            if (tid < 3) {
                nw[tid] = e0;
                nw[3 * St + tid] = e1;
                if (ia + 1 < rp.tnumb[ty]) {  // the next atom's end points, used one move later
                    const double4 b0 = bead(beads, rp, a + 1, s0 % P), b1 = bead(beads, rp, a + 1, (s0 + St) % P);
                    e0 = tid == 0 ? b0.x : (tid == 1 ? b0.y : b0.z);
                    e1 = tid == 0 ? b1.x : (tid == 1 ? b1.y : b1.z);
                }
            }
            if (kk == 0) __syncthreads();  // the batch's draws are visible to every lane
            double cum = 0.0;
            bool accepted = false;
            for (int level = 0; level < L; ++level) {
                const int lss = 1 << (L - level);
                const int half = lss >> 1;
                const int nmid = 1 << level;
                // new midpoints: slot = half + q*lss, from the current new end points
                for (int i = tid; i < 3 * nmid; i += nth) {
                    int q = i / 3, id = i - 3 * q;
                    int slot = half + q * lss;
                    const double v = 0.5 * (nw[3 * (slot - half) + id] + nw[3 * (slot + half) + id]) +
                                     g[3 * (nmid - 1 + q) + id];
                    nw[3 * slot + id] = v;
                    reinterpret_cast<R*>(nwf + slot)[id] = (R)v;
                }
                __syncthreads();
                // Delta V at the new midpoints: items (midpoint, partner)
                const int np = N - 1;
                const int items = nmid * np;
                const bool rot_a = FAM == 1 ? (rp.nl_atom >= 0 && a >= rp.nl_atom) : (a == rp.lin_atom);
                // an atom with the linear rotor last among its partners: the atom-atom terms and
                // the rotor terms in one index space, rotor terms last, so a wave runs one
                // table's code except at the seam
                const bool split = FAM == 0 && !rot_a && rp.lin_atom == N - 1;
                auto item_sum = [&](auto half_c, auto mi_c) -> double {
                    constexpr bool H = decltype(half_c)::value;
                    // the minimum image as its own instance of the loop, so the plain one is unchanged
                    const double* bx = decltype(mi_c)::value ? box : nullptr;
                    // (that instance takes every pair through the general branch)
                    const bool spl = !decltype(mi_c)::value && split;
                    double acc = 0.0;
                    // item i = (midpoint qa, partner index jp) with den partners per midpoint in
                    // the atom range; a lane's items advance by nth, so (qa, jp) is carried
                    // along with one division per level
                    const int den = spl ? np - 1 : np;
                    const int n1 = spl ? nmid * den : items;  // items of the atom range
                    const int dd = den > 0 ? den : 1;
                    const int dq = nth / dd, dj = nth - dq * dd;
                    int qa = tid / dd, jp = tid - qa * dd;
                    for (int i = tid; i < items; i += nth) {
                        int q, j;
                        bool rj;
                        if (i >= n1) {  // the rotor terms after the atom range (split)
                            rj = true;
                            q = i - n1;
                            j = N - 1;
                        } else {
                            q = qa;
                            j = jp + (jp >= a);
                            rj = spl ? false
                                       : (FAM == 1 ? (rp.nl_atom >= 0 && j >= rp.nl_atom) : (j == rp.lin_atom));
                        }
                        jp += dj;
                        qa += dq;
                        if (jp >= dd) {
                            jp -= dd;
                            ++qa;
                        }
                        const int slot = half + q * lss;
                        const P4 pj = win[j * W + slot], pa = win[a * W + slot];
                        const P4 pn = nwf[slot];
                        const R xn = pn.x, yn = pn.y, zn = pn.z;
                        if constexpr (FAM == 1) {
                            const Frame& fa = fr[(rot_a ? a - rp.nl_atom : 0) * W + slot];
                            const Frame& fj = fr[(rj ? j - rp.nl_atom : 0) * W + slot];
                            acc += pair_nl<R>(l1, kn, l3, rot_a, rj, xn, yn, zn, pj.x, pj.y, pj.z, fa, fj, bx) -
                                   pair_nl<R>(l1, kn, l3, rot_a, rj, pa.x, pa.y, pa.z, pj.x, pj.y, pj.z, fa, fj, bx);
                        } else if (spl) {
                            if (rj) {
                                const P4 ax = nhf[slot];
                                const double nx = ax.x, ny = ax.y, nz = ax.z;
                                acc += (double)pair_vr<R, H>(l1, kn, l2, false, true, xn, yn, zn, pj.x, pj.y, pj.z, nx, ny, nz, bx) -
                                       (double)pair_vr<R, H>(l1, kn, l2, false, true, pa.x, pa.y, pa.z, pj.x, pj.y, pj.z, nx, ny, nz, bx);
                            } else {
                                acc += (double)pair_vr<R, H>(l1, kn, l2, false, false, xn, yn, zn, pj.x, pj.y, pj.z, 0.0, 0.0, 1.0, bx) -
                                       (double)pair_vr<R, H>(l1, kn, l2, false, false, pa.x, pa.y, pa.z, pj.x, pj.y, pj.z, 0.0, 0.0, 1.0, bx);
                            }
                        } else {
                            const P4 ax = nhf[slot];
                            const double nx = ax.x, ny = ax.y, nz = ax.z;
                            acc += (double)pair_vr<R, H>(l1, kn, l2, rot_a, rj, xn, yn, zn, pj.x, pj.y, pj.z, nx, ny, nz, bx) -
                                   (double)pair_vr<R, H>(l1, kn, l2, rot_a, rj, pa.x, pa.y, pa.z, pj.x, pj.y, pj.z, nx, ny, nz, bx);
                        }
                    }
                    return acc;
                };
                using T1 = std::integral_constant<bool, true>;
                using F1 = std::integral_constant<bool, false>;
                double acc;
                if constexpr (MI)
                    acc = half_knots ? (box ? item_sum(T1{}, T1{}) : item_sum(T1{}, F1{}))
                                     : (box ? item_sum(F1{}, T1{}) : item_sum(F1{}, F1{}));
                else
                    acc = half_knots ? item_sum(T1{}, F1{}) : item_sum(F1{}, F1{});
                acc = block_sum(acc, red);
                evals += 2ull * (unsigned long long)items;
                // sum over all points up to this level; level acceptance of mc_piqmc.cc:263-273
                const double prev = cum;
                cum += acc;
                double deltav = (cum - 2.0 * prev) * (rp.tau * (double)half);
                accepted = deltav < 0.0 || deltav < g[3 * nm + level];
                if (!accepted) break;
            }
            tot[ty] += 1;
            if (accepted) {  // the new midpoints into the window and, in FP64, to HBM
                acc_n[ty] += 1;
                for (int i = tid; i < St - 1; i += nth) {
                    const int slot = 1 + i;
                    const double x = nw[3 * slot], y = nw[3 * slot + 1], z = nw[3 * slot + 2];
                    P4 q;
                    q.x = (R)x;
                    q.y = (R)y;
                    q.z = (R)z;
                    q.w = (R)0;
                    win[a * W + slot] = q;
                    beads[rp.bead_off + (long long)a * P + (s0 + slot) % P] = make_double4(x, y, z, 0.0);
                }
            }
            __syncthreads();
        }
    }
    if (tid == 0) {
        for (int ty = run.type_begin; ty < tend; ++ty) {
            atomicAdd(&mut[wk.x].total[ty][1], tot[ty]);
            atomicAdd(&mut[wk.x].accep[ty][1], acc_n[ty]);
        }
        atomicAdd(&mut[wk.x].pair_evals[0], evals);
    }
}

// the linear family keeps four waves per SIMD; the non-linear family needs more registers
template <typename R, bool MI>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_bisect(
    const Rep* __restrict__ reps, Mut* __restrict__ mut, double4* __restrict__ beads, const RotBead* __restrict__ rotb,
    Tables T, const int2* __restrict__ work, Run run) {
    bisect_body<R, 0, MI>(reps, mut, beads, rotb, nullptr, T, Tables3{}, work, run);
}

template <typename R, bool MI>
__global__ __launch_bounds__(256) void k_bisect_nl(const Rep* __restrict__ reps, Mut* __restrict__ mut,
                                                   double4* __restrict__ beads, const RotBead* __restrict__ rotb,
                                                   const Frame* __restrict__ frames, Tables T, Tables3 T3,
                                                   const int2* __restrict__ work, Run run) {
    bisect_body<R, 1, MI>(reps, mut, beads, rotb, frames, T, T3, work, run);
}

// ------------------------------------------------------------------------- linear rotor
// The rotor chain of a replica is split into parts of contiguous slices [k0, k1) (even
// length), one workgroup each.  A part builds the rotor-partner geometry of its translational
// slices in LDS once per launch and keeps it for every sweep of the launch (translations do
// not move during rotor sweeps).  Between half-sweeps a part needs one orientation from each
// neighbour: slice k0-1 (odd, moved by the part before) before its even half and slice k1
// (even, moved by the part after) before its odd half.  Each part publishes the boundary slice
// it just moved to rotb and then raises its flag (count of half-sweeps done); the reader spins
// on the flag with a bounded count, so every wave exits.  Boundary values and flags are
// written and read only with device-coherent atomics, and the writer waits for its value
// stores to complete before it stores the flag: the exchange needs no cache write-back or
// invalidation, which on a multi-XCD part would flush the L2 holding everyone's tables.
// Results do not depend on the split: draws are counter-based per (sweep, slice).
// (RotPart: pimc_plan.hpp)

// Polls of a numbered word before a wait gives up and sets the part's error bit (each poll
// sleeps between loads, so this bounds a wait at seconds, far beyond any normal exchange)
constexpr unsigned ROT_SPIN_LIMIT = 1u << 22;

// ------------------------------------------------------------------------- centroid, split
// MCMolecularMove (mc_piqmc.cc:54-102) with a replica's slices split over W workgroups
// (parts [k0, k1) of 0..P-1, RotPart reused): atoms in turn, each part sums Delta V over its
// slices and all partners and publishes the partial as two 64-bit words that carry the atom's
// number (base + ia + 1) above 32 bits of the double; lanes 0..2W-1 poll every part's words
// (bounded spin), and every part then adds the W partials in part order -- the same sum, so the
// same decision -- and moves its own slices of the atom.  Partner positions at a part's slices
// change only inside that part, so no bead data crosses parts.  Partials are double-buffered by
// atom parity: a part can overwrite a buffer only after every part has published the following
// atom, that is after every part has read it.
template <typename R, int FAM, bool MI>
__global__ __launch_bounds__(256) void k_centroid_parts(const Rep* __restrict__ reps, Mut* __restrict__ mut,
                                                        double4* __restrict__ beads, const RotBead* __restrict__ rotb,
                                                        const Frame* __restrict__ frames, Tables T, Tables3 T3,
                                                        const RotPart* __restrict__ parts, unsigned* __restrict__ flags,
                                                        unsigned long long* __restrict__ slots, int nslots, Run run) {
    extern __shared__ double smem[];  // the 1D spline knots
    __shared__ double red[4];
    __shared__ double part_sum[64];
    __shared__ double s_dv;
    __shared__ int s_err;
    const RotPart pt = parts[blockIdx.x];
    if (pt.r < 0) return;  // padding of the XCD-ordered grid
    const int r = pt.r;
    const Rep& rp = reps[r];
    const int ty = run.type_begin;
    if (ty >= rp.ntypes) return;
    const uint32_t pass = mut[r].pass;
    const int N = rp.natoms, P = rp.P, np = N - 1;
    const int t0 = pt.k0, nt = pt.k1 - pt.k0, W = pt.W;
    const Lut1D<R> l1 = lut1d<R>(T, rp.pot1d_id);
    const Lut2D<R> l2 = lut2d<R>(T, rp.pot2d_id);
    vec4_t<R>* kn = reinterpret_cast<vec4_t<R>*>(smem);
    for (int i = threadIdx.x; i < l1.n; i += blockDim.x) kn[i] = l1.k[i];
    __syncthreads();
    unsigned* myflag = flags + pt.fbase + pt.w;
    const unsigned base = __hip_atomic_load(myflag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const double* box = MI && rp.minimage ? rp.box : nullptr;  // MINIMAGE (not with a rotating linear rotor)
    if (threadIdx.x == 0) s_err = 0;
    unsigned long long acc_n = 0;
    for (int ia = 0; ia < rp.tnumb[ty]; ++ia) {
        const int a = rp.tbegin[ty] + ia;
        uint4 w0 = draw(run.seed, rp.key, pass, K_CENTROID, 0, (uint32_t)a, 0, 0);
        uint4 w1 = draw(run.seed, rp.key, pass, K_CENTROID, 0, (uint32_t)a, 0, 1);
        const double dx = rp.mcstep[ty] * (u53(w0.x, w0.y) - 0.5);
        const double dy = rp.mcstep[ty] * (u53(w0.z, w0.w) - 0.5);
        const double dz = rp.mcstep[ty] * (u53(w1.x, w1.y) - 0.5);
        const int items = nt * np;
        const bool rot_a = FAM == 1 ? (rp.nl_atom >= 0 && a >= rp.nl_atom) : (a == rp.lin_atom);
        // an atom with the linear rotor last among its partners: atom-atom and rotor terms in
        // separate loops (no wave runs both tables' code)
        const bool split = FAM == 0 && !rot_a && rp.lin_atom == N - 1;
        // the minimum image as its own instance of the loops, so the plain one is unchanged
        auto items_sum = [&](auto mi_c) -> double {
            const double* bx = decltype(mi_c)::value ? box : nullptr;
            double acc = 0.0;
            if (split) {  // items: nt*(np-1) atom-atom terms first, then the nt rotor terms
                const int np1 = np - 1, n1 = nt * np1;
                // item i = (slice t0 + q, partner index jp); a lane's items advance by blockDim,
                // so (q, jp) is carried along with one division per atom
                const int dd = np1 > 0 ? np1 : 1, nth = (int)blockDim.x;
                const int dq = nth / dd, dj = nth - dq * dd;
                int qa = (int)threadIdx.x / dd, jpa = (int)threadIdx.x - qa * dd;
                for (int i = threadIdx.x; i < items; i += blockDim.x) {
                    const bool rj = i >= n1;
                    const int it = t0 + (rj ? i - n1 : qa), jp = rj ? 0 : jpa;
                    const int j = rj ? N - 1 : jp + (jp >= a);
                    jpa += dj;
                    qa += dq;
                    if (jpa >= dd) {
                        jpa -= dd;
                        ++qa;
                    }
                    const double4 xa = bead(beads, rp, a, it), xj = bead(beads, rp, j, it);
                    double nx = 0.0, ny = 0.0, nz = 1.0;
                    if (rj) rotor_axis(rotb, rp, it, nx, ny, nz);
                    R vn, vo;
                    if (rj) {
                        vn = pair_v<R>(l1, kn, l2, false, true, xa.x + dx, xa.y + dy, xa.z + dz, xj.x, xj.y, xj.z, nx, ny, nz,
                                       bx);
                        vo = pair_v<R>(l1, kn, l2, false, true, xa.x, xa.y, xa.z, xj.x, xj.y, xj.z, nx, ny, nz, bx);
                    } else {
                        vn = pair_v<R>(l1, kn, l2, false, false, xa.x + dx, xa.y + dy, xa.z + dz, xj.x, xj.y, xj.z, 0.0,
                                       0.0, 1.0, bx);
                        vo = pair_v<R>(l1, kn, l2, false, false, xa.x, xa.y, xa.z, xj.x, xj.y, xj.z, 0.0, 0.0, 1.0, bx);
                    }
                    acc += (double)vn - (double)vo;
                }
            }
            for (int i = split ? items : (int)threadIdx.x; i < items; i += blockDim.x) {
                const int it = t0 + i / np, jp = i % np;
                const int j = jp + (jp >= a);
                const bool rot_j = FAM == 1 ? (rp.nl_atom >= 0 && j >= rp.nl_atom) : (j == rp.lin_atom);
                const double4 xa = bead(beads, rp, a, it), xj = bead(beads, rp, j, it);
                if constexpr (FAM == 1) {
                    const Lut3D<R> l3 = lut3d<R>(T3, rp.pot3d_id);
                    const Frame& fa = frame_at(frames, rp, rot_a ? a : rp.nl_atom, it);
                    const Frame& fj = frame_at(frames, rp, rot_j ? j : rp.nl_atom, it);
                    const double vn = pair_nl<R>(l1, kn, l3, rot_a, rot_j, xa.x + dx, xa.y + dy, xa.z + dz, xj.x, xj.y,
                                                 xj.z, fa, fj, bx);
                    const double vo = pair_nl<R>(l1, kn, l3, rot_a, rot_j, xa.x, xa.y, xa.z, xj.x, xj.y, xj.z, fa, fj, bx);
                    acc += vn - vo;
                } else {
                    double nx, ny, nz;
                    rotor_axis(rotb, rp, it, nx, ny, nz);
                    const R vn = pair_v<R>(l1, kn, l2, rot_a, rot_j, xa.x + dx, xa.y + dy, xa.z + dz, xj.x, xj.y, xj.z,
                                           nx, ny, nz, bx);
                    const R vo = pair_v<R>(l1, kn, l2, rot_a, rot_j, xa.x, xa.y, xa.z, xj.x, xj.y, xj.z, nx, ny, nz, bx);
                    acc += (double)vn - (double)vo;
                }
            }
            return acc;
        };
        double acc;
        if constexpr (MI)
            acc = box ? items_sum(std::integral_constant<bool, true>{}) : items_sum(std::integral_constant<bool, false>{});
        else
            acc = items_sum(std::integral_constant<bool, false>{});
        const double partial = block_sum(acc, red);
        unsigned long long* buf = slots + ((size_t)(ia & 1) * nslots + pt.fbase) * 2;
        if (W == 1) {
            if (threadIdx.x == 0) s_dv = partial;
        } else {
            const unsigned tag = base + (unsigned)ia + 1u;
            if (threadIdx.x < 2) {
                const unsigned long long bits = (unsigned long long)__double_as_longlong(partial);
                __hip_atomic_store(buf + 2 * pt.w + threadIdx.x,
                                   ((unsigned long long)tag << 32) | (threadIdx.x ? bits >> 32 : bits & 0xffffffffull),
                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            if (threadIdx.x < 2 * W) {  // word 2q + 1 is the high half of part q's partial
                unsigned long long wv;
                unsigned spins = 0;
                for (;;) {
                    wv = __hip_atomic_load(buf + threadIdx.x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    if ((unsigned)(wv >> 32) == tag) break;
                    if (++spins > ROT_SPIN_LIMIT) {
                        s_err = 8;
                        break;
                    }
                    __builtin_amdgcn_s_sleep(2);
                }
                reinterpret_cast<unsigned*>(part_sum)[threadIdx.x] = (unsigned)wv;
            }
            __syncthreads();
            if (threadIdx.x == 0) {
                double s = 0.0;
                for (int q = 0; q < W; ++q) s += part_sum[q];
                s_dv = s;
            }
        }
        __syncthreads();
        if (s_err) break;  // a part did not arrive: the later atoms' waits would fail as well
        const double deltav = s_dv;
        bool accepted = false;
        if (deltav < 0.0) accepted = true;
        else if (exp(-deltav * rp.tau) > u53(w1.z, w1.w)) accepted = true;
        if (accepted) {
            acc_n += 1;
            for (int it = t0 + (int)threadIdx.x; it < t0 + nt; it += blockDim.x) {
                const long long idx = rp.bead_off + (long long)a * P + it;
                const double4 b = beads[idx];
                beads[idx] = make_double4(b.x + dx, b.y + dy, b.z + dz, 0.0);
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if (W > 1)  // every part of the replica ran the same atoms
            __hip_atomic_store(myflag, base + (unsigned)rp.tnumb[ty], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (s_err) atomicOr(&mut[r].err, s_err);
        if (pt.w == 0) {
            atomicAdd(&mut[r].total[ty][0], (unsigned long long)rp.tnumb[ty]);
            atomicAdd(&mut[r].accep[ty][0], acc_n);
        }
        atomicAdd(&mut[r].pair_evals[1], 2ull * (unsigned long long)rp.tnumb[ty] * nt * np);
    }
}

// Sum over a slice's partners of V(r, cos theta) with the rotor axis n: cos = -(n . dr)/r
// (PotRotEnergy mc_piqmc.cc:1967-2044).  Items i = gl, gl+G, ... of the slice's geometry block,
// loaded in batches of eight so their table lookups overlap.
template <typename R>
__device__ __forceinline__ double slice_energy(const Lut2D<R>& l2, const vec4_t<R>* pk, int items, int gl, int G,
                                               double nx, double ny, double nz) {
    constexpr int B = 8;  // partner look-ups in flight per lane
    double acc = 0.0;
    const R fx = (R)nx, fy = (R)ny, fz = (R)nz;
    if constexpr (is_fp32<R>()) {
        if (l2.uniform) {  // all cells of a batch first, then all loads, then the lerps
            for (int base = gl; base < items; base += B * G) {
                int cell[B];
                float wr[B], wc[B];
#pragma unroll
                for (int b = 0; b < B; ++b) {
                    const int i = base + b * G;
                    const float4 d = i < items ? pk[i] : make_float4(0.f, 0.f, 1.f, 1.f);
                    lpot2d_u_cell(l2, d.w, -(fx * d.x + fy * d.y + fz * d.z), cell[b], wr[b], wc[b]);
                }
                float4 q[B];
#pragma unroll
                for (int b = 0; b < B; ++b) q[b] = l2.vc[cell[b]];
#pragma unroll
                for (int b = 0; b < B; ++b)
                    if (base + b * G < items) acc += (double)lpot2d_u_lerp(q[b].x, q[b].y, q[b].z, q[b].w, wr[b], wc[b]);
            }
            acc = group_sum(acc, G);
            return acc;
        }
    }
    for (int base = gl; base < items; base += B * G) {
        vec4_t<R> d[B];
#pragma unroll
        for (int b = 0; b < B; ++b) {
            int i = base + b * G;
            if (i < items) d[b] = pk[i];
            else { d[b].x = d[b].y = d[b].z = (R)0; d[b].w = (R)1; }
        }
#pragma unroll
        for (int b = 0; b < B; ++b) {
            R c;
            if constexpr (is_fp32<R>()) c = -(fx * d[b].x + fy * d[b].y + fz * d[b].z);
            else c = -((fx * d[b].x + fy * d[b].y + fz * d[b].z) / d[b].w);
            R v = lpot2d<R>(l2, d[b].w, c);
            if (base + b * G < items) acc += (double)v;
        }
    }
    acc = group_sum(acc, G);
    return acc;
}

// LDS (local index kl = k - k0 + 1; kl = 0 and ns+1 are the neighbours' boundary slices):
// phi, cth, nv[3], E, Lk over ns+2 entries (double), then the geometry block
// geo[(it - k0*RR) * np + jp] (vec4).  E[kl] is the partner energy of slice kl at its current
// axis and Lk[kl] the density factor rho(n_kl . n_kl+1) of the link kl -> kl+1: the "old"
// values of a move are the "new" values of the last accepted move that changed them (for a
// link to a neighbour part, the value that part published with its slice).
template <typename R>
__global__ __launch_bounds__(256) void k_rotor(const Rep* __restrict__ reps, Mut* __restrict__ mut,
                                               const double4* __restrict__ beads, RotBead* __restrict__ rotb,
                                               Tables T, Tables3 T3, const RotPart* __restrict__ parts,
                                               unsigned* __restrict__ flags, unsigned long long* __restrict__ xrec, Run run) {
    extern __shared__ double smem[];
    __shared__ unsigned long long red_acc[4];
    __shared__ int s_err;
    const RotPart pt = parts[blockIdx.x];
    if (pt.r < 0) return;  // padding of the XCD-ordered grid
    const int r = pt.r;
    const Rep& rp = reps[r];
    if (rp.rotkind != 1 || rp.rot_atom < 0) return;
    const int Prot = rp.Prot, RR = rp.RotRatio, P = rp.P, N = rp.natoms;
    int nsweep, j0, jstride;
    if (run.ref_order) {
        if (run.u >= P) return;
        nsweep = 1; j0 = run.u; jstride = 0;
    } else {
        if (run.u >= rp.S) return;
        nsweep = P / rp.S; j0 = run.u; jstride = rp.S;
    }
    const int k0 = pt.k0, ns = pt.k1 - pt.k0, W = pt.W;
    const int np = N - 1;
    const int items = RR * np;
    double* phi = smem;
    double* cth = phi + (ns + 2);
    double* nv = cth + (ns + 2);
    double* E = nv + 3 * (ns + 2);
    double* Lk = E + (ns + 2);
    vec4_t<R>* geo = reinterpret_cast<vec4_t<R>*>(smem + ((7 * (ns + 2) + 3) & ~3));
    const int ar = rp.rot_atom;
    const long long rbase = rp.rot_off + (long long)(ar - rp.tbegin[rp.imtype]) * Prot;
    unsigned* myflag = flags + pt.fbase + pt.w;  // half-sweeps done by the launches before this one
    if (threadIdx.x == 0) s_err = 0;
    // own slices and both neighbours' boundary slices
    for (int kl = threadIdx.x; kl < ns + 2; kl += blockDim.x) {
        int k = k0 + kl - 1;
        if (k < 0) k += Prot;
        if (k >= Prot) k -= Prot;
        RotBead b = rotb[rbase + k];
        phi[kl] = b.phi;
        cth[kl] = b.cth;
        nv[3 * kl] = b.nx;
        nv[3 * kl + 1] = b.ny;
        nv[3 * kl + 2] = b.nz;
    }
    // rotor-partner geometry of the part's translational slices (beads of one atom are
    // contiguous in slices, so consecutive lanes take consecutive slices)
    const int nit = ns * RR, it0 = k0 * RR;
    for (int i = threadIdx.x; i < nit * np; i += blockDim.x) {
        const int jp = i / nit, itl = i - jp * nit;
        const int j = jp + (jp >= ar);
        double4 xa = bead(beads, rp, ar, it0 + itl), xj = bead(beads, rp, j, it0 + itl);
        R dx = (R)(xa.x - xj.x), dy = (R)(xa.y - xj.y), dz = (R)(xa.z - xj.z);
        R rr = sqrt(dx * dx + dy * dy + dz * dz);
        vec4_t<R> v;
        if constexpr (is_fp32<R>()) {  // unit vector: a move's cos(theta) is then one dot product
            R ir = (R)1 / rr;
            v.x = dx * ir; v.y = dy * ir; v.z = dz * ir;
        } else {
            v.x = dx; v.y = dy; v.z = dz;
        }
        v.w = rr;
        geo[itl * np + jp] = v;
    }
    const uint32_t pass = mut[r].pass;
    const int rtype = rp.imtype;
    const Lut2D<R> l2 = lut2d<R>(T, rp.pot2d_id);
    const TabRot trt = T.tr[rp.rotdens];
    // lane groups of G lanes (power of two <= 64) per slice move
    const int maxmoves = (ns + 1) / 2;
    int G = 1;
    while (G < 64 && G * 2 * maxmoves <= (int)blockDim.x) G *= 2;
    const int groups = blockDim.x / G;
    const int g = threadIdx.x / G, gl = threadIdx.x % G;
    const unsigned base = __hip_atomic_load(myflag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    for (int kl = 1 + g; kl <= ns; kl += groups) {
        double e = slice_energy<R>(l2, geo + (long long)(kl - 1) * items, items, gl, G, nv[3 * kl], nv[3 * kl + 1],
                                   nv[3 * kl + 2]);
        if (gl == 0) E[kl] = e;
    }
    for (int kl = threadIdx.x; kl <= ns; kl += blockDim.x) {
        double p = nv[3 * kl] * nv[3 * kl + 3] + nv[3 * kl + 1] * nv[3 * kl + 4] + nv[3 * kl + 2] * nv[3 * kl + 5];
        Lk[kl] = rot_rho(T, trt, p);
    }
    __syncthreads();
    unsigned long long nacc = 0, nev = (threadIdx.x == 0) ? (unsigned long long)ns * items : 0ull;
    int err = 0;
    for (int sw = 0; sw < nsweep; ++sw) {
        const uint32_t jsweep = (uint32_t)(j0 + sw * jstride);
        for (int parity = 0; parity < 2; ++parity) {
            const unsigned h = 2u * (unsigned)sw + (unsigned)parity;  // half-sweeps done before this one
            // the neighbour's boundary slice this half reads: k0-1 (even half), k1 (odd half).
            // Split parts receive it inside the boundary move below.  A single part is its own
            // neighbour: local slices 0 and ns are one slice, as are 1 and ns + 1, and links 0 and
            // ns one link, so the move that changes one copy writes the other (below)
            const int nmoves = (ns - parity + 1) / 2;  // k0 is even, so local slice kl has parity (kl - 1) & 1
            // the move that reads the neighbour's slice: kl = 1 (even half) or kl = ns (odd half)
            const bool receive = W > 1 && !(sw == 0 && parity == 0);
            const int kl_b = parity == 0 ? 1 : ns;
            for (int mv = g; mv < nmoves; mv += groups) {
                const int kl = 2 * mv + parity + 1;
                const int k = k0 + kl - 1;
                const bool bnd = receive && kl == kl_b;
                // the lane whose link factor needs the neighbour's slice: rho(n_k-1 . n') on the
                // lane of kl - 1 in the even half, rho(n' . n_k+1) on the lane of kl + 1 in the odd
                const int dep = bnd ? (G == 1 ? 0 : parity) : -1;
                // both draws of the move up front (independent, so they interleave)
                const uint4 w0 = draw(run.seed, rp.key, pass, K_ROTLIN, jsweep, (uint32_t)ar, (uint32_t)k, 0);
                const uint4 w1 = draw(run.seed, rp.key, pass, K_ROTLIN, jsweep, (uint32_t)ar, (uint32_t)k, 1);
                double cost = cth[kl] + rp.rtstep * (u53(w0.x, w0.y) - 0.5);
                double ph = phi[kl] + rp.rtstep * (u53(w0.z, w0.w) - 0.5);
                if (cost > 1.0) cost = 2.0 - cost;
                if (cost < -1.0) cost = -2.0 - cost;
                double sint = sqrt(1.0 - cost * cost);
                double nnx = sint * cos(ph), nny = sint * sin(ph), nnz = cost;
                // new link factors rho(n_k-1 . n') and rho(n' . n_k+1), on two lanes when G >= 2;
                // their knots are loaded before the partner gathers, so the loads overlap (the
                // factor that needs the neighbour's slice waits for it, after the gathers)
                double rho0 = 0.0, rho1 = 0.0, pa = 0.0, pb = 0.0;
                RhoCell ca{}, cb{};
                auto link_a = [&]() {
                    const int kn = gl == 0 ? kl - 1 : kl + 1;
                    pa = gl == 0 ? nv[3 * kn] * nnx + nv[3 * kn + 1] * nny + nv[3 * kn + 2] * nnz
                                 : nnx * nv[3 * kn] + nny * nv[3 * kn + 1] + nnz * nv[3 * kn + 2];
                    ca = rot_rho_load(T, trt, pa);
                };
                auto link_b = [&]() {
                    pb = nnx * nv[3 * kl + 3] + nny * nv[3 * kl + 4] + nnz * nv[3 * kl + 5];
                    cb = rot_rho_load(T, trt, pb);
                };
                if (gl < 2 && gl != dep) link_a();
                if (G == 1 && !(bnd && parity == 1)) link_b();
                double enew = slice_energy<R>(l2, geo + (long long)(kl - 1) * items, items, gl, G, nnx, nny, nnz);
                if (bnd) {
                    // The even half reads what the previous part published after its odd half of
                    // the last sweep, the odd half what the next part published after its even
                    // half of this sweep: the orientation and the density of the shared link (the
                    // same product of the same operands this part would form), as eight words
                    // that each carry the publishing half's number (base + h) above 32 bits of a
                    // value.  Lanes of this move's group poll the words themselves, so the data
                    // comes with its check; the neighbour cannot publish that record again before
                    // this part's next half.  Only this move waits, after its partner gathers.
                    const int nb = parity == 0 ? (pt.w + W - 1) % W : (pt.w + 1) % W;
                    const unsigned long long* rec = xrec + ((size_t)(pt.fbase + nb) * 2 + (parity == 0 ? 1 : 0)) * 8;
                    const unsigned want = base + h;
                    const int klr = parity == 0 ? 0 : ns + 1;
                    for (int w = gl; w < 8; w += G) {
                        unsigned long long wv = 0;
                        unsigned spins = 0;
                        for (;;) {
                            wv = __hip_atomic_load(rec + w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                            if ((unsigned)(wv >> 32) == want) break;
                            if (++spins > ROT_SPIN_LIMIT) {
                                s_err = 2;
                                break;
                            }
                            __builtin_amdgcn_s_sleep(2);
                        }
                        const int i = w >> 1;  // doubles: the orientation, then the link density
                        double* dst = i < 3 ? &nv[3 * klr + i] : &Lk[parity == 0 ? 0 : ns];
                        reinterpret_cast<unsigned*>(dst)[w & 1] = (unsigned)wv;
                    }
                    __builtin_amdgcn_wave_barrier();
                    if (gl == dep) link_a();
                    if (G == 1 && parity == 1) link_b();
                }
                if (gl < 2) {
                    if (gl == 0) rho0 = rot_rho_eval(T, trt, pa, ca);
                    else rho1 = rot_rho_eval(T, trt, pa, ca);
                }
                if (G == 1) rho1 = rot_rho_eval(T, trt, pb, cb);
                if (G > 1) rho1 = __shfl_xor(rho1, 1, G);
                if (gl == 0) {
                    nev += (unsigned long long)items;
                    double dold = Lk[kl - 1] * Lk[kl];
                    if (fabs(dold) < RZERO) dold = 0.0;
                    double dnew = rho0 * rho1;
                    if (fabs(dnew) < RZERO) dnew = 0.0;
                    if (dold < 0.0 || dnew < 0.0) {
                        err = 1;  // "Negative rot density" (mc_piqmc.cc:850,886): the move is not taken
                    } else {
                        double rd = (dold > RZERO) ? dnew / dold : 1.0;
                        rd *= exp(-rp.tau * (enew - E[kl]));
                        const bool accepted = rd > 1.0 || rd > u53(w1.x, w1.y);
                        if (accepted) {
                            ++nacc;
                            cth[kl] = cost;
                            phi[kl] = ph;
                            nv[3 * kl] = nnx;
                            nv[3 * kl + 1] = nny;
                            nv[3 * kl + 2] = nnz;
                            E[kl] = enew;
                            Lk[kl - 1] = rho0;
                            Lk[kl] = rho1;
                            if (W == 1 && (kl == 1 || kl == ns)) {  // the other copy of the slice and link
                                const int km = kl == 1 ? ns + 1 : 0;
                                nv[3 * km] = nnx;
                                nv[3 * km + 1] = nny;
                                nv[3 * km + 2] = nnz;
                                Lk[kl == 1 ? ns : 0] = kl == 1 ? rho0 : rho1;
                            }
                        }
                    }
                }
            }
            __syncthreads();
            // a wait that gave up: leave at once instead of waiting again at every later half
            // (every lane reads s_err after the barrier; it is written again only in the next half)
            if (s_err) {
                sw = nsweep;
                break;
            }
            // publish the boundary slice this half moved (k0 after the even half, k1-1 after
            // the odd half), then the count of half-sweeps done
            if (W > 1 && threadIdx.x < 8) {
                // orientation and the density of the link to the neighbour's adjacent slice
                // (Lk[0] after the even half, for the previous part; Lk[ns] after the odd half,
                // for the next), two words per double; no wait for the stores, the reader
                // checks every word's number
                const int kl = parity == 0 ? 1 : ns;
                const int i = threadIdx.x >> 1;
                const double v = i < 3 ? nv[3 * kl + i] : Lk[parity == 0 ? 0 : ns];
                const unsigned long long bits = (unsigned long long)__double_as_longlong(v);
                const unsigned half = (threadIdx.x & 1) ? (unsigned)(bits >> 32) : (unsigned)bits;
                __hip_atomic_store(xrec + ((size_t)(pt.fbase + pt.w) * 2 + parity) * 8 + threadIdx.x,
                                   ((unsigned long long)(base + h + 1u) << 32) | half, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
            }
        }
    }
    if (W > 1 && threadIdx.x == 0)  // every part of the chain ran the same halves
        __hip_atomic_store(myflag, base + 2u * (unsigned)nsweep, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    for (int kl = 1 + threadIdx.x; kl <= ns; kl += blockDim.x) {
        RotBead b = rotb[rbase + k0 + kl - 1];
        b.phi = phi[kl];
        b.cth = cth[kl];
        b.nx = nv[3 * kl];
        b.ny = nv[3 * kl + 1];
        b.nz = nv[3 * kl + 2];
        rotb[rbase + k0 + kl - 1] = b;
    }
    unsigned long long s = block_sum(nacc, red_acc);
    unsigned long long ne = block_sum(nev, red_acc);
    if (err) atomicOr(&mut[r].err, 1);
    if (threadIdx.x == 0) {
        if (s_err) atomicOr(&mut[r].err, s_err);
        atomicAdd(&mut[r].total[rtype][2], (unsigned long long)nsweep * ns);
        atomicAdd(&mut[r].accep[rtype][2], s);
        atomicAdd(&mut[r].pair_evals[2], ne);
    }
}

// ------------------------------------------------------------------------- non-linear rotors
// MCRotations3D (mc_piqmc.cc:692-779) with MCRot3Dstep (:938-1199): even slices, then odd;
// at each slice every rotor in turn.  One workgroup per replica and a group of G lanes per
// slice: the density factors rho(t0,t1), rho(t1,t2) (old) and rho(t0,t1'), rho(t1',t2) (new)
// on four lanes, the old and new sums of PotRotE3D (mc_piqmc.cc:2046-2151) over the slice's
// translational slices and partners across the group.  Old energy sums are recomputed for
// every move as the reference does: with two rotors a partner rotor may have moved since.  The
// old density factors are cached per link (Lk[q][k] = rho(slice k, slice k+1) of rotor q): they
// are the new factors of the last accepted move on that link, from the same frames, so a move
// evaluates only its two new links.  Orientations live in rotb and their frames in `frames`; a
// group keeps the frames of all rotors at its slice (and the proposal's) in LDS, so a rotor sees
// the moves of the rotors before it there.
// LDS: Lk [nrot*Prot] doubles, then the groups' frames [groups*(nrot+1)].
template <typename R, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_rotor3d(const Rep* __restrict__ reps, Mut* __restrict__ mut,
                                                 const double4* __restrict__ beads, RotBead* __restrict__ rotb,
                                                 Frame* __restrict__ frames, Tables T, Tables3 T3, const int* __restrict__ replist,
                                                 Run run) {
    extern __shared__ double smem[];
    __shared__ unsigned long long red_acc[BLOCK / 64];  // one partial per wave
    __shared__ int s_err;
    const int r = replist[blockIdx.x];
    const Rep& rp = reps[r];
    if (rp.rotkind != 2 || rp.rot3d_id < 0) return;
    const int Prot = rp.Prot, RR = rp.RotRatio, P = rp.P, N = rp.natoms, nrot = rp.nrot;
    int nsweep, j0, jstride;
    if (run.ref_order) {
        if (run.u >= P) return;
        nsweep = 1; j0 = run.u; jstride = 0;
    } else {
        if (run.u >= rp.S) return;
        nsweep = P / rp.S; j0 = run.u; jstride = rp.S;
    }
    const int np = N - 1;
    const int items = RR * np;
    const long long rb0 = rp.rot_off;  // rotor q, slice k at rb0 + q*Prot + k
    const double4* tab = T3.r3 + (long long)rp.rot3d_id * 3 * ROT3D_SIZE;
    const Lut1D<R> l1 = lut1d<R>(T, rp.pot1d_id);
    const Lut3D<R> l3 = lut3d<R>(T3, rp.pot3d_id);
    const int maxmoves = (Prot + 1) / 2;
    int G = 1;
    while (G < 64 && G * 2 * maxmoves <= (int)blockDim.x) G *= 2;
    const int groups = blockDim.x / G;
    const int g = threadIdx.x / G, gl = threadIdx.x % G;
    double* Lk = smem;
    Frame* gf = reinterpret_cast<Frame*>(smem + ((nrot * Prot + 1) & ~1)) + (long long)g * (nrot + 1);
    const uint32_t pass = mut[r].pass;
    const int rtype = rp.imtype;
    const double twopi = 2.0 * PI_D;
    if (threadIdx.x == 0) s_err = 0;
    unsigned long long nacc = 0;
    int err = 0;
    for (int i = threadIdx.x; i < nrot * Prot; i += blockDim.x) {
        const int q = i / Prot, k = i - q * Prot;
        const long long b = rb0 + (long long)q * Prot;
        if (rotden_rho(frames[b + k], frames[b + (k + 1 == Prot ? 0 : k + 1)], tab, Lk[i])) err = 1;
    }
    __syncthreads();
    for (int sw = 0; sw < nsweep; ++sw) {
        const uint32_t jsweep = (uint32_t)(j0 + sw * jstride);
        for (int parity = 0; parity < 2; ++parity) {
            const int nmoves = (Prot - parity + 1) / 2;
            for (int mv = g; mv < nmoves; mv += groups) {
                const int k = 2 * mv + parity;
                int k0 = k - 1, k2 = k + 1;
                if (k0 < 0) k0 += Prot;
                if (k2 >= Prot) k2 -= Prot;
                for (int q = gl; q < nrot; q += G) gf[q] = frames[rb0 + (long long)q * Prot + k];
                __builtin_amdgcn_wave_barrier();
                for (int q = 0; q < nrot; ++q) {
                    const int gatom = rp.nl_atom + q;
                    const long long ik = rb0 + (long long)q * Prot + k;
                    const RotBead cur = rotb[ik];
                    const uint4 w0 = draw(run.seed, rp.key, pass, K_ROT3D, jsweep, (uint32_t)gatom, (uint32_t)k, 0);
                    const uint4 w1 = draw(run.seed, rp.key, pass, K_ROT3D, jsweep, (uint32_t)gatom, (uint32_t)k, 1);
                    const double rand1 = u53(w0.x, w0.y), rand2 = u53(w0.z, w0.w);
                    const double rand3 = u53(w1.x, w1.y), rand4 = u53(w1.z, w1.w);
                    double cost = cur.cth + rp.rtstep * (rand1 - 0.5);
                    double phi = cur.phi + twopi * (rp.rtstep * (rand2 - 0.5));
                    double chi = cur.chi + twopi * (rp.rtstep * (rand3 - 0.5));
                    if (phi < 0.0) phi = twopi + phi;
                    if (chi < 0.0) chi = twopi + chi;
                    phi = fmod(phi, twopi);
                    chi = fmod(chi, twopi);
                    if (cost > 1.0) cost = 2.0 - cost;
                    if (cost < -1.0) cost = -2.0 - cost;
                    if (gl == 0) gf[nrot] = make_frame(phi, cost, chi);
                    __builtin_amdgcn_wave_barrier();
                    // density factors: old (t0,t1) and (t1,t2) from the link cache, new (t0,t1')
                    // and (t1',t2) on lanes 0 and 1
                    double d[4] = {Lk[q * Prot + k0], Lk[q * Prot + k], 0.0, 0.0};
                    for (int c = 2 + gl; c < 4; c += G) {
                        const Frame& fa = (c == 3) ? gf[nrot] : frames[rb0 + (long long)q * Prot + k0];
                        const Frame& fb = (c == 2) ? gf[nrot] : frames[rb0 + (long long)q * Prot + k2];
                        if (rotden_rho(fa, fb, tab, d[c])) err = 1;
                    }
                    if (G > 1)
                        for (int c = 2; c < 4; ++c) d[c] = __shfl(d[c], (c - 2) % G, G);
                    // PotRotE3D at the old and the proposed orientation
                    double eo = 0.0, en = 0.0;
                    for (int i = gl; i < items; i += G) {
                        const int it = k * RR + i / np, jp = i % np;
                        const int j = jp + (jp >= gatom);
                        const bool rj = j >= rp.nl_atom;
                        const double4 xa = bead(beads, rp, gatom, it), xj = bead(beads, rp, j, it);
                        const Frame& fj = gf[rj ? j - rp.nl_atom : q];
                        eo += pair_nl<R>(l1, l1.k, l3, true, rj, xa.x, xa.y, xa.z, xj.x, xj.y, xj.z, gf[q], fj);
                        en += pair_nl<R>(l1, l1.k, l3, true, rj, xa.x, xa.y, xa.z, xj.x, xj.y, xj.z, gf[nrot], fj);
                    }
                    for (int o = G >> 1; o > 0; o >>= 1) {  // the two sums' shuffles interleave
                        eo += __shfl_xor(eo, o, G);
                        en += __shfl_xor(en, o, G);
                    }
                    double dold = d[0] * d[1];
                    if (fabs(dold) < RZERO) dold = 0.0;
                    if (dold < 0.0) dold = fabs(dold);
                    double dnew = d[2] * d[3];
                    if (fabs(dnew) < RZERO) dnew = 0.0;
                    if (dnew < 0.0) dnew = fabs(dnew);
                    double rd = (dold > RZERO) ? dnew / dold : 1.0;
                    rd *= exp(-rp.tau * (en - eo));
                    const bool accepted = rd > 1.0 || rd > rand4;
                    if (accepted && gl == 0) {
                        ++nacc;
                        const double sint = sqrt(1.0 - cost * cost);
                        RotBead b = cur;
                        b.cth = cost;
                        b.phi = phi;
                        b.chi = chi;
                        b.nx = sint * cos(phi);
                        b.ny = sint * sin(phi);
                        b.nz = cost;
                        rotb[ik] = b;
                        frames[ik] = gf[nrot];
                        gf[q] = gf[nrot];
                        Lk[q * Prot + k0] = d[2];
                        Lk[q * Prot + k] = d[3];
                    }
                    __builtin_amdgcn_wave_barrier();
                }
            }
            __syncthreads();
        }
    }
    if (err) s_err = 1;
    unsigned long long s = block_sum(nacc, red_acc);
    if (threadIdx.x == 0) {
        if (s_err) atomicOr(&mut[r].err, 4);
        const unsigned long long moves = (unsigned long long)nsweep * Prot * nrot;
        atomicAdd(&mut[r].total[rtype][2], moves);
        atomicAdd(&mut[r].accep[rtype][2], s);
        atomicAdd(&mut[r].pair_evals[2], 2ull * moves * (unsigned long long)items);
    }
}

// MCRotations3D with a chain split over W >= 2 workgroups by slice range [k0, k1) (even
// length), as k_rotor splits linear chains: a half-sweep is a parallel step over the part's
// slices with the moves of k_rotor3d.  Before its even half a part needs the frames of slice
// k0-1 (the previous part's last, odd) and before its odd half those of slice k1 (the next
// part's first, even).  After its even half a part publishes its first slice's frames, after its
// odd half its last slice's, each with the density of its link to the neighbour, as a record of
// 64-bit words that each carry the publishing half's number above 32 bits of a double; the
// receiver's lanes poll the words themselves, so the data arrives with its check, and write the
// halves into LDS.  A record is rewritten only after the neighbour has read it (its next half
// needs this part's).  Bounded waits set error bit 2.
// Record (part, parity) at pub + part * pub_stride + parity * pub_stride / 2: 2 words per double
// of the nrot frames, then of the nrot link densities.
// LDS: Lk[q][kl], kl = 0..ns (link kl joins local slices kl and kl+1; local slice 0 is k0-1 and
// ns+1 is k1), the neighbours' frames fnb[2][nrot] (k0-1, then k1), then the groups' frames.
template <typename R>
__global__ __launch_bounds__(256) void k_rotor3d_parts(const Rep* __restrict__ reps, Mut* __restrict__ mut,
                                                       const double4* __restrict__ beads, RotBead* __restrict__ rotb,
                                                       Frame* __restrict__ frames, Tables T, Tables3 T3,
                                                       const RotPart* __restrict__ parts, unsigned* __restrict__ flags,
                                                       unsigned long long* __restrict__ pub, int pub_stride, Run run) {
    extern __shared__ double smem[];
    __shared__ unsigned long long red_acc[4];
    __shared__ int s_err;
    const RotPart pt = parts[blockIdx.x];
    if (pt.r < 0) return;  // padding of the XCD-ordered grid
    const int r = pt.r;
    const Rep& rp = reps[r];
    if (rp.rotkind != 2 || rp.rot3d_id < 0) return;
    const int Prot = rp.Prot, RR = rp.RotRatio, P = rp.P, N = rp.natoms, nrot = rp.nrot;
    int nsweep, j0, jstride;
    if (run.ref_order) {
        if (run.u >= P) return;
        nsweep = 1; j0 = run.u; jstride = 0;
    } else {
        if (run.u >= rp.S) return;
        nsweep = P / rp.S; j0 = run.u; jstride = rp.S;
    }
    const int k0 = pt.k0, ns = pt.k1 - pt.k0, W = pt.W;
    const int L = ns + 1;  // links per rotor
    const int np = N - 1;
    const int items = RR * np;
    const long long rb0 = rp.rot_off;
    const double4* tab = T3.r3 + (long long)rp.rot3d_id * 3 * ROT3D_SIZE;
    const Lut1D<R> l1 = lut1d<R>(T, rp.pot1d_id);
    const Lut3D<R> l3 = lut3d<R>(T3, rp.pot3d_id);
    const int maxmoves = (ns + 1) / 2;
    int G = 1;
    while (G < 64 && G * 2 * maxmoves <= (int)blockDim.x) G *= 2;
    const int groups = blockDim.x / G;
    const int g = threadIdx.x / G, gl = threadIdx.x % G;
    const int nat = rp.nl_atom;  // atoms come first: 0 .. nat-1
    const int items_a = RR * nat, items_r = RR * (nrot - 1);
    double* Lk = smem;
    double* Ea = Lk + nrot * L;  // rotor-atom energy of every own slice at its current orientation
    Frame* fnb = reinterpret_cast<Frame*>(smem + ((nrot * L + nrot * ns + 1) & ~1));
    Frame* gf = fnb + 2 * nrot + (long long)g * (nrot + 1);
    unsigned* myflag = flags + pt.fbase + pt.w;  // half-sweeps done by the launches before this one
    constexpr int FD = (int)(sizeof(Frame) / sizeof(double));  // doubles per frame
    const int nwords = 2 * nrot * (FD + 1);                      // words of one record
    const int half_rec = pub_stride / 2;
    unsigned long long* mypub = pub + (long long)(pt.fbase + pt.w) * pub_stride;   // + parity * half_rec
    const unsigned long long* prevlast = pub + (long long)(pt.fbase + (pt.w + W - 1) % W) * pub_stride + half_rec;
    const unsigned long long* nextfirst = pub + (long long)(pt.fbase + (pt.w + 1) % W) * pub_stride;
    const unsigned base = __hip_atomic_load(myflag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint32_t pass = mut[r].pass;
    const int rtype = rp.imtype;
    const double twopi = 2.0 * PI_D;
    auto own = [&](int q, int k) -> const Frame& { return frames[rb0 + (long long)q * Prot + k]; };
    if (threadIdx.x == 0) s_err = 0;
    unsigned long long nacc = 0;
    int err = 0;
    // at launch start no neighbour has moved its boundary slice yet (its first move on it waits
    // for this part), so the boundary frames come straight from memory
    for (int i = threadIdx.x; i < 2 * nrot; i += blockDim.x) {
        const int side = i / nrot, q = i - side * nrot;
        int k = side ? k0 + ns : k0 - 1;
        if (k < 0) k += Prot;
        if (k >= Prot) k -= Prot;
        fnb[i] = own(q, k);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nrot * L; i += blockDim.x) {
        const int q = i / L, kl = i - q * L;
        const Frame& fa = kl == 0 ? fnb[q] : own(q, k0 + kl - 1);
        const Frame& fb = kl == ns ? fnb[nrot + q] : own(q, k0 + kl);
        if (rotden_rho(fa, fb, tab, Lk[i])) err = 1;
    }
    // rotor-atom sums at the current orientations: atoms do not move during the sweep, so a
    // move's old sum is this one or the new sum of the last move accepted at the slice (the same
    // lanes add the same terms in the same order either way)
    for (int kl = 1 + g; kl <= ns; kl += groups) {
        const int k = k0 + kl - 1;
        for (int q = 0; q < nrot; ++q) {
            const int gatom = rp.nl_atom + q;
            const Frame& fq = own(q, k);
            double e = 0.0;
            for (int i = gl; i < items_a; i += G) {
                const int it = k * RR + i / nat, a = i % nat;
                const double4 xa = bead(beads, rp, gatom, it), xj = bead(beads, rp, a, it);
                e += (double)vcord_v<R>(l3, fq, xa.x, xa.y, xa.z, xj.x, xj.y, xj.z);
            }
            e = group_sum(e, G);
            if (gl == 0) Ea[q * ns + kl - 1] = e;
        }
    }
    __syncthreads();
    for (int sw = 0; sw < nsweep; ++sw) {
        const uint32_t jsweep = (uint32_t)(j0 + sw * jstride);
        for (int parity = 0; parity < 2; ++parity) {
            const unsigned h = 2u * (unsigned)sw + (unsigned)parity;  // half-sweeps done before this one
            if (!(sw == 0 && parity == 0)) {
                // even half: the previous part's record of its odd half of the last sweep; odd
                // half: the next part's record of its even half of this sweep.  The neighbour's
                // boundary frames, and the densities of the shared links that it evaluated from
                // the same frames in the same order; each lane polls its own words.
                const unsigned long long* src = parity == 0 ? prevlast : nextfirst;
                unsigned* dstf = reinterpret_cast<unsigned*>(fnb + (parity == 0 ? 0 : nrot));
                const unsigned want = base + h;
                for (int w = threadIdx.x; w < nwords; w += blockDim.x) {
                    unsigned long long wv;
                    unsigned spins = 0;
                    for (;;) {
                        wv = __hip_atomic_load(src + w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        if ((unsigned)(wv >> 32) == want) break;
                        if (++spins > ROT_SPIN_LIMIT) {
                            s_err = 2;
                            break;
                        }
                        __builtin_amdgcn_s_sleep(2);
                    }
                    const int i = w >> 1;  // the double: frames first, then link densities
                    if (i < nrot * FD) dstf[w] = (unsigned)wv;
                    else reinterpret_cast<unsigned*>(&Lk[(i - nrot * FD) * L + (parity == 0 ? 0 : ns)])[w & 1] = (unsigned)wv;
                }
                __syncthreads();
                // a wait that gave up: no later wait of this launch can succeed either, so the
                // part leaves at once (s_err is read by every lane after the barrier, and no lane
                // writes it again before the barrier that closes this half)
                if (s_err) {
                    sw = nsweep;
                    break;
                }
            }
            const int nmoves = (ns - parity + 1) / 2;  // k0 is even: local slice kl has parity (kl - 1) & 1
            for (int mv = g; mv < nmoves; mv += groups) {
                const int kl = 2 * mv + parity + 1;
                const int k = k0 + kl - 1;
                for (int q = gl; q < nrot; q += G) gf[q] = own(q, k);
                __builtin_amdgcn_wave_barrier();
                for (int q = 0; q < nrot; ++q) {
                    const int gatom = rp.nl_atom + q;
                    const long long ik = rb0 + (long long)q * Prot + k;
                    const RotBead cur = rotb[ik];
                    const uint4 w0 = draw(run.seed, rp.key, pass, K_ROT3D, jsweep, (uint32_t)gatom, (uint32_t)k, 0);
                    const uint4 w1 = draw(run.seed, rp.key, pass, K_ROT3D, jsweep, (uint32_t)gatom, (uint32_t)k, 1);
                    const double rand1 = u53(w0.x, w0.y), rand2 = u53(w0.z, w0.w);
                    const double rand3 = u53(w1.x, w1.y), rand4 = u53(w1.z, w1.w);
                    double cost = cur.cth + rp.rtstep * (rand1 - 0.5);
                    double phi = cur.phi + twopi * (rp.rtstep * (rand2 - 0.5));
                    double chi = cur.chi + twopi * (rp.rtstep * (rand3 - 0.5));
                    if (phi < 0.0) phi = twopi + phi;
                    if (chi < 0.0) chi = twopi + chi;
                    phi = fmod(phi, twopi);
                    chi = fmod(chi, twopi);
                    if (cost > 1.0) cost = 2.0 - cost;
                    if (cost < -1.0) cost = -2.0 - cost;
                    if (gl == 0) gf[nrot] = make_frame(phi, cost, chi);
                    __builtin_amdgcn_wave_barrier();
                    double d[4] = {Lk[q * L + kl - 1], Lk[q * L + kl], 0.0, 0.0};
                    for (int c = 2 + gl; c < 4; c += G) {
                        const Frame& fa = (c == 3) ? gf[nrot] : (kl == 1 ? fnb[q] : own(q, k - 1));
                        const Frame& fb = (c == 2) ? gf[nrot] : (kl == ns ? fnb[nrot + q] : own(q, k + 1));
                        if (rotden_rho(fa, fb, tab, d[c])) err = 1;
                    }
                    if (G > 1)
                        for (int c = 2; c < 4; ++c) d[c] = __shfl(d[c], (c - 2) % G, G);
                    // rotor-atom terms at the proposal (the old sum is cached), then the terms of
                    // the other rotors at both orientations (a partner rotor may have moved)
                    double ea = 0.0, eo = 0.0, en = 0.0;
                    for (int i = gl; i < items_a; i += G) {
                        const int it = k * RR + i / nat, a = i % nat;
                        const double4 xa = bead(beads, rp, gatom, it), xj = bead(beads, rp, a, it);
                        ea += (double)vcord_v<R>(l3, gf[nrot], xa.x, xa.y, xa.z, xj.x, xj.y, xj.z);
                    }
                    for (int i = gl; i < items_r; i += G) {
                        const int it = k * RR + i / (nrot - 1), pp = i % (nrot - 1);
                        const int p = pp + (pp >= q);
                        const double4 xa = bead(beads, rp, gatom, it), xj = bead(beads, rp, rp.nl_atom + p, it);
                        // the two orientations in turn, not interleaved: one caleng's sites
                        // live at a time, which keeps the kernel within 256 registers (two
                        // waves per SIMD, so two parts of a water chain share a CU)
#pragma unroll 1
                        for (int o = 0; o < 2; ++o) {
                            const double e = caleng_v(xa.x, xa.y, xa.z, xj.x, xj.y, xj.z, gf[o ? nrot : q], gf[p]);
                            if (o) en += e;
                            else eo += e;
                        }
                    }
                    for (int o = G >> 1; o > 0; o >>= 1) {  // the sums' shuffles interleave
                        ea += __shfl_xor(ea, o, G);
                        eo += __shfl_xor(eo, o, G);
                        en += __shfl_xor(en, o, G);
                    }
                    const double ea_old = Ea[q * ns + kl - 1];
                    eo += ea_old;
                    en += ea;
                    double dold = d[0] * d[1];
                    if (fabs(dold) < RZERO) dold = 0.0;
                    if (dold < 0.0) dold = fabs(dold);
                    double dnew = d[2] * d[3];
                    if (fabs(dnew) < RZERO) dnew = 0.0;
                    if (dnew < 0.0) dnew = fabs(dnew);
                    double rd = (dold > RZERO) ? dnew / dold : 1.0;
                    rd *= exp(-rp.tau * (en - eo));
                    const bool accepted = rd > 1.0 || rd > rand4;
                    if (accepted && gl == 0) {
                        ++nacc;
                        const double sint = sqrt(1.0 - cost * cost);
                        RotBead b = cur;
                        b.cth = cost;
                        b.phi = phi;
                        b.chi = chi;
                        b.nx = sint * cos(phi);
                        b.ny = sint * sin(phi);
                        b.nz = cost;
                        rotb[ik] = b;
                        frames[ik] = gf[nrot];
                        gf[q] = gf[nrot];
                        Lk[q * L + kl - 1] = d[2];
                        Lk[q * L + kl] = d[3];
                        Ea[q * ns + kl - 1] = ea;
                    }
                    __builtin_amdgcn_wave_barrier();
                }
            }
            __syncthreads();
            // publish the boundary slice this half moved: its frames and the boundary link
            // densities (Lk[q][0] after the even half, Lk[q][ns] after the odd), two words per
            // double numbered base + h + 1; no wait for the stores, the reader checks every word
            {
                const int ks = parity == 0 ? k0 : k0 + ns - 1;
                unsigned long long* dst = mypub + (parity == 0 ? 0 : half_rec);
                const unsigned long long tag = (unsigned long long)(base + h + 1u) << 32;
                for (int w = threadIdx.x; w < nwords; w += blockDim.x) {
                    const int i = w >> 1;
                    const double v = i < nrot * FD ? reinterpret_cast<const double*>(&own(i / FD, ks))[i % FD]
                                                   : Lk[(i - nrot * FD) * L + (parity == 0 ? 0 : ns)];
                    const unsigned long long bits = (unsigned long long)__double_as_longlong(v);
                    __hip_atomic_store(dst + w, tag | ((w & 1) ? (bits >> 32) : (bits & 0xffffffffull)), __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
                }
            }
        }
    }
    if (threadIdx.x == 0)  // every part of the chain ran the same halves
        __hip_atomic_store(myflag, base + 2u * (unsigned)nsweep, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (err) atomicOr(&mut[r].err, 4);
    unsigned long long sacc = block_sum(nacc, red_acc);
    if (threadIdx.x == 0) {
        if (s_err) atomicOr(&mut[r].err, s_err);
        const unsigned long long moves = (unsigned long long)nsweep * ns * nrot;
        atomicAdd(&mut[r].total[rtype][2], moves);
        atomicAdd(&mut[r].accep[rtype][2], sacc);
        // per move the proposal's rotor-atom terms and both orientations' rotor-rotor terms, and
        // the cached sums once per launch
        atomicAdd(&mut[r].pair_evals[2], moves * (unsigned long long)(items_a + 2 * items_r) +
                                             (unsigned long long)ns * nrot * items_a);
    }
}

// ------------------------------------------------------------------------- restart
// Rotation matrices of every non-linear rotor slice of replica r from its stored angles, with
// the code the moves use when they accept (make_frame), so a resumed chain sees the same bits.
__global__ void k_frames_from_angles(const Rep* __restrict__ reps, const RotBead* __restrict__ rotb,
                                     Frame* __restrict__ frames, int r, int n) {
    const Rep& rp = reps[r];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const RotBead& b = rotb[rp.rot_off + i];
    frames[rp.rot_off + i] = make_frame(b.phi, b.cth, b.chi);
}

// ------------------------------------------------------------------------- rotational correlation
// GetRCF (mc_estim.cc:1098-1139), one sample: for every time origin it0 and offset itc of the
// first rotor of the rotating type, p = n(it0).n(it0+itc) with the unit axes of MCCosine;
// rcf[0][itc] += p and rcf[l][itc] += P_l(p) for l = 1..9.  (The reference adds 1.0 for l >= 1:
// its Legendre call is commented out; here P_l comes from the three-term recurrence.)
// Grid (replica, offset block of 256): a lane per offset, accumulated in FP64.
__global__ __launch_bounds__(256) void k_rcf(const Rep* __restrict__ reps, const RotBead* __restrict__ rotb,
                                             const int* __restrict__ list, double* __restrict__ rcf) {
    const Rep& rp = reps[list[blockIdx.x]];
    const int Prot = rp.Prot;
    const RotBead* b = rotb + rp.rot_off;
    double* out = rcf + rp.rcf_off;
    for (int itc = blockIdx.y * blockDim.x + threadIdx.x; itc < Prot; itc += gridDim.y * blockDim.x) {
        double acc[RCF_L];
        for (int l = 0; l < RCF_L; ++l) acc[l] = 0.0;
        int t1 = itc;
        for (int it0 = 0; it0 < Prot; ++it0) {
            const double p = b[it0].nx * b[t1].nx + b[it0].ny * b[t1].ny + b[it0].nz * b[t1].nz;
            acc[0] += p;
            double pm = 1.0, pc = p;  // P_0, P_1
            acc[1] += pc;
            for (int l = 1; l + 1 < RCF_L; ++l) {
                const double pn = ((2 * l + 1) * p * pc - l * pm) / (l + 1);
                pm = pc;
                pc = pn;
                acc[l + 1] += pc;
            }
            if (++t1 == Prot) t1 = 0;
        }
        for (int l = 0; l < RCF_L; ++l) out[(long long)l * Prot + itc] += acc[l];
    }
}

// ------------------------------------------------------------------------- delayed acceptance
// MCRotations3D (mc_piqmc.cc:692-779) with the one-stage test of MCRot3Dstep (:1166-1178),
// min(1, rho_new/rho_old * exp(-tau dV)), split into three factors (multi-stage delayed
// acceptance, Christen & Fox 2005), each a Metropolis test of its own:
//   stage 0  min(1, rho_new/rho_old), the density factor, which the move has in hand anyway;
//   stage 1  min(1, exp(-tau (dVs + dVx))), Vs the surrogate network summed over the slice's
//            network pairs and Vx the exact energy of the other pairs;
//   stage 2  min(1, exp(-tau (dV_net - dVs))), the exact network-pair sums.
// Each factor is in detailed balance with its own part of the target, so their product is in
// detailed balance with rho * exp(-tau V) for any surrogate: the stationary distribution is the
// one-stage move's.  A proposal is charged the network only when it passes stage 0, and the exact
// network-pair sums only when it passes stage 1.  Stage 0 takes its uniform from counter word 3
// = 2 (second half), stage 1 the one-stage test's (rand4), stage 2 counter word 3 = 2 (first).
// Network pairs: rotor-atom pairs when the replica has atoms (rotor-rotor pairs of two rotors are
// then exact in stage 1), rotor-rotor pairs when it has none.
//
// Chains are split into parts as k_rotor3d_parts splits them, with the same boundary records
// between half-sweeps (a part of W = 1 is a whole chain and trades with itself).  Within a half,
// rotor q in turn (a move of rotor q at slice k sees the moves of the rotors before it at k, as
// in the one-stage kernel); each wave takes its moves in chunks of DA_MV, lane j the chunk's move
// j: (1) proposal, new link densities and the exact stage-1 terms; (2) stage 0, then the network
// rows of the moves that passed it, listed first, 32 per pass through the matrix cores, summed
// per move in row order; (3) stage 1; (4) stage 2 for the moves that passed -- the whole wave
// sums one move's rotor-atom pairs, or each lane its move's rotor-rotor pairs; (5) updates.  The
// surrogate and exact rotor-atom sums of every own slice at its current orientation are cached
// (atoms do not move during the sweep): a move evaluates only its proposal, and an accepted move
// leaves its new sums as the cached ones.
// LDS: the packed network; Lk [nrot][ns+1], Vs and Ex [nrot][ns] (doubles); the neighbours'
// boundary frames [2 nrot]; per wave DA_MV proposal frames and 32 row outputs (DA_MV and the
// LDS size: pimc_plan.hpp).

// Body-frame features of a rotor-rotor pair for the network: the distance of the centres, the
// cosines of both body z axes (the dipole axis of the TIP4P geometry) with the unit separation
// from rotor q to rotor p, and the cosine between the two z axes.
__device__ __forceinline__ void mlp_rotor_pair_features(const Frame& fq, const Frame& fp, double qx, double qy,
                                                        double qz, double px, double py, double pz, float out[MLP_IN]) {
    const double rx = px - qx, ry = py - qy, rz = pz - qz;
    const double rn = sqrt(rx * rx + ry * ry + rz * rz);
    const double inv = rn > 0.0 ? 1.0 / rn : 0.0;
    out[0] = (float)rn;
    out[1] = (float)((fq.m[6] * rx + fq.m[7] * ry + fq.m[8] * rz) * inv);
    out[2] = (float)((fp.m[6] * rx + fp.m[7] * ry + fp.m[8] * rz) * inv);
    out[3] = (float)(fq.m[6] * fp.m[6] + fq.m[7] * fp.m[7] + fq.m[8] * fp.m[8]);
}

template <typename R>
__global__ __launch_bounds__(256) void k_rotor3d_da(const Rep* __restrict__ reps, Mut* __restrict__ mut,
                                                    const double4* __restrict__ beads, RotBead* __restrict__ rotb,
                                                    Frame* __restrict__ frames, Tables T, Tables3 T3,
                                                    const RotPart* __restrict__ parts, unsigned* __restrict__ flags,
                                                    unsigned long long* __restrict__ pub, int pub_stride, Run run,
                                                    const MlpPacked* __restrict__ Wg) {
    extern __shared__ uint4 mlp_lds[];
    MlpPacked* W = reinterpret_cast<MlpPacked*>(mlp_lds);
    __shared__ unsigned long long red_a[4], red_p[4], red_n[4];
    __shared__ int s_err;
    const RotPart pt = parts[blockIdx.x];
    if (pt.r < 0) return;  // padding of the XCD-ordered grid
    const int r = pt.r;
    const Rep& rp = reps[r];
    if (rp.rotkind != 2 || rp.rot3d_id < 0) return;
    const int Prot = rp.Prot, RR = rp.RotRatio, P = rp.P, nrot = rp.nrot;
    const int nat = rp.nl_atom;              // atoms come first: 0 .. nat-1
    const bool net_atoms = nat > 0;          // network pairs: rotor-atom, else rotor-rotor
    int nsweep, j0, jstride;
    if (run.ref_order) {
        if (run.u >= P) return;
        nsweep = 1; j0 = run.u; jstride = 0;
    } else {
        if (run.u >= rp.S) return;
        nsweep = P / rp.S; j0 = run.u; jstride = rp.S;
    }
    const int k0 = pt.k0, ns = pt.k1 - pt.k0, NP = pt.W;
    const int L = ns + 1;
    const long long rb0 = rp.rot_off;
    const double4* tab = T3.r3 + (long long)rp.rot3d_id * 3 * ROT3D_SIZE;
    const Lut1D<R> l1 = lut1d<R>(T, rp.pot1d_id);
    const Lut3D<R> l3 = lut3d<R>(T3, rp.pot3d_id);
    const int wv = threadIdx.x >> 6, l = threadIdx.x & 63, nw = blockDim.x >> 6;
    double* Lk = reinterpret_cast<double*>(W + 1);
    double* Vs = Lk + nrot * L;   // cached surrogate sums (rotor-atom network)
    double* Ex = Vs + nrot * ns;  // cached exact rotor-atom sums
    Frame* fnb = reinterpret_cast<Frame*>(Lk + ((nrot * L + 2 * nrot * ns + 1) & ~1));
    Frame* prop = fnb + 2 * nrot + wv * DA_MV;
    float* rowo = reinterpret_cast<float*>(fnb + 2 * nrot + nw * DA_MV) + wv * 32;
    unsigned* myflag = flags + pt.fbase + pt.w;
    constexpr int FD = (int)(sizeof(Frame) / sizeof(double));
    const int nwords = 2 * nrot * (FD + 1);
    const int half_rec = pub_stride / 2;
    unsigned long long* mypub = pub + (long long)(pt.fbase + pt.w) * pub_stride;
    const unsigned long long* prevlast = pub + (long long)(pt.fbase + (pt.w + NP - 1) % NP) * pub_stride + half_rec;
    const unsigned long long* nextfirst = pub + (long long)(pt.fbase + (pt.w + 1) % NP) * pub_stride;
    const unsigned base = __hip_atomic_load(myflag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint32_t pass = mut[r].pass;
    const int rtype = rp.imtype;
    const double twopi = 2.0 * PI_D;
    const double ys = (double)Wg->ys;
    // network rows per move: RR x atoms (proposal only; the old sum is cached), or RR x other
    // rotors x (old, new)
    const int rpm = net_atoms ? RR * nat : RR * (nrot - 1) * 2;
    const int items_x = RR * nat;  // exact rotor-atom pairs of a slice
    auto own = [&](int q, int k) -> const Frame& { return frames[rb0 + (long long)q * Prot + k]; };
    if (threadIdx.x == 0) s_err = 0;
    unsigned long long nacc = 0, npass = 0, nnet = 0;
    int err = 0;
    mlp_stage(Wg, W);
    for (int i = threadIdx.x; i < 2 * nrot; i += blockDim.x) {
        const int side = i / nrot, q = i - side * nrot;
        int k = side ? k0 + ns : k0 - 1;
        if (k < 0) k += Prot;
        if (k >= Prot) k -= Prot;
        fnb[i] = own(q, k);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nrot * L; i += blockDim.x) {
        const int q = i / L, kl = i - q * L;
        const Frame& fa = kl == 0 ? fnb[q] : own(q, k0 + kl - 1);
        const Frame& fb = kl == ns ? fnb[nrot + q] : own(q, k0 + kl);
        if (rotden_rho(fa, fb, tab, Lk[i])) err = 1;
    }
    // The network's rows for slices of rotor q, summed per slice in row order.  Entry j < nm of
    // the list is the move of lane src (src_j on lane j): its slice is that lane's kl_j and, for
    // the proposal rows, its frame prop[src]; sums of (old, new) rows return in so/sn on lane j.
    auto net_rows = [&](int q, int nm, int kl_j, int src_j, bool proposal, double& so, double& sn) {
        so = 0.0;
        sn = 0.0;
        const int total = nm * rpm;
        for (int b = 0; b < total; b += 32) {
            const int row = b + (l & 31);
            // the row's move and its slice, read from the move's lane with every lane active
            const int jj = min(row, total - 1) / rpm;
            const int sl = __shfl(src_j, jj);
            const int klr = __shfl(kl_j, sl);
            float f[MLP_IN] = {0.0f, 0.0f, 0.0f, 0.0f};
            if (row < total) {
                const int idx = row - jj * rpm;
                const int k = k0 + klr - 1;
                const int gq = rp.nl_atom + q;
                if (net_atoms) {
                    const int s = idx / nat, a = idx - s * nat;
                    const int it = k * RR + s;
                    const double4 xa = bead(beads, rp, gq, it), xj = bead(beads, rp, a, it);
                    const Frame& fq = proposal ? prop[sl] : own(q, k);
                    mlp_pair_features(fq.m, fq.zn, xa.x, xa.y, xa.z, xj.x, xj.y, xj.z, f);
                } else {
                    const int which = idx & 1, rest = idx >> 1;
                    const int s = rest / (nrot - 1), pp = rest - s * (nrot - 1);
                    const int p = pp + (pp >= q);
                    const int it = k * RR + s;
                    const double4 xq = bead(beads, rp, gq, it), xp = bead(beads, rp, rp.nl_atom + p, it);
                    const Frame& fq = (proposal && which) ? prop[sl] : own(q, k);
                    mlp_rotor_pair_features(fq, own(p, k), xq.x, xq.y, xq.z, xp.x, xp.y, xp.z, f);
                }
#pragma unroll
                for (int c = 0; c < MLP_IN; ++c) f[c] = (f[c] - W->xm[c]) * W->xinv[c];
            }
            // an opaque zero offset keeps the weight reads inside the loop (hoisted out they
            // would hold ~130 registers of weights per lane)
            int z = 0;
            __asm__ volatile("" : "+s"(z));
            const float o = mlp_rows32(reinterpret_cast<const MlpPacked*>(reinterpret_cast<const char*>(W) + z), f);
            if (l < 32) rowo[l] = o;
            __builtin_amdgcn_wave_barrier();
            if (l < nm) {  // the rows of move l that fall in this batch, in row order
                const int r0 = max(l * rpm, b), r1 = min((l + 1) * rpm, b + 32);
                for (int rr = r0; rr < r1; ++rr) {
                    const int idx = rr - l * rpm;
                    const double v = (double)rowo[rr - b];
                    if (net_atoms || (idx & 1)) sn += v;
                    else so += v;
                }
            }
            __builtin_amdgcn_wave_barrier();
        }
    };
    // exact rotor-atom sum of rotor q at local slice kl with frame f, over the whole wave
    auto exact_atoms = [&](int q, int kl, const Frame& f) -> double {
        const int k = k0 + kl - 1, gq = rp.nl_atom + q;
        double e = 0.0;
        for (int i = l; i < items_x; i += 64) {
            const int s = i / nat, a = i - s * nat;
            const int it = k * RR + s;
            const double4 xa = bead(beads, rp, gq, it), xj = bead(beads, rp, a, it);
            e += (double)vcord_v<R>(l3, f, xa.x, xa.y, xa.z, xj.x, xj.y, xj.z);
        }
        return wave_sum(e);
    };
    // cached sums of every own slice at its current orientation (rotor-atom network only)
    if (net_atoms) {
        for (int q = 0; q < nrot; ++q) {
            for (int c0 = wv * DA_MV; c0 < ns; c0 += nw * DA_MV) {
                const int nm = min(DA_MV, ns - c0);
                double so, sn;
                net_rows(q, nm, c0 + l + 1, l, false, so, sn);
                if (l < nm) Vs[q * ns + c0 + l] = sn;
                for (int j = 0; j < nm; ++j) {
                    const double e = exact_atoms(q, c0 + j + 1, own(q, k0 + c0 + j));
                    if (l == 0) Ex[q * ns + c0 + j] = e;
                }
            }
        }
    }
    __syncthreads();
    for (int sw = 0; sw < nsweep; ++sw) {
        const uint32_t jsweep = (uint32_t)(j0 + sw * jstride);
        for (int parity = 0; parity < 2; ++parity) {
            const unsigned h = 2u * (unsigned)sw + (unsigned)parity;
            if (!(sw == 0 && parity == 0)) {
                if (NP == 1) {
                    // one part: the boundary slices are its own, moved in the last half
                    if (threadIdx.x < nrot) {
                        const int q = threadIdx.x;
                        if (parity == 0) {
                            fnb[q] = own(q, Prot - 1);
                            Lk[q * L] = Lk[q * L + ns];
                        } else {
                            fnb[nrot + q] = own(q, 0);
                            Lk[q * L + ns] = Lk[q * L];
                        }
                    }
                } else {
                    const unsigned long long* src = parity == 0 ? prevlast : nextfirst;
                    unsigned* dstf = reinterpret_cast<unsigned*>(fnb + (parity == 0 ? 0 : nrot));
                    const unsigned want = base + h;
                    for (int w = threadIdx.x; w < nwords; w += blockDim.x) {
                        unsigned long long wvv;
                        unsigned spins = 0;
                        for (;;) {
                            wvv = __hip_atomic_load(src + w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                            if ((unsigned)(wvv >> 32) == want) break;
                            if (++spins > ROT_SPIN_LIMIT) {
                                s_err = 2;
                                break;
                            }
                            __builtin_amdgcn_s_sleep(2);
                        }
                        const int i = w >> 1;
                        if (i < nrot * FD) dstf[w] = (unsigned)wvv;
                        else reinterpret_cast<unsigned*>(&Lk[(i - nrot * FD) * L + (parity == 0 ? 0 : ns)])[w & 1] = (unsigned)wvv;
                    }
                }
                __syncthreads();
                if (s_err) {  // a wait gave up: leave at once (see k_rotor3d_parts)
                    sw = nsweep;
                    break;
                }
            }
            const int nmoves = (ns - parity + 1) / 2;  // k0 is even: local slice kl has parity (kl - 1) & 1
            const int mine = nmoves > wv ? (nmoves - wv + nw - 1) / nw : 0;  // moves wv, wv + nw, ...
            for (int q = 0; q < nrot; ++q) {
                const int gq = rp.nl_atom + q;
                for (int c0 = 0; c0 < mine; c0 += DA_MV) {
                    const int nm = min(DA_MV, mine - c0);
                    // (1) proposal of move l
                    const int mv = wv + (c0 + l) * nw;
                    const int kl = 2 * mv + parity + 1, k = k0 + kl - 1;
                    const long long ik = rb0 + (long long)q * Prot + k;
                    double cost = 0.0, phi = 0.0, chi = 0.0, ratio = 1.0, d2 = 0.0, d3 = 0.0, dvx = 0.0;
                    double rand4 = 0.0, rand5 = 0.0, rand6 = 0.0;
                    RotBead cur{};
                    if (l < nm) {
                        cur = rotb[ik];
                        const uint4 w0 = draw(run.seed, rp.key, pass, K_ROT3D, jsweep, (uint32_t)gq, (uint32_t)k, 0);
                        const uint4 w1 = draw(run.seed, rp.key, pass, K_ROT3D, jsweep, (uint32_t)gq, (uint32_t)k, 1);
                        const uint4 w2 = draw(run.seed, rp.key, pass, K_ROT3D, jsweep, (uint32_t)gq, (uint32_t)k, 2);
                        const double rand1 = u53(w0.x, w0.y), rand2 = u53(w0.z, w0.w), rand3 = u53(w1.x, w1.y);
                        rand4 = u53(w1.z, w1.w);
                        rand5 = u53(w2.x, w2.y);
                        rand6 = u53(w2.z, w2.w);
                        cost = cur.cth + rp.rtstep * (rand1 - 0.5);
                        phi = cur.phi + twopi * (rp.rtstep * (rand2 - 0.5));
                        chi = cur.chi + twopi * (rp.rtstep * (rand3 - 0.5));
                        if (phi < 0.0) phi = twopi + phi;
                        if (chi < 0.0) chi = twopi + chi;
                        phi = fmod(phi, twopi);
                        chi = fmod(chi, twopi);
                        if (cost > 1.0) cost = 2.0 - cost;
                        if (cost < -1.0) cost = -2.0 - cost;
                        const Frame fn = make_frame(phi, cost, chi);
                        prop[l] = fn;
                        const Frame& fa = kl == 1 ? fnb[q] : own(q, k - 1);
                        const Frame& fb = kl == ns ? fnb[nrot + q] : own(q, k + 1);
                        if (rotden_rho(fa, fn, tab, d2)) err = 1;
                        if (rotden_rho(fn, fb, tab, d3)) err = 1;
                        double dold = Lk[q * L + kl - 1] * Lk[q * L + kl];
                        if (fabs(dold) < RZERO) dold = 0.0;
                        if (dold < 0.0) dold = fabs(dold);
                        double dnew = d2 * d3;
                        if (fabs(dnew) < RZERO) dnew = 0.0;
                        if (dnew < 0.0) dnew = fabs(dnew);
                        ratio = (dold > RZERO) ? dnew / dold : 1.0;
                        if (net_atoms) {  // rotor-rotor pairs exact in stage 1
                            const Frame& fo = own(q, k);
                            for (int p = 0; p < nrot; ++p) {
                                if (p == q) continue;
                                for (int s = 0; s < RR; ++s) {
                                    const int it = k * RR + s;
                                    const double4 xa = bead(beads, rp, gq, it), xp = bead(beads, rp, rp.nl_atom + p, it);
                                    const Frame& fp = own(p, k);
                                    dvx += caleng_v(xa.x, xa.y, xa.z, xp.x, xp.y, xp.z, fn, fp) -
                                           caleng_v(xa.x, xa.y, xa.z, xp.x, xp.y, xp.z, fo, fp);
                                }
                            }
                        }
                    }
                    __builtin_amdgcn_wave_barrier();
                    // (2) stage 0, the density factor alone; the network rows of the moves that
                    // pass it, listed first in lane order (slot: a lane's place in that order,
                    // src: the lane whose move fills this lane's place)
                    const bool pass0 = l < nm && (ratio > 1.0 || ratio > rand6);
                    const unsigned long long m0 = __ballot(pass0);
                    const int n0 = __popcll(m0);
                    const int below = __popcll(m0 & ((1ull << l) - 1ull));
                    const int slot = pass0 ? below : n0 + (l - below);
                    const int src = __builtin_amdgcn_ds_permute(slot << 2, l);
                    double so = 0.0, sn = 0.0;
                    if (n0) {
                        net_rows(q, n0, kl, src, true, so, sn);
                        so = __shfl(so, slot);
                        sn = __shfl(sn, slot);
                    }
                    // (3) stage 1, the surrogate factor
                    bool pass1 = false;
                    double dvs = 0.0;
                    if (pass0) {
                        if (net_atoms) so = Vs[q * ns + kl - 1];
                        dvs = ys * (sn - so);
                        const double rd1 = exp(-rp.tau * (dvs + dvx));
                        pass1 = rd1 > 1.0 || rd1 > rand4;
                    }
                    // (4) stage 2: exact network-pair sums of the moves that passed
                    double dve = 0.0, ex_new = 0.0;
                    if (net_atoms) {
                        unsigned long long todo = __ballot(pass1);
                        while (todo) {
                            const int jj = __ffsll((long long)todo) - 1;
                            todo &= todo - 1;
                            const int klj = __shfl(kl, jj);
                            const double e = exact_atoms(q, klj, prop[jj]);
                            if (l == jj) {
                                ex_new = e;
                                dve = e - Ex[q * ns + klj - 1];
                            }
                        }
                    } else if (pass1) {
                        const Frame& fo = own(q, k);
                        for (int p = 0; p < nrot; ++p) {
                            if (p == q) continue;
                            for (int s = 0; s < RR; ++s) {
                                const int it = k * RR + s;
                                const double4 xa = bead(beads, rp, gq, it), xp = bead(beads, rp, rp.nl_atom + p, it);
                                const Frame& fp = own(p, k);
                                dve += caleng_v(xa.x, xa.y, xa.z, xp.x, xp.y, xp.z, prop[l], fp) -
                                       caleng_v(xa.x, xa.y, xa.z, xp.x, xp.y, xp.z, fo, fp);
                            }
                        }
                    }
                    // (5) stage 2 and the update
                    if (pass0) {
                        nnet += (unsigned long long)rpm;
                        if (pass1) {
                            ++npass;
                            const double rd2 = exp(-rp.tau * (dve - dvs));
                            if (rd2 > 1.0 || rd2 > rand5) {
                                ++nacc;
                                const double sint = sqrt(1.0 - cost * cost);
                                RotBead b = cur;
                                b.cth = cost;
                                b.phi = phi;
                                b.chi = chi;
                                b.nx = sint * cos(phi);
                                b.ny = sint * sin(phi);
                                b.nz = cost;
                                rotb[ik] = b;
                                frames[ik] = prop[l];
                                Lk[q * L + kl - 1] = d2;
                                Lk[q * L + kl] = d3;
                                if (net_atoms) {
                                    Vs[q * ns + kl - 1] = sn;
                                    Ex[q * ns + kl - 1] = ex_new;
                                }
                            }
                        }
                    }
                    __builtin_amdgcn_wave_barrier();
                }
                __syncthreads();  // rotor q + 1 sees rotor q's moves at the same slices
            }
            // publish the boundary slice this half moved (k_rotor3d_parts' record)
            if (NP > 1) {
                const int ks = parity == 0 ? k0 : k0 + ns - 1;
                unsigned long long* dst = mypub + (parity == 0 ? 0 : half_rec);
                const unsigned long long tag = (unsigned long long)(base + h + 1u) << 32;
                for (int w = threadIdx.x; w < nwords; w += blockDim.x) {
                    const int i = w >> 1;
                    const double v = i < nrot * FD ? reinterpret_cast<const double*>(&own(i / FD, ks))[i % FD]
                                                   : Lk[(i - nrot * FD) * L + (parity == 0 ? 0 : ns)];
                    const unsigned long long bits = (unsigned long long)__double_as_longlong(v);
                    __hip_atomic_store(dst + w, tag | ((w & 1) ? (bits >> 32) : (bits & 0xffffffffull)), __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
                }
            }
        }
    }
    if (threadIdx.x == 0 && NP > 1)  // every part of the chain ran the same halves
        __hip_atomic_store(myflag, base + 2u * (unsigned)nsweep, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (err) atomicOr(&mut[r].err, 4);
    const unsigned long long sa = block_sum(nacc, red_a);
    const unsigned long long sp = block_sum(npass, red_p);
    const unsigned long long sn_ = block_sum(nnet, red_n);
    if (threadIdx.x == 0) {
        if (s_err) atomicOr(&mut[r].err, s_err);
        const unsigned long long moves = (unsigned long long)nsweep * ns * nrot;
        atomicAdd(&mut[r].total[rtype][2], moves);
        atomicAdd(&mut[r].accep[rtype][2], sa);
        atomicAdd(&mut[r].da[0], moves);
        atomicAdd(&mut[r].da[1], sp);
        atomicAdd(&mut[r].da[2], sa);
        // exact network-pair evaluations of the moves that reached stage 2
        atomicAdd(&mut[r].pair_evals[2], (net_atoms ? 1ull : 2ull) * sp * (unsigned long long)(net_atoms ? items_x : RR * (nrot - 1)));
        atomicAdd(&mut[r].surrogate_evals, sn_);
    }
}

// Training rows of the surrogate from the current configuration of replica r, for the pairs the
// delayed-acceptance move gives the network.  With atoms: each (rotor q, rotor slice k,
// translational slice it of k, atom a) gives the body-frame features of the pair and its exact
// energy (vcord on the 3D table in FP64, PotRotE3D's atom term); row order q, k, it, a.  Without
// atoms: each (rotor q, slice k, translational slice it, other rotor p) gives the rotor-pair
// features and the TIP4P energy of caleng (PotRotE3D's rotor term).
__global__ void k_surrogate_collect(const Rep* __restrict__ reps, const double4* __restrict__ beads,
                                    const Frame* __restrict__ frames, Tables3 T3, int r, float* __restrict__ X,
                                    float* __restrict__ y, int nrows) {
    const Rep& rp = reps[r];
    const int RR = rp.RotRatio, nat = rp.nl_atom, nrot = rp.nrot, Prot = rp.Prot;
    const int npart = nat > 0 ? nat : nrot - 1;  // partners per translational slice
    const Lut3D<double> l3 = lut3d<double>(T3, rp.pot3d_id);
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nrows; i += gridDim.x * blockDim.x) {
        const int per_q = Prot * RR * npart;
        const int q = i / per_q, rem0 = i - q * per_q;
        const int k = rem0 / (RR * npart), rem = rem0 - k * (RR * npart);
        const int it = k * RR + rem / npart, jp = rem % npart;
        const int gq = nat + q;
        const double4 xa = bead(beads, rp, gq, it);
        const Frame& f = frames[rp.rot_off + (long long)q * Prot + k];
        float fe[MLP_IN];
        double e;
        if (nat > 0) {
            const double4 xj = bead(beads, rp, jp, it);
            mlp_pair_features(f.m, f.zn, xa.x, xa.y, xa.z, xj.x, xj.y, xj.z, fe);
            e = (double)vcord_v<double>(l3, f, xa.x, xa.y, xa.z, xj.x, xj.y, xj.z);
        } else {
            const int p = jp + (jp >= q);
            const double4 xp = bead(beads, rp, nat + p, it);
            const Frame& fp = frames[rp.rot_off + (long long)p * Prot + k];
            mlp_rotor_pair_features(f, fp, xa.x, xa.y, xa.z, xp.x, xp.y, xp.z, fe);
            e = caleng_v(xa.x, xa.y, xa.z, xp.x, xp.y, xp.z, f, fp);
        }
        for (int c = 0; c < MLP_IN; ++c) X[(long long)i * MLP_IN + c] = fe[c];
        y[i] = (float)e;
    }
}

// ------------------------------------------------------------------------- end of pass
__global__ void k_pass_end(Mut* mut, int n) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < n) mut[r].pass += 1;
}

// ------------------------------------------------------------------------- estimators
// One workgroup per (replica, tile of est_ts slices).  All atoms' beads on the tile's slices and
// one slice past it are staged in LDS once; lanes then take every pair at every slice of the
// tile, the pairs enumerated as (a, a + d mod N) for d = 1 .. N/2 so that a tile has the same
// work at any N.  g(r) is counted in LDS and added to the replica's bins once per tile.
// part[row_off + tile] = (pair potential of the tile, its kinetic links / (4 beta lambda)).
// The potential and the bins follow GetPotEnergy_Densities (mc_estim.cc:500-687), the kinetic
// term GetKinEnergy (mc_estim.cc:876-937).
template <typename R, int FAM, bool MI>
__global__ __launch_bounds__(256) void k_est_tiles(const Rep* __restrict__ reps, const double4* __restrict__ beads,
                                                   const RotBead* __restrict__ rotb, const Frame* __restrict__ frames,
                                                   Tables T, Tables3 T3, const int2* __restrict__ tiles,
                                                   unsigned* __restrict__ hist, double2* __restrict__ part) {
    extern __shared__ double smem[];
    __shared__ double red[16];
    __shared__ unsigned g1s[BINSR];
    const int2 tk = tiles[blockIdx.x];
    const Rep& rp = reps[tk.x];
    const int N = rp.natoms, P = rp.P, TS = rp.est_ts, lts = __ffs(TS) - 1;
    const int s0 = tk.y * TS, nts = min(TS, P - s0), W = nts + 1;
    double* lx = smem;
    double* ly = lx + N * W;
    double* lz = ly + N * W;
    double* ax = lz + N * W;  // rotor axis at the tile's slices
    double* ay = ax + TS;
    double* az = ay + TS;
    int* kind = reinterpret_cast<int*>(az + TS);  // type | molecule kind << 8
    for (int i = threadIdx.x; i < N * W; i += blockDim.x) {
        const int a = i / W, w = i - a * W;
        const int it = s0 + w < P ? s0 + w : s0 + w - P;
        const double4 b = bead(beads, rp, a, it);
        lx[i] = b.x;
        ly[i] = b.y;
        lz[i] = b.z;
    }
    for (int w = threadIdx.x; w < nts; w += blockDim.x) rotor_axis(rotb, rp, s0 + w, ax[w], ay[w], az[w]);
    for (int a = threadIdx.x; a < N; a += blockDim.x) {
        const int t = type_of(rp, a);
        kind[a] = t | (rp.tmol[t] << 8);
    }
    unsigned* g1 = hist ? hist + rp.hist_off : nullptr;
    unsigned* g2h = g1 ? g1 + BINSR : nullptr;
    unsigned* g3 = (g1 && rp.h3_off >= 0) ? g1 + rp.h3_off : nullptr;
    for (int i = threadIdx.x; i < BINSR; i += blockDim.x) g1s[i] = 0u;
    __syncthreads();
    const double* box = MI && rp.minimage ? rp.box : nullptr;  // MINIMAGE (not with a rotating linear rotor)
    double kin = 0.0;
    const Lut1D<R> l1 = lut1d<R>(T, rp.pot1d_id);
    const Lut2D<R> l2 = lut2d<R>(T, rp.pot2d_id);
    const double drad = MAX_RADIUS / (double)BINSR;
    const double dth = 3.14159265358979323846 / (double)(BINST - 1);
    const double dch = 2.0 * 3.14159265358979323846 / (double)(BINSC - 1);
    const int npairs = N * (N - 1) / 2;
    const float invN = 1.0f / (float)N;
    double pot = 0.0;
    // kinetic links (a, s) - (a, s + 1), then the pairs; the minimum image as its own instance
    // of the loops, so the plain one is unchanged
    auto tile_sums = [&](auto mi_c) {
        const double* bx = decltype(mi_c)::value ? box : nullptr;
        for (int i = threadIdx.x; i < N * nts; i += blockDim.x) {
            const int a = i / nts, w = i - a * nts;
            const int j = a * W + w;
            const double dx = nearest_image(lx[j] - lx[j + 1], bx, 0), dy = nearest_image(ly[j] - ly[j + 1], bx, 1),
                         dz = nearest_image(lz[j] - lz[j + 1], bx, 2);
            kin += (dx * dx + dy * dy + dz * dz) / (4.0 * rp.beta * rp.lambda[kind[a] & 255]);
        }
        for (int i = threadIdx.x; i < (npairs << lts); i += blockDim.x) {
            const int w = i & (TS - 1);
            if (w >= nts) continue;
            const int p = i >> lts;
            // p = N (d - 1) + a; the float quotient is exact while N^2 < 2^23
            const int d1 = N < 2896 ? (int)(((float)p + 0.5f) * invN) : p / N;
            const int a0 = p - d1 * N;
            int a1 = a0 + d1 + 1;
            if (a1 >= N) a1 -= N;
            const int k0 = kind[a0], k1 = kind[a1];
            const int t0 = k0 & 255, t1 = k1 & 255, m0 = k0 >> 8, m1 = k1 >> 8;
            const int j0 = a0 * W + w, j1 = a1 * W + w;
            const double ex = nearest_image(lx[j0] - lx[j1], bx, 0), ey = nearest_image(ly[j0] - ly[j1], bx, 1),
                         ez = nearest_image(lz[j0] - lz[j1], bx, 2);
            const R dx = (R)ex, dy = (R)ey, dz = (R)ez;
            const R rr = sqrt(dx * dx + dy * dy + dz * dz);
            if constexpr (FAM == 1) {
                if (m0 == 2 || m1 == 2) {  // mc_estim.cc:582-672: vcord or caleng
                    const Lut3D<R> l3 = lut3d<R>(T3, rp.pot3d_id);
                    const int it = s0 + w;
                    const Frame& f0 = frame_at(frames, rp, m0 == 2 ? a0 : rp.nl_atom, it);
                    const Frame& f1 = frame_at(frames, rp, m1 == 2 ? a1 : rp.nl_atom, it);
                    pot += pair_nl<R>(l1, l1.k, l3, m0 == 2, m1 == 2, lx[j0], ly[j0], lz[j0], lx[j1], ly[j1], lz[j1], f0, f1);
                    if (g3 && m0 != m1) {  // bin_3Ddensity of vcord's FP64 geometry (mc_estim.cc:637,1252-1273)
                        double rad, the, chi;
                        if (m0 == 2) vcord_geometry(f0, lx[j0], ly[j0], lz[j0], lx[j1], ly[j1], lz[j1], rad, the, chi);
                        else vcord_geometry(f1, lx[j1], ly[j1], lz[j1], lx[j0], ly[j0], lz[j0], rad, the, chi);
                        const int br = (int)floor(rad / drad);
                        if (br < BINSR && br >= 0) {
                            const int bt = (int)floor(the / dth);
                            if (bt < BINST && bt >= 0) {
                                const int bc = (int)floor(chi / dch);
                                if (bc < BINSC && bc >= 0) atomicAdd(&g3[(br * BINST + bt) * BINSC + bc], 1u);
                            }
                        }
                    }
                    continue;
                }
            }
            if (m0 == 1 || m1 == 1) {  // axis to the partner: from the rotor to the atom
                const double nx = ax[w], ny = ay[w], nz = az[w];
                R cost = ((R)nx * dx + (R)ny * dy + (R)nz * dz) / rr;
                if (m0 == 1) cost = -cost;
                pot += (double)lpot2d<R>(l2, rr, cost);
                if (g2h) {  // bins from FP64 geometry, as bin_2Ddensity gets it (mc_estim.cc:1233-1250)
                    const double rd = sqrt(ex * ex + ey * ey + ez * ez);
                    double cd = (nx * ex + ny * ey + nz * ez) / rd;
                    if (m0 == 1) cd = -cd;
                    const int br = (int)floor(rd / drad);
                    if (br < BINSR && br >= 0) {
                        const int bt = (int)floor(acos(cd) / dth);
                        if (bt < BINST && bt >= 0) atomicAdd(&g2h[br * BINST + bt], 1u);
                    }
                }
            } else if (t0 == t1 && m0 == 0) {
                pot += (double)spot1d<R>(l1, l1.k, rr);
                if (g1) {  // bin_1Ddensity (mc_estim.cc:1275-1286) on the FP64 distance
                    const int br = (int)floor(sqrt(ex * ex + ey * ey + ez * ez) / drad);
                    if (br < BINSR && br >= 0) atomicAdd(&g1s[br], 1u);
                }
            }
        }
    };
    if constexpr (MI) {
        if (box) tile_sums(std::integral_constant<bool, true>{});
        else tile_sums(std::integral_constant<bool, false>{});
    } else {
        tile_sums(std::integral_constant<bool, false>{});
    }
    const double sp = block_sum(pot, red);
    const double sk = block_sum(kin, red);
    if (g1)
        for (int i = threadIdx.x; i < BINSR; i += blockDim.x)
            if (g1s[i]) atomicAdd(&g1[i], g1s[i]);
    if (threadIdx.x == 0) part[rp.row_off + tk.y] = make_double2(sp, sk);
}

// out (if not null) receives 8 doubles per replica; otherwise the sample ring.
__global__ __launch_bounds__(256) void k_est_final(const Rep* __restrict__ reps, Mut* __restrict__ mut,
                                                   const RotBead* __restrict__ rotb, const Frame* __restrict__ frames,
                                                   Tables T, Tables3 T3,
                                                   const int* __restrict__ replist, const double2* __restrict__ rows,
                                                   double* __restrict__ samples, int sample_cap, double* __restrict__ out,
                                                   unsigned* __restrict__ hist) {
    __shared__ double red[16];
    const int r = replist[blockIdx.x];
    const Rep& rp = reps[r];
    const int N = rp.natoms, P = rp.P;
    // rotational energy, GetRotEnergy mc_estim.cc:939-987
    double srot = 0.0, esq = 0.0, term = 0.0;
    if (rp.rotkind == 1 && rp.rot_atom >= 0 && rp.rotdens >= 0) {
        const long long rbase = rp.rot_off + (long long)(rp.rot_atom - rp.tbegin[rp.imtype]) * rp.Prot;
        double a1 = 0, a2 = 0, a3 = 0;
        for (int k = threadIdx.x; k < rp.Prot; k += blockDim.x) {
            RotBead b0 = rotb[rbase + k], b1 = rotb[rbase + (k + 1) % rp.Prot];
            double p0 = b0.nx * b1.nx + b0.ny * b1.ny + b0.nz * b1.nz;
            double rd = rot_spline(T, rp.rotdens, p0, 0);
            double de = rot_spline(T, rp.rotdens, p0, 1);
            if (fabs(rd) > RZERO) a1 += de / rd;
            a2 += (de / rd) * (de / rd);
            a3 += rot_spline(T, rp.rotdens, p0, 2) / rd;
        }
        srot = block_sum(a1, red);
        term = block_sum(a2, red);
        esq = block_sum(a3, red);
    }
    if (rp.rotkind == 2 && rp.rot3d_id >= 0) {  // GetRotE3D mc_estim.cc:989-1096, per rotor
        const double4* tab = T3.r3 + (long long)rp.rot3d_id * 3 * ROT3D_SIZE;
        // relative Euler angle bins (mc_estim.cc:1046-1058): theta by pi/49, phi and chi by 2 pi/99
        unsigned* hr = (hist && rp.hrel_off >= 0) ? hist + rp.hist_off + rp.hrel_off : nullptr;
        const double dth = PI_D / (double)(BINST - 1), dch = 2.0 * PI_D / (double)(BINSC - 1);
        for (int q = 0; q < rp.nrot; ++q) {
            const long long fb = rp.rot_off + (long long)q * rp.Prot;
            double a1 = 0, a2 = 0, a3 = 0;
            for (int k = threadIdx.x; k < rp.Prot; k += blockDim.x) {
                double rho, er, es, rel[3];
                rotden_frames(frames[fb + k], frames[fb + (k + 1) % rp.Prot], tab, rho, er, es, hr ? rel : nullptr);
                if (hr) {
                    const int bt = (int)floor(rel[1] / dth);
                    if (bt < BINST && bt >= 0) atomicAdd(&hr[bt], 1u);
                    const int bp = (int)floor(rel[0] / dch);
                    if (bp < BINSC && bp >= 0) atomicAdd(&hr[BINST + bp], 1u);
                    const int bc = (int)floor(rel[2] / dch);
                    if (bc < BINSC && bc >= 0) atomicAdd(&hr[BINST + BINSC + bc], 1u);
                }
                a1 += er;
                a2 += es;
                a3 += er * er;
            }
            const double n = (double)rp.Prot;
            srot += block_sum(a1, red) / n;
            esq += block_sum(a2, red) / (n * n);
            term += block_sum(a3, red) / (n * n);
        }
    }
    if (threadIdx.x == 0) {
        double r2avr = 0.0, spot = 0.0;
        const int ntiles = (P + rp.est_ts - 1) / rp.est_ts;
        for (int a = 0; a < ntiles; ++a) {
            double2 v = rows[rp.row_off + a];
            spot += v.x;
            r2avr += v.y;
        }
        spot /= (double)P;
        double skin = (double)P * rp.T * (0.5 * (double)(3 * N) - r2avr);
        double n3 = (double)(3 * N);
        double a = n3 * 0.5 / rp.tau - skin - spot - srot;
        double sCv = -0.5 * n3 / (rp.beta * rp.tau) - a * a + (2.0 / rp.beta) * (0.5 * n3 / rp.tau - skin) + term - esq;
        double b = n3 * 0.5 / rp.tau - skin;
        double sCvt = -0.5 * n3 / (rp.beta * rp.tau) - b * b + (2.0 / rp.beta) * (0.5 * n3 / rp.tau - skin);
        double sCvr = -srot * srot + term - esq;
        double rec[8] = {skin, spot, srot, esq, sCv, sCvt, sCvr, term};
        double* dst = nullptr;
        if (out) dst = out + 8 * r;
        else if (samples) {
            unsigned ns = mut[r].nsamples;
            if ((int)ns < sample_cap) dst = samples + rp.sample_off + 8ll * ns;
            mut[r].nsamples = ns + 1;
        }
        if (dst)
            for (int q = 0; q < 8; ++q) dst[q] = rec[q];
        atomicAdd(&mut[r].pair_evals[3], (unsigned long long)N * (N - 1) / 2 * P);
    }
}

// ------------------------------------------------------------------------- PotEnergy(atom,pos,it)
template <typename R>
__global__ void k_pot_slice(const Rep* __restrict__ reps, const double4* __restrict__ beads,
                            const RotBead* __restrict__ rotb, const Frame* __restrict__ frames, Tables T, Tables3 T3, int nq,
                            const int* __restrict__ qrep,
                            const int* __restrict__ qatom, const int* __restrict__ qit, const double* __restrict__ qpos,
                            double* __restrict__ out) {
    int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nq) return;
    const Rep& rp = reps[qrep[q]];
    int a = qatom[q], it = qit[q];
    const Lut1D<R> l1 = lut1d<R>(T, rp.pot1d_id);
    const Lut2D<R> l2 = lut2d<R>(T, rp.pot2d_id);
    double nx, ny, nz;
    rotor_axis(rotb, rp, it, nx, ny, nz);
    const double* box = rp.minimage ? rp.box : nullptr;
    double s = 0.0;
    for (int j = 0; j < rp.natoms; ++j) {
        if (j == a) continue;
        double4 xj = bead(beads, rp, j, it);
        if (rp.fam == 1) {
            const bool ra = rp.nl_atom >= 0 && a >= rp.nl_atom, rj = rp.nl_atom >= 0 && j >= rp.nl_atom;
            const Lut3D<R> l3 = lut3d<R>(T3, rp.pot3d_id);
            s += pair_nl<R>(l1, l1.k, l3, ra, rj, qpos[3 * q], qpos[3 * q + 1], qpos[3 * q + 2], xj.x, xj.y, xj.z,
                            frame_at(frames, rp, ra ? a : rp.nl_atom, it), frame_at(frames, rp, rj ? j : rp.nl_atom, it),
                            box);
        } else {
            s += (double)pair_v<R>(l1, l1.k, l2, a == rp.lin_atom, j == rp.lin_atom, qpos[3 * q], qpos[3 * q + 1],
                                   qpos[3 * q + 2], xj.x, xj.y, xj.z, nx, ny, nz, box);
        }
    }
    out[q] = s;
}

}  // namespace pimc
