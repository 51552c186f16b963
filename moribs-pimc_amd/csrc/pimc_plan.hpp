// pimc_plan.hpp -- host-side checks and launch plans that need no device: the validation of a
// replica description, and the split of linear-rotor chains into workgroups.  Plain C++ so
// that the library can run them without a GPU (pimc_check_replica, pimc_plan_rotor_parts).
#pragma once
#include "../../include/moribs_pimc.h"
#include <algorithm>
#include <functional>
#include <string>
#include <vector>

namespace pimc {

// A rotor chain's slice range [k0, k1) run by one workgroup: part w of the W parts of replica
// r, whose exchange flags start at fbase.  Also the slice ranges of the split centroid move.
struct RotPart {
    int r, k0, k1, w, W, fbase, pad0, pad1;
};

// The input checks of MCInit and IOReadParams that apply to the sampled path (mc_setup.cc:384-
// 398, mc_input.cc:210-330), with the reference's messages, plus the limits of this library.
// An empty string means the description is accepted.
inline std::string check_replica(const pimc_replica_desc& d) {
    if (d.ntypes < 1 || d.ntypes > PIMC_MAX_TYPES) return "ntypes out of range";
    if (d.P < 2 || d.P > 65535) return "P out of range";
    int natoms = 0, nmolt = 0, natt = 0;
    bool rotating = false;
    for (int t = 0; t < d.ntypes; ++t) {
        if (d.numb[t] < 1) return "every type needs at least one atom";
        if (d.molecule[t] < 0 || d.molecule[t] > 2) return "unknown molecule kind";
        if (d.molecule[t] == 0 && ++natt > 1) return "No more then one atom/molecule type: densities and potential energy";
        if (d.molecule[t] != 0 && t != d.ntypes - 1) return "Molecules should follow atoms in input file";
        if (d.molecule[t] == 2) {
            ++nmolt;
            if (d.numb[t] > 2) return "GetRotE3D supports at most two non-linear rotors";
            if (d.ntypes > 1 && d.pot3d < 0) return "non-linear rotor without a 3D potential table";
            if (d.Prot > 0 && d.rot3d < 0) return "rotating non-linear rotor without rho/E/E^2 tables";
        }
        if (d.molecule[t] == 1) {
            ++nmolt;
            if (d.numb[t] != 1) return "No more than one linear dopant molecule so far";
        }
        if (d.molecule[t] != 0 && d.Prot > 0) rotating = true;
        if (d.levels[t] < 1 || d.levels[t] > 8) return "bisection levels must be 1..8";
        if ((1 << d.levels[t]) >= d.P) return "Segment size is larger then a number of time slices";
        natoms += d.numb[t];
    }
    if (natoms > 4096) return "too many atoms";
    if (nmolt > 1) return "one molecule type per replica";
    if (d.Prot > 0 && d.P % d.Prot) return "NumbTimes is not proportional to NumbRotTimes";
    if (d.Prot == 1) return "rotational sampling needs at least two rotational slices";
    // the checkerboard sweeps (mc_piqmc.cc:490-512, :728-760) move even slices together, then odd
    // ones: with an odd count slices 0 and Prot-1 are both even and neighbours
    if (rotating && d.Prot % 2) return "the checkerboard rotor sweep needs an even number of rotor slices";
    return std::string();
}

// A linear-rotor chain to split: replica r, Prot rotor slices, RotRatio translational slices per
// rotor slice, natoms atoms (the rotor's partners are the natoms - 1 others).
struct RotorChain {
    int r, Prot, RotRatio, natoms;
};

// LDS of a part of ns slices in k_rotor's layout: seven doubles per local slice (two
// neighbours' included) padded to four, then the partner geometry, one vec4 per translational
// slice and partner.
inline size_t rotor_part_lds(int ns, long long per_slice) {
    return (size_t)((7 * (ns + 2) + 3) & ~3) * sizeof(double) + (size_t)((long long)ns * per_slice);
}

constexpr size_t ROTOR_LDS_MAX = 160 * 1024;

// Splits each chain into parts of an even number of slices (the last part may be shorter but is
// even too, since Prot is even), sized so a part's LDS stays within lds_kb and, when
// part_slices_max > 0, at most that many slices.  Parts of a chain wait on one another and are
// all placed on one XCD, so a chain may have at most as many parts as one XCD holds resident at
// the launch's LDS size: cus_per_xcd * blocks_per_cu(lds).  A chain above that is planned again
// with fewer, larger parts; when even those do not fit (a part beyond 160 KB, or one that cannot
// be resident with its neighbours) the plan fails with a message.  parts come out grouped by
// chain in the order given, fbase the index of the chain's first part; lds is the launch's LDS.
inline int plan_linear_rotor(const std::vector<RotorChain>& chains, size_t geo_bytes, int lds_kb, int part_slices_max,
                             int cus_per_xcd, const std::function<int(size_t)>& blocks_per_cu,
                             std::vector<RotPart>& parts, size_t& lds, std::string& err) {
    const long long budget = (long long)lds_kb * 1024;
    const size_t n = chains.size();
    std::vector<int> ns(n), W(n);
    std::vector<long long> per_slice(n);
    auto balance = [](int prot, int w, int& s, int& Wout) {  // even slices per part, W parts
        s = ((prot + w - 1) / w + 1) & ~1;
        if (s >= prot) s = prot;
        Wout = (prot + s - 1) / s;
    };
    for (size_t i = 0; i < n; ++i) {
        const RotorChain& c = chains[i];
        per_slice[i] = (long long)c.RotRatio * (c.natoms - 1) * (long long)geo_bytes;
        int s = (int)std::max<long long>(2, (budget - 16 * (long long)sizeof(double)) /
                                                (7 * (long long)sizeof(double) + per_slice[i])) & ~1;
        if (part_slices_max > 0) s = std::max(2, std::min(s, part_slices_max & ~1));
        if (s >= c.Prot) s = c.Prot;
        balance(c.Prot, (c.Prot + s - 1) / s, ns[i], W[i]);
    }
    for (int iter = 0;; ++iter) {
        lds = 0;
        for (size_t i = 0; i < n; ++i) lds = std::max(lds, rotor_part_lds(ns[i], per_slice[i]));
        if (lds > ROTOR_LDS_MAX) {
            err = "rotor chain exceeds LDS: a part of the linear-rotor sweep needs " + std::to_string(lds) + " bytes";
            return -1;
        }
        if (n == 0) break;
        const long long cap = (long long)cus_per_xcd * std::max(0, blocks_per_cu(lds));
        bool changed = false;
        for (size_t i = 0; i < n; ++i) {
            if (W[i] <= cap) continue;
            if (cap < 1) {
                err = "linear-rotor sweep: parts of " + std::to_string(lds) + " bytes of LDS cannot be resident";
                return -1;
            }
            int s, w;
            balance(chains[i].Prot, (int)std::min<long long>(cap, W[i] - 1), s, w);
            ns[i] = s;
            W[i] = w;
            changed = true;
        }
        if (!changed) break;
        if (iter > 64) {
            err = "linear-rotor sweep: no plan of parts fits one XCD";
            return -1;
        }
    }
    parts.clear();
    for (size_t i = 0; i < n; ++i) {
        const int fbase = (int)parts.size();
        for (int w = 0; w < W[i]; ++w) {
            RotPart q{};
            q.r = chains[i].r;
            q.k0 = w * ns[i];
            q.k1 = std::min(chains[i].Prot, (w + 1) * ns[i]);
            q.w = w;
            q.W = W[i];
            q.fbase = fbase;
            parts.push_back(q);
        }
    }
    return 0;
}

// LDS of a part of ns slices of k_rotor3d_parts with nrot rotors: the link densities and the
// cached rotor-atom sums (nrot x (2 ns + 1) doubles, padded to two), the neighbours' boundary
// frames (2 nrot), then per lane group of G lanes the frames of every rotor at its slice plus
// the proposal's.  frame_bytes is sizeof(Frame); G follows the kernel's rule for 256 lanes.
inline size_t rotor3d_part_lds(int ns, int nrot, size_t frame_bytes) {
    const int maxmoves = (ns + 1) / 2;
    int G = 1;
    while (G < 64 && G * 2 * maxmoves <= 256) G *= 2;
    return (size_t)((nrot * (2 * ns + 1) + 1) & ~1) * sizeof(double) + (size_t)2 * nrot * frame_bytes +
           (size_t)(256 / G) * (nrot + 1) * frame_bytes;
}

// LDS of a part of ns slices of k_rotor3d_da (delayed acceptance, 256 lanes = 4 waves): the
// packed network (mlp_bytes), the link densities and the two cached sums per own slice
// (nrot x (3 ns + 1) doubles, padded to two), the neighbours' boundary frames and DA_MV proposal
// frames per wave, then 32 row outputs per wave.
constexpr int DA_MV = 64;  // moves a wave takes at once in the delayed-acceptance sweep: one per lane
inline size_t rotor3d_da_part_lds(int ns, int nrot, size_t frame_bytes, size_t mlp_bytes) {
    return mlp_bytes + (size_t)((nrot * (ns + 1) + 2 * nrot * ns + 1) & ~1) * sizeof(double) +
           (size_t)(2 * nrot + 4 * DA_MV) * frame_bytes + (size_t)4 * 32 * sizeof(float);
}

// Split of the non-linear rotor chains: one count W for the launch, a power of two up to 64, the
// largest whose parts hold at least 16 slices of the longest chain and fit LDS (lds_for(slices)),
// such that the parts of all chains are resident at once (ncu * blocks_per_cu) and the W parts
// of one chain, which wait on one another on one XCD, fit that XCD (cus_per_xcd *
// blocks_per_cu).  Returns 1 when whole chains are better left to one workgroup each.
inline int plan_rotor3d_split(int protmax, int nchains, int ncu, int cus_per_xcd,
                              const std::function<size_t(int)>& lds_for,
                              const std::function<int(size_t)>& blocks_per_cu) {
    int W = 1;
    for (int w = 2; w <= 64; w *= 2) {
        const int ns = ((protmax + w - 1) / w + 1) & ~1;
        const size_t lds = lds_for(ns);
        if (ns < 16 || lds > ROTOR_LDS_MAX) break;
        const long long occ = std::max(0, blocks_per_cu(lds));
        if ((long long)w * nchains > occ * ncu || (long long)w > occ * cus_per_xcd) break;
        W = w;
    }
    return W;
}

}  // namespace pimc
