// Fused docking search kernels for gfx950: pose energy + analytic gradient, Vina-style BFGS, Monte Carlo chains.
// Implementation header: every translation unit that instantiates the kernels for one torsion bucket (dock_km*.hip)
// includes it; dock.hip holds the launch settings and the dispatch over buckets.
//
// Same algorithm as the reference's fused search kernel (src/adock/mckernel.py:21-447): torsion-tree forward
// kinematics, trilinear grid lookup with the quadratic out-of-box penalty, intra-ligand pair terms, gradient in pose
// space (translation, world-frame torque, torsions), BFGS with Armijo backtracking, MC mutation + Metropolis.
//
// MI355X mapping (one wave64 = one chain; W chains per workgroup, independent):
//   * lane a holds heavy atom a, and with APL = 2 atoms per lane also atom a + 64 (ligands of 65-128 heavy atoms):
//     reference coordinates, radius, type flags, grid-slot pointer and the bit mask of torsions that move the atom
//     live in registers for the whole chain;
//   * lane 7+k computes sin/cos of torsion k once per evaluation and stores them in the torsion's LDS record next to
//     its reference axis; forward kinematics and the torque projection read the records as LDS broadcasts;
//   * intra-ligand pairs are strided over lanes, two rounds of 64 per iteration, branch-free; each pair writes its
//     gradient to its own LDS slot and every atom sums the slots of its pairs from a per-atom list built once per
//     ligand on the host (no atomics, no cross-lane traffic in the pair loop);
//   * every reduction (energy, 3 forces, 3 torques, one per torsion, BFGS dot products) is a DPP row scan + row
//     broadcast, no LDS round trip;
//   * BFGS vectors are lane-distributed (lane i = component i) and lane i keeps row i of the inverse Hessian in
//     registers, so H.g and the rank-2 update are register FMAs against readlane-broadcast operands.
//   Gradient lane layout: 0-2 translation, 3-5 torque, 6 always zero, 7+k torsion k (state layout t(3) q(4) theta:
//   torsion k sits on the same lane in the state and in the gradient; the zero lane is inert in BFGS). So one wave
//   holds at most 57 torsions; the buckets go to KM = 56.
#pragma once
#include <atomic>
#include <map>
#include <mutex>
#include <tuple>
#include <type_traits>

#include "common.hpp"
#include "kernels.hpp"

namespace adock {

// launch settings shared by every bucket's translation unit (defined in dock.hip)
extern std::atomic<int> g_waves_per_block;     // 0: adaptive
extern std::atomic<int> g_line_search_width;   // 0: adaptive
extern std::atomic<int> g_exact_staging;       // 0: read from memory, 1: staged, 2: staged with per-atom masks

namespace {

// Identity that makes the compiler hold v in a VGPR (an empty asm with a VGPR operand; no instruction is emitted).
// The latency-bound dense kernels run one wave per SIMD with VGPRs to spare while their uniform values overflow the
// SGPRs: the grid and scoring constants the pair loop and the lookup read on every evaluation are pinned in VGPRs
// there instead of being spilled to VGPR lanes and reloaded with v_readlane (single complex: 121 -> 93 SGPR spills,
// search kernel 117.5 -> 115.4 ms). The values are unchanged.
DEVI float in_vgpr(float v) {
    asm volatile("" : "+v"(v));
    return v;
}

// All-ones where bit u of t is set, else zero, hidden from the optimiser (an empty asm with a VGPR operand) so that
// bit_select below stays bit operations and is not turned back into a select.
DEVI unsigned bit_mask(unsigned t, int u) {
    unsigned m = (unsigned)((int)(t << (31 - u)) >> 31);
    asm("" : "+v"(m));
    return m;
}

// m ? a : b for a lane-varying m of 0 / all-ones bits, as one bit-field insert. Written as a float select, the
// compiler may turn it into a divergent branch around the loads and arithmetic feeding `a` (it did in the
// forward-kinematics loop: +8 branches and +24 waits per evaluation); bit operations stay branch-free.
DEVI float bit_select(unsigned m, float a, float b) {
    return __uint_as_float((__float_as_uint(a) & m) | (__float_as_uint(b) & ~m));
}

// Pair words: i | j << SH | (hydrophobic pair) << CLS | (h-bond pair) << CLS + 1. One atom per lane (64-atom
// ligands): 16-bit words, 7-bit indices (64 = the dummy atom). Two atoms per lane: 32-bit words, 8-bit indices (128 =
// the dummy atom).
template <int APL> struct PairWord;
template <> struct PairWord<1> {
    using T = unsigned short;
    static constexpr int SH = 7, CLS = 14;
    static constexpr unsigned M = 127u;
};
template <> struct PairWord<2> {
    using T = unsigned;
    static constexpr int SH = 8, CLS = 16;
    static constexpr unsigned M = 255u;
};
template <int APL> DEVI unsigned pw_first(unsigned pw) { return pw & PairWord<APL>::M; }
// a bit-field extract held opaque (an empty asm), so the scaled LDS address takes one shift-add after it: left to
// itself the compiler folds extract and scaling into a shift, a mask and an add
template <int APL> DEVI unsigned pw_second(unsigned pw) {
    unsigned j = __builtin_amdgcn_ubfe(pw, PairWord<APL>::SH, PairWord<APL>::SH);
    asm("" : "+v"(j));
    return j;
}
template <int APL> DEVI int pw_class(unsigned pw) { return (int)((pw >> PairWord<APL>::CLS) & 3u); }
template <int APL> constexpr unsigned pw_hyd() { return 1u << PairWord<APL>::CLS; }
template <int APL> constexpr unsigned pw_hb() { return 2u << PairWord<APL>::CLS; }
template <int APL> constexpr unsigned pw_pad() { return 64u * APL; }   // (dummy atom, atom 0)

// bit mask of the torsions moving an atom: 32 bits up to KM = 32, 64 above
template <int KM> using TMask = std::conditional_t<(KM > 32), unsigned long long, unsigned>;

// Per-wave LDS layout, in bytes (AMX = 64 APL atom slots, P128 = P_MAX rounded up to a multiple of 128, KM + 1
// torsion records):
//   X4  [AMX + 1]   float4  world xyz + radius of each atom; entry AMX is a dummy atom far beyond the cutoff (padding
//                           pairs point at it, so they contribute exactly zero without a validity flag)
//   TR0, TR1, TR2 [KM + 1] float4  torsion k's rotation about its reference axis as an affine map (row i: R_i0,
//                                  R_i1, R_i2, b_i; p -> R p + b), rebuilt every evaluation (identity past nt)
//   TU  [KM + 1]    float4  current (world) axis unit vector u_k
//   TM  [KM + 1]    float4  current axis moment m_k = p_k x u_k (torque about the axis: u.(X x F) + m.F)
//   RD  [128]       float   the evaluation's 7 + KM wave sums (wave_sum_to_lanes); between evaluations, BFGS's
//                           lane-distributed vectors for broadcast reads
//   FS  [P128 + 4]  float   slot of pair q: (dE/dr) / r, so dE/dX of either atom is FS[q] (X_self - X_partner);
//                           slot P128 stays zero
//   flg [AMX]       int     atom type flags
//   pl  [P128 + 8]  word    pair words (PairWord<APL>), strided over lanes (pair q on lane q % 64 in round q / 64);
//                           padding pairs (and entry P128) are (dummy, 0)
// The per-atom pair-slot lists stay in memory (the host-built packed ligand, read through the caches; a complex's
// chains share them): atom a's entries are fs | px << 16, the byte offsets of the pair's slot in FS and of its partner
// atom in X4, so an entry costs two address additions (each list padded to 8 entries; the row ends with 8 padding
// entries: the zero slot and the dummy atom). Each lane walks one segment (start and count in its lane and slot): its
// atom's list, or with one atom per lane, for the longest lists, its head, while a lane past the atoms (a helper) walks
// the tail and hands its partial sum to the atom's lane (ds_bpermute). The longest walk of the wave sets the trip count.
// Dense variant (the latency-bound speculative kernel, one atom per lane): FS is replaced by FD [64 DEG_MAX + 1]
// float4, row k column a ^ k = atom a's k-th pair: its separation X_a - X_partner (sign applied by the writer) and
// slot (dE/dr)/r (unused slots stay zero, the last one takes the padding pairs' writes), and the lists by pdst [P128]
// u32 (destination slot of pair q's first atom | second atom << 16). The per-atom sum reads one permuted row per entry (no list indirection, no bank conflicts);
// the XOR skew spreads the writes of one atom's pairs over the banks.
__host__ __device__ inline int pad128(int P_MAX) { return (P_MAX + 127) & ~127; }
// row length of one ligand's slot lists in the packed arrays (AM atoms, each list padded to a multiple of 8 entries,
// then one group of 8 padding entries)
__host__ __device__ inline int inc_row(int P_MAX, int AM) { return 2 * pad128(P_MAX) + 8 * AM + 8; }
__host__ __device__ inline int wave_lds_bytes(int KM, int P_MAX, int APL = 1) {
    const int P = pad128(P_MAX);
    const int pw = APL == 1 ? 2 : 4;
    return 16 * (64 * APL + 1) + 5 * 16 * (KM + 1) + 4 * 128 + 4 * (P + 4) + 4 * 64 * APL + pw * (P + 8);
}
__host__ __device__ inline int pad512(int P_MAX) { return (P_MAX + 511) & ~511; }
__host__ __device__ inline int wave_lds_bytes_dense(int KM, int P_MAX, int DEG_MAX) {
    const int P = pad512(P_MAX);   // pair words and destinations padded to whole 8-round chunks
    return 16 * (64 + 1) + 5 * 16 * (KM + 1) + 4 * 128 + 16 * (64 * DEG_MAX + 1) + 4 * 64 + 2 * (P + 8) + 4 * P;
}
constexpr int LDS_PER_WORKGROUP = 160 * 1024;   // gfx950: 160 KB of LDS per CU, all of it addressable by one workgroup

template <int APL>
struct WaveLDS {
    float4* X4;
    float4* TR0; float4* TR1; float4* TR2; float4* TU; float4* TM;
    float* RD;            // the evaluation's wave reduction (64 floats)
    float* FS;
    int* flg;
    typename PairWord<APL>::T* pl;
    float4* FD;           // dense variant
    unsigned* pdst;
    int P128, trash;
    float4* RN4;          // exact mode: staged receptor atoms (xyz, radius) of the chain's list, RNF their type flags
    int* RNF;
    unsigned* MK;         // exact mode: per 32 staged atoms and atom slot, one bit mask per lane; NULL: no masks
    int nrs;              // staging capacity (0: none)
};

template <int KM, int APL>
struct ChainCtx {
    // lane = atom (slot s: atom lane + 64 s)
    float rx[APL], ry[APL], rz[APL], rad[APL];
    int flg[APL];
    TMask<KM> tmask[APL];
    bool atom[APL];
    const float* gp[APL];
    int is[APL], ic8[APL];   // this lane's segment of slot lists: incg[is .. is + ic8), a multiple of 8 entries
    // one atom per lane: the lane whose partial sum this lane adds (split atom: its helper) or serves (helper lane:
    // the atom), as a byte offset for ds_bpermute; seg: the atom whose list the segment belongs to; k0: the segment's
    // first entry within that atom's list; zc: the atom whose dense row is all zeros (the padding read)
    int pb, seg, k0, zc;
    bool helper;             // a lane past the atoms that walks the second half of a split atom's list
    bool split;              // uniform: the ligand has split lists
    const unsigned* incg;    // the ligand's row of slot lists (memory); incg[incpad ..] are 8 padding entries
    int incpad;
    // lane = torsion k: its axis atoms; lane 7 + k: torsion k's reference axis (point, unit vector)
    int axa, axb;
    float rpx, rpy, rpz, rux, ruy, ruz;
    // uniform
    int na, nt, np, np2, D8;
    // one atom per lane: bit r set when round r of the compact pair loop (pairs 128 r .. 128 r + 127) holds a
    // hydrophobic (rhyd) or h-bond (rhb) pair, so the loop tests a bit instead of a ballot per round (a 64-atom
    // ligand's at most 2,016 pairs take 16 rounds; the library admits P_MAX up to 4,096 = 32 rounds there)
    unsigned rhyd, rhb;
    float rg;
    int ny, nz;
    float o0, o1, o2, lo0, lo1, lo2, hi0, hi1, hi2, inv_s, pen, fmx0, fmx1, fmx2;
    int exact;
    const float* rrx; const int* rrt; int rn; const int* nb; int nbn;
    bool staged;          // exact mode: the receptor atoms of the loop sit in LDS (w.RN4 / w.RNF)
    bool masked;          // ... and each atom walks only the staged atoms its masks (w.MK) mark
    // masks: built from each atom's position (mx0..2) when mvalid; rebuilt once any atom has moved EXACT_MARGIN
    mutable bool mvalid;
    mutable float mx0[APL], mx1[APL], mx2[APL];
    const float* sfrad; const int* sfhyd; const int* sfdon; const int* sfacc;
};

template <int KM, int APL, bool DENSE = false>
DEVI WaveLDS<APL> carve_lds(char* base, int P_MAX, int DEG_MAX = 0) {
    static_assert(!DENSE || APL == 1, "the dense layout holds one atom per lane");
    constexpr int AMX = 64 * APL;
    WaveLDS<APL> w;
    w.P128 = DENSE ? pad512(P_MAX) : pad128(P_MAX);
    float4* f = reinterpret_cast<float4*>(base);
    w.X4 = f; f += AMX + 1;
    w.TR0 = f; f += KM + 1;
    w.TR1 = f; f += KM + 1;
    w.TR2 = f; f += KM + 1;
    w.TU = f; f += KM + 1;
    w.TM = f; f += KM + 1;
    w.RD = reinterpret_cast<float*>(f); f += 32;
    w.FS = nullptr;
    w.FD = nullptr;
    w.RN4 = nullptr;
    w.RNF = nullptr;
    w.MK = nullptr;
    w.nrs = 0;
    w.pdst = nullptr;
    w.trash = 64 * DEG_MAX;
    if constexpr (DENSE) { w.FD = f; f += 64 * DEG_MAX + 1; }
    else { w.FS = reinterpret_cast<float*>(f); f += (w.P128 + 4) / 4; }
    w.flg = reinterpret_cast<int*>(f);
    w.pl = reinterpret_cast<typename PairWord<APL>::T*>(w.flg + AMX);
    if constexpr (DENSE) w.pdst = reinterpret_cast<unsigned*>(w.pl + w.P128 + 8);
    return w;
}

// exact-mode staging area of a wave, after the wave's regular region: nrs float4 atoms + nrs int flags (+ with masks
// nrs / 32 rows of APL x 64 lane masks); nrs is a multiple of 64
__host__ __device__ inline int stage_bytes(int nrs, int masks, int APL = 1) { return nrs * (masks ? 20 + 8 * APL : 20); }
template <int APL>
DEVI void attach_stage(WaveLDS<APL>& w, char* base, int nrs, int masks) {
    if (nrs <= 0) return;
    w.RN4 = reinterpret_cast<float4*>(base);
    w.RNF = reinterpret_cast<int*>(w.RN4 + nrs);
    w.MK = masks ? reinterpret_cast<unsigned*>(w.RNF + nrs) : nullptr;
    w.nrs = nrs;
}

// Masks of the exact loop cover the receptor atoms within cutoff + EXACT_MARGIN (+ 0.01 A against rounding) of where
// the atom was when they were built; while no atom has moved more than EXACT_MARGIN since, every receptor atom within
// the cutoff of an atom is marked, and the masked atoms are visited in list order: the sums equal the full loop's.
constexpr float EXACT_MARGIN = 1.0f;

// one receptor atom's exact term for an atom (radius rad, flags flg) with r2 < cutoff^2 (offset rx, ry, rz); every
// exact loop uses it
DEVI void exact_add(float rad, int flg, const SFConst& sf, float rx, float ry, float rz, float r2, float rad_b,
                    int flg_b, float& el, float& gx, float& gy, float& gz) {
    const float r = fast_sqrt(r2);
    float de;
    el += pair_e(r - rad - rad_b, pair_flags(flg, flg_b), sf, de);
    const float sc = de * fast_rcp(fmaxf(r, 1e-6f));
    gx += sc * rx; gy += sc * ry; gz += sc * rz;
}

template <int KM, int APL, bool DENSE = false, bool EX = false>
DEVI ChainCtx<KM, APL> setup_chain(int l, int lane, const LigArgs& L, const GridArgs& G, const SFArgs& S,
                                   const RecArgs& R, const WaveLDS<APL>& w) {
    using PW = PairWord<APL>;
    constexpr int AMX = 64 * APL;
    ChainCtx<KM, APL> c;
    const int AM = L.AM;
    c.na = __builtin_amdgcn_readfirstlane(L.na[l]);
    c.nt = min(__builtin_amdgcn_readfirstlane(L.nt[l]), KM);
    c.np = min(__builtin_amdgcn_readfirstlane(L.np[l]), L.P_MAX);
    c.np2 = (c.np + 127) & ~127;
    c.rg = L.rg[l];
    const int ri = __builtin_amdgcn_readfirstlane(G.rec_idx[l]);
    const int nx = (int)G.gdims[3 * ri];
    c.ny = (int)G.gdims[3 * ri + 1]; c.nz = (int)G.gdims[3 * ri + 2];
    c.o0 = G.origin[3 * ri]; c.o1 = G.origin[3 * ri + 1]; c.o2 = G.origin[3 * ri + 2];
    c.lo0 = G.lo[3 * ri]; c.lo1 = G.lo[3 * ri + 1]; c.lo2 = G.lo[3 * ri + 2];
    c.hi0 = G.hi[3 * ri]; c.hi1 = G.hi[3 * ri + 1]; c.hi2 = G.hi[3 * ri + 2];
    c.inv_s = 1.f / G.spacing; c.pen = G.pen;
    c.fmx0 = (float)(nx - 1) - 1e-4f; c.fmx1 = (float)(c.ny - 1) - 1e-4f; c.fmx2 = (float)(c.nz - 1) - 1e-4f;
    if constexpr (DENSE) {
        c.o0 = in_vgpr(c.o0); c.o1 = in_vgpr(c.o1); c.o2 = in_vgpr(c.o2);
        c.lo0 = in_vgpr(c.lo0); c.lo1 = in_vgpr(c.lo1); c.lo2 = in_vgpr(c.lo2);
        c.hi0 = in_vgpr(c.hi0); c.hi1 = in_vgpr(c.hi1); c.hi2 = in_vgpr(c.hi2);
        c.inv_s = in_vgpr(c.inv_s); c.pen = in_vgpr(c.pen);
        c.fmx0 = in_vgpr(c.fmx0); c.fmx1 = in_vgpr(c.fmx1); c.fmx2 = in_vgpr(c.fmx2);
    }
    const long long n3 = (long long)nx * c.ny * c.nz;
    const float* grid_r = G.grid + G.goffs[ri];
#pragma unroll
    for (int s = 0; s < APL; ++s) {
        const int a = lane + 64 * s;
        c.atom[s] = a < c.na;
        c.rx[s] = c.ry[s] = c.rz[s] = 0.f; c.rad[s] = 0.f; c.flg[s] = 0; c.tmask[s] = 0;
        c.gp[s] = grid_r;
        if (c.atom[s]) {
            const long long ia = (long long)l * AM + a;
            c.rx[s] = L.ref[3 * ia]; c.ry[s] = L.ref[3 * ia + 1]; c.rz[s] = L.ref[3 * ia + 2];
            const int t = L.typ[ia];
            c.rad[s] = S.rad[t];
            c.flg[s] = S.hyd[t] | (S.don[t] << 1) | (S.acc[t] << 2);
            c.gp[s] += (long long)L.gslot[ia] * n3;
            for (int k = 0; k < c.nt; ++k)
                c.tmask[s] |= (TMask<KM>)(L.tm[((long long)l * KM + k) * AM + a] ? 1u : 0u) << k;
        }
    }
    // torsion records start as identities / zeros; the maps of torsions < nt and the current axes are rewritten every
    // evaluation. Lane k holds torsion k's axis atoms, lane 7 + k its reference axis.
    c.axa = c.axb = 0;
    if (lane < c.nt) {
        c.axa = L.ax[((long long)l * KM + lane) * 2];
        c.axb = L.ax[((long long)l * KM + lane) * 2 + 1];
    }
    c.rpx = c.rpy = c.rpz = c.rux = c.ruy = c.ruz = 0.f;
    {
        const int tk = lane - 7;
        if (tk >= 0 && tk < c.nt) {
            const float* axp = L.axp + ((long long)l * KM + tk) * 3;
            const float* axu = L.axu + ((long long)l * KM + tk) * 3;
            c.rpx = axp[0]; c.rpy = axp[1]; c.rpz = axp[2];
            c.rux = axu[0]; c.ruy = axu[1]; c.ruz = axu[2];
        }
    }
    if (lane <= KM) {
        w.TR0[lane] = make_float4(1.f, 0.f, 0.f, 0.f);
        w.TR1[lane] = make_float4(0.f, 1.f, 0.f, 0.f);
        w.TR2[lane] = make_float4(0.f, 0.f, 1.f, 0.f);
        w.TU[lane] = make_float4(0.f, 0.f, 0.f, 0.f);
        w.TM[lane] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
#pragma unroll
    for (int s = 0; s < APL; ++s) w.flg[lane + 64 * s] = c.flg[s];
    if (lane == 0) w.X4[AMX] = make_float4(1e18f, 1e18f, 1e18f, 0.f);
    if constexpr (DENSE) {
        for (int i = lane; i <= w.trash; i += WAVE) w.FD[i] = make_float4(0.f, 0.f, 0.f, 0.f);
        for (int q = lane; q < w.P128; q += WAVE) w.pdst[q] = (unsigned)w.trash | ((unsigned)w.trash << 16);
    } else if (lane == 0) {
        w.FS[w.P128] = 0.f;
    }
    if (lane < 8) w.pl[w.P128 + lane] = (typename PW::T)pw_pad<APL>();
    wave_lds_sync();
    // pair words; padding (and self-pairs): the dummy atom
    const int* pr = L.pr + (long long)l * L.P_MAX * 2;
    for (int q = lane; q < w.P128; q += WAVE) {
        unsigned word = pw_pad<APL>();
        if (q < c.np) {
            const int i = pr[2 * q], j = pr[2 * q + 1];
            if (i != j)
                word = (unsigned)i | ((unsigned)j << PW::SH) | ((unsigned)pair_flags(w.flg[i], w.flg[j]) << PW::CLS);
        }
        w.pl[q] = (typename PW::T)word;
    }
    wave_lds_sync();
    c.rhyd = c.rhb = 0u;
    if constexpr (APL == 1) {
        for (int q = 0; q < c.np2; q += WAVE) {
            const unsigned pw = w.pl[q + lane];
            if (__ballot((pw & pw_hyd<1>()) != 0u) != 0ull) c.rhyd |= 1u << (q >> 7);
            if (__ballot((pw & pw_hb<1>()) != 0u) != 0ull) c.rhb |= 1u << (q >> 7);
        }
    }
    // each atom's pair-slot list (host-built): read from memory in the evaluations, or (dense variant) turned into the
    // destination slots of its pairs (the atom is the pair's second one when its partner is the pair's first)
    c.incg = L.inc + (long long)l * inc_row(L.P_MAX, AM);
    c.incpad = inc_row(L.P_MAX, AM) - 8;
    // ist: start | count << 16 | partner << 24 per lane (atom slot); a helper lane holds no atom and a non-empty
    // segment, which continues its partner atom's list
    int d8 = 0;
    c.pb = 4 * lane; c.seg = lane; c.k0 = 0; c.zc = lane; c.helper = false;
#pragma unroll
    for (int s = 0; s < APL; ++s) {
        const int a = lane + 64 * s;
        const int e = L.ist[(long long)l * AM + a];
        c.is[s] = e & 0xffff;
        c.ic8[s] = (e >> 16) & 0xff;
        if constexpr (APL == 1) {
            const int pl = (e >> 24) & 63;
            c.helper = !c.atom[0] && c.ic8[0] > 0;
            c.pb = 4 * pl;
            c.seg = c.helper ? pl : lane;
            if (c.helper) c.k0 = max(0, c.is[0] - (L.ist[(long long)l * AM + pl] & 0xffff));
        }
        if constexpr (DENSE) c.ic8[s] = max(0, min(c.ic8[s], L.DEG_MAX - c.k0));
        d8 = max(d8, c.ic8[s]);
        if constexpr (DENSE) {
            const unsigned* src = c.incg + c.is[s];
            unsigned short* p16 = reinterpret_cast<unsigned short*>(w.pdst);
            for (int k = 0; k < c.ic8[s]; ++k) {
                const unsigned v = src[k];
                const unsigned q = (v & 0xffffu) >> 2, p = v >> 20;
                if (q < (unsigned)c.np) {
                    const unsigned side = pw_first<APL>(w.pl[q]) == p ? 1u : 0u;
                    const int kk = c.k0 + k;
                    p16[2 * q + side] = (unsigned short)(kk * 64 + (c.seg ^ kk));
                }
            }
        }
    }
    c.split = APL == 1 && __ballot(c.helper) != 0ull;
    // split lists exist only when lanes past the atoms do, so atom 63 is absent and its dense row reads zero
    if (c.split) c.zc = 63;
    c.D8 = (int)wave_max_f((float)d8);
    c.exact = EX ? 1 : 0;
    c.nb = nullptr; c.nbn = 0; c.staged = false; c.masked = false;
    c.mvalid = false;
#pragma unroll
    for (int s = 0; s < APL; ++s) c.mx0[s] = c.mx1[s] = c.mx2[s] = 0.f;
    if constexpr (EX) {
        c.rrx = R.rxyz + (long long)ri * R.RN_MAX * 3;
        c.rrt = R.rtyp + (long long)ri * R.RN_MAX;
        c.rn = R.rn[ri];
        c.sfrad = S.rad; c.sfhyd = S.hyd; c.sfdon = S.don; c.sfacc = S.acc;
    }
    wave_lds_sync();
    return c;
}

// energy of pair word pw (0 for padding) at atom coordinates + radii xi, xj, and its gradient slot (dE/dX of the
// first atom; the second gets minus it)
template <int APL>
DEVI float pair_math(float4 xi, float4 xj, unsigned pw, const SFConst& sf, float4& s, bool hyd, bool hb) {
    const float rx = xi.x - xj.x, ry = xi.y - xj.y, rz = xi.z - xj.z;
    const float r2 = rx * rx + ry * ry + rz * rz;
    const float rinv = fast_rsq(fmaxf(r2, 1e-12f));
    const float rr = r2 * rinv;
    float de;
    const float e = pair_e_nb(rr - xi.w - xj.w, pw_class<APL>(pw), sf, de, hyd, hb);
    const float in = rr < sf.cut ? 1.f : 0.f;
    const float sc = in * de * rinv;
    s = make_float4(rx, ry, rz, sc);   // separation X_i - X_j and the slot (dE/dr)/r
    return in * e;
}

template <int APL>
DEVI float pair_slot(const WaveLDS<APL>& w, unsigned pw, const SFConst& sf, float4& s, bool hyd, bool hb) {
    return pair_math<APL>(w.X4[pw_first<APL>(pw)], w.X4[pw_second<APL>(pw)], pw, sf, s, hyd, hb);
}

// Exact (pairwise) inter energy and force of atom slot s at (X0, X1, X2): over the chain's receptor neighbour list or
// the whole receptor, from the LDS-staged copy (with or without the per-atom masks) or from memory. Every variant
// visits the in-cutoff atoms in list order, so all give the same bits.
template <int KM, int APL>
DEVI void exact_inter(const ChainCtx<KM, APL>& c, const SFConst& sf, const WaveLDS<APL>& w, int lane, int s,
                      float X0, float X1, float X2, float& el, float& gx, float& gy, float& gz) {
    const float cut2 = sf.cut * sf.cut;
    const int nloop = c.nb ? c.nbn : c.rn;
    if (c.masked) {
        // the marked atoms of each 32, in list order (masks kept fresh by refresh_masks)
        for (int cb = 0; cb < nloop; cb += 32) {
            unsigned m = w.MK[((cb >> 5) * APL + s) * WAVE + lane];
            while (m != 0u) {
                const int bb = cb + __builtin_ctz(m);
                m &= m - 1u;
                const float4 a = w.RN4[bb];
                const float rx = X0 - a.x, ry = X1 - a.y, rz = X2 - a.z;
                const float r2 = rx * rx + ry * ry + rz * rz;
                if (r2 >= cut2) continue;
                exact_add(c.rad[s], c.flg[s], sf, rx, ry, rz, r2, a.w, w.RNF[bb], el, gx, gy, gz);
            }
        }
    } else if (c.staged) {
        // same terms, in the same order, from the atoms staged in LDS (broadcast reads)
        for (int bb = 0; bb < nloop; ++bb) {
            const float4 a = w.RN4[bb];
            const float rx = X0 - a.x, ry = X1 - a.y, rz = X2 - a.z;
            const float r2 = rx * rx + ry * ry + rz * rz;
            if (r2 >= cut2) continue;
            exact_add(c.rad[s], c.flg[s], sf, rx, ry, rz, r2, a.w, w.RNF[bb], el, gx, gy, gz);
        }
    } else {
        for (int bb = 0; bb < nloop; ++bb) {
            const int b = c.nb ? c.nb[bb] : bb;
            const float rx = X0 - c.rrx[3 * b], ry = X1 - c.rrx[3 * b + 1], rz = X2 - c.rrx[3 * b + 2];
            const float r2 = rx * rx + ry * ry + rz * rz;
            if (r2 >= cut2) continue;
            const int tb = c.rrt[b];
            exact_add(c.rad[s], c.flg[s], sf, rx, ry, rz, r2, c.sfrad[tb],
                      c.sfhyd[tb] | (c.sfdon[tb] << 1) | (c.sfacc[tb] << 2), el, gx, gy, gz);
        }
    }
}

// Exact mode with masks: rebuild every slot's masks (one broadcast pass over the staged list) at the chain's first
// evaluation and whenever any atom has moved more than EXACT_MARGIN from where its masks were built.
template <int KM, int APL>
DEVI void refresh_masks(const ChainCtx<KM, APL>& c, const WaveLDS<APL>& w, int lane, const SFConst& sf,
                        const float* X0, const float* X1, const float* X2) {
    bool moved = false;
#pragma unroll
    for (int s = 0; s < APL; ++s) {
        const float ddx = X0[s] - c.mx0[s], ddy = X1[s] - c.mx1[s], ddz = X2[s] - c.mx2[s];
        moved |= c.atom[s] && ddx * ddx + ddy * ddy + ddz * ddz > EXACT_MARGIN * EXACT_MARGIN;
    }
    if (c.mvalid && __ballot(moved) == 0ull) return;
    const int nloop = c.nb ? c.nbn : c.rn;
    const float lim = sf.cut + EXACT_MARGIN + 0.01f, lim2 = lim * lim;
    for (int cb = 0; cb < nloop; cb += 32) {
        const int ne = min(32, nloop - cb);
        unsigned m[APL];
#pragma unroll
        for (int s = 0; s < APL; ++s) m[s] = 0u;
        for (int b = 0; b < ne; ++b) {
            const float4 a = w.RN4[cb + b];
#pragma unroll
            for (int s = 0; s < APL; ++s) {
                const float rx = X0[s] - a.x, ry = X1[s] - a.y, rz = X2[s] - a.z;
                if (rx * rx + ry * ry + rz * rz < lim2) m[s] |= 1u << b;
            }
        }
#pragma unroll
        for (int s = 0; s < APL; ++s) w.MK[((cb >> 5) * APL + s) * WAVE + lane] = m[s];
    }
#pragma unroll
    for (int s = 0; s < APL; ++s) { c.mx0[s] = X0[s]; c.mx1[s] = X1[s]; c.mx2[s] = X2[s]; }
    c.mvalid = true;
}

// Energy and lane-distributed gradient of the lane-distributed state x. Returns E (uniform). EX: exact pairwise
// inter energy over the receptor atoms (or a neighbour list) instead of the grids; a compile-time switch, so the grid
// kernels carry none of its state. ea / Xx..Xz: atom slot 0's inter energy and world position (feature output).
template <int KM, int APL, bool DENSE = false, bool EX = false>
DEVI float eval_eg(const ChainCtx<KM, APL>& c, const SFConst& sf, const WaveLDS<APL>& w, int lane, float x,
                   float& gout, float& ea, float& Xx, float& Xy, float& Xz) {
    const float t0 = bcast(x, 0), t1 = bcast(x, 1), t2 = bcast(x, 2);
    float qw = bcast(x, 3), qx = bcast(x, 4), qy = bcast(x, 5), qz = bcast(x, 6);
    const float nq = fast_rsq(qw * qw + qx * qx + qy * qy + qz * qz);
    qw *= nq; qx *= nq; qy *= nq; qz *= nq;
    const float R00 = 1.f - 2.f * (qy * qy + qz * qz), R01 = 2.f * (qx * qy - qz * qw), R02 = 2.f * (qx * qz + qy * qw);
    const float R10 = 2.f * (qx * qy + qz * qw), R11 = 1.f - 2.f * (qx * qx + qz * qz), R12 = 2.f * (qy * qz - qx * qw);
    const float R20 = 2.f * (qx * qz - qy * qw), R21 = 2.f * (qy * qz + qx * qw), R22 = 1.f - 2.f * (qx * qx + qy * qy);
    // lane 7 + k: torsion k's rotation about its reference axis (Rodrigues: R = c I + s [u]x + (1 - c) u u^T) as the
    // affine map p -> R p + (a - R a)
    {
        float sn, cs;
        fast_sincos(x, sn, cs);
        const int tk = lane - 7;
        if (tk >= 0 && tk < c.nt) {
            const float ux = c.rux, uy = c.ruy, uz = c.ruz, t = 1.f - cs;
            const float r00 = cs + t * ux * ux, r01 = t * ux * uy - sn * uz, r02 = t * ux * uz + sn * uy;
            const float r10 = t * ux * uy + sn * uz, r11 = cs + t * uy * uy, r12 = t * uy * uz - sn * ux;
            const float r20 = t * ux * uz - sn * uy, r21 = t * uy * uz + sn * ux, r22 = cs + t * uz * uz;
            const float ax = c.rpx, ay = c.rpy, az = c.rpz;
            w.TR0[tk] = make_float4(r00, r01, r02, ax - (r00 * ax + r01 * ay + r02 * az));
            w.TR1[tk] = make_float4(r10, r11, r12, ay - (r10 * ax + r11 * ay + r12 * az));
            w.TR2[tk] = make_float4(r20, r21, r22, az - (r20 * ax + r21 * ay + r22 * az));
        }
    }
    wave_lds_sync();
    // forward kinematics, leaf-to-root: maps read 4 at a time, applied where the atom's mask bit is set (maps past nt
    // are identities); each map read serves every atom slot of the lane
    float px[APL], py[APL], pz[APL];
#pragma unroll
    for (int s = 0; s < APL; ++s) { px[s] = c.rx[s]; py[s] = c.ry[s]; pz[s] = c.rz[s]; }
    for (int kb = 0; kb < c.nt; kb += 4) {
        float4 a[4], b[4], d[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) { a[u] = w.TR0[kb + u]; b[u] = w.TR1[kb + u]; d[u] = w.TR2[kb + u]; }
#pragma unroll
        for (int s = 0; s < APL; ++s) {
            const unsigned t4 = (unsigned)(c.tmask[s] >> kb);
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const unsigned m = bit_mask(t4, u);
                const float nx = a[u].x * px[s] + a[u].y * py[s] + a[u].z * pz[s] + a[u].w;
                const float ny = b[u].x * px[s] + b[u].y * py[s] + b[u].z * pz[s] + b[u].w;
                const float nz = d[u].x * px[s] + d[u].y * py[s] + d[u].z * pz[s] + d[u].w;
                px[s] = bit_select(m, nx, px[s]); py[s] = bit_select(m, ny, py[s]); pz[s] = bit_select(m, nz, pz[s]);
            }
        }
    }
    float X0[APL], X1[APL], X2[APL];
    // grid corners: issue the 8 gathers now, consume them after the pair work (latency overlaps the LDS work)
    float v000[APL], v001[APL], v010[APL], v011[APL], v100[APL], v101[APL], v110[APL], v111[APL];
    float wx[APL], wy[APL], wz[APL], in0[APL], in1[APL], in2[APL];
#pragma unroll
    for (int s = 0; s < APL; ++s) {
        X0[s] = R00 * px[s] + R01 * py[s] + R02 * pz[s] + t0;
        X1[s] = R10 * px[s] + R11 * py[s] + R12 * pz[s] + t1;
        X2[s] = R20 * px[s] + R21 * py[s] + R22 * pz[s] + t2;
        v000[s] = v001[s] = v010[s] = v011[s] = v100[s] = v101[s] = v110[s] = v111[s] = 0.f;
        wx[s] = wy[s] = wz[s] = in0[s] = in1[s] = in2[s] = 0.f;
        if (c.atom[s] && !EX) {
            const float f0 = (X0[s] - c.o0) * c.inv_s, f1 = (X1[s] - c.o1) * c.inv_s, f2 = (X2[s] - c.o2) * c.inv_s;
            const float fc0 = fminf(fmaxf(f0, 0.f), c.fmx0), fc1 = fminf(fmaxf(f1, 0.f), c.fmx1),
                        fc2 = fminf(fmaxf(f2, 0.f), c.fmx2);
            in0[s] = (f0 > 0.f && f0 < c.fmx0) ? c.inv_s : 0.f;
            in1[s] = (f1 > 0.f && f1 < c.fmx1) ? c.inv_s : 0.f;
            in2[s] = (f2 > 0.f && f2 < c.fmx2) ? c.inv_s : 0.f;
            const int i0 = (int)floorf(fc0), i1 = (int)floorf(fc1), i2 = (int)floorf(fc2);
            wx[s] = fc0 - (float)i0; wy[s] = fc1 - (float)i1; wz[s] = fc2 - (float)i2;
            const int b000 = (i0 * c.ny + i1) * c.nz + i2;
            const int bx = b000 + c.ny * c.nz;
            const float* G = c.gp[s];
            v000[s] = G[b000]; v001[s] = G[b000 + 1]; v010[s] = G[b000 + c.nz]; v011[s] = G[b000 + c.nz + 1];
            v100[s] = G[bx]; v101[s] = G[bx + 1]; v110[s] = G[bx + c.nz]; v111[s] = G[bx + c.nz + 1];
        }
        w.X4[lane + 64 * s] = make_float4(X0[s], X1[s], X2[s], c.rad[s]);
    }
    wave_lds_sync();
    // current torsion axes (lane k < nt) from the world coordinates of the axis atoms: unit vector and moment
    if (lane < c.nt) {
        const float4 A = w.X4[c.axa], B = w.X4[c.axb];
        float ux = B.x - A.x, uy = B.y - A.y, uz = B.z - A.z;
        const float nn = fast_rsq(fmaxf(ux * ux + uy * uy + uz * uz, 1e-12f));
        ux *= nn; uy *= nn; uz *= nn;
        w.TU[lane] = make_float4(ux, uy, uz, 0.f);
        w.TM[lane] = make_float4(A.y * uz - A.z * uy, A.z * ux - A.x * uz, A.x * uy - A.y * ux, 0.f);
    }
    // intra-ligand pairs, branch-free: each pair writes its gradient slot. Dense variant: chunks of 8 rounds of 64
    // with every pair word, destination and coordinate read issued before the math (one LDS round trip per chunk),
    // +slot into the first atom's row, -slot into the second's. Compact variant: two rounds per iteration, pair lists
    // h-bond pairs first, then hydrophobic ones, so most iterations skip those terms (uniform test over the wave).
    float ep_l = 0.f;
    if constexpr (DENSE) {
        for (int cb = 0; cb < c.np; cb += 8 * WAVE) {
            unsigned pw[8], dd[8];
            float4 xi[8], xj[8];
#pragma unroll
            for (int r = 0; r < 8; ++r) { pw[r] = w.pl[cb + r * WAVE + lane]; dd[r] = w.pdst[cb + r * WAVE + lane]; }
#pragma unroll
            for (int r = 0; r < 8; ++r) { xi[r] = w.X4[pw_first<1>(pw[r])]; xj[r] = w.X4[pw_second<1>(pw[r])]; }
#pragma unroll
            for (int r = 0; r < 8; ++r) {
                // pairs are sorted h-bond, hydrophobic, other: most rounds skip both slope terms (uniform test)
                const bool hyd = __ballot((pw[r] & pw_hyd<1>()) != 0u) != 0ull;
                const bool hb = __ballot((pw[r] & pw_hb<1>()) != 0u) != 0ull;
                float4 s0;
                ep_l += pair_math<1>(xi[r], xj[r], pw[r], sf, s0, hyd, hb);
                w.FD[dd[r] & 0xffffu] = s0;
                w.FD[dd[r] >> 16] = make_float4(-s0.x, -s0.y, -s0.z, s0.w);
            }
        }
    } else {
        for (int r = 0; r * 2 * WAVE < c.np2; ++r) {
            const int q0 = r * 2 * WAVE + lane;
            const unsigned pw0 = w.pl[q0], pw1 = w.pl[q0 + WAVE];
            bool hyd, hb;
            if constexpr (APL == 1) {
                hyd = (c.rhyd >> r) & 1u;
                hb = (c.rhb >> r) & 1u;
            } else {
                hyd = __ballot(((pw0 | pw1) & pw_hyd<APL>()) != 0u) != 0ull;
                hb = __ballot(((pw0 | pw1) & pw_hb<APL>()) != 0u) != 0ull;
            }
            float4 s0, s1;
            ep_l += pair_slot<APL>(w, pw0, sf, s0, hyd, hb);
            ep_l += pair_slot<APL>(w, pw1, sf, s1, hyd, hb);
            w.FS[q0] = s0.w;
            w.FS[q0 + WAVE] = s1.w;
        }
    }
    wave_lds_sync();
    // each atom sums its slots: 8 list entries per 16-byte read, then 8 independent slot reads in flight (padding
    // entries and lanes past their list read the zero slot)
    float pfx[APL], pfy[APL], pfz[APL];
#pragma unroll
    for (int s = 0; s < APL; ++s) { pfx[s] = pfy[s] = pfz[s] = 0.f; }
    if constexpr (DENSE) {
        for (int k = 0; k < c.D8; k += 8) {
            // the lane's segment: entries k0 + k.. of atom seg's row; past it, the zero row of atom zc
            const bool in = k < c.ic8[0];
            const int kk = in ? c.k0 + k : k;
            const int xk = (in ? c.seg : c.zc) ^ kk;
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const float4 f = w.FD[(kk + u) * 64 + (xk ^ u)];
                pfx[0] = fmaf(f.w, f.x, pfx[0]); pfy[0] = fmaf(f.w, f.y, pfy[0]); pfz[0] = fmaf(f.w, f.z, pfz[0]);
            }
        }
    } else {
        // each entry: the byte offsets of the pair's slot (dE/dr)/r and of the partner atom; lanes past their list
        // read the row's trailing padding (zero slot, dummy atom). The atom adds fma(slot, X_self - X_partner): the
        // dense rows hold the same separation and slot, so both layouts give the same bits
        const char* fsb = reinterpret_cast<const char*>(w.FS);
        const char* x4b = reinterpret_cast<const char*>(w.X4);
#pragma unroll
        for (int s = 0; s < APL; ++s) {
            // the segment's atom position (a helper lane walks another atom's list)
            float sx = X0[s], sy = X1[s], sz = X2[s];
            if constexpr (APL == 1) {
                const float4 t = w.X4[c.seg];
                sx = t.x; sy = t.y; sz = t.z;
            }
            for (int k = 0; k < c.D8; k += 8) {
                const uint4* r4 = reinterpret_cast<const uint4*>(c.incg + (k < c.ic8[s] ? c.is[s] + k : c.incpad));
                const uint4 va = r4[0], vb = r4[1];
                const unsigned ws[8] = {va.x, va.y, va.z, va.w, vb.x, vb.y, vb.z, vb.w};
                float sc[8];
                float4 xp[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    sc[u] = *reinterpret_cast<const float*>(fsb + (ws[u] & 0xffffu));
                    xp[u] = *reinterpret_cast<const float4*>(x4b + (ws[u] >> 16));
                }
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    pfx[s] = fmaf(sc[u], sx - xp[u].x, pfx[s]);
                    pfy[s] = fmaf(sc[u], sy - xp[u].y, pfy[s]);
                    pfz[s] = fmaf(sc[u], sz - xp[u].z, pfz[s]);
                }
            }
        }
    }
    if constexpr (APL == 1) {
        // split lists: each atom adds its helper's partial sum (lanes without one add -0, which changes no value);
        // helper lanes end with zero
        if (c.split) {
            const float hx = c.helper ? pfx[0] : -0.f, hy = c.helper ? pfy[0] : -0.f, hz = c.helper ? pfz[0] : -0.f;
            const float ox = __int_as_float(__builtin_amdgcn_ds_bpermute(c.pb, __float_as_int(hx)));
            const float oy = __int_as_float(__builtin_amdgcn_ds_bpermute(c.pb, __float_as_int(hy)));
            const float oz = __int_as_float(__builtin_amdgcn_ds_bpermute(c.pb, __float_as_int(hz)));
            pfx[0] = (c.helper ? 0.f : pfx[0]) + ox;
            pfy[0] = (c.helper ? 0.f : pfy[0]) + oy;
            pfz[0] = (c.helper ? 0.f : pfz[0]) + oz;
        }
    }
    if constexpr (EX) {
        if (c.masked) refresh_masks<KM, APL>(c, w, lane, sf, X0, X1, X2);
    }
    float el_all = 0.f;   // slot 0 assigns, later slots add (no extra add on the one-atom-per-lane path)
    float fx[APL], fy[APL], fz[APL];
#pragma unroll
    for (int s = 0; s < APL; ++s) {
        float el = 0.f, gx = 0.f, gy = 0.f, gz = 0.f;
        if (c.atom[s]) {
            if constexpr (!EX) {
                const float c00 = v000[s] + wz[s] * (v001[s] - v000[s]), c01 = v010[s] + wz[s] * (v011[s] - v010[s]);
                const float c10 = v100[s] + wz[s] * (v101[s] - v100[s]), c11 = v110[s] + wz[s] * (v111[s] - v110[s]);
                const float c0 = c00 + wy[s] * (c01 - c00), c1 = c10 + wy[s] * (c11 - c10);
                el = c0 + wx[s] * (c1 - c0);
                gx = (c1 - c0) * in0[s];
                gy = ((c01 - c00) + wx[s] * ((c11 - c10) - (c01 - c00))) * in1[s];
                const float e0 = (v001[s] - v000[s]) + wy[s] * ((v011[s] - v010[s]) - (v001[s] - v000[s]));
                const float e1 = (v101[s] - v100[s]) + wy[s] * ((v111[s] - v110[s]) - (v101[s] - v100[s]));
                gz = (e0 + wx[s] * (e1 - e0)) * in2[s];
            } else {
                exact_inter<KM, APL>(c, sf, w, lane, s, X0[s], X1[s], X2[s], el, gx, gy, gz);
            }
            const float l0 = fmaxf(c.lo0 - X0[s], 0.f), h0 = fmaxf(X0[s] - c.hi0, 0.f);
            const float l1 = fmaxf(c.lo1 - X1[s], 0.f), h1 = fmaxf(X1[s] - c.hi1, 0.f);
            const float l2 = fmaxf(c.lo2 - X2[s], 0.f), h2 = fmaxf(X2[s] - c.hi2, 0.f);
            el += c.pen * (l0 * l0 + h0 * h0 + l1 * l1 + h1 * h1 + l2 * l2 + h2 * h2);
            gx += 2.f * c.pen * (h0 - l0); gy += 2.f * c.pen * (h1 - l1); gz += 2.f * c.pen * (h2 - l2);
        }
        if (s == 0) { ea = el; Xx = X0[0]; Xy = X1[0]; Xz = X2[0]; el_all = el; }
        else el_all += el;
        fx[s] = gx + pfx[s]; fy[s] = gy + pfy[s]; fz[s] = gz + pfz[s];
    }
    // every gradient component and the energy are reduced over the wave together, so the DPP steps of the 7 + KM
    // independent sums interleave instead of running one dependent chain after another
    float v[7 + KM];
    v[6] = el_all + ep_l;
#pragma unroll
    for (int s = 0; s < APL; ++s) {
        const float dX = X0[s] - t0, dY = X1[s] - t1, dZ = X2[s] - t2;
        float part[6] = {fx[s], fy[s], fz[s], dY * fz[s] - dZ * fy[s], dZ * fx[s] - dX * fz[s],
                         dX * fy[s] - dY * fx[s]};
#pragma unroll
        for (int i = 0; i < 6; ++i) v[i] = s == 0 ? part[i] : v[i] + part[i];
        // torque about each current axis from the atoms it moves: u.((X - p) x F) = u.(X x F) + m.F with m = p x u;
        // records in groups of 8 (records past nt are zero)
        const float Tx = X1[s] * fz[s] - X2[s] * fy[s], Ty = X2[s] * fx[s] - X0[s] * fz[s],
                    Tz = X0[s] * fy[s] - X1[s] * fx[s];
#pragma unroll
        for (int kb = 0; kb < KM; kb += 8) {
            if (kb < c.nt) {
                const unsigned t8 = (unsigned)(c.tmask[s] >> kb);
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const float4 cu = w.TU[kb + u], cm = w.TM[kb + u];
                    const float tq = bit_select(bit_mask(t8, u), cu.x * Tx + cu.y * Ty + cu.z * Tz + cm.x * fx[s] +
                                                                     cm.y * fy[s] + cm.z * fz[s], 0.f);
                    v[7 + kb + u] = s == 0 ? tq : v[7 + kb + u] + tq;
                }
            } else if (s == 0) {
#pragma unroll
                for (int u = 0; u < 8; ++u) v[7 + kb + u] = 0.f;
            }
        }
    }
    const float r = wave_sum_to_lanes<7 + KM>(v, w.RD, lane);
    gout = (lane < 6 || (lane >= 7 && lane < 7 + c.nt)) ? r : 0.f;
    return bcast(r, 6);
}

// x (+) alpha p on the lane-distributed state
DEVI float conf_inc(float x, float p, float alpha, int lane) {
    float qw = bcast(x, 3), qx = bcast(x, 4), qy = bcast(x, 5), qz = bcast(x, 6);
    quat_inc(qw, qx, qy, qz, alpha * bcast(p, 3), alpha * bcast(p, 4), alpha * bcast(p, 5));
    const float lin = x + alpha * p;
    return lane == 3 ? qw : lane == 4 ? qx : lane == 5 ? qy : lane == 6 ? qz : lin;
}

// The reference's MC mutation (one entity: translation by amp_pos * ball, rotation by amp_rot / max(rg, 1) * ball, or
// one torsion set uniformly in [-pi, pi)), drawn from the counter RNG slots 0, 1 and 10.. of (chain, step).
// entity the plain move perturbs: 0 translation, 1 rotation, 2 + k torsion k (uniform, slot 0)
template <class Ctx>
DEVI int plain_entity(const Ctx& c, unsigned long long seed, int cid, int sg) {
    const int nm = 2 + c.nt;
    return min((int)(urand(seed, cid, sg, 0) * (float)nm), nm - 1);
}
template <class Ctx>
DEVI float entity_move(float x, const Ctx& c, int lane, unsigned long long seed, int cid, int sg, int which,
                       float amp_pos, float amp_rot) {
    if (which == 0) {
        float v0, v1, v2;
        rand_ball(seed, cid, sg, 10, v0, v1, v2);
        const float v = lane == 0 ? v0 : lane == 1 ? v1 : v2;
        if (lane < 3) x += amp_pos * v;
    } else if (which == 1) {
        float v0, v1, v2;
        rand_ball(seed, cid, sg, 10, v0, v1, v2);
        const float sc = amp_rot / fmaxf(c.rg, 1.0f);
        float qw = bcast(x, 3), qx = bcast(x, 4), qy = bcast(x, 5), qz = bcast(x, 6);
        quat_inc(qw, qx, qy, qz, sc * v0, sc * v1, sc * v2);
        x = lane == 3 ? qw : lane == 4 ? qx : lane == 5 ? qy : lane == 6 ? qz : x;
    } else {
        const float th = (2.f * urand(seed, cid, sg, 1) - 1.f) * 3.14159265f;
        if (lane == 7 + which - 2) x = th;
    }
    return x;
}
template <class Ctx>
DEVI float plain_move(float x, const Ctx& c, int lane, unsigned long long seed, int cid, int sg, float amp_pos,
                      float amp_rot) {
    return entity_move(x, c, lane, seed, cid, sg, plain_entity(c, seed, cid, sg), amp_pos, amp_rot);
}

// Register budget of the throughput chain kernels (per-chain, persistent, learned step): at most 168 VGPRs, so three
// waves share a SIMD. They issue VALU about 87 % of the time at two waves; a third wave covers the LDS and gather
// latency that idles the rest, worth more than the few spills (mostly chain set-up) the budget costs: the
// multi-receptor batch runs 520 -> 614 docked ligands/s.
// Kept to the one-atom-per-lane buckets up to 32 torsions: past them (longer reductions, two atom slots per lane) the
// budget would spill 0.4-0.9 KB per lane, and those kernels keep the compiler's own allocation.
template <int KM, int APL> constexpr int three_waves() { return (APL == 1 && KM <= 32) ? 3 : 1; }
#define ADOCK_WAVES(KM, APL) __attribute__((amdgpu_waves_per_eu(three_waves<KM, APL>(), three_waves<KM, APL>() == 3 ? 3 : 8)))
// the dense LDS tile (up to ~40 KB per wave) leaves room for few waves per CU: the compiler's own register allocation
#define ADOCK_WAVES_D(KM, APL, DENSE)                                                                                  \
    __attribute__((amdgpu_waves_per_eu(DENSE ? 1 : three_waves<KM, APL>(),                                            \
                                       (!DENSE && three_waves<KM, APL>() == 3) ? 3 : 8)))

// BFGS stops once |g| < 1e-5 (the reference's sqrt(g.g) >= 1e-5 test): sqrtf is correctly rounded and monotonic, so
// the test equals g.g >= the smallest float whose square root rounds to at least 1e-5f, 0x2edbe6fd (~9.9999987e-11;
// its predecessor's square root rounds to 9.999999e-06), without the square root
#define GG_STOP __uint_as_float(0x2edbe6fdu)

// A lane-distributed vector (component j in lane j) made readable by every lane: written to this wave's LDS at
// construction, component j then read as a broadcast (all lanes the same address; the compiler merges neighbouring
// components into 16-byte reads).
template <int DV>
struct BcastVec {
    const float4* a;
    DEVI BcastVec(float* buf, int lane, float v) : a(reinterpret_cast<const float4*>(buf)) {
        buf[lane] = v;
        wave_lds_sync();
    }
    DEVI float operator[](int j) const {
        const float4 q = a[j >> 2];
        return (j & 3) == 0 ? q.x : (j & 3) == 1 ? q.y : (j & 3) == 2 ? q.z : q.w;
    }
};

// The evaluation at the pose bfgs returns (KEEP): its lane-distributed gradient and this lane's atom energy. Every pass
// of the loop ends on an evaluation of the pose it keeps, so the last evaluation is that of the returned pose, except
// when the start pose is restored after a worse end (valid = false: its evaluation, the first, is not kept).
struct LastEval { float g = 0.f, ea = 0.f; bool valid = false; };

// Vina-style BFGS on the lane-distributed state x (in/out); returns the final energy. One evaluation site: the first
// pass evaluates x itself, every later pass evaluates one line-search trial.
template <int KM, int APL, bool DENSE = false, bool EX = false, bool KEEP = false>
DEVI float bfgs(const ChainCtx<KM, APL>& c, const SFConst& sf, const WaveLDS<APL>& w, int lane, float& x, int iters,
                int& nev, LastEval* last = nullptr) {
    constexpr int DV = 7 + KM;
    float h[DV];
#pragma unroll
    for (int j = 0; j < DV; ++j) h[j] = (j == lane) ? 1.f : 0.f;
    const float xo = x;
    float g = 0.f, p = 0.f, f0 = 0.f, f_orig = 0.f, pg = 0.f, alpha = 1.f;
    float kg = 0.f, kea = 0.f;
    int step = 0, trial = -1;
    while (true) {
        const float xe = trial < 0 ? x : conf_inc(x, p, alpha, lane);
        float ge, ea_, x0_, x1_, x2_;
        const float fe = eval_eg<KM, APL, DENSE, EX>(c, sf, w, lane, xe, ge, ea_, x0_, x1_, x2_);
        ++nev;
        if constexpr (KEEP) { kg = ge; kea = ea_; }
        if (trial < 0) {
            f0 = fe; f_orig = fe; g = ge;
            if (iters <= 0) break;
        } else {
            const bool ok = fe - f0 < 1e-4f * alpha * pg;
            if (!ok) alpha *= 0.5f;
            if (!ok && trial < 9) { ++trial; continue; }
            const float y = ge - g;
            f0 = fe; x = xe; g = ge;
            // |g|^2, y.p and y.y reduced together (the same DPP sequence per sum as three separate reductions)
            float red[3] = {ge * ge, y * p, y * y};
            wave_sum_multi<3>(red);
            const float gg = red[0], yp = red[1];
            if (!(gg >= GG_STOP)) break;
            if (step == 0) {
                const float yy = red[2];
                if (fabsf(yy) > EPSF) {
                    const float dg = alpha * yp / yy;
                    asm volatile("");   // keeps this once-per-run reset a branch, not selects on every iteration
#pragma unroll
                    for (int j = 0; j < DV; ++j) h[j] = (j == lane) ? dg : 0.f;
                }
            }
            if (!(alpha * yp < EPSF)) {
                // the vectors' components reach every lane as LDS broadcast reads (16 bytes per read) instead of a
                // v_readlane each: the same operands in the same order, so the same bits, with a quarter of the
                // instructions on the VALU
                float mhy = 0.f;
                {
                    const BcastVec<DV> yb(w.RD, lane, y);
#pragma unroll
                    for (int j = 0; j < DV; ++j) mhy += h[j] * yb[j];
                }
                mhy = -mhy;
                const float yhy = -wave_sum(y * mhy);
                const float r = 1.f / (alpha * yp);
                const float c2 = alpha * alpha * (r * r * yhy + r);
                const float ar = alpha * r;
                const BcastVec<DV> pb(w.RD, lane, p), mb(w.RD + 64, lane, mhy);
                // H += ar (mhy p^T + p mhy^T) + c2 p p^T, as two FMAs per entry: row i gets pj A_i + mj B_i
                const float A = fmaf(ar, mhy, c2 * p), B = ar * p;
#pragma unroll
                for (int j = 0; j < DV; ++j) h[j] = fmaf(pb[j], A, fmaf(mb[j], B, h[j]));
            }
            if (++step >= iters) break;
        }
        float s = 0.f;
        {
            const BcastVec<DV> gb(w.RD, lane, g);
#pragma unroll
            for (int j = 0; j < DV; ++j) s += h[j] * gb[j];
        }
        p = -s;
        pg = wave_sum(p * g);
        alpha = 1.f;
        trial = 0;
    }
    const bool restored = !(f0 <= f_orig);
    if (restored) { f0 = f_orig; x = xo; }
    if constexpr (KEEP) *last = LastEval{kg, kea, !restored};
    return f0;
}

// ---- speculative line search: the NS waves of one workgroup share one chain ----
// Every wave holds the same chain state and runs the same BFGS bookkeeping (identical results). At each line search
// the waves evaluate the trials t = tb + wave (step 2^-t) of one batch in parallel and all take the first one that
// passes the Armijo test (after ten failures: trial 9 with the step halved once more), so the accepted points,
// energies and evaluation counts equal the one-wave search bit for bit while a step costs about one evaluation of
// latency instead of two. Worth it when the chains leave SIMDs idle (a few hundred chains on 256 CUs).
// Exchange area, two buffers x (waves <= 4): each wave publishes its trial (pose, gradient, energy, Armijo outcome)
// in its own slot; one barrier per batch, and alternating buffers make a slot safe to rewrite two batches later (every
// wave has passed the intervening barrier, so it has finished reading).
struct SpecLDS { float* x; float* g; float* fe; int* ok; };   // x [2][4][64], g [2][4][64], fe [2][4], ok [2][4]
constexpr int SPEC_LDS_BYTES = 4 * (2 * 4 * 64 * 2 + 2 * 4 * 2);

DEVI SpecLDS carve_spec(char* base) {
    float* f = reinterpret_cast<float*>(base);
    return SpecLDS{f, f + 512, f + 1024, reinterpret_cast<int*>(f + 1032)};
}

template <int KM, int NS, bool DENSE, bool EX>
DEVI float bfgs_spec(const ChainCtx<KM, 1>& c, const SFConst& sf, const WaveLDS<1>& w, const SpecLDS& xs, int lane,
                     int wid,
                     float& x, int iters, int& nev, int& buf) {
    constexpr int DV = 7 + KM;
    float h[DV];
#pragma unroll
    for (int j = 0; j < DV; ++j) h[j] = (j == lane) ? 1.f : 0.f;
    const float xo = x;
    float g, ea_, x0_, x1_, x2_;
    float f0 = eval_eg<KM, 1, DENSE, EX>(c, sf, w, lane, x, g, ea_, x0_, x1_, x2_);
    ++nev;
    const float f_orig = f0;
    int step = 0;
    while (iters > 0) {
        float s = 0.f;
#pragma unroll
        for (int j = 0; j < DV; ++j) s += h[j] * bcast(g, j);
        const float p = -s;
        const float pg = wave_sum(p * g);
        float xa = x, ga = 0.f, fa = 0.f, alpha = 1.f;
        for (int tb = 0; tb < 10; tb += NS) {
            const int t = tb + wid;
            bool ok = false;
            float xe = x, fe = 0.f, ge = 0.f;
            if (t < 10) {
                const float at = ldexpf(1.f, -t);
                xe = conf_inc(x, p, at, lane);
                fe = eval_eg<KM, 1, DENSE, EX>(c, sf, w, lane, xe, ge, ea_, x0_, x1_, x2_);
                ok = fe - f0 < 1e-4f * at * pg;
            }
            const int slot = buf * 4 + wid;
            xs.x[slot * 64 + lane] = xe;
            xs.g[slot * 64 + lane] = ge;
            if (lane == 0) { xs.fe[slot] = fe; xs.ok[slot] = ok ? 1 : 0; }
            __syncthreads();
            int first = -1;
#pragma unroll
            for (int u = NS - 1; u >= 0; --u) first = xs.ok[buf * 4 + u] ? u : first;
            first = __builtin_amdgcn_readfirstlane(first);
            const int win = first >= 0 ? first : (tb + NS >= 10 ? 9 - tb : -1);
            const int rd = buf * 4 + (win < 0 ? 0 : win);
            buf ^= 1;
            if (win < 0) {
                nev += NS;
                continue;
            }
            nev += win + 1;
            xa = xs.x[rd * 64 + lane]; ga = xs.g[rd * 64 + lane]; fa = xs.fe[rd];
            alpha = ldexpf(1.f, -(tb + win));
            if (first < 0) alpha *= 0.5f;
            break;
        }
        const float y = ga - g;
        f0 = fa; x = xa; g = ga;
        float red[3] = {g * g, y * p, y * y};
        wave_sum_multi<3>(red);
        const float gg = red[0], yp = red[1];
        if (!(gg >= GG_STOP)) break;
        if (step == 0) {
            const float yy = red[2];
            if (fabsf(yy) > EPSF) {
                const float dg = alpha * yp / yy;
                asm volatile("");   // keeps this once-per-run reset a branch, not selects on every iteration
#pragma unroll
                for (int j = 0; j < DV; ++j) h[j] = (j == lane) ? dg : 0.f;
            }
        }
        if (!(alpha * yp < EPSF)) {
            float mhy = 0.f;
#pragma unroll
            for (int j = 0; j < DV; ++j) mhy += h[j] * bcast(y, j);
            mhy = -mhy;
            const float yhy = -wave_sum(y * mhy);
            const float r = 1.f / (alpha * yp);
            const float c2 = alpha * alpha * (r * r * yhy + r);
            const float ar = alpha * r;
            const float A = fmaf(ar, mhy, c2 * p), B = ar * p;   // as in bfgs: two FMAs per entry
#pragma unroll
            for (int j = 0; j < DV; ++j) h[j] = fmaf(bcast(p, j), A, fmaf(bcast(mhy, j), B, h[j]));
        }
        if (++step >= iters) break;
    }
    if (!(f0 <= f_orig)) { f0 = f_orig; x = xo; }
    return f0;
}

// One MC chain on the calling wave (all 64 lanes active), or on the NS waves of a workgroup with the speculative line
// search (wave wid; wave 0 writes the results). w: this wave's LDS region.
template <int KM, int APL, int NS = 1, bool DENSE = false, bool EX = false>
DEVI int run_chain(int ch, int lane, const WaveLDS<APL>& w, const LigArgs& L, const GridArgs& G, const SFArgs& S,
                   const RecArgs& R, const int* __restrict__ chain_lig, const McArgs& M, int wid = 0,
                   SpecLDS xs = SpecLDS{}) {
    constexpr int SD = 7 + KM;
    SFConst sf = make_sf(S.prm);
    if constexpr (DENSE) {
        sf.g1o = in_vgpr(sf.g1o); sf.g1s = in_vgpr(sf.g1s); sf.g1d = in_vgpr(sf.g1d); sf.g2o = in_vgpr(sf.g2o);
        sf.g2s = in_vgpr(sf.g2s); sf.g2d = in_vgpr(sf.g2d); sf.g1b = in_vgpr(sf.g1b); sf.g2b = in_vgpr(sf.g2b);
        sf.hbh = in_vgpr(sf.hbh); sf.bbh = in_vgpr(sf.bbh); sf.w1 = in_vgpr(sf.w1); sf.w2 = in_vgpr(sf.w2);
        sf.w3 = in_vgpr(sf.w3); sf.w4 = in_vgpr(sf.w4); sf.w5 = in_vgpr(sf.w5); sf.repo = in_vgpr(sf.repo);
        sf.hinv = in_vgpr(sf.hinv); sf.binv = in_vgpr(sf.binv); sf.cut = in_vgpr(sf.cut);
    }
    const int l = __builtin_amdgcn_readfirstlane(chain_lig[ch]);
    ChainCtx<KM, APL> c = setup_chain<KM, APL, DENSE, EX>(l, lane, L, G, S, R, w);
    if (M.nbr) {
        c.nb = M.nbr + (long long)ch * M.NB_MAX;
        c.nbn = min(max(__builtin_amdgcn_readfirstlane(M.nbr_n[ch]), 0), M.NB_MAX);
    }
    if constexpr (EX) {
        // the chain's receptor atoms (neighbour list or whole receptor) into LDS once, when they fit the staging area
        const int nloop = c.nb ? c.nbn : c.rn;
        c.staged = w.nrs > 0 && nloop <= w.nrs;
        c.masked = c.staged && w.MK != nullptr;
        c.mvalid = false;
        if (c.staged) {
            for (int bb = lane; bb < nloop; bb += WAVE) {
                const int b = c.nb ? c.nb[bb] : bb;
                const int t = c.rrt[b];
                w.RN4[bb] = make_float4(c.rrx[3 * b], c.rrx[3 * b + 1], c.rrx[3 * b + 2], c.sfrad[t]);
                w.RNF[bb] = c.sfhyd[t] | (c.sfdon[t] << 1) | (c.sfacc[t] << 2);
            }
            wave_lds_sync();
        }
    }
    const int cid = __builtin_amdgcn_readfirstlane(M.chain_id[ch]);
    const int S_ = 7 + c.nt;
    const int iters = M.bfgs_iters > 0 ? M.bfgs_iters : (25 + c.na) / 3;
    float* gcur = M.cur + (long long)ch * SD;
    float* gbest = M.best + (long long)ch * SD;
    float cur = lane < S_ ? gcur[lane] : 0.f;
    float best = lane < S_ ? gbest[lane] : 0.f;
    float eb = M.best_e[ch];
    float ec = M.cur_e[ch];
    int nev = 0;
    const unsigned long long seed = M.seed;
    if (!M.init_opt && ec < eb) { eb = ec; best = cur; }
    int xbuf = 0;   // exchange buffer of the speculative line search; alternates over the whole chain
    // st = -1 is the optional initial optimisation of cur; st >= 0 are MC steps
    for (int st = M.init_opt ? -1 : 0; st < M.mc_steps; ++st) {
        float x = cur;
        int sg = 0;
        if (st >= 0) {
            sg = M.step_offset + st;
            x = plain_move(x, c, lane, seed, cid, sg, M.amp_pos, M.amp_rot);
        }
        float en;
        if constexpr (NS == 1) en = bfgs<KM, APL, DENSE, EX>(c, sf, w, lane, x, iters, nev);
        else en = bfgs_spec<KM, NS, DENSE, EX>(c, sf, w, xs, lane, wid, x, iters, nev, xbuf);
        if (st < 0) {
            ec = en; cur = x;
            if (ec < eb) { eb = ec; best = cur; }
        } else {
            const bool accept = (en < ec) || (urand(seed, cid, sg, 2) < expf((ec - en) / M.temp));
            if (accept) {
                ec = en; cur = x;
                if (en < eb) { eb = en; best = x; }
            }
        }
    }
    if (wid == 0) {
        if (lane < S_) { gcur[lane] = cur; gbest[lane] = best; }
        if (lane == 0) {
            M.cur_e[ch] = ec; M.best_e[ch] = eb;
            if (M.eval_count) atomicAdd(M.eval_count, (unsigned long long)nev);
            if (M.lig_evals && l < M.n_lig_evals) atomicAdd(M.lig_evals + l, (unsigned long long)nev);
        }
    }
    return l;
}

template <int KM, int APL, bool EX>
__global__ __launch_bounds__(256) ADOCK_WAVES(KM, APL) void mc_kernel(LigArgs L, GridArgs G, SFArgs S, RecArgs R,
                                                  const int* __restrict__ chain_lig, int n_chains, McArgs M) {
    extern __shared__ __attribute__((aligned(16))) char lds[];
    const int lane = lane_id();
    const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x / WAVE);
    const int ch = blockIdx.x * (blockDim.x / WAVE) + wid;
    if (ch >= n_chains) return;
    const int wl = wave_lds_bytes(KM, L.P_MAX, APL), wb = wl + stage_bytes(M.NBS, M.NBM, APL);
    WaveLDS<APL> w = carve_lds<KM, APL>(lds + wid * wb, L.P_MAX);
    attach_stage(w, lds + wid * wb + wl, M.NBS, M.NBM);
    run_chain<KM, APL, 1, false, EX>(ch, lane, w, L, G, S, R, chain_lig, M);
}

// One chain per workgroup of NS waves (speculative line search); grid = n_chains.
template <int KM, int NS, bool EX>
__global__ __launch_bounds__(256) void mc_spec_kernel(LigArgs L, GridArgs G, SFArgs S, RecArgs R,
                                                       const int* __restrict__ chain_lig, int n_chains, McArgs M) {
    extern __shared__ __attribute__((aligned(16))) char lds[];
    const int lane = lane_id();
    const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x / WAVE);
    const int ch = blockIdx.x;
    if (ch >= n_chains) return;
    const int wl = wave_lds_bytes_dense(KM, L.P_MAX, L.DEG_MAX), wb = wl + stage_bytes(M.NBS, M.NBM);
    WaveLDS<1> w = carve_lds<KM, 1, true>(lds + wid * wb, L.P_MAX, L.DEG_MAX);
    attach_stage(w, lds + wid * wb + wl, M.NBS, M.NBM);
    run_chain<KM, 1, NS, true, EX>(ch, lane, w, L, G, S, R, chain_lig, M, wid, carve_spec(lds + NS * wb));
}

// Persistent variant: a grid sized to the chip's resident capacity; every wave pulls chain indices (from `order`, e.g.
// longest chains first) with an atomic counter until its queue is empty, then serves the other queues in turn.
// Workgroup b starts on queue b % nq: the dispatcher deals workgroups to the 8 XCDs in turn, so with nq = 8 each queue's
// chains (the host puts the complexes of a few receptors in each) run on one XCD and their grids share its L2. Every
// queue is clamped to [0, n) and a wave leaves a queue for good once it is drained, so the loop ends. Per complex it
// counts finished chains and stamps the 100 MHz real-time clock when the last one ends (times[0] = earliest wave
// start).
// DENSE: the pair gradients of each wave in the dense LDS tile (per-atom rows of DEG_MAX float4 entries, one atom per
// lane) instead of one scalar slot per pair with the per-atom lists read from memory; the same bits either way, so
// the layout is a launch-time choice (the dense tile needs more LDS per wave and no list indirection, but issues
// more VALU instructions: the batches measured so far run faster compact).
template <int KM, int APL, bool EX, bool DENSE = false>
__global__ __launch_bounds__(256) ADOCK_WAVES_D(KM, APL, DENSE) void mc_persistent_kernel(LigArgs L, GridArgs G, SFArgs S, RecArgs R,
                                                             const int* __restrict__ chain_lig, int n_chains, McArgs M,
                                                             QueueArgs Q) {
    static_assert(!DENSE || (APL == 1 && !EX), "the dense tile holds one atom per lane, grid mode");
    extern __shared__ __attribute__((aligned(16))) char lds[];
    const int lane = lane_id();
    const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x / WAVE);
    const int wb = DENSE ? wave_lds_bytes_dense(KM, L.P_MAX, L.DEG_MAX) : wave_lds_bytes(KM, L.P_MAX, APL);
    const WaveLDS<APL> w = carve_lds<KM, APL, DENSE>(lds + wid * wb, L.P_MAX, L.DEG_MAX);
    if (Q.times && lane == 0) atomicMin(Q.times, (unsigned long long)__builtin_amdgcn_s_memrealtime());
    const int nq = Q.qstart && Q.nq > 0 ? Q.nq : 1;
    int q = (int)(blockIdx.x % (unsigned)nq);
    for (int left = nq; left > 0;) {
        const int qs = nq > 1 ? min(max(Q.qstart[q], 0), n_chains) : 0;
        const int qe = nq > 1 ? min(max(Q.qstart[q + 1], qs), n_chains) : n_chains;
        int idx = 0;
        if (lane == 0) idx = atomicAdd(Q.counter + q, 1);
        idx = bcast_i(idx, 0);
        if (idx >= qe - qs) {
            q = q + 1 == nq ? 0 : q + 1;
            --left;
            continue;
        }
        const int pos = qs + idx;
        const int ch = Q.order ? __builtin_amdgcn_readfirstlane(Q.order[pos]) : pos;
        const int l = run_chain<KM, APL, 1, DENSE, EX>(ch, lane, w, L, G, S, R, chain_lig, M);
        if (Q.done && lane == 0 && l < Q.n_done) {
            const int old = atomicAdd(Q.done + l, 1);
            if (old + 1 == Q.lig_chains[l] && Q.times) Q.times[1 + l] = __builtin_amdgcn_s_memrealtime();
        }
    }
}

// MULTI = false: one state per wave. MULTI = true: each wave evaluates per_wave consecutive states and sets the
// ligand up again only when chain_lig changes (what the setup leaves in LDS - the dummy atom, the zero slot, identity
// maps past nt, pair words, slot lists - survives evaluations, as in a chain); it holds more registers, so it pays
// only when there are enough states for several per wave.
template <int KM, int APL, bool EX, bool MULTI>
__global__ __launch_bounds__(256) void eg_kernel(LigArgs L, GridArgs G, SFArgs S, RecArgs R,
                                                  const int* __restrict__ chain_lig, int n_chains,
                                                  const float* __restrict__ state, float* __restrict__ e_out,
                                                  float* __restrict__ g_out, int per_wave) {
    extern __shared__ __attribute__((aligned(16))) char lds[];
    constexpr int SD = 7 + KM, DM = 6 + KM;
    const int lane = lane_id();
    const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x / WAVE);
    const long long wave = (long long)blockIdx.x * (blockDim.x / WAVE) + wid;
    const long long first = MULTI ? wave * per_wave : wave;
    if (first >= n_chains) return;
    const WaveLDS<APL> w = carve_lds<KM, APL>(lds + wid * wave_lds_bytes(KM, L.P_MAX, APL), L.P_MAX);
    const SFConst sf = make_sf(S.prm);
    if constexpr (!MULTI) {
        const int ch = (int)first;
        const ChainCtx<KM, APL> c = setup_chain<KM, APL, false, EX>(__builtin_amdgcn_readfirstlane(chain_lig[ch]), lane,
                                                                    L, G, S, R, w);
        const float x = lane < SD ? state[(long long)ch * SD + lane] : 0.f;
        float g, ea_, x0_, x1_, x2_;
        const float e = eval_eg<KM, APL, false, EX>(c, sf, w, lane, x, g, ea_, x0_, x1_, x2_);
        if (lane < 6) g_out[(long long)ch * DM + lane] = g;
        else if (lane >= 7 && lane < 7 + c.nt) g_out[(long long)ch * DM + lane - 1] = g;
        if (lane == 0) e_out[ch] = e;
        return;
    }
    const int last = (int)min(first + per_wave, (long long)n_chains);
    ChainCtx<KM, APL> c;
    int l_set = -1;
    for (int ch = (int)first; ch < last; ++ch) {
        const int l = __builtin_amdgcn_readfirstlane(chain_lig[ch]);
        if (l != l_set) {
            c = setup_chain<KM, APL, false, EX>(l, lane, L, G, S, R, w);
            l_set = l;
        }
        const float x = lane < SD ? state[(long long)ch * SD + lane] : 0.f;
        float g, ea_, x0_, x1_, x2_;
        const float e = eval_eg<KM, APL, false, EX>(c, sf, w, lane, x, g, ea_, x0_, x1_, x2_);
        if (lane < 6) g_out[(long long)ch * DM + lane] = g;
        else if (lane >= 7 && lane < 7 + c.nt) g_out[(long long)ch * DM + lane - 1] = g;
        if (lane == 0) e_out[ch] = e;
    }
}

inline int cu_count() {
    static int cus = 0;
    if (!cus) {
        int dev = 0;
        (void)hipGetDevice(&dev);
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        if (cus <= 0) cus = 256;
    }
    return cus;
}

// chains per workgroup when not set explicitly: spread small launches one wave per CU (latency), pack large ones 4
// per workgroup (a workgroup's waves share the CU's LDS and instruction cache)
inline int auto_waves(int n_chains) {
    const int cus = cu_count();
    const int w = n_chains / (4 * cus);
    return w >= 4 ? 4 : w >= 2 ? 2 : 1;
}
inline int waves_for(int n_chains) {
    const int w = g_waves_per_block.load(std::memory_order_relaxed);
    return w > 0 ? w : auto_waves(n_chains);
}

// chains per workgroup actually launched: the requested count, lowered until the workgroup's LDS fits
inline int fit_waves(int W, int KM, int P_MAX, int extra = 0, int APL = 1) {
    while (W > 1 && (size_t)W * (wave_lds_bytes(KM, P_MAX, APL) + extra) > (size_t)LDS_PER_WORKGROUP) --W;
    return W;
}

// exact mode: receptor atoms staged per wave: the longest neighbour list (or the whole receptor), padded to 64, at
// most STAGE_MAX; a chain whose list is longer than the staging area reads its atoms from memory
constexpr int STAGE_MAX = 1024;
inline int exact_stage(const McArgs& M, const RecArgs& R) {
    const int n = ((M.nbr ? M.NB_MAX : R.RN_MAX) + WAVE - 1) & ~(WAVE - 1);
    return g_exact_staging && n > 0 ? min(n, STAGE_MAX) : 0;
}

// waves per chain of the speculative line search: 4 while the chains fit one per CU, 2 up to two per CU, else 1;
// lowered until the workgroup's LDS fits
inline int line_search_width(int n_chains, int KM, int P_MAX, int DEG_MAX, int extra = 0) {
    int ns = g_line_search_width;
    if (ns <= 0) ns = n_chains <= cu_count() ? 4 : n_chains <= 2 * cu_count() ? 2 : 1;
    while (ns > 1 && (size_t)ns * (wave_lds_bytes_dense(KM, P_MAX, DEG_MAX) + extra) + SPEC_LDS_BYTES >
                         (size_t)LDS_PER_WORKGROUP)
        ns >>= 1;
    return ns;
}

// dynamic LDS above 64 KB is requested from the runtime once per kernel and size
inline void allow_lds(const void* fn, size_t bytes) {
    if (bytes <= 64 * 1024) return;
    static std::mutex mu;
    static std::map<const void*, size_t> granted;
    std::lock_guard<std::mutex> lock(mu);
    size_t& g = granted[fn];
    if (bytes > g) {
        (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
        g = bytes;
    }
}

// chains per workgroup, at most W0, that keep the most chains resident per CU at wb LDS bytes per wave, from the
// runtime's occupancy calculator: once the per-wave LDS is large, four waves per workgroup leave LDS unused that one
// or two would fill. Ties go to the larger workgroup (equal chains share its instruction stream: the widened single
// complex runs 4 % faster), or with prefer_small to the smaller one, whose slot frees as soon as its own chains end
// (uneven chains: the exact polish of many ligands runs 8 % faster). Cached per (device, kernel, W0, wb).
inline int packed_waves(const void* fn, int W0, size_t wb, bool prefer_small = false) {
    if (W0 <= 1) return W0;
    static std::mutex mu;
    static std::map<std::tuple<int, const void*, int, size_t>, int> cache;
    int dev = 0;
    (void)hipGetDevice(&dev);
    const auto key = std::make_tuple(dev, fn, W0, wb);
    {
        std::lock_guard<std::mutex> lock(mu);
        const auto it = cache.find(key);
        if (it != cache.end()) return it->second;
    }
    int best_w = W0, best_res = -1;
    for (int w = W0; w >= 1; --w) {
        if ((size_t)w * wb > (size_t)LDS_PER_WORKGROUP) continue;
        allow_lds(fn, (size_t)w * wb);
        int per_cu = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, w * WAVE, (size_t)w * wb) != hipSuccess) continue;
        if (per_cu * w > best_res || (prefer_small && per_cu * w == best_res)) { best_res = per_cu * w; best_w = w; }
    }
    std::lock_guard<std::mutex> lock(mu);
    cache[key] = best_w;
    return best_w;
}

template <int KM, int APL, bool EX>
int launch_mc_tx(const LigArgs& L, const GridArgs& G, const SFArgs& S, const RecArgs& R, const int* chain_lig,
                 int n_chains, const McArgs& M0, hipStream_t stream) {
    McArgs M = M0;
    M.NBS = EX ? exact_stage(M, R) : 0;
    M.NBM = M.NBS > 0 && g_exact_staging == 2 ? 1 : 0;
    const int extra = stage_bytes(M.NBS, M.NBM, APL);
    // the speculative line search runs the dense one-atom-per-lane layout
    const int ns = APL == 1 ? line_search_width(n_chains, KM, L.P_MAX, L.DEG_MAX, extra) : 1;
    if constexpr (APL == 1) {
        if (ns > 1) {
            const size_t shm = (size_t)ns * (wave_lds_bytes_dense(KM, L.P_MAX, L.DEG_MAX) + extra) + SPEC_LDS_BYTES;
            if (ns == 4) {
                allow_lds(reinterpret_cast<const void*>(mc_spec_kernel<KM, 4, EX>), shm);
                hipLaunchKernelGGL((mc_spec_kernel<KM, 4, EX>), dim3(n_chains), dim3(4 * WAVE), shm, stream, L, G, S,
                                   R, chain_lig, n_chains, M);
            } else {
                allow_lds(reinterpret_cast<const void*>(mc_spec_kernel<KM, 2, EX>), shm);
                hipLaunchKernelGGL((mc_spec_kernel<KM, 2, EX>), dim3(n_chains), dim3(2 * WAVE), shm, stream, L, G, S,
                                   R, chain_lig, n_chains, M);
            }
            return (int)hipGetLastError();
        }
    }
    const int wave_bytes = wave_lds_bytes(KM, L.P_MAX, APL) + extra;
    int W = fit_waves(waves_for(n_chains), KM, L.P_MAX, extra, APL);
    if (g_waves_per_block.load(std::memory_order_relaxed) <= 0)
        W = packed_waves(reinterpret_cast<const void*>(mc_kernel<KM, APL, EX>), W, wave_bytes, EX);
    const size_t shm = (size_t)W * wave_bytes;
    const unsigned nblk = (unsigned)((n_chains + W - 1) / W);
    allow_lds(reinterpret_cast<const void*>(mc_kernel<KM, APL, EX>), shm);
    hipLaunchKernelGGL((mc_kernel<KM, APL, EX>), dim3(nblk), dim3(W * WAVE), shm, stream, L, G, S, R, chain_lig,
                       n_chains, M);
    return (int)hipGetLastError();
}

template <int KM, int APL, bool EX, bool DENSE = false>
int launch_mc_persistent_tx(const LigArgs& L, const GridArgs& G, const SFArgs& S, const RecArgs& R,
                            const int* chain_lig, int n_chains, const McArgs& M, const QueueArgs& Q, int blocks,
                            hipStream_t stream) {
    if constexpr (APL == 1 && !EX && !DENSE) {
        if (Q.dense) return launch_mc_persistent_tx<KM, 1, false, true>(L, G, S, R, chain_lig, n_chains, M, Q, blocks,
                                                                        stream);
    }
    const int wave_bytes = DENSE ? wave_lds_bytes_dense(KM, L.P_MAX, L.DEG_MAX) : wave_lds_bytes(KM, L.P_MAX, APL);
    const int req = Q.waves > 0 ? Q.waves : g_waves_per_block.load(std::memory_order_relaxed);
    int W = req > 0 ? req : 4;
    while (W > 1 && (size_t)W * wave_bytes > (size_t)LDS_PER_WORKGROUP) --W;
    if (req <= 0)
        W = packed_waves(reinterpret_cast<const void*>(mc_persistent_kernel<KM, APL, EX, DENSE>), W, wave_bytes);
    const size_t shm = (size_t)W * wave_bytes;
    if (blocks <= 0) {
        // resident capacity, queried once per (device, kernel, W, LDS size) so later launches (and graph captures)
        // only launch
        static std::mutex mu;
        static std::map<long long, int> cap;
        int dev = 0;
        (void)hipGetDevice(&dev);
        const long long key = ((long long)dev << 41) | ((long long)DENSE << 40) | ((long long)W << 32) | (long long)shm;
        std::lock_guard<std::mutex> lock(mu);
        auto it = cap.find(key);
        if (it == cap.end()) {
            int cus = 0, per_cu = 0;
            (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
            (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, mc_persistent_kernel<KM, APL, EX, DENSE>,
                                                               W * WAVE, shm);
            it = cap.emplace(key, cus * (per_cu > 0 ? per_cu : 1)).first;
        }
        blocks = it->second;
    }
    const int need = (n_chains + W - 1) / W;
    if (blocks > need) blocks = need;
    allow_lds(reinterpret_cast<const void*>(mc_persistent_kernel<KM, APL, EX, DENSE>), shm);
    hipLaunchKernelGGL((mc_persistent_kernel<KM, APL, EX, DENSE>), dim3(blocks), dim3(W * WAVE), shm, stream, L, G, S, R,
                       chain_lig, n_chains, M, Q);
    return (int)hipGetLastError();
}

constexpr int EG_PER_WAVE = 8, EG_MULTI_WAVES_PER_CU = 16;

template <int KM, int APL, bool EX>
int launch_eg_tx(const LigArgs& L, const GridArgs& G, const SFArgs& S, const RecArgs& R, const int* chain_lig,
                 int n_chains, const float* state, float* e_out, float* g_out, hipStream_t stream) {
    const int W = fit_waves(waves_for(n_chains), KM, L.P_MAX, 0, APL);
    const size_t shm = (size_t)W * wave_lds_bytes(KM, L.P_MAX, APL);
    // from EG_MULTI_WAVES_PER_CU x CUs waves' worth of EG_PER_WAVE states, several states per wave (65,536 poses of one
    // complex: 250 -> 284 M evaluations/s)
    if ((long long)n_chains >= (long long)EG_MULTI_WAVES_PER_CU * cu_count() * EG_PER_WAVE) {
        const long long waves = ((long long)n_chains + EG_PER_WAVE - 1) / EG_PER_WAVE;
        const unsigned nblk = (unsigned)((waves + W - 1) / W);
        allow_lds(reinterpret_cast<const void*>(eg_kernel<KM, APL, EX, true>), shm);
        hipLaunchKernelGGL((eg_kernel<KM, APL, EX, true>), dim3(nblk), dim3(W * WAVE), shm, stream, L, G, S, R,
                           chain_lig, n_chains, state, e_out, g_out, EG_PER_WAVE);
        return (int)hipGetLastError();
    }
    const unsigned nblk = (unsigned)((n_chains + W - 1) / W);
    allow_lds(reinterpret_cast<const void*>(eg_kernel<KM, APL, EX, false>), shm);
    hipLaunchKernelGGL((eg_kernel<KM, APL, EX, false>), dim3(nblk), dim3(W * WAVE), shm, stream, L, G, S, R,
                       chain_lig, n_chains, state, e_out, g_out, 1);
    return (int)hipGetLastError();
}

// resident waves per CU of the grid-mode chain kernels at a launch's bucket and per-wave LDS: the per-chain kernel
// (mc_kernel, as launched for many chains) or the persistent one, from the runtime's occupancy calculator
template <int KM, int APL>
int mc_resident_waves_t(int P_MAX, int persistent) {
    const int wave_bytes = wave_lds_bytes(KM, P_MAX, APL);
    const int req = g_waves_per_block.load(std::memory_order_relaxed);
    int W = fit_waves(persistent ? (req > 0 ? req : 4) : waves_for(1 << 20), KM, P_MAX, 0, APL);
    const void* fn = persistent ? reinterpret_cast<const void*>(mc_persistent_kernel<KM, APL, false>)
                                : reinterpret_cast<const void*>(mc_kernel<KM, APL, false>);
    if (req <= 0) W = packed_waves(fn, W, wave_bytes);
    const size_t shm = (size_t)W * wave_bytes;
    allow_lds(fn, shm);
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, W * WAVE, shm) != hipSuccess) return -1;
    return per_cu * W;
}

// move record: features before the move [128], target [23], mask [23], energy gain, entity moved (-1: a network
// displacement move), the chain's best energy minus its current one before the move (<= 0)
constexpr int NFEAT = NET_IN, NOUT = NET_OUT, LOG_W = NFEAT + 2 * NOUT + 3;
// Counter-RNG slots of the learned step. The plain move uses slots 0 (entity), 1 (torsion angle), 2 (acceptance) and
// 10 .. 105 (unit-ball tries 10 + 3k, k < 32); the step counter sits 8 bits up, so every slot must stay below 256 or
// it aliases a slot of another step. The network move draws its coin at SLOT_NET_PLAIN and one normal per lane from
// the pair SLOT_NET_NOISE + 2 lane (+ 1): all disjoint from the plain move's slots and below 256.
constexpr int SLOT_NET_PLAIN = 108, SLOT_NET_ENTITY = 109, SLOT_NET_NOISE = 110;
static_assert(SLOT_NET_PLAIN > 10 + 3 * 31 + 2 && SLOT_NET_ENTITY < SLOT_NET_NOISE, "net coins overlap the plain move");
static_assert(SLOT_NET_NOISE + 2 * (WAVE - 1) + 1 < 256, "net noise slots alias the next step's counter");


DEVI int wave_count(bool p) { return __popcll(__ballot(p)); }

// standard normal from two counter-RNG draws (Box-Muller), per lane
DEVI float rnorm(unsigned long long seed, int cid, int sg, int slot) {
    const float u1 = fmaxf(urand(seed, cid, sg, slot), 1e-7f), u2 = urand(seed, cid, sg, slot + 1);
    return sqrtf(-2.f * logf(u1)) * cosf(6.28318531f * u2);
}

// The 128 proposal-network features of the pose x (evaluated: energy E, lane-distributed gradient g, per-atom
// energy ea): box-normalised translation, canonical quaternion, sin/cos of the torsions, squashed gradient, energy
// summaries, a 16-bin histogram of atom energies, per-octant (ligand frame) mean / max atom energy, ligand size
// descriptors and the mean energy of the atoms each torsion moves. Written as f[lane], f[64 + lane].
template <int KM>
DEVI void write_features(float* f, const ChainCtx<KM, 1>& c, int lane, float x, float E, float g, float ea) {
    float lo_v = 0.f, hi_v = 0.f;
    const float cen[3] = {0.5f * (c.lo0 + c.hi0), 0.5f * (c.lo1 + c.hi1), 0.5f * (c.lo2 + c.hi2)};
    const float hal[3] = {fmaxf(0.5f * (c.hi0 - c.lo0), 1e-3f), fmaxf(0.5f * (c.hi1 - c.lo1), 1e-3f),
                          fmaxf(0.5f * (c.hi2 - c.lo2), 1e-3f)};
    if (lane < 3) lo_v = (x - cen[lane == 0 ? 0 : lane == 1 ? 1 : 2]) / hal[lane == 0 ? 0 : lane == 1 ? 1 : 2];
    float qw = bcast(x, 3), qx = bcast(x, 4), qy = bcast(x, 5), qz = bcast(x, 6);
    const float qn = fast_rsq(qw * qw + qx * qx + qy * qy + qz * qz) * (qw < 0.f ? -1.f : 1.f);
    if (lane >= 3 && lane < 7) lo_v = x * qn;
    const float xc = __shfl(x, lane >= 23 && lane < 39 ? lane - 16 : 0);
    if (lane >= 7 && lane < 23 && lane - 7 < c.nt) lo_v = sinf(x);
    if (lane >= 23 && lane < 39 && lane - 23 < c.nt) lo_v = cosf(xc);
    const int gsrc = lane >= 39 && lane < 45 ? lane - 39 : lane >= 45 && lane < 61 ? lane - 38 : 0;
    const float gv = __shfl(g, gsrc);
    if (lane >= 39 && lane < 61) lo_v = tanhf(0.1f * gv);
    const float esum = wave_sum(c.atom[0] ? ea : 0.f);
    const float inv_na = 1.f / (float)max(c.na, 1);
    const int nout = wave_count(c.atom[0] && ea > 1.f);
    if (lane == 61) lo_v = tanhf(0.1f * E);
    if (lane == 62) lo_v = tanhf(0.1f * esum);
    if (lane == 63) lo_v = (float)nout * inv_na;
    // 16-bin histogram of atom energies
    const float edges[15] = {-1.2f, -0.9f, -0.7f, -0.5f, -0.35f, -0.2f, -0.1f, 0.f, 0.1f, 0.25f, 0.5f, 1.f, 2.f, 5.f, 20.f};
    int bin = 0;
#pragma unroll
    for (int b = 0; b < 15; ++b) bin += ea > edges[b] ? 1 : 0;
#pragma unroll
    for (int b = 0; b < 16; ++b) {
        const int cnt = wave_count(c.atom[0] && bin == b);
        if (lane == b) hi_v = (float)cnt * inv_na;
    }
    // ligand-frame octants of the reference coordinates
    const int oct = (c.rx[0] > 0.f ? 1 : 0) | (c.ry[0] > 0.f ? 2 : 0) | (c.rz[0] > 0.f ? 4 : 0);
#pragma unroll
    for (int o = 0; o < 8; ++o) {
        const bool in = c.atom[0] && oct == o;
        const int cnt = wave_count(in);
        const float sm = wave_sum(in ? ea : 0.f);
        const float mx = wave_max_f(in ? ea : -INFINITY);
        if (lane == 16 + o) hi_v = cnt ? tanhf(sm / (float)cnt) : 0.f;
        if (lane == 24 + o) hi_v = cnt ? tanhf(mx) : 0.f;
    }
    if (lane == 32) hi_v = (float)c.na / 64.f;
    if (lane == 33) hi_v = (float)c.nt / 16.f;
    if (lane == 34) hi_v = c.rg / 5.f;
    if (lane == 35) hi_v = (float)c.np / 1024.f;
    for (int k = 0; k < c.nt && k < 16; ++k) {
        const bool in = c.atom[0] && ((c.tmask[0] >> k) & 1u);
        const int cnt = wave_count(in);
        const float sm = wave_sum(in ? ea : 0.f);
        if (lane == 48 + k) hi_v = cnt ? tanhf(sm / (float)cnt) : 0.f;
    }
    f[lane] = lo_v;
    f[64 + lane] = hi_v;
}

// Move proposed by the network output y (lane-distributed like the state): translation + y_t + sig_t N(0,1),
// rotation normalise(y_q) * q followed by a random rotation vector sig_r N(0, I), torsions + y_theta + sig_th N(0,1).
template <class Ctx>
DEVI float net_move(float x, float y, const Ctx& c, int lane, unsigned long long seed, int cid, int sg,
                    float sig_t, float sig_r, float sig_th) {
    const float n = rnorm(seed, cid, sg, SLOT_NET_NOISE + 2 * lane);
    float out = x;
    if (lane < 3) out = x + y + sig_t * n;
    if (lane >= 7 && lane < 7 + c.nt) out = x + y + sig_th * n;
    float aw = bcast(y, 3), ax = bcast(y, 4), ay = bcast(y, 5), az = bcast(y, 6);
    const float an = aw * aw + ax * ax + ay * ay + az * az;
    if (an > 1e-6f) {
        const float r = fast_rsq(an);
        aw *= r; ax *= r; ay *= r; az *= r;
    } else { aw = 1.f; ax = ay = az = 0.f; }
    const float bw = bcast(x, 3), bx = bcast(x, 4), by = bcast(x, 5), bz = bcast(x, 6);
    float qw = aw * bw - ax * bx - ay * by - az * bz;
    float qx = aw * bx + ax * bw + ay * bz - az * by;
    float qy = aw * by - ax * bz + ay * bw + az * bx;
    float qz = aw * bz + ax * by - ay * bx + az * bw;
    quat_inc(qw, qx, qy, qz, sig_r * bcast(n, 3), sig_r * bcast(n, 4), sig_r * bcast(n, 5));
    out = lane == 3 ? qw : lane == 4 ? qx : lane == 5 ? qy : lane == 6 ? qz : out;
    return out;
}

// Entity-choice network move (net_mode 1): the network's outputs score the entities (translation: output 0,
// rotation: output 3, torsion k < 16: output 7 + k, later torsions 0), one is drawn with probability proportional to
// exp(score / tau) (slot SLOT_NET_ENTITY), and the reference's perturbation of that entity is applied with its own
// slots. The scores are the network's estimates of each entity's expected energy gain.
template <class Ctx>
DEVI int net_entity(float y, const Ctx& c, int lane, unsigned long long seed, int cid, int sg, float tau) {
    const int ne = 2 + c.nt;
    const int src = lane == 0 ? 0 : lane == 1 ? 3 : lane - 2 + 7;
    const float sc = __shfl(y, src < 64 ? src : 0);
    const bool valid = lane < ne;
    const float scv = (lane >= 2 && lane - 2 >= 16) ? 0.f : sc;
    const float mx = wave_max_f(valid ? scv : -INFINITY);
    const float w = valid ? expf((scv - mx) / fmaxf(tau, 1e-6f)) : 0.f;
    const float total = wave_sum(w);
    // inclusive prefix of the weights over the entity lanes, then the first entity whose prefix passes u * total
    float pre = w;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const float o = __shfl_up(pre, d);
        if (lane >= d) pre += o;
    }
    const float u = urand(seed, cid, sg, SLOT_NET_ENTITY) * total;
    const unsigned long long hit = __ballot(valid && pre > u);
    return hit ? __builtin_ctzll(hit) : ne - 1;
}

// target of an accepted move cur -> x: translation delta, rotation dq = q_x conj(q_cur) (w >= 0), wrapped torsion
// deltas; mask 1 on the components that exist
template <class Ctx>
DEVI void move_target(float cur, float x, const Ctx& c, int lane, float& t, float& m) {
    float aw = bcast(x, 3), ax = bcast(x, 4), ay = bcast(x, 5), az = bcast(x, 6);
    float bw = bcast(cur, 3), bx = -bcast(cur, 4), by = -bcast(cur, 5), bz = -bcast(cur, 6);
    const float na = fast_rsq(aw * aw + ax * ax + ay * ay + az * az), nb = fast_rsq(bw * bw + bx * bx + by * by + bz * bz);
    aw *= na; ax *= na; ay *= na; az *= na; bw *= nb; bx *= nb; by *= nb; bz *= nb;
    float qw = aw * bw - ax * bx - ay * by - az * bz;
    float qx = aw * bx + ax * bw + ay * bz - az * by;
    float qy = aw * by - ax * bz + ay * bw + az * bx;
    float qz = aw * bz + ax * by - ay * bx + az * bw;
    const float sg = qw < 0.f ? -1.f : 1.f;
    t = 0.f; m = 0.f;
    if (lane < 3) { t = x - cur; m = 1.f; }
    if (lane == 3) { t = qw * sg; m = 1.f; }
    if (lane == 4) { t = qx * sg; m = 1.f; }
    if (lane == 5) { t = qy * sg; m = 1.f; }
    if (lane == 6) { t = qz * sg; m = 1.f; }
    if (lane >= 7 && lane < 7 + c.nt) {
        float d = x - cur;
        d = d - 6.28318531f * rintf(d * 0.159154943f);
        t = d; m = 1.f;
    }
}

// One move record at *log_count (atomic; records past log_cap are dropped): the features before the move (fin, 128
// floats, memory or LDS), the move cur -> x as target and mask, the energy gain ec - en, the entity moved and
// eb - ec (so gain + eb - ec > 0 marks a move that found a new chain best).
template <class Ctx>
DEVI void log_move(const StepArgs& P, const float* fin, float cur, float x, const Ctx& c, int lane, float gain,
                   int ent, float best_gap) {
    int slot = 0;
    if (lane == 0) slot = atomicAdd(P.log_count, 1);
    slot = bcast_i(slot, 0);
    if (slot >= P.log_cap) return;
    float* rec = P.log_buf + (long long)slot * LOG_W;
    rec[lane] = fin[lane];
    rec[64 + lane] = fin[64 + lane];
    float t, m;
    move_target(cur, x, c, lane, t, m);
    if (lane < NOUT) { rec[NFEAT + lane] = t; rec[NFEAT + NOUT + lane] = m; }
    if (lane == 0) {
        rec[NFEAT + 2 * NOUT] = gain;
        rec[NFEAT + 2 * NOUT + 1] = (float)ent;
        rec[NFEAT + 2 * NOUT + 2] = best_gap;
    }
}

// the move of one step: the reference's random mutation (plain), the network's displacement (net_mode 0) or the
// reference's mutation of the entity the network picks (net_mode 1); ent = the entity moved, -1 for a displacement
template <class Ctx>
DEVI float step_move(float cur, bool plain, float y, const Ctx& c, int lane, unsigned long long seed, int cid, int sg,
                     const McArgs& M, const StepArgs& P, int& ent) {
    if (plain || P.net_mode == 1) {
        ent = plain ? plain_entity(c, seed, cid, sg) : net_entity(y, c, lane, seed, cid, sg, P.tau);
        return entity_move(cur, c, lane, seed, cid, sg, ent, M.amp_pos, M.amp_rot);
    }
    ent = -1;
    return net_move(cur, y, c, lane, seed, cid, sg, P.sig_t, P.sig_r, P.sig_th);
}

// One MC step of every chain (launched once per step): optional features at the current pose (feat), a move from the
// network (Y) or the reference's random mutation (fallback with probability p_plain, or when Y is not finite), BFGS,
// Metropolis, logging of accepted moves (features before the move, target, mask, energy gain) and the features of the
// new current pose. init = 1: BFGS of the start pose only (the reference's init_opt). With use_net = 0 and init_opt
// followed by steps 0..S-1 this reproduces the fused plain kernel bit for bit.
template <int KM>
__global__ __launch_bounds__(256) ADOCK_WAVES(KM, 1) void mc_step_kernel(LigArgs L, GridArgs G, SFArgs S, RecArgs R,
                                                       const int* __restrict__ chain_lig, int n_chains, McArgs M,
                                                       StepArgs P) {
    extern __shared__ __attribute__((aligned(16))) char lds[];
    constexpr int SD = 7 + KM;
    const int lane = lane_id();
    const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x / WAVE);
    const int ch = blockIdx.x * (blockDim.x / WAVE) + wid;
    if (ch >= n_chains) return;
    const WaveLDS<1> w = carve_lds<KM, 1>(lds + wid * wave_lds_bytes(KM, L.P_MAX), L.P_MAX);
    const SFConst sf = make_sf(S.prm);
    const int l = __builtin_amdgcn_readfirstlane(chain_lig[ch]);
    const ChainCtx<KM, 1> c = setup_chain<KM, 1>(l, lane, L, G, S, R, w);
    const int cid = __builtin_amdgcn_readfirstlane(M.chain_id[ch]);
    const int S_ = 7 + c.nt;
    const int iters = M.bfgs_iters > 0 ? M.bfgs_iters : (25 + c.na) / 3;
    float* gcur = M.cur + (long long)ch * SD;
    float* gbest = M.best + (long long)ch * SD;
    float cur = lane < S_ ? gcur[lane] : 0.f;
    float best = lane < S_ ? gbest[lane] : 0.f;
    float eb = M.best_e[ch], ec = M.cur_e[ch];
    int nev = 0;
    const unsigned long long seed = M.seed;
    const int sg = M.step_offset;
    if (P.init) {
        float x = cur;
        ec = bfgs<KM, 1>(c, sf, w, lane, x, iters, nev);
        cur = x;
        if (ec < eb) { eb = ec; best = cur; }
    } else {
        bool plain = !P.use_net || !P.Y || urand(seed, cid, sg, SLOT_NET_PLAIN) < P.p_plain;
        float y = 0.f;
        if (!plain) {
            y = lane < NOUT ? P.Y[(long long)ch * NOUT + lane] : 0.f;
            plain = wave_count(!isfinite(y)) > 0;
        }
        int ent = -1;
        float x = step_move(cur, plain, y, c, lane, seed, cid, sg, M, P, ent);
        const float en = bfgs<KM, 1>(c, sf, w, lane, x, iters, nev);
        const bool accept = (en < ec) || (urand(seed, cid, sg, 2) < expf((ec - en) / M.temp));
        if ((accept || P.log_all) && P.log_buf && P.feat)
            log_move(P, P.feat + (long long)ch * NFEAT, cur, x, c, lane, ec - en, ent, eb - ec);
        if (accept) {
            ec = en; cur = x;
            if (en < eb) { eb = en; best = x; }
        }
    }
    if (P.feat) {
        float g, ea, a0, a1, a2;
        const float e = eval_eg<KM, 1>(c, sf, w, lane, cur, g, ea, a0, a1, a2);
        ++nev;
        write_features(P.feat + (long long)ch * NFEAT, c, lane, cur, e, g, ea);
    }
    if (lane < S_) { gcur[lane] = cur; gbest[lane] = best; }
    if (lane == 0) {
        M.cur_e[ch] = ec; M.best_e[ch] = eb;
        if (M.eval_count) atomicAdd(M.eval_count, (unsigned long long)nev);
    }
}

template <int KM>
int launch_mc_step_t(const LigArgs& L, const GridArgs& G, const SFArgs& S, const RecArgs& R, const int* chain_lig,
                     int n_chains, const McArgs& M, const StepArgs& P, hipStream_t stream) {
    const int W = fit_waves(waves_for(n_chains), KM, L.P_MAX);
    const size_t shm = (size_t)W * wave_lds_bytes(KM, L.P_MAX);
    const unsigned nblk = (unsigned)((n_chains + W - 1) / W);
    allow_lds(reinterpret_cast<const void*>(mc_step_kernel<KM>), shm);
    hipLaunchKernelGGL(mc_step_kernel<KM>, dim3(nblk), dim3(W * WAVE), shm, stream, L, G, S, R, chain_lig, n_chains,
                       M, P);
    return (int)hipGetLastError();
}


// ---- the proposal network inside the chain loop ----
// One chain's forward pass on its wave's own MFMAs, the operation of mlp_infer_kernel for one row: every row of the
// A operand is this chain's activation vector (BF16, the wave's LDS), B a 16-output tile of W^T read from the
// L2-resident weights ([out, in]: each lane's 8 consecutive K values are one 16-byte load). Each output is the same
// K-ordered chain of v_mfma_f32_16x16x32_bf16, the same bias add, ReLU and BF16 rounding as in the batched kernel, so
// the move equals the batched inference's bit for bit. D[i][j] sits on lane j + 16 (i / 4), element i % 4; with all
// rows equal, lane j of every 16-lane quarter holds output j of the tile. 208 MFMAs per forward pass.
struct NetLDS { float* feat; unsigned short* a0; unsigned short* a1; };   // feat [128] FP32, a0 / a1 [256] BF16
constexpr int NET_LDS_BYTES = 4 * NET_IN + 2 * 2 * NET_HID;

DEVI NetLDS carve_net(char* base) {
    float* f = reinterpret_cast<float*>(base);
    unsigned short* a = reinterpret_cast<unsigned short*>(f + NET_IN);
    return NetLDS{f, a, a + NET_HID};
}

// out[0 .. N) = act(W in + b) for the wave's vector in[0 .. K) (LDS, BF16), out written as BF16 by lanes 0-15
template <int K, int N>
DEVI void net_hidden(const unsigned short* in, unsigned short* out, const unsigned short* __restrict__ W,
                     const float* __restrict__ b, int lane) {
    const int li = lane & 15, lk = (lane >> 4) * 8;
#pragma unroll 2
    for (int t = 0; t < N / 16; ++t) {
        const int j = t * 16 + li;
        const unsigned short* wr = W + (long long)j * K + lk;
        float4_t acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int k0 = 0; k0 < K; k0 += 32) {
            const bf16x8_t a = *reinterpret_cast<const bf16x8_t*>(in + k0 + lk);
            const bf16x8_t bv = *reinterpret_cast<const bf16x8_t*>(wr + k0);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bv, acc, 0, 0, 0);
        }
        const float v = fmaxf(acc[0] + b[j], 0.f);
        if (lane < 16) out[j] = f2bf(v);
    }
    wave_lds_sync();
}

// the network's move for the features in n.feat: component j on lane j (j < 23), 0 on the other lanes
DEVI float net_forward(const NetLDS& n, const NetArgs& N, int lane) {
    n.a0[lane] = f2bf(n.feat[lane]);
    n.a0[64 + lane] = f2bf(n.feat[64 + lane]);
    wave_lds_sync();
    net_hidden<NET_IN, NET_HID>(n.a0, n.a1, N.W1, N.b1, lane);
    net_hidden<NET_HID, NET_HID>(n.a1, n.a0, N.W2, N.b2, lane);
    // last layer: outputs 0-15 and 16-31 (23 valid; rows past 22 of W3 do not exist and enter as zeros)
    const int li = lane & 15, lk = (lane >> 4) * 8;
    const bool hi_ok = 16 + li < NET_OUT;
    const unsigned short* w0 = N.W3 + (long long)li * NET_HID + lk;
    const unsigned short* w1 = N.W3 + (long long)(hi_ok ? 16 + li : 0) * NET_HID + lk;
    float4_t acc0 = {0.f, 0.f, 0.f, 0.f}, acc1 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int k0 = 0; k0 < NET_HID; k0 += 32) {
        const bf16x8_t a = *reinterpret_cast<const bf16x8_t*>(n.a0 + k0 + lk);
        const bf16x8_t b0 = *reinterpret_cast<const bf16x8_t*>(w0 + k0);
        bf16x8_t b1 = {};
        if (hi_ok) b1 = *reinterpret_cast<const bf16x8_t*>(w1 + k0);
        acc0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b0, acc0, 0, 0, 0);
        acc1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b1, acc1, 0, 0, 0);
    }
    // lane j < 16: output j (first tile); lane 16 + i: output 16 + i (second tile, quarter 1)
    float y = 0.f;
    if (lane < 16) y = acc0[0] + N.b3[lane];
    else if (lane < NET_OUT) y = acc1[0] + N.b3[lane];
    return y;
}

// The learned search in one launch: each wave runs its chain's init BFGS and mc_steps MC steps with the chain set up
// once; before a step whose coin (slot SLOT_NET_PLAIN) calls for the network, the wave runs the network on the
// features of its current pose (net_forward), and after an accepted move it takes the new pose's features from the
// evaluation BFGS ended on (no extra evaluation; a fresh one only when BFGS restored its start). Moves, BFGS,
// Metropolis test and move log are those of mc_step_kernel, so a launch equals the step-by-step search (init step,
// then adock_mlp_infer + one step launch per MC step) bit for bit in poses, energies and logged records (in another
// order); it skips the step kernel's extra evaluation per step (eval_count differs). With use_net = 0 and no log it
// is mc_kernel's search. P.feat (may be NULL) receives the features of every chain's final current pose.
template <int KM>
DEVI void learned_chain(int ch, int lane, const WaveLDS<1>& w, const NetLDS& nl, const LigArgs& L, const GridArgs& G,
                        const SFArgs& S, const RecArgs& R, const int* __restrict__ chain_lig, const McArgs& M,
                        const StepArgs& P, const NetArgs& N) {
    constexpr int SD = 7 + KM;
    const SFConst sf = make_sf(S.prm);
    const int l = __builtin_amdgcn_readfirstlane(chain_lig[ch]);
    const ChainCtx<KM, 1> c = setup_chain<KM, 1>(l, lane, L, G, S, R, w);
    const int cid = __builtin_amdgcn_readfirstlane(M.chain_id[ch]);
    const int S_ = 7 + c.nt;
    const int iters = M.bfgs_iters > 0 ? M.bfgs_iters : (25 + c.na) / 3;
    const bool use_net = P.use_net != 0;
    const bool feat_on = use_net || P.log_buf != nullptr || P.feat != nullptr;
    float* gcur = M.cur + (long long)ch * SD;
    float* gbest = M.best + (long long)ch * SD;
    float cur = lane < S_ ? gcur[lane] : 0.f;
    float best = lane < S_ ? gbest[lane] : 0.f;
    float eb = M.best_e[ch], ec = M.cur_e[ch];
    int nev = 0;
    const unsigned long long seed = M.seed;
    // One BFGS call site for the start optimisation (st = -1), the MC steps and the evaluation that gives the features
    // of the current pose when no kept evaluation has them (fresh: BFGS with no iterations evaluates cur once), so the
    // chain body holds one copy of the evaluation code.
    bool fresh = feat_on && !M.init_opt;
    // the network's output at the current pose: a rejected move leaves the pose and its features as they were, so
    // the next network step reuses it instead of running the network again (the same values)
    float ynet = 0.f;
    bool ynet_ok = false;
    if (!M.init_opt && ec < eb) { eb = ec; best = cur; }
    for (int st = M.init_opt ? -1 : 0; st < M.mc_steps || fresh;) {
        const int sg = M.step_offset + (st < 0 ? 0 : st);
        float x = cur;
        bool plain = true;
        int ent = -1;
        if (!fresh && st >= 0) {
            plain = !use_net || urand(seed, cid, sg, SLOT_NET_PLAIN) < P.p_plain;
            float y = 0.f;
            if (!plain) {
                if (!ynet_ok) { ynet = net_forward(nl, N, lane); ynet_ok = true; }
                y = ynet;
                plain = wave_count(!isfinite(y)) > 0;
            }
            x = step_move(cur, plain, y, c, lane, seed, cid, sg, M, P, ent);
        }
        LastEval le;
        const float en = bfgs<KM, 1, false, false, true>(c, sf, w, lane, x, fresh ? 0 : iters, nev, &le);
        bool moved = false;   // cur is a new pose (its features are due)
        if (fresh) {
            fresh = false;
            moved = true;
        } else if (st < 0) {
            ec = en; cur = x;
            if (ec < eb) { eb = ec; best = cur; }
            moved = true;
            ++st;
        } else {
            const bool accept = (en < ec) || (urand(seed, cid, sg, 2) < expf((ec - en) / M.temp));
            if ((accept || P.log_all) && P.log_buf) log_move(P, nl.feat, cur, x, c, lane, ec - en, ent, eb - ec);
            if (accept) {
                ec = en; cur = x;
                if (en < eb) { eb = en; best = x; }
                moved = true;
            }
            ++st;
        }
        if (moved) ynet_ok = false;
        if (moved && feat_on) {
            if (le.valid) {
                write_features(nl.feat, c, lane, cur, en, le.g, le.ea);
                wave_lds_sync();
            } else {
                fresh = true;   // BFGS restored its start: evaluate cur before the next move
            }
        }
    }
    if (P.feat) {
        P.feat[(long long)ch * NFEAT + lane] = nl.feat[lane];
        P.feat[(long long)ch * NFEAT + 64 + lane] = nl.feat[64 + lane];
    }
    if (lane < S_) { gcur[lane] = cur; gbest[lane] = best; }
    if (lane == 0) {
        M.cur_e[ch] = ec; M.best_e[ch] = eb;
        if (M.eval_count) atomicAdd(M.eval_count, (unsigned long long)nev);
        if (M.lig_evals && l < M.n_lig_evals) atomicAdd(M.lig_evals + l, (unsigned long long)nev);
    }
}

// One chain per wave (Q.counter NULL: grid = the chains), or persistent (grid = the resident capacity): every wave
// pulls positions of Q.order (NULL: identity), e.g. longest chains first, from the counter Q.counter[0] (0 at launch)
// until all n_chains are taken, so a wave that ends a short chain starts the next one instead of leaving its slot
// idle behind its workgroup's longest chain.
template <int KM>
__global__ __launch_bounds__(256) ADOCK_WAVES(KM, 1) void mc_learned_kernel(LigArgs L, GridArgs G, SFArgs S, RecArgs R,
                                                          const int* __restrict__ chain_lig, int n_chains, McArgs M,
                                                          StepArgs P, NetArgs N, QueueArgs Q) {
    extern __shared__ __attribute__((aligned(16))) char lds[];
    const int lane = lane_id();
    const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x / WAVE);
    const int wl = wave_lds_bytes(KM, L.P_MAX);
    const WaveLDS<1> w = carve_lds<KM, 1>(lds + wid * (wl + NET_LDS_BYTES), L.P_MAX);
    const NetLDS nl = carve_net(lds + wid * (wl + NET_LDS_BYTES) + wl);
    // one call site of the chain body for both launch forms (grid form: one pass with the wave's own index)
    for (int pass = 0;; ++pass) {
        int idx = 0;
        if (Q.counter) {
            if (lane == 0) idx = atomicAdd(Q.counter, 1);
            idx = bcast_i(idx, 0);
        } else {
            if (pass > 0) break;
            idx = blockIdx.x * (blockDim.x / WAVE) + wid;
        }
        if (idx >= n_chains) break;
        const int ch = Q.order ? __builtin_amdgcn_readfirstlane(Q.order[idx]) : idx;
        learned_chain<KM>(ch, lane, w, nl, L, G, S, R, chain_lig, M, P, N);
    }
}

template <int KM>
int launch_mc_learned_t(const LigArgs& L, const GridArgs& G, const SFArgs& S, const RecArgs& R, const int* chain_lig,
                        int n_chains, const McArgs& M, const StepArgs& P, const NetArgs& N, const QueueArgs& Q,
                        hipStream_t stream) {
    const int wave_bytes = wave_lds_bytes(KM, L.P_MAX) + NET_LDS_BYTES;
    const void* fn = reinterpret_cast<const void*>(mc_learned_kernel<KM>);
    int W = fit_waves(waves_for(n_chains), KM, L.P_MAX, NET_LDS_BYTES);
    if (g_waves_per_block.load(std::memory_order_relaxed) <= 0) W = packed_waves(fn, W, wave_bytes);
    const size_t shm = (size_t)W * wave_bytes;
    allow_lds(fn, shm);
    unsigned nblk = (unsigned)((n_chains + W - 1) / W);
    if (Q.counter) {
        // resident capacity, queried once per (device, W, LDS size)
        static std::mutex mu;
        static std::map<long long, int> cap;
        int dev = 0;
        (void)hipGetDevice(&dev);
        const long long key = ((long long)dev << 40) | ((long long)W << 32) | (long long)shm;
        std::lock_guard<std::mutex> lock(mu);
        auto it = cap.find(key);
        if (it == cap.end()) {
            int per_cu = 0;
            (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, mc_learned_kernel<KM>, W * WAVE, shm);
            it = cap.emplace(key, cu_count() * (per_cu > 0 ? per_cu : 1)).first;
        }
        nblk = std::min(nblk, (unsigned)it->second);
    }
    hipLaunchKernelGGL(mc_learned_kernel<KM>, dim3(nblk), dim3(W * WAVE), shm, stream, L, G, S, R, chain_lig, n_chains,
                       M, P, N, Q);
    return (int)hipGetLastError();
}

}  // namespace

// The external entry points of one torsion bucket (declared in dock_buckets.hpp): one translation unit per bucket
// expands this macro, so the buckets compile in parallel. Ligands of up to 64 heavy atoms run one atom per lane,
// larger ones (AM 96 or 128) two; the learned step kernel takes up to 64.
#define ADOCK_BUCKET_DEFINE(KM)                                                                                        \
    int launch_mc_km##KM(const LigArgs& L, const GridArgs& G, const SFArgs& S, const RecArgs& R, int exact,            \
                         const int* cl, int n, const McArgs& M, hipStream_t st) {                                      \
        if (L.AM > 64)                                                                                                 \
            return exact ? launch_mc_tx<KM, 2, true>(L, G, S, R, cl, n, M, st)                                         \
                         : launch_mc_tx<KM, 2, false>(L, G, S, R, cl, n, M, st);                                       \
        return exact ? launch_mc_tx<KM, 1, true>(L, G, S, R, cl, n, M, st)                                             \
                     : launch_mc_tx<KM, 1, false>(L, G, S, R, cl, n, M, st);                                           \
    }                                                                                                                  \
    int launch_eg_km##KM(const LigArgs& L, const GridArgs& G, const SFArgs& S, const RecArgs& R, int exact,            \
                         const int* cl, int n, const float* state, float* e, float* g, hipStream_t st) {               \
        if (L.AM > 64)                                                                                                 \
            return exact ? launch_eg_tx<KM, 2, true>(L, G, S, R, cl, n, state, e, g, st)                               \
                         : launch_eg_tx<KM, 2, false>(L, G, S, R, cl, n, state, e, g, st);                             \
        return exact ? launch_eg_tx<KM, 1, true>(L, G, S, R, cl, n, state, e, g, st)                                   \
                     : launch_eg_tx<KM, 1, false>(L, G, S, R, cl, n, state, e, g, st);                                 \
    }                                                                                                                  \
    int launch_mc_persistent_km##KM(const LigArgs& L, const GridArgs& G, const SFArgs& S, const RecArgs& R, int exact, \
                                    const int* cl, int n, const McArgs& M, const QueueArgs& Q, int blocks,             \
                                    hipStream_t st) {                                                                  \
        if (L.AM > 64)                                                                                                 \
            return exact ? launch_mc_persistent_tx<KM, 2, true>(L, G, S, R, cl, n, M, Q, blocks, st)                   \
                         : launch_mc_persistent_tx<KM, 2, false>(L, G, S, R, cl, n, M, Q, blocks, st);                 \
        return exact ? launch_mc_persistent_tx<KM, 1, true>(L, G, S, R, cl, n, M, Q, blocks, st)                       \
                     : launch_mc_persistent_tx<KM, 1, false>(L, G, S, R, cl, n, M, Q, blocks, st);                     \
    }                                                                                                                  \
    int launch_mc_step_km##KM(const LigArgs& L, const GridArgs& G, const SFArgs& S, const RecArgs& R, const int* cl,   \
                              int n, const McArgs& M, const StepArgs& P, hipStream_t st) {                             \
        return L.AM > 64 ? -1 : launch_mc_step_t<KM>(L, G, S, R, cl, n, M, P, st);                                     \
    }                                                                                                                  \
    int launch_mc_learned_km##KM(const LigArgs& L, const GridArgs& G, const SFArgs& S, const RecArgs& R,               \
                                 const int* cl, int n, const McArgs& M, const StepArgs& P, const NetArgs& N,           \
                                 const QueueArgs& Q, hipStream_t st) {                                                 \
        return L.AM > 64 ? -1 : launch_mc_learned_t<KM>(L, G, S, R, cl, n, M, P, N, Q, st);                            \
    }                                                                                                                  \
    int mc_resident_waves_km##KM(int P_MAX, int persistent, int APL) {                                                 \
        return APL > 1 ? mc_resident_waves_t<KM, 2>(P_MAX, persistent) : mc_resident_waves_t<KM, 1>(P_MAX, persistent); \
    }

}  // namespace adock
