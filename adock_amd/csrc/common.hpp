// Device helpers shared by the docking kernels: wave64 reductions and broadcasts, the Vina-family pair term, the
// counter-based RNG, quaternion increment. Written for CDNA (gfx950): 64-lane wavefronts, DPP row operations,
// v_readlane broadcasts into scalar registers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define DEVI __device__ __forceinline__

namespace adock {

constexpr int WAVE = 64;
constexpr float EPSF = 1.1920929e-07f;
constexpr float LOG2E = 1.4426950408889634f;

DEVI int lane_id() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)); }

// value of v held by lane `src` (uniform src), broadcast to the whole wave through a scalar register
DEVI float bcast(float v, int src) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), src)); }
DEVI int bcast_i(int v, int src) { return __builtin_amdgcn_readlane(v, src); }

template <int CTRL, int ROW_MASK = 0xf>
DEVI float dpp_f(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROW_MASK, 0xf, false));
}

// Sum over all 64 lanes, result uniform. Inclusive row scan with row_shr 1/2/4/8 leaves each 16-lane row's sum in
// its last lane; row_bcast:15 and row_bcast:31 carry rows 0 and 0+1 upward, so lane 63 ends with the total.
// Must be called with all 64 lanes active.
DEVI float wave_sum(float v) {
    v += dpp_f<0x111>(v);
    v += dpp_f<0x112>(v);
    v += dpp_f<0x114>(v);
    v += dpp_f<0x118>(v);
    v += dpp_f<0x142, 0xa>(v);
    v += dpp_f<0x143, 0xc>(v);
    return bcast(v, 63);
}

// Sums of N independent per-lane values over the wave, all results uniform; the DPP steps of the N sums interleave.
template <int N>
DEVI void wave_sum_multi(float (&v)[N]) {
#pragma unroll
    for (int i = 0; i < N; ++i) v[i] += dpp_f<0x111>(v[i]);
#pragma unroll
    for (int i = 0; i < N; ++i) v[i] += dpp_f<0x112>(v[i]);
#pragma unroll
    for (int i = 0; i < N; ++i) v[i] += dpp_f<0x114>(v[i]);
#pragma unroll
    for (int i = 0; i < N; ++i) v[i] += dpp_f<0x118>(v[i]);
#pragma unroll
    for (int i = 0; i < N; ++i) v[i] += dpp_f<0x142, 0xa>(v[i]);
#pragma unroll
    for (int i = 0; i < N; ++i) v[i] += dpp_f<0x143, 0xc>(v[i]);
#pragma unroll
    for (int i = 0; i < N; ++i) v[i] = bcast(v[i], 63);
}

// Sums of N per-lane values over the wave; value i's total lands in lane i (lanes past N get 0). red: this wave's own
// LDS, 4 * ceil(N / 4) floats. gfx950's lane-half swaps fold two registers per step: v_permlane32_swap of values 2p
// and 2p + 1 gives [2p low half, 2p+1 low half] and [2p high half, 2p+1 high half], whose sum holds value 2p summed to
// 32 lanes in lanes 0-31 and value 2p + 1 in lanes 32-63; v_permlane16_swap of two such registers (odd 16-lane rows
// of the first with even rows of the second) and a sum leave four values one per row (rows 0..3: values 4q, 4q + 2,
// 4q + 1, 4q + 3); a 4-step row scan puts each row's total in its last lane, which writes it to red. About 2.5 N VALU
// instructions against 7 N for N separate 64-lane scans and broadcasts. Must be called with all 64 lanes active.
DEVI void wave_lds_sync();
template <int N>
DEVI float wave_sum_to_lanes(const float (&v)[N], float* red, int lane) {
    constexpr int NP = (N + 3) & ~3;
    float s[NP / 2];
#pragma unroll
    for (int p = 0; p < NP / 2; ++p) {
        const float a = 2 * p < N ? v[2 * p] : 0.f, b = 2 * p + 1 < N ? v[2 * p + 1] : 0.f;
        const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(a), __float_as_uint(b), false, false);
        s[p] = __uint_as_float(r[0]) + __uint_as_float(r[1]);
    }
    float u[NP / 4];
#pragma unroll
    for (int q = 0; q < NP / 4; ++q) {
        const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(s[2 * q]), __float_as_uint(s[2 * q + 1]),
                                                        false, false);
        u[q] = __uint_as_float(r[0]) + __uint_as_float(r[1]);
    }
#pragma unroll
    for (int q = 0; q < NP / 4; ++q) u[q] += dpp_f<0x111>(u[q]);
#pragma unroll
    for (int q = 0; q < NP / 4; ++q) u[q] += dpp_f<0x112>(u[q]);
#pragma unroll
    for (int q = 0; q < NP / 4; ++q) u[q] += dpp_f<0x114>(u[q]);
#pragma unroll
    for (int q = 0; q < NP / 4; ++q) u[q] += dpp_f<0x118>(u[q]);
    if ((lane & 15) == 15) {
        const int vi = ((lane >> 3) & 2) | (lane >> 5);   // row 0, 1, 2, 3 -> value offset 0, 2, 1, 3
#pragma unroll
        for (int q = 0; q < NP / 4; ++q) red[4 * q + vi] = u[q];
    }
    wave_lds_sync();
    return lane < N ? red[lane] : 0.f;
}

template <int CTRL, int ROW_MASK = 0xf>
DEVI float dpp_max_step(float v) {
    const float o = __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(-INFINITY), __float_as_int(v), CTRL,
                                                               ROW_MASK, 0xf, false));
    return fmaxf(v, o);
}
// maximum over the 64 lanes (all active), uniform result
DEVI float wave_max_f(float v) {
    v = dpp_max_step<0x111>(v);
    v = dpp_max_step<0x112>(v);
    v = dpp_max_step<0x114>(v);
    v = dpp_max_step<0x118>(v);
    v = dpp_max_step<0x142, 0xa>(v);
    v = dpp_max_step<0x143, 0xc>(v);
    return bcast(v, 63);
}

// Orders this wave's LDS writes before its later LDS reads (one wave owns the region).
DEVI void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Proposal network shape (features -> hidden -> hidden -> move components; the last layer's 23 outputs padded to two
// 16-wide MFMA tiles) and the BF16 operand types of gfx950's v_mfma_f32_16x16x32_bf16.
constexpr int NET_IN = 128, NET_HID = 256, NET_OUT = 23, NET_OUTP = 32;
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float float4_t __attribute__((ext_vector_type(4)));

// float32 -> bfloat16, round to nearest even (NaN kept quiet)
DEVI unsigned short f2bf(float f) {
    unsigned u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x7fffffu)) return (unsigned short)((u >> 16) | 0x40);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

DEVI float fast_exp2(float x) { return __builtin_amdgcn_exp2f(x); }
DEVI float fast_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
DEVI float fast_rsq(float x) { return __builtin_amdgcn_rsqf(x); }
DEVI float fast_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
// sin / cos through the hardware units, which take revolutions: the argument is reduced to [-1/2, 1/2] revolution
// first (v_sin_f32 / v_cos_f32 are accurate there; error ~1e-7, far below the 1e-4 parity bar)
DEVI void fast_sincos(float x, float& s, float& c) {
    float r = x * 0.15915494309189535f;
    r -= rintf(r);
    s = __builtin_amdgcn_sinf(r);
    c = __builtin_amdgcn_cosf(r);
}

// Scoring-function constants in evaluation-friendly form. Built from the 15-float search-order vector
// (g1_off, g1_w, g2_off, g2_w, rep_off, hyd_good, hyd_bad, hb_good, hb_bad, cutoff, w_g1, w_g2, w_rep, w_hyd, w_hb).
struct SFConst {
    float g1o, g1s, g1d;   // offset, sqrt(log2 e)/width, dE factor -2 w / (width * sqrt(log2 e))
    float g2o, g2s, g2d;
    float g1b, g2b;        // -offset * sqrt(log2 e)/width: u = d * g1s + g1b in one FMA
    float hbh, bbh;        // bad * inv: ramp v = bad * inv - d * inv in one FMA
    float w1, w2, w3, w4, w5;
    float repo;
    float hb_bad, hinv;    // hydrophobic: bad, 1/(bad - good)
    float bb_bad, binv;    // hbond
    float cut;
};

DEVI SFConst make_sf(const float* __restrict__ prm) {
    SFConst s;
    const float rl = 1.2011224087864498f;   // sqrt(log2 e)
    s.g1o = prm[0]; s.g1s = rl / prm[1]; s.g2o = prm[2]; s.g2s = rl / prm[3];
    s.repo = prm[4];
    s.hb_bad = prm[6]; s.hinv = 1.f / (prm[6] - prm[5]);
    s.bb_bad = prm[8]; s.binv = 1.f / (prm[8] - prm[7]);
    s.cut = prm[9];
    s.w1 = prm[10]; s.w2 = prm[11]; s.w3 = prm[12]; s.w4 = prm[13]; s.w5 = prm[14];
    s.g1d = -2.f * s.w1 / (prm[1] * rl);
    s.g2d = -2.f * s.w2 / (prm[3] * rl);
    s.g1b = -s.g1o * s.g1s; s.g2b = -s.g2o * s.g2s;
    s.hbh = s.hb_bad * s.hinv; s.bbh = s.bb_bad * s.binv;
    return s;
}

// atom flag bits: 1 hydrophobic, 2 donor, 4 acceptor
DEVI int pair_flags(int fi, int fj) {
    const int hyd = fi & fj & 1;
    const int hb = ((fi >> 1) & (fj >> 2) & 1) | ((fi >> 2) & (fj >> 1) & 1);
    return hyd | (hb << 1);
}

// Pair energy and dE/dr at surface distance d = r - R_i - R_j (same terms as the reference pair_e).
DEVI float pair_e(float d, int pf, const SFConst& s, float& de) {
    const float u1 = (d - s.g1o) * s.g1s, g1 = fast_exp2(-u1 * u1);
    const float u2 = (d - s.g2o) * s.g2s, g2 = fast_exp2(-u2 * u2);
    float ep = s.w1 * g1 + s.w2 * g2;
    float dd = s.g1d * u1 * g1 + s.g2d * u2 * g2;
    const float dr = d - s.repo;
    if (dr < 0.f) { ep += s.w3 * dr * dr; dd += s.w3 * 2.f * dr; }
    if (pf & 1) {
        const float v = (s.hb_bad - d) * s.hinv;
        if (v >= 1.f) ep += s.w4;
        else if (v > 0.f) { ep += s.w4 * v; dd -= s.w4 * s.hinv; }
    }
    if (pf & 2) {
        const float v = (s.bb_bad - d) * s.binv;
        if (v >= 1.f) ep += s.w5;
        else if (v > 0.f) { ep += s.w5 * v; dd -= s.w5 * s.binv; }
    }
    de = dd;
    return ep;
}

// Branch-free form of pair_e for SIMT loops: same terms (the slope terms clamp v to [0, 1]; their derivative is
// non-zero only strictly inside the ramp). pf: bit 0 hydrophobic pair, bit 1 h-bond pair. HYD / HB = false leave the
// term out for every lane (callers pass false when no lane holds such a pair).
DEVI float pair_e_nb(float d, int pf, const SFConst& s, float& de, bool HYD = true, bool HB = true) {
    const float u1 = fmaf(d, s.g1s, s.g1b), g1 = fast_exp2(-u1 * u1);
    const float u2 = fmaf(d, s.g2s, s.g2b), g2 = fast_exp2(-u2 * u2);
    const float dr = fminf(d - s.repo, 0.f);
    float ep = s.w1 * g1 + s.w2 * g2 + s.w3 * dr * dr;
    float dd = s.g1d * u1 * g1 + s.g2d * u2 * g2 + 2.f * s.w3 * dr;
    if (HYD) {
        const float vh = fmaf(-d, s.hinv, s.hbh);
        const float hh = (pf & 1) ? 1.f : 0.f;
        ep += hh * s.w4 * fminf(fmaxf(vh, 0.f), 1.f);
        dd -= hh * ((vh > 0.f && vh < 1.f) ? s.w4 * s.hinv : 0.f);
    }
    if (HB) {
        const float vb = fmaf(-d, s.binv, s.bbh);
        const float hb = (pf & 2) ? 1.f : 0.f;
        ep += hb * s.w5 * fminf(fmaxf(vb, 0.f), 1.f);
        dd -= hb * ((vb > 0.f && vb < 1.f) ? s.w5 * s.binv : 0.f);
    }
    de = dd;
    return ep;
}

// SplitMix64 finaliser and the per-(chain, step, slot) uniform draw in [0, 1) with 24 bits.
DEVI uint64_t smix(uint64_t x) {
    x += 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}
DEVI float urand(uint64_t seed, int chain, int step, int slot) {
    const uint64_t h = smix(seed ^ smix(((uint64_t)(uint32_t)chain << 32) ^ ((uint64_t)(int64_t)step << 8) ^
                                       (uint64_t)(int64_t)slot));
    return (float)(h >> 40) * (1.0f / 16777216.0f);
}
// rejection-sampled point in the unit ball; slots slot + 3k .. slot + 3k + 2 for try k < 32
DEVI void rand_ball(uint64_t seed, int chain, int step, int slot, float& a, float& b, float& c) {
    for (int k = 0; k < 32; ++k) {
        a = 2.f * urand(seed, chain, step, slot + 3 * k) - 1.f;
        b = 2.f * urand(seed, chain, step, slot + 3 * k + 1) - 1.f;
        c = 2.f * urand(seed, chain, step, slot + 3 * k + 2) - 1.f;
        if (a * a + b * b + c * c < 1.f) return;
    }
    a = b = c = 0.f;
}

// (w,x,y,z) <- quat(rotation vector r) * (w,x,y,z), normalised
DEVI void quat_inc(float& w, float& x, float& y, float& z, float rx, float ry, float rz) {
    const float ang = fast_sqrt(rx * rx + ry * ry + rz * rz);
    float hw, k;
    if (ang > 1e-6f) {
        float sn, cs;
        fast_sincos(0.5f * ang, sn, cs);
        hw = cs; k = sn * fast_rcp(ang);
    } else { hw = 1.f; k = 0.5f; }
    const float aw = hw, ax = rx * k, ay = ry * k, az = rz * k;
    const float nw = aw * w - ax * x - ay * y - az * z;
    const float nx = aw * x + ax * w + ay * z - az * y;
    const float ny = aw * y - ax * z + ay * w + az * x;
    const float nz = aw * z + ax * y - ay * x + az * w;
    const float nn = fast_rsq(nw * nw + nx * nx + ny * ny + nz * nz);
    w = nw * nn; x = nx * nn; y = ny * nn; z = nz * nn;
}

}  // namespace adock
