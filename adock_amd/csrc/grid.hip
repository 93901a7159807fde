// Receptor scoring-grid builder for gfx950: FP32 VALU pairwise kernel.
//
// Computes, for every grid point p = origin + h (i, j, k) and every requested ligand XS type q, the sum over receptor
// atoms a with |p - x_a|^2 < cutoff^2 of the Vina-family terms (gauss1, gauss2, repulsion, hydrophobic, hbond), either
// weighted into one value per type (out [T, nx, ny, nz], z fastest) or as 5 raw channels (out [5, T, nx, ny, nz]).
// Same values as the reference's build_grid kernel (src/adock/grid.py:14-71).
//
// Second kernel, the fitted (linear-in-weights) scoring function's builder (src/adock/fit/gridfit.py:13-58): per
// (point, receptor atom, ligand type) six basis values of the surface distance d = r - R_rec - R_lig (gauss1, gauss2,
// capped repulsion, hydrophobic and h-bond ramps) and of r (ionic 1/max(r, r_min)^2 - 1/cutoff^2), weighted by a
// per-(receptor ext type, ligand type) table W[n_rt, n_lt, 6] that is staged in LDS for the pass's ligand types.
//
// MI355X design: one workgroup owns an 8x8x4 tile of grid points, one point per thread; each wave holds a 4x4x4 cube of
// it, so the atoms in cutoff of at least one lane of a wave (the lanes that run the type loop) come from a sphere only
// 1 A wider than the cutoff. The receptor is streamed through LDS in 256-atom chunks; each chunk is culled against
// the tile's bounding box grown by the cutoff and compacted in receptor order with wave ballots (deterministic
// summation order). Every thread then walks the compacted list with LDS broadcast reads. The in-cutoff test uses the
// same float32 operations and fused-multiply-add nesting as the CPU oracle so both see the same atom set.
#include "common.hpp"
#include "kernels.hpp"

namespace adock {
namespace {

constexpr int TI = 8, TJ = 8, TK = 4;          // tile: 256 points, four 4x4x4 wave cubes

// thread -> point of the tile: wave w takes the cube at i = 4 (w >> 1), j = 4 (w & 1); lanes run z fastest
__device__ __forceinline__ void tile_point(int tid, int& ti, int& tj, int& tk) {
    const int w = tid >> 6, l = tid & 63;
    ti = ((w >> 1) << 2) + (l >> 4);
    tj = ((w & 1) << 2) + ((l >> 2) & 3);
    tk = l & 3;
}
constexpr int GB = TI * TJ * TK;
constexpr int TQ = 8;                          // ligand types per pass

template <bool PER_TERM>
__global__ __launch_bounds__(GB) void build_grid_kernel(
    const float* __restrict__ rxyz, const int* __restrict__ rtype, int n_rec, const float* __restrict__ radius,
    const int* __restrict__ hyd, const int* __restrict__ don, const int* __restrict__ acc,
    const int* __restrict__ lig_types, int n_lt, int q0, float ox, float oy, float oz, float spacing, int nx, int ny,
    int nz, const float* __restrict__ prm, float* __restrict__ out, int tiles_j, int tiles_k) {
    __shared__ float4 s_atom[GB];      // x, y, z, radius
    __shared__ int s_flag[GB];         // 1 hydrophobic, 2 donor, 4 acceptor
    __shared__ int s_wcount[GB / WAVE];
    const int tid = threadIdx.x;
    const int lane = tid & (WAVE - 1), wid = tid / WAVE;
    int tile = blockIdx.x;
    const int bk = tile % tiles_k; tile /= tiles_k;
    const int bj = tile % tiles_j;
    const int bi = tile / tiles_j;
    int ti, tj, tk;
    tile_point(tid, ti, tj, tk);
    const int i = bi * TI + ti, j = bj * TJ + tj, k = bk * TK + tk;
    const bool valid = (i < nx) && (j < ny) && (k < nz);
    const float x = fmaf((float)i, spacing, ox), y = fmaf((float)j, spacing, oy), z = fmaf((float)k, spacing, oz);
    const float cut = prm[0];
    const float cut2 = cut * cut;
    // tile bounding box grown by the cutoff (+1e-3 A so the cull is conservative under rounding)
    const float gr = cut + 1e-3f;
    const float blo0 = fmaf((float)(bi * TI), spacing, ox) - gr, bhi0 = fmaf((float)(bi * TI + TI - 1), spacing, ox) + gr;
    const float blo1 = fmaf((float)(bj * TJ), spacing, oy) - gr, bhi1 = fmaf((float)(bj * TJ + TJ - 1), spacing, oy) + gr;
    const float blo2 = fmaf((float)(bk * TK), spacing, oz) - gr, bhi2 = fmaf((float)(bk * TK + TK - 1), spacing, oz) + gr;

    // scoring constants (grid order: cutoff, g1_off, g1_w, g2_off, g2_w, rep_off, hyd_good, hyd_bad, hb_good, hb_bad,
    // w_g1..w_hb)
    const float rl = 1.2011224087864498f;
    const float g1o = prm[1], g1s = rl / prm[2], g2o = prm[3], g2s = rl / prm[4], repo = prm[5];
    const float hbad = prm[7], hinv = 1.f / (prm[7] - prm[6]);
    const float bbad = prm[9], binv = 1.f / (prm[9] - prm[8]);
    const float w1 = prm[10], w2 = prm[11], w3 = prm[12], w4 = prm[13], w5 = prm[14];

    const int nq = min(TQ, n_lt - q0);
    float rq[TQ];
    int fq[TQ];
#pragma unroll
    for (int q = 0; q < TQ; ++q) {
        const int t = q < nq ? lig_types[q0 + q] : 0;
        rq[q] = radius[t];
        fq[q] = hyd[t] | (don[t] << 1) | (acc[t] << 2);
    }
    float a0[TQ], a1[TQ], a2[TQ], a3[TQ], a4[TQ];
#pragma unroll
    for (int q = 0; q < TQ; ++q) { a0[q] = a1[q] = a2[q] = a3[q] = a4[q] = 0.f; }

    for (int base = 0; base < n_rec; base += GB) {
        const int a = base + tid;
        float4 at = make_float4(0.f, 0.f, 0.f, 0.f);
        int fl = 0;
        bool keep = false;
        if (a < n_rec) {
            at.x = rxyz[3 * a]; at.y = rxyz[3 * a + 1]; at.z = rxyz[3 * a + 2];
            const int t = rtype[a];
            at.w = radius[t];
            fl = hyd[t] | (don[t] << 1) | (acc[t] << 2);
            keep = at.x >= blo0 && at.x <= bhi0 && at.y >= blo1 && at.y <= bhi1 && at.z >= blo2 && at.z <= bhi2;
        }
        const uint64_t m = __ballot(keep);
        if (lane == 0) s_wcount[wid] = __popcll(m);
        __syncthreads();
        int off = 0, tot = 0;
#pragma unroll
        for (int w = 0; w < GB / WAVE; ++w) {
            const int c = s_wcount[w];
            off += (w < wid) ? c : 0;
            tot += c;
        }
        if (keep) {
            const int pos = off + __popcll(m & ((1ull << lane) - 1ull));
            s_atom[pos] = at;
            s_flag[pos] = fl;
        }
        __syncthreads();
        for (int b = 0; b < tot; ++b) {
            const float4 r = s_atom[b];
            const float dx = r.x - x, dy = r.y - y, dz = r.z - z;
            const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
            if (r2 < cut2) {
                const float rr = fast_sqrt(r2);
                const int fa = s_flag[b];
                const float rmr = rr - r.w;
#pragma unroll
                for (int q = 0; q < TQ; ++q) {
                    if (q >= nq) break;
                    const float d = rmr - rq[q];
                    const float u1 = (d - g1o) * g1s, u2 = (d - g2o) * g2s;
                    a0[q] += fast_exp2(-u1 * u1);
                    a1[q] += fast_exp2(-u2 * u2);
                    const float dr = fminf(d - repo, 0.f);
                    a2[q] += dr * dr;
                    const int pf = pair_flags(fa, fq[q]);
                    const float hy = fminf(fmaxf((hbad - d) * hinv, 0.f), 1.f);
                    const float hb = fminf(fmaxf((bbad - d) * binv, 0.f), 1.f);
                    a3[q] += (pf & 1) ? hy : 0.f;
                    a4[q] += (pf & 2) ? hb : 0.f;
                }
            }
        }
        __syncthreads();
    }
    if (!valid) return;
    const long long npts = (long long)nx * ny * nz;
    const long long p = ((long long)i * ny + j) * nz + k;
#pragma unroll
    for (int q = 0; q < TQ; ++q) {
        if (q >= nq) break;
        if (PER_TERM) {
            out[(long long)(0 * n_lt + q0 + q) * npts + p] = a0[q];
            out[(long long)(1 * n_lt + q0 + q) * npts + p] = a1[q];
            out[(long long)(2 * n_lt + q0 + q) * npts + p] = a2[q];
            out[(long long)(3 * n_lt + q0 + q) * npts + p] = a3[q];
            out[(long long)(4 * n_lt + q0 + q) * npts + p] = a4[q];
        } else {
            out[(long long)(q0 + q) * npts + p] = w1 * a0[q] + w2 * a1[q] + w3 * a2[q] + w4 * a3[q] + w5 * a4[q];
        }
    }
}

// Fitted function: W[rt, q, k] of the pass's TQ ligand types sits in LDS (rt-major, 6 floats per (rt, q)); receptor
// ext types are at most FIT_RT_MAX (12 blocks of 18 XS types). Atoms whose type lies outside [0, n_rt) are skipped.
constexpr int FIT_RT_MAX = 216;
constexpr int FIT_NB = 6;

// Fitted-function constants from prm (cutoff, g1_off, g1_w, g2_off, g2_w, rep_off, hyd_good, hyd_bad, hb_good, hb_bad,
// ion_rmin, rep_cap) in evaluation form: Gaussian widths folded with sqrt(log2 e) so exp(-u^2) = exp2(-(u')^2).
struct FitConst {
    float cut, cut2, g1o, g1s, g2o, g2s, repo, hbad, hinv, bbad, binv, rmin, cap, icut2;
};
DEVI FitConst make_fit(const float* __restrict__ prm) {
    FitConst f;
    const float rl = 1.2011224087864498f;   // sqrt(log2 e)
    f.cut = prm[0];
    f.cut2 = f.cut * f.cut;
    f.g1o = prm[1]; f.g1s = rl / prm[2]; f.g2o = prm[3]; f.g2s = rl / prm[4]; f.repo = prm[5];
    f.hbad = prm[7]; f.hinv = 1.f / (prm[7] - prm[6]);
    f.bbad = prm[9]; f.binv = 1.f / (prm[9] - prm[8]);
    f.rmin = prm[10]; f.cap = prm[11]; f.icut2 = 1.f / f.cut2;
    return f;
}
// ionic basis 1 / max(r, r_min)^2 - 1 / cutoff^2 (depends on r only)
DEVI float fit_ion(const FitConst& f, float rr) {
    const float rc = fmaxf(rr, f.rmin);
    return 1.f / (rc * rc) - f.icut2;
}
// sum_k w[k] basis_k at surface distance d (gauss1, gauss2, capped repulsion, hydrophobic and h-bond ramps) + w[5] ion
DEVI float fit_pair(const FitConst& f, const float* wq, float d, float ion) {
    const float u1 = (d - f.g1o) * f.g1s, u2 = (d - f.g2o) * f.g2s;
    float sv = wq[0] * fast_exp2(-u1 * u1);
    sv = fmaf(wq[1], fast_exp2(-u2 * u2), sv);
    const float dr = fminf(d - f.repo, 0.f);
    sv = fmaf(wq[2], fminf(dr * dr, f.cap), sv);
    sv = fmaf(wq[3], fminf(fmaxf((f.hbad - d) * f.hinv, 0.f), 1.f), sv);
    sv = fmaf(wq[4], fminf(fmaxf((f.bbad - d) * f.binv, 0.f), 1.f), sv);
    return fmaf(wq[5], ion, sv);
}

// Several receptor sets in one launch: blockIdx.y = set, whose atoms start set_atoms atoms further on, whose origin is
// origins[3 set ..] (NULL: ox, oy, oz) and whose grids start set_out floats further on (one set: 0, NULL, 0).
__global__ __launch_bounds__(GB) void build_grid_fit_kernel(
    const float* __restrict__ rxyz, const int* __restrict__ rext, int n_rec, const float* __restrict__ rad_rec,
    const float* __restrict__ rad_lig, int n_lt, const float* __restrict__ W, int n_rt, int q0, float ox, float oy,
    float oz, float spacing, int nx, int ny, int nz, const float* __restrict__ prm, float* __restrict__ out,
    int tiles_j, int tiles_k, long long set_atoms, const float* __restrict__ origins, long long set_out) {
    __shared__ float4 s_atom[GB];      // x, y, z, radius
    __shared__ int s_rt[GB];           // receptor ext type
    __shared__ int s_wcount[GB / WAVE];
    __shared__ float s_w[FIT_RT_MAX * TQ * FIT_NB];
    const int set = blockIdx.y;
    rxyz += (long long)set * set_atoms * 3;
    rext += (long long)set * set_atoms;
    out += (long long)set * set_out;
    if (origins) { ox = origins[3 * set]; oy = origins[3 * set + 1]; oz = origins[3 * set + 2]; }
    const int tid = threadIdx.x;
    const int lane = tid & (WAVE - 1), wid = tid / WAVE;
    int tile = blockIdx.x;
    const int bk = tile % tiles_k; tile /= tiles_k;
    const int bj = tile % tiles_j;
    const int bi = tile / tiles_j;
    int ti, tj, tk;
    tile_point(tid, ti, tj, tk);
    const int i = bi * TI + ti, j = bj * TJ + tj, k = bk * TK + tk;
    const bool valid = (i < nx) && (j < ny) && (k < nz);
    const float x = fmaf((float)i, spacing, ox), y = fmaf((float)j, spacing, oy), z = fmaf((float)k, spacing, oz);
    const FitConst f = make_fit(prm);
    const float gr = f.cut + 1e-3f;
    const float blo0 = fmaf((float)(bi * TI), spacing, ox) - gr, bhi0 = fmaf((float)(bi * TI + TI - 1), spacing, ox) + gr;
    const float blo1 = fmaf((float)(bj * TJ), spacing, oy) - gr, bhi1 = fmaf((float)(bj * TJ + TJ - 1), spacing, oy) + gr;
    const float blo2 = fmaf((float)(bk * TK), spacing, oz) - gr, bhi2 = fmaf((float)(bk * TK + TK - 1), spacing, oz) + gr;

    const int nq = min(TQ, n_lt - q0);
    const int n_w = n_rec > 0 ? min(n_rt, FIT_RT_MAX) * TQ * FIT_NB : 0;
    for (int e = tid; e < n_w; e += GB) {
        const int kk = e % FIT_NB, q = (e / FIT_NB) % TQ, rt = e / (FIT_NB * TQ);
        s_w[e] = q < nq ? W[((long long)rt * n_lt + q0 + q) * FIT_NB + kk] : 0.f;
    }
    float rq[TQ], acc_e[TQ];
#pragma unroll
    for (int q = 0; q < TQ; ++q) {
        rq[q] = q < nq ? rad_lig[q0 + q] : 0.f;
        acc_e[q] = 0.f;
    }

    for (int base = 0; base < n_rec; base += GB) {
        const int a = base + tid;
        float4 at = make_float4(0.f, 0.f, 0.f, 0.f);
        int rt = 0;
        bool keep = false;
        if (a < n_rec) {
            at.x = rxyz[3 * a]; at.y = rxyz[3 * a + 1]; at.z = rxyz[3 * a + 2];
            rt = rext[a];
            const bool typed = rt >= 0 && rt < n_rt;
            at.w = typed ? rad_rec[rt] : 0.f;
            keep = typed && at.x >= blo0 && at.x <= bhi0 && at.y >= blo1 && at.y <= bhi1 && at.z >= blo2 &&
                   at.z <= bhi2;
        }
        const uint64_t m = __ballot(keep);
        if (lane == 0) s_wcount[wid] = __popcll(m);
        __syncthreads();
        int off = 0, tot = 0;
#pragma unroll
        for (int w = 0; w < GB / WAVE; ++w) {
            const int c = s_wcount[w];
            off += (w < wid) ? c : 0;
            tot += c;
        }
        if (keep) {
            const int pos = off + __popcll(m & ((1ull << lane) - 1ull));
            s_atom[pos] = at;
            s_rt[pos] = rt;
        }
        __syncthreads();
        for (int b = 0; b < tot; ++b) {
            const float4 r = s_atom[b];
            const float dx = r.x - x, dy = r.y - y, dz = r.z - z;
            const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
            if (r2 < f.cut2) {
                const float rr = fast_sqrt(r2);
                const float ion = fit_ion(f, rr);
                const float rmr = rr - r.w;
                const float* wb = s_w + s_rt[b] * (TQ * FIT_NB);
#pragma unroll
                for (int q = 0; q < TQ; ++q) {
                    if (q >= nq) break;
                    acc_e[q] += fit_pair(f, wb + q * FIT_NB, rmr - rq[q], ion);
                }
            }
        }
        __syncthreads();
    }
    if (!valid) return;
    const long long npts = (long long)nx * ny * nz;
    const long long p = ((long long)i * ny + j) * nz + k;
#pragma unroll
    for (int q = 0; q < TQ; ++q) {
        if (q >= nq) break;
        out[(long long)(q0 + q) * npts + p] = acc_e[q];
    }
}

// Exact pairwise fitted inter energy of docked poses (src/adock/search.py:73-89 with FittedPotential.energy,
// src/adock/fit/terms.py:339-344): out[l, m] = sum over ligand atoms a of pose (l, m) with a type in [0, n_lt) and
// receptor atoms b of receptor rec_idx[l] with an ext type in [0, n_rt) and |x_a - x_b|^2 < cutoff^2 of
// sum_k W[rt_b, lt_a, k] basis_k, with the same basis as the fitted grid (a grid point is an atom here).
//
// One wave per pose, one lane per ligand atom (poses with more than 64 atom slots take several 64-atom slices), four
// independent poses per workgroup. Each wave streams the receptor in 64-atom chunks, keeps the atoms inside its own
// pose's box grown by the cutoff (ballot-compacted in receptor order into the wave's LDS slice) and walks them with LDS
// broadcast reads; the six pair weights come from W in memory (all lanes of a wave read the same receptor type row,
// so the reads hit a few cache lines). Per-lane sums, then a fixed-order wave sum: the result is deterministic. A
// complex whose rec_idx lies outside [0, n_sets) gets NaN.
constexpr int XP = 4;                          // poses (waves) per workgroup

__global__ __launch_bounds__(XP * WAVE) void inter_exact_fit_kernel(
    const float* __restrict__ X, const int* __restrict__ lig_type, int M, int A, const int* __restrict__ rec_idx,
    const float* __restrict__ rxyz, const int* __restrict__ rext, int n_sets, int NR,
    const float* __restrict__ rad_rec, const float* __restrict__ rad_lig, int n_lt, const float* __restrict__ W,
    int n_rt, const float* __restrict__ prm, float* __restrict__ out, int blocks_m) {
    __shared__ float4 s_atom[XP][WAVE];   // x, y, z, radius
    __shared__ int s_rt[XP][WAVE];
    const int lane = threadIdx.x & (WAVE - 1), wid = threadIdx.x / WAVE;
    const int l = blockIdx.x / blocks_m;
    const int m = (blockIdx.x % blocks_m) * XP + wid;
    if (m >= M) return;                   // waves are independent: no workgroup barrier below
    const int ri = rec_idx[l];
    if (ri < 0 || ri >= n_sets) {
        if (lane == 0) out[(long long)l * M + m] = __builtin_nanf("");
        return;
    }
    const FitConst f = make_fit(prm);
    const float gr = f.cut + 1e-3f;
    const float* rx = rxyz + (long long)ri * NR * 3;
    const int* re = rext + (long long)ri * NR;
    float4* sa = s_atom[wid];
    int* st = s_rt[wid];
    float acc = 0.f;
    for (int a0 = 0; a0 < A; a0 += WAVE) {
        const int a = a0 + lane;
        const int lt = a < A ? lig_type[(long long)l * A + a] : -1;
        const bool act = lt >= 0 && lt < n_lt;
        float x = 0.f, y = 0.f, z = 0.f;
        if (act) {
            const float* p = X + (((long long)l * M + m) * A + a) * 3;
            x = p[0]; y = p[1]; z = p[2];
        }
        // the pose slice's box grown by the cutoff (+1e-3 A: the cull is conservative under rounding); empty when no
        // lane is active, then no receptor atom passes
        const float blo0 = -wave_max_f(act ? -x : -INFINITY) - gr, bhi0 = wave_max_f(act ? x : -INFINITY) + gr;
        const float blo1 = -wave_max_f(act ? -y : -INFINITY) - gr, bhi1 = wave_max_f(act ? y : -INFINITY) + gr;
        const float blo2 = -wave_max_f(act ? -z : -INFINITY) - gr, bhi2 = wave_max_f(act ? z : -INFINITY) + gr;
        const float rq = act ? rad_lig[lt] : 0.f;
        const float* wl = W + (long long)(act ? lt : 0) * FIT_NB;
        const long long wrow = (long long)n_lt * FIT_NB;
        for (int base = 0; base < NR; base += WAVE) {
            const int b = base + lane;
            float4 at = make_float4(0.f, 0.f, 0.f, 0.f);
            int rt = 0;
            bool keep = false;
            if (b < NR) {
                at.x = rx[3 * b]; at.y = rx[3 * b + 1]; at.z = rx[3 * b + 2];
                rt = re[b];
                const bool typed = rt >= 0 && rt < n_rt;
                at.w = typed ? rad_rec[rt] : 0.f;
                keep = typed && at.x >= blo0 && at.x <= bhi0 && at.y >= blo1 && at.y <= bhi1 && at.z >= blo2 &&
                       at.z <= bhi2;
            }
            const uint64_t mk = __ballot(keep);
            if (keep) {
                const int pos = __popcll(mk & ((1ull << lane) - 1ull));
                sa[pos] = at;
                st[pos] = rt;
            }
            wave_lds_sync();
            const int cnt = __popcll(mk);
            if (act) {
                for (int c = 0; c < cnt; ++c) {
                    const float4 r = sa[c];
                    const float dx = r.x - x, dy = r.y - y, dz = r.z - z;
                    const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
                    if (r2 < f.cut2) {
                        const float rr = fast_sqrt(r2);
                        acc += fit_pair(f, wl + st[c] * wrow, rr - r.w - rq, fit_ion(f, rr));
                    }
                }
            }
            wave_lds_sync();
        }
    }
    acc = wave_sum(acc);
    if (lane == 0) out[(long long)l * M + m] = acc;
}

}  // namespace

int build_grid_fit_rt_max() { return FIT_RT_MAX; }

int launch_build_grid_fit(const float* rxyz, const int* rext, int n_rec, const float* rad_rec, const float* rad_lig,
                          int n_lt, const float* W, int n_rt, float ox, float oy, float oz, float spacing, int nx,
                          int ny, int nz, const float* prm, float* out, hipStream_t stream) {
    const int tiles_i = (nx + TI - 1) / TI, tiles_j = (ny + TJ - 1) / TJ, tiles_k = (nz + TK - 1) / TK;
    const long long nblk = (long long)tiles_i * tiles_j * tiles_k;
    for (int q0 = 0; q0 < n_lt; q0 += TQ)
        hipLaunchKernelGGL(build_grid_fit_kernel, dim3((unsigned)nblk), dim3(GB), 0, stream, rxyz, rext, n_rec,
                           rad_rec, rad_lig, n_lt, W, n_rt, q0, ox, oy, oz, spacing, nx, ny, nz, prm, out, tiles_j,
                           tiles_k, 0ll, (const float*)nullptr, 0ll);
    return (int)hipGetLastError();
}

int launch_build_grid_fit_sets(const float* rxyz, const int* rext, int n_sets, int n_rec, const float* rad_rec,
                               const float* rad_lig, int n_lt, const float* W, int n_rt, const float* origins,
                               float spacing, int nx, int ny, int nz, const float* prm, float* out,
                               hipStream_t stream) {
    const int tiles_i = (nx + TI - 1) / TI, tiles_j = (ny + TJ - 1) / TJ, tiles_k = (nz + TK - 1) / TK;
    const long long nblk = (long long)tiles_i * tiles_j * tiles_k;
    const long long set_out = (long long)n_lt * nx * ny * nz;
    for (int q0 = 0; q0 < n_lt; q0 += TQ)
        hipLaunchKernelGGL(build_grid_fit_kernel, dim3((unsigned)nblk, (unsigned)n_sets), dim3(GB), 0, stream, rxyz,
                           rext, n_rec, rad_rec, rad_lig, n_lt, W, n_rt, q0, 0.f, 0.f, 0.f, spacing, nx, ny, nz, prm,
                           out, tiles_j, tiles_k, (long long)n_rec, origins, set_out);
    return (int)hipGetLastError();
}

int launch_inter_exact_fit(const float* X, const int* lig_type, int L, int M, int A, const int* rec_idx,
                           const float* rxyz, const int* rext, int n_sets, int NR, const float* rad_rec,
                           const float* rad_lig, int n_lt, const float* W, int n_rt, const float* prm, float* out,
                           hipStream_t stream) {
    if (L == 0 || M == 0) return 0;
    const int blocks_m = (M + XP - 1) / XP;
    const long long nblk = (long long)L * blocks_m;
    hipLaunchKernelGGL(inter_exact_fit_kernel, dim3((unsigned)nblk), dim3(XP * WAVE), 0, stream, X, lig_type, M, A,
                       rec_idx, rxyz, rext, n_sets, NR, rad_rec, rad_lig, n_lt, W, n_rt, prm, out, blocks_m);
    return (int)hipGetLastError();
}

int launch_build_grid(const float* rxyz, const int* rtype, int n_rec, const float* radius, const int* hyd,
                      const int* don, const int* acc, const int* lig_types, int n_lt, float ox, float oy, float oz,
                      float spacing, int nx, int ny, int nz, const float* prm, int per_term, float* out,
                      hipStream_t stream) {
    const int tiles_i = (nx + TI - 1) / TI, tiles_j = (ny + TJ - 1) / TJ, tiles_k = (nz + TK - 1) / TK;
    const long long nblk = (long long)tiles_i * tiles_j * tiles_k;
    for (int q0 = 0; q0 < n_lt; q0 += TQ) {
        if (per_term)
            hipLaunchKernelGGL(build_grid_kernel<true>, dim3((unsigned)nblk), dim3(GB), 0, stream, rxyz, rtype, n_rec,
                               radius, hyd, don, acc, lig_types, n_lt, q0, ox, oy, oz, spacing, nx, ny, nz, prm, out,
                               tiles_j, tiles_k);
        else
            hipLaunchKernelGGL(build_grid_kernel<false>, dim3((unsigned)nblk), dim3(GB), 0, stream, rxyz, rtype, n_rec,
                               radius, hyd, don, acc, lig_types, n_lt, q0, ox, oy, oz, spacing, nx, ny, nz, prm, out,
                               tiles_j, tiles_k);
    }
    return (int)hipGetLastError();
}

}  // namespace adock
