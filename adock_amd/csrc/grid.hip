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

__global__ __launch_bounds__(GB) void build_grid_fit_kernel(
    const float* __restrict__ rxyz, const int* __restrict__ rext, int n_rec, const float* __restrict__ rad_rec,
    const float* __restrict__ rad_lig, int n_lt, const float* __restrict__ W, int n_rt, int q0, float ox, float oy,
    float oz, float spacing, int nx, int ny, int nz, const float* __restrict__ prm, float* __restrict__ out,
    int tiles_j, int tiles_k) {
    __shared__ float4 s_atom[GB];      // x, y, z, radius
    __shared__ int s_rt[GB];           // receptor ext type
    __shared__ int s_wcount[GB / WAVE];
    __shared__ float s_w[FIT_RT_MAX * TQ * FIT_NB];
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
    // prm: cutoff, g1_off, g1_w, g2_off, g2_w, rep_off, hyd_good, hyd_bad, hb_good, hb_bad, ion_rmin, rep_cap
    const float cut = prm[0];
    const float cut2 = cut * cut;
    const float gr = cut + 1e-3f;
    const float blo0 = fmaf((float)(bi * TI), spacing, ox) - gr, bhi0 = fmaf((float)(bi * TI + TI - 1), spacing, ox) + gr;
    const float blo1 = fmaf((float)(bj * TJ), spacing, oy) - gr, bhi1 = fmaf((float)(bj * TJ + TJ - 1), spacing, oy) + gr;
    const float blo2 = fmaf((float)(bk * TK), spacing, oz) - gr, bhi2 = fmaf((float)(bk * TK + TK - 1), spacing, oz) + gr;
    const float rl = 1.2011224087864498f;   // sqrt(log2 e): exp(-u^2) = exp2(-(u rl)^2)
    const float g1o = prm[1], g1s = rl / prm[2], g2o = prm[3], g2s = rl / prm[4], repo = prm[5];
    const float hbad = prm[7], hinv = 1.f / (prm[7] - prm[6]);
    const float bbad = prm[9], binv = 1.f / (prm[9] - prm[8]);
    const float rmin = prm[10], cap = prm[11], icut2 = 1.f / cut2;

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
            if (r2 < cut2) {
                const float rr = fast_sqrt(r2);
                const float rc = fmaxf(rr, rmin);
                const float ion = 1.f / (rc * rc) - icut2;
                const float rmr = rr - r.w;
                const float* wb = s_w + s_rt[b] * (TQ * FIT_NB);
#pragma unroll
                for (int q = 0; q < TQ; ++q) {
                    if (q >= nq) break;
                    const float* wq = wb + q * FIT_NB;
                    const float d = rmr - rq[q];
                    const float u1 = (d - g1o) * g1s, u2 = (d - g2o) * g2s;
                    float sv = wq[0] * fast_exp2(-u1 * u1);
                    sv = fmaf(wq[1], fast_exp2(-u2 * u2), sv);
                    const float dr = fminf(d - repo, 0.f);
                    sv = fmaf(wq[2], fminf(dr * dr, cap), sv);
                    sv = fmaf(wq[3], fminf(fmaxf((hbad - d) * hinv, 0.f), 1.f), sv);
                    sv = fmaf(wq[4], fminf(fmaxf((bbad - d) * binv, 0.f), 1.f), sv);
                    sv = fmaf(wq[5], ion, sv);
                    acc_e[q] += sv;
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
                           tiles_k);
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
