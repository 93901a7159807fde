// Stage A rigid enumeration for gfx950: for every (complex, conformer, rotation) the lowest per-atom-capped
// nearest-grid energy over the complex's translation lattice, and the lattice index that attains it. Same arithmetic
// and tie rule as the reference's stagea_kernel (src/adock/mckernel.py:451-521): rotation by the normalised
// quaternion, round-half-even grid index clamped to the grid, quadratic out-of-box penalty, min(atom energy, cap)
// summed in atom order, argmin with ties to the lowest translation index.
//
// MI355X mapping: one 256-thread workgroup per (l, c, q); the rotated heavy atoms (xyz + grid-slot) sit in LDS as
// float4 and are read as broadcasts; threads stride over translations and keep a private (min, index); the
// workgroup argmin is a wave butterfly (__shfl_xor) plus one LDS step across the 4 waves. The work is
// scattered L2/MALL grid gathers (one per atom and translation), so many small workgroups keep the CUs' memory
// pipelines full.
#include "common.hpp"
#include "kernels.hpp"

namespace adock {

namespace {

constexpr int SA_THREADS = 256;
constexpr int SA_WAVES = SA_THREADS / WAVE;

DEVI bool sa_better(float e, int i, float be, int bi) { return e < be || (e == be && i < bi); }

__global__ __launch_bounds__(SA_THREADS) void stage_a_kernel(const float* __restrict__ conf,
                                                              const int* __restrict__ gslot,
                                                              const int* __restrict__ na, int AM, int C,
                                                              const float* __restrict__ quats, int Q,
                                                              const float* __restrict__ trans,
                                                              const int* __restrict__ ntrans, int NT_MAX, GridArgs G,
                                                              float cap, float* __restrict__ best_e,
                                                              int* __restrict__ best_i) {
    __shared__ float4 XS[SA_THREADS];   // rotated heavy atoms, AM <= SA_THREADS
    __shared__ float wbe[SA_WAVES];
    __shared__ int wbi[SA_WAVES];
    const long long b = blockIdx.x;
    const int q = (int)(b % Q), c = (int)((b / Q) % C);
    const int l = (int)(b / ((long long)Q * C));
    const int n_at = min(na[l], min(AM, SA_THREADS));
    const int tid = threadIdx.x;
    if (tid < n_at) {
        float qw = quats[4 * q], qx = quats[4 * q + 1], qy = quats[4 * q + 2], qz = quats[4 * q + 3];
        const float nq = 1.f / sqrtf(qw * qw + qx * qx + qy * qy + qz * qz);
        qw *= nq; qx *= nq; qy *= nq; qz *= nq;
        const float* p = conf + (((long long)l * C + c) * AM + tid) * 3;
        const float px = p[0], py = p[1], pz = p[2];
        const float x = (1 - 2 * (qy * qy + qz * qz)) * px + 2 * (qx * qy - qz * qw) * py + 2 * (qx * qz + qy * qw) * pz;
        const float y = 2 * (qx * qy + qz * qw) * px + (1 - 2 * (qx * qx + qz * qz)) * py + 2 * (qy * qz - qx * qw) * pz;
        const float z = 2 * (qx * qz - qy * qw) * px + 2 * (qy * qz + qx * qw) * py + (1 - 2 * (qx * qx + qy * qy)) * pz;
        XS[tid] = make_float4(x, y, z, __int_as_float(gslot[(long long)l * AM + tid]));
    }
    __syncthreads();
    const int ri = G.rec_idx[l];
    const float* Gp = G.grid + G.goffs[ri];
    const int nx = (int)G.gdims[3 * ri], ny = (int)G.gdims[3 * ri + 1], nz = (int)G.gdims[3 * ri + 2];
    const long long n3 = (long long)nx * ny * nz;
    const float ox = G.origin[3 * ri], oy = G.origin[3 * ri + 1], oz = G.origin[3 * ri + 2];
    const float lx = G.lo[3 * ri], ly = G.lo[3 * ri + 1], lz = G.lo[3 * ri + 2];
    const float hx = G.hi[3 * ri], hy = G.hi[3 * ri + 1], hz = G.hi[3 * ri + 2];
    const float inv = 1.f / G.spacing, pen = G.pen;
    float be = 3.4e38f;
    int bi = 0;
    const int nT = ntrans[l];
    for (int ti = tid; ti < nT; ti += SA_THREADS) {
        const float* T = trans + ((long long)l * NT_MAX + ti) * 3;
        const float T0 = T[0], T1 = T[1], T2 = T[2];
        float e = 0.f;
#pragma unroll 4
        for (int a = 0; a < n_at; ++a) {
            const float4 p = XS[a];
            const float x = p.x + T0, y = p.y + T1, z = p.z + T2;
            const int i = min(max((int)rintf((x - ox) * inv), 0), nx - 1);
            const int j = min(max((int)rintf((y - oy) * inv), 0), ny - 1);
            const int k = min(max((int)rintf((z - oz) * inv), 0), nz - 1);
            float ea = Gp[(long long)__float_as_int(p.w) * n3 + ((long long)i * ny + j) * nz + k];
            const float dx0 = fmaxf(lx - x, 0.f), dx1 = fmaxf(x - hx, 0.f);
            const float dy0 = fmaxf(ly - y, 0.f), dy1 = fmaxf(y - hy, 0.f);
            const float dz0 = fmaxf(lz - z, 0.f), dz1 = fmaxf(z - hz, 0.f);
            ea += pen * (dx0 * dx0 + dx1 * dx1 + dy0 * dy0 + dy1 * dy1 + dz0 * dz0 + dz1 * dz1);
            e += fminf(ea, cap);
        }
        if (e < be) { be = e; bi = ti; }
    }
#pragma unroll
    for (int m = 1; m < WAVE; m <<= 1) {
        const float oe = __shfl_xor(be, m);
        const int oi = __shfl_xor(bi, m);
        if (sa_better(oe, oi, be, bi)) { be = oe; bi = oi; }
    }
    const int w = tid / WAVE;
    if ((tid & (WAVE - 1)) == 0) { wbe[w] = be; wbi[w] = bi; }
    __syncthreads();
    if (tid == 0) {
        float e = wbe[0];
        int i = wbi[0];
        for (int u = 1; u < SA_WAVES; ++u)
            if (sa_better(wbe[u], wbi[u], e, i)) { e = wbe[u]; i = wbi[u]; }
        best_e[b] = e;
        best_i[b] = i;
    }
}

}  // namespace

int launch_stage_a(const float* conf, const int* gslot, const int* na, int AM, int L, int C, const float* quats,
                   int Q, const float* trans, const int* ntrans, int NT_MAX, const GridArgs& G, float cap,
                   float* best_e, int* best_i, hipStream_t stream) {
    const long long n = (long long)L * C * Q;
    if (n == 0) return 0;
    hipLaunchKernelGGL(stage_a_kernel, dim3((unsigned)n), dim3(SA_THREADS), 0, stream, conf, gslot, na, AM, C, quats, Q,
                       trans, ntrans, NT_MAX, G, cap, best_e, best_i);
    return (int)hipGetLastError();
}

}  // namespace adock
