// Learned Monte Carlo proposal network (128 -> 256 -> 256 -> 23, ReLU): inference on MFMA, training kernels on VALU.
//
// Inference (mlp_infer_kernel): BF16 weights and activations, FP32 accumulation, gfx950's v_mfma_f32_16x16x32_bf16
// (twice the depth of the CDNA3 16x16x16 form: half the MFMAs per layer). One workgroup (4 wavefronts) owns a 64-row
// tile of the batch: the features are converted to BF16 into LDS, each wave computes a quarter of every layer's
// output tiles for all 64 rows (A fragments from LDS, B fragments straight from the L2-resident weights, stored
// [out, in] so each lane's 8 consecutive K values are one 16-byte load, each fetched once per workgroup), bias + ReLU
// is applied on the accumulators and the BF16 activations go back to LDS for the next layer. Fragment layouts of the 16x16x32 BF16 MFMA: A[i][k] on lane i + 16 (k/8), element k%8;
// B[k][j] on lane j + 16 (k/8), element k%8; D[i][j] on lane j + 16 (i/4), element i%4.
//
// Training: FP32 tiled GEMM (64x64 tiles, 16-deep K slices through LDS, 4x4 outputs per thread) with either operand
// transposed, plus bias / ReLU / loss / Adam elementwise kernels; the host composes forward and backward passes.
#include "common.hpp"
#include "kernels.hpp"

namespace adock {
namespace {

constexpr int IN = NET_IN, HID = NET_HID, OUT = NET_OUT, OUTP = NET_OUTP;
constexpr int ROWS = 64;

// LDS element of (row, col) in a BF16 activation tile with K columns: each row's 16-byte chunks are XOR-swizzled by
// the row (chunk c of row i at c ^ (i % 16)), so the 16 rows one A-fragment read touches (same K offset, row pitch a
// multiple of 256 B) land on 16 different 4-bank groups instead of one. K / 8 >= 16 chunks.
DEVI int swz(int row, int col, int K) { return row * K + ((((col >> 3) ^ (row & 15)) << 3) | (col & 7)); }

// One layer for the workgroup's 64-row tile: out[64 x N] = act(in[64 x K] * W^T + b), W [N, K] bf16 (row = output),
// in / out LDS bf16 [ROWS][K] / [ROWS][N] (swizzled). The N / 16 output tiles are dealt over the 4 waves (tile t to
// wave t % 4) and each wave computes its tiles for all four 16-row strips, so every B fragment (16 outputs x 32 K,
// one 16-byte load per lane from the L2-resident weights) is fetched once per workgroup and feeds 4 MFMAs. Each
// output is the same K-ordered MFMA chain as in a one-strip-per-wave mapping (the in-loop inference of the chain
// kernels relies on that: the same bits).
template <int K, int N, bool RELU>
DEVI void layer_tiles(const unsigned short* __restrict__ in, unsigned short* out, float* outf,
                      const unsigned short* __restrict__ W, const float* __restrict__ b, int wave, int lane,
                      int nvalid) {
    constexpr int NT = N / 16, TPW = (NT + 3) / 4;
    static_assert(K % 32 == 0 && K / 8 >= 16, "layer depth: a multiple of the MFMA's 32, at least 16 chunks");
    float4_t acc[TPW][4];
#pragma unroll
    for (int u = 0; u < TPW; ++u)
#pragma unroll
        for (int st = 0; st < 4; ++st) acc[u][st] = float4_t{0.f, 0.f, 0.f, 0.f};
    const int li = lane & 15, lk = (lane >> 4) * 8;
#pragma unroll 2
    for (int k0 = 0; k0 < K; k0 += 32) {
        bf16x8_t a[4];
#pragma unroll
        for (int st = 0; st < 4; ++st)
            a[st] = *reinterpret_cast<const bf16x8_t*>(in + swz(st * 16 + li, k0 + lk, K));
#pragma unroll
        for (int u = 0; u < TPW; ++u) {
            const int t = wave + 4 * u;
            if (t < NT) {
                const int j = t * 16 + li;
                bf16x8_t bv = {};
                if (j < nvalid) bv = *reinterpret_cast<const bf16x8_t*>(W + (long long)j * K + k0 + lk);
#pragma unroll
                for (int st = 0; st < 4; ++st)
                    acc[u][st] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[st], bv, acc[u][st], 0, 0, 0);
            }
        }
    }
#pragma unroll
    for (int u = 0; u < TPW; ++u) {
        const int t = wave + 4 * u;
        if (t >= NT) continue;
        const int j = t * 16 + li;
        const float bj = j < nvalid ? b[j] : 0.f;
#pragma unroll
        for (int st = 0; st < 4; ++st)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = st * 16 + (lane >> 4) * 4 + r;
                float v = acc[u][st][r] + bj;
                if (RELU) v = fmaxf(v, 0.f);
                if (out) out[swz(i, j, N)] = f2bf(v);
                if (outf && j < nvalid) outf[i * OUTP + j] = v;
            }
    }
}

__global__ __launch_bounds__(256) void mlp_infer_kernel(int B, const float* __restrict__ X,
                                                         const unsigned short* __restrict__ W1, const float* __restrict__ b1,
                                                         const unsigned short* __restrict__ W2, const float* __restrict__ b2,
                                                         const unsigned short* __restrict__ W3, const float* __restrict__ b3,
                                                         float* __restrict__ Y) {
    __shared__ __attribute__((aligned(16))) unsigned short s_a[ROWS * HID];
    __shared__ __attribute__((aligned(16))) unsigned short s_b[ROWS * HID];
    float* s_y = reinterpret_cast<float*>(s_b);   // layer 3 reads s_a only: s_b is free for the FP32 output tile
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const long long row0 = (long long)blockIdx.x * ROWS;
    for (int e = tid; e < ROWS * IN; e += 256) {
        const long long r = row0 + e / IN;
        s_a[swz(e / IN, e % IN, IN)] = f2bf(r < B ? X[r * IN + e % IN] : 0.f);
    }
    __syncthreads();
    layer_tiles<IN, HID, true>(s_a, s_b, nullptr, W1, b1, wave, lane, HID);
    __syncthreads();
    layer_tiles<HID, HID, true>(s_b, s_a, nullptr, W2, b2, wave, lane, HID);
    __syncthreads();
    layer_tiles<HID, OUTP, false>(s_a, nullptr, s_y, W3, b3, wave, lane, OUT);
    __syncthreads();
    for (int e = tid; e < ROWS * OUT; e += 256) {
        const long long r = row0 + e / OUT;
        if (r < B) Y[r * OUT + e % OUT] = s_y[(e / OUT) * OUTP + e % OUT];
    }
}

// C[M,N] = alpha op(A) op(B) + beta C, row-major; op(A) = A [M,K] (lda) or A^T with A [K,M]; same for B.
// SPLIT: workgroup z sums k in [z kc, (z + 1) kc) only and stores its partial tile to P + z M N (row-major [M, N]);
// splitk_reduce_kernel adds the partials in z order (deterministic) and applies alpha and beta. For the weight gradients
// (M, N <= 256, K = the batch): 16 output tiles alone leave most of the chip idle.
template <bool TA, bool TB, bool SPLIT = false>
__global__ __launch_bounds__(256) void gemm_f32_kernel(int M, int N, int K, float alpha, const float* __restrict__ A,
                                                        int lda, const float* __restrict__ Bm, int ldb, float beta,
                                                        float* __restrict__ C, int ldc, int kc = 0,
                                                        float* __restrict__ P = nullptr) {
    constexpr int BM = 64, BN = 64, BK = 16;
    __shared__ float sA[BK][BM + 4];
    __shared__ float sB[BK][BN + 4];
    const int tid = threadIdx.x;
    const int tx = tid % 16, ty = tid / 16;
    const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
    const int kb = SPLIT ? blockIdx.z * kc : 0;
    const int ke = SPLIT ? min(K, kb + kc) : K;
    float acc[4][4] = {};
    for (int k0 = kb; k0 < ke; k0 += BK) {
        for (int e = tid; e < BK * BM; e += 256) {
            const int kk = e / BM, mm = e % BM;
            const int m = m0 + mm, k = k0 + kk;
            float v = 0.f;
            if (m < M && k < ke) v = TA ? A[(long long)k * lda + m] : A[(long long)m * lda + k];
            sA[kk][mm] = v;
        }
        for (int e = tid; e < BK * BN; e += 256) {
            const int kk = e / BN, nn = e % BN;
            const int n = n0 + nn, k = k0 + kk;
            float v = 0.f;
            if (n < N && k < ke) v = TB ? Bm[(long long)n * ldb + k] : Bm[(long long)k * ldb + n];
            sB[kk][nn] = v;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < BK; ++kk) {
            float a[4], bb[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) a[i] = sA[kk][ty * 4 + i];
#pragma unroll
            for (int j = 0; j < 4; ++j) bb[j] = sB[kk][tx * 4 + j];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = fmaf(a[i], bb[j], acc[i][j]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int m = m0 + ty * 4 + i;
        if (m >= M) continue;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int n = n0 + tx * 4 + j;
            if (n >= N) continue;
            if constexpr (SPLIT) {
                P[((long long)blockIdx.z * M + m) * N + n] = acc[i][j];
            } else {
                float* c = C + (long long)m * ldc + n;
                *c = alpha * acc[i][j] + (beta != 0.f ? beta * *c : 0.f);
            }
        }
    }
}

__global__ void splitk_reduce_kernel(int M, int N, int S, const float* __restrict__ P, float alpha, float beta,
                                     float* __restrict__ C, int ldc) {
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long long)M * N) return;
    const long long mn = (long long)M * N;
    float acc = 0.f;
    for (int z = 0; z < S; ++z) acc += P[z * mn + e];
    float* c = C + (e / N) * ldc + e % N;
    *c = alpha * acc + (beta != 0.f ? beta * *c : 0.f);
}

// Y[M,N] += b (broadcast over rows), then ReLU if relu
__global__ void bias_act_kernel(int M, int N, float* __restrict__ Y, const float* __restrict__ b, int relu) {
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long long)M * N) return;
    float v = Y[e] + b[e % N];
    Y[e] = relu ? fmaxf(v, 0.f) : v;
}

// dH[e] *= (H[e] > 0)  (ReLU backward on the post-activation)
__global__ void relu_bwd_kernel(long long n, float* __restrict__ dH, const float* __restrict__ H) {
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n && !(H[e] > 0.f)) dH[e] = 0.f;
}

// db[j] = sum_i dY[i, j]  (one thread per column, rows strided; N small)
__global__ void colsum_kernel(int M, int N, const float* __restrict__ dY, float* __restrict__ db) {
    const int j = blockIdx.x;
    float s = 0.f;
    for (int i = threadIdx.x; i < M; i += blockDim.x) s += dY[(long long)i * N + j];
    s = wave_sum(s);
    __shared__ float part[4];
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        float t = 0.f;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += part[w];
        db[j] = t;
    }
}

// masked mean-squared error: dY = 2 m (Y - T) / norm; loss partial sums into loss[0] (atomic)
__global__ void mse_kernel(long long n, const float* __restrict__ Y, const float* __restrict__ T,
                           const float* __restrict__ Mk, float inv_norm, float* __restrict__ dY,
                           float* __restrict__ loss) {
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    float l = 0.f;
    if (e < n) {
        const float m = Mk ? Mk[e] : 1.f;
        const float d = (Y[e] - T[e]) * m;
        dY[e] = 2.f * d * inv_norm;
        l = d * d * inv_norm;
    }
    l = wave_sum(l);
    if (loss && (threadIdx.x & 63) == 0 && l != 0.f) atomicAdd(loss, l);
}

// Adam (bias-corrected), weight decay off
__global__ void adam_kernel(long long n, float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                            float* __restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2) {
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const float gi = g[e];
    const float mi = b1 * m[e] + (1.f - b1) * gi;
    const float vi = b2 * v[e] + (1.f - b2) * gi * gi;
    m[e] = mi; v[e] = vi;
    w[e] -= lr * (mi * c1) / (sqrtf(vi * c2) + eps);
}

__global__ void to_bf16_kernel(long long n, const float* __restrict__ x, unsigned short* __restrict__ y) {
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) y[e] = f2bf(x[e]);
}

inline unsigned nb(long long n, int t = 256) { return (unsigned)((n + t - 1) / t); }

}  // namespace

int launch_mlp_infer(int B, const float* X, const unsigned short* W1, const float* b1, const unsigned short* W2,
                     const float* b2, const unsigned short* W3, const float* b3, float* Y, hipStream_t s) {
    hipLaunchKernelGGL(mlp_infer_kernel, dim3(nb(B, ROWS)), dim3(256), 0, s, B, X, W1, b1, W2, b2, W3, b3, Y);
    return (int)hipGetLastError();
}

int launch_gemm_f32(int ta, int tb, int M, int N, int K, float alpha, const float* A, int lda, const float* Bm,
                    int ldb, float beta, float* C, int ldc, hipStream_t s) {
    const dim3 grid((N + 63) / 64, (M + 63) / 64);
    if (!ta && !tb) hipLaunchKernelGGL((gemm_f32_kernel<false, false>), grid, dim3(256), 0, s, M, N, K, alpha, A, lda, Bm, ldb, beta, C, ldc);
    else if (!ta && tb) hipLaunchKernelGGL((gemm_f32_kernel<false, true>), grid, dim3(256), 0, s, M, N, K, alpha, A, lda, Bm, ldb, beta, C, ldc);
    else if (ta && !tb) hipLaunchKernelGGL((gemm_f32_kernel<true, false>), grid, dim3(256), 0, s, M, N, K, alpha, A, lda, Bm, ldb, beta, C, ldc);
    else hipLaunchKernelGGL((gemm_f32_kernel<true, true>), grid, dim3(256), 0, s, M, N, K, alpha, A, lda, Bm, ldb, beta, C, ldc);
    return (int)hipGetLastError();
}

int launch_gemm_f32_splitk(int ta, int tb, int M, int N, int K, float alpha, const float* A, int lda, const float* Bm,
                           int ldb, float beta, float* C, int ldc, int splits, float* work, hipStream_t s) {
    const int kc = ((K + splits - 1) / splits + 15) & ~15;   // whole 16-deep K steps per split
    const int S = (K + kc - 1) / kc;
    const dim3 grid((N + 63) / 64, (M + 63) / 64, S);
    if (!ta && !tb) hipLaunchKernelGGL((gemm_f32_kernel<false, false, true>), grid, dim3(256), 0, s, M, N, K, alpha, A, lda, Bm, ldb, beta, C, ldc, kc, work);
    else if (!ta && tb) hipLaunchKernelGGL((gemm_f32_kernel<false, true, true>), grid, dim3(256), 0, s, M, N, K, alpha, A, lda, Bm, ldb, beta, C, ldc, kc, work);
    else if (ta && !tb) hipLaunchKernelGGL((gemm_f32_kernel<true, false, true>), grid, dim3(256), 0, s, M, N, K, alpha, A, lda, Bm, ldb, beta, C, ldc, kc, work);
    else hipLaunchKernelGGL((gemm_f32_kernel<true, true, true>), grid, dim3(256), 0, s, M, N, K, alpha, A, lda, Bm, ldb, beta, C, ldc, kc, work);
    hipLaunchKernelGGL(splitk_reduce_kernel, dim3(nb((long long)M * N, 256)), dim3(256), 0, s, M, N, S, work, alpha,
                       beta, C, ldc);
    return (int)hipGetLastError();
}

int launch_bias_act(int M, int N, float* Y, const float* b, int relu, hipStream_t s) {
    hipLaunchKernelGGL(bias_act_kernel, dim3(nb((long long)M * N)), dim3(256), 0, s, M, N, Y, b, relu);
    return (int)hipGetLastError();
}
int launch_relu_bwd(long long n, float* dH, const float* H, hipStream_t s) {
    hipLaunchKernelGGL(relu_bwd_kernel, dim3(nb(n)), dim3(256), 0, s, n, dH, H);
    return (int)hipGetLastError();
}
int launch_colsum(int M, int N, const float* dY, float* db, hipStream_t s) {
    hipLaunchKernelGGL(colsum_kernel, dim3(N), dim3(256), 0, s, M, N, dY, db);
    return (int)hipGetLastError();
}
int launch_mse(long long n, const float* Y, const float* T, const float* Mk, float inv_norm, float* dY, float* loss,
               hipStream_t s) {
    hipLaunchKernelGGL(mse_kernel, dim3(nb(n)), dim3(256), 0, s, n, Y, T, Mk, inv_norm, dY, loss);
    return (int)hipGetLastError();
}
int launch_adam(long long n, float* w, const float* g, float* m, float* v, float lr, float b1, float b2, float eps,
                int t, hipStream_t s) {
    const float c1 = 1.f / (1.f - powf(b1, (float)t)), c2 = 1.f / (1.f - powf(b2, (float)t));
    hipLaunchKernelGGL(adam_kernel, dim3(nb(n)), dim3(256), 0, s, n, w, g, m, v, lr, b1, b2, eps, c1, c2);
    return (int)hipGetLastError();
}
int launch_to_bf16(long long n, const float* x, unsigned short* y, hipStream_t s) {
    hipLaunchKernelGGL(to_bf16_kernel, dim3(nb(n)), dim3(256), 0, s, n, x, y);
    return (int)hipGetLastError();
}

}  // namespace adock
