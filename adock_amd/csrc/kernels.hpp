// Argument bundles and launch functions shared by the kernel translation units and the C ABI layer.
#pragma once
#include <hip/hip_runtime.h>

namespace adock {
struct LigArgs {
    const float* ref; const int* typ; const unsigned char* tm; const int* ax; const float* axp; const float* axu;
    const int* pr; const int* na; const int* nt; const int* np; const float* rg; int P_MAX; const int* gslot;
    const unsigned* inc; const int* ist; int DEG_MAX;
    int AM = 64;   // heavy-atom bucket of the packed arrays: 64 (one atom per lane) or 96 / 128 (two per lane)
};
struct GridArgs {
    const float* grid; const float* origin; const float* lo; const float* hi; const int* rec_idx;
    const long long* goffs; const long long* gdims; int T; float spacing; float pen;
};
struct SFArgs { const float* prm; const float* rad; const int* hyd; const int* don; const int* acc; };
struct RecArgs { const float* rxyz; const int* rtyp; const int* rn; int RN_MAX; };
struct McArgs {
    const int* chain_id; const int* nbr; const int* nbr_n; int NB_MAX;
    float* cur; float* cur_e; float* best; float* best_e;
    int init_opt, mc_steps, bfgs_iters; float temp, amp_pos, amp_rot; unsigned long long seed; int step_offset;
    unsigned long long* eval_count;
    unsigned long long* lig_evals = nullptr;   // per-complex evaluation counts [n_lig_evals] (NULL: none)
    int n_lig_evals = 0;
    int NBS = 0;   // exact mode: receptor atoms staged in each wave's LDS (set by the launcher; 0 = read from memory)
    int NBM = 0;   // exact mode: per-atom masks over the staged atoms (set by the launcher)
};
// work queues of the persistent chain kernel: order [n] (chain indices, may be NULL = identity) holds nq consecutive
// queues, queue q = order[qstart[q] .. qstart[q + 1]) (qstart NULL or nq = 0: one queue over all n); counter [max(nq,
// 1)] (0 at launch), done [L] (finished chains per complex, 0 at launch), lig_chains [L] (chains per complex), times
// [1 + L] (times[0] = ~0 at launch; 100 MHz real-time stamps). done / lig_chains / times may be NULL.
struct QueueArgs {
    const int* order; const int* qstart; int nq; int* counter; int* done; const int* lig_chains;
    unsigned long long* times;
    int n_done = 0x7fffffff;   // entries of done / lig_chains (times: 1 + n_done); complexes at or past it are not counted
    int waves = 0;             // chains (wavefronts) per workgroup of this launch; 0: the process default
    int dense = 0;             // pair-gradient LDS tile: 0 compact (a slot per pair), 1 dense rows (one atom per lane,
                               // grid mode; other launches keep the compact one)
};
// learned-proposal MC step: Y [N,23] network output (NULL: plain moves), feat [N,128] features (in: at the current
// pose; out: at the new current pose; NULL: not computed), log_buf [cap, 128+23+23+2] move records (accepted moves,
// or every move with log_all). net_mode 0: the network's output is a displacement; 1: it scores the entities and the
// reference's mutation of the drawn one (probability ~ exp(score / tau)) is applied.
struct StepArgs {
    const float* Y; float* feat; float* log_buf; int* log_count; int log_cap;
    int use_net; float p_plain, sig_t, sig_r, sig_th; int init;
    int net_mode = 0; float tau = 0.1f; int log_all = 0;
};
// proposal network weights for inference inside the chain loop: BF16 [out, in] (W1 [256,128], W2 [256,256],
// W3 [23,256]) and FP32 biases
struct NetArgs {
    const unsigned short* W1; const float* b1; const unsigned short* W2; const float* b2; const unsigned short* W3;
    const float* b3;
};
int launch_build_grid(const float*, const int*, int, const float*, const int*, const int*, const int*, const int*,
                      int, float, float, float, float, int, int, int, const float*, int, float*, hipStream_t);
int launch_build_grid_fit(const float*, const int*, int, const float*, const float*, int, const float*, int, float,
                          float, float, float, int, int, int, const float*, float*, hipStream_t);
int build_grid_fit_rt_max();
int launch_build_grid_fit_sets(const float*, const int*, int, int, const float*, const float*, int, const float*, int,
                               const float*, float, int, int, int, const float*, float*, hipStream_t);
int launch_inter_exact_fit(const float*, const int*, int, int, int, const int*, const float*, const int*, int, int,
                           const float*, const float*, int, const float*, int, const float*, float*, hipStream_t);
int launch_mc(int KM, const LigArgs&, const GridArgs&, const SFArgs&, const RecArgs&, int, const int*, int,
              const McArgs&, hipStream_t);
int launch_eg(int KM, const LigArgs&, const GridArgs&, const SFArgs&, const RecArgs&, int, const int*, int,
              const float*, float*, float*, hipStream_t);
int launch_mc_persistent(int KM, const LigArgs&, const GridArgs&, const SFArgs&, const RecArgs&, int, const int*, int,
                         const McArgs&, const QueueArgs&, int blocks, hipStream_t);
int launch_batch_reset(int* counter, int n_counters, int* done, int n_done, unsigned long long* times,
                       int n_times, hipStream_t);
int launch_mlp_infer(int B, const float* X, const unsigned short* W1, const float* b1, const unsigned short* W2,
                     const float* b2, const unsigned short* W3, const float* b3, float* Y, hipStream_t s);
int launch_gemm_f32(int ta, int tb, int M, int N, int K, float alpha, const float* A, int lda, const float* Bm,
                    int ldb, float beta, float* C, int ldc, hipStream_t s);
int launch_gemm_f32_splitk(int ta, int tb, int M, int N, int K, float alpha, const float* A, int lda, const float* Bm,
                           int ldb, float beta, float* C, int ldc, int splits, float* work, hipStream_t s);
int launch_bias_act(int M, int N, float* Y, const float* b, int relu, hipStream_t s);
int launch_relu_bwd(long long n, float* dH, const float* H, hipStream_t s);
int launch_colsum(int M, int N, const float* dY, float* db, hipStream_t s);
int launch_mse(long long n, const float* Y, const float* T, const float* Mk, float inv_norm, float* dY, float* loss,
               hipStream_t s);
int launch_adam(long long n, float* w, const float* g, float* m, float* v, float lr, float b1, float b2, float eps,
                int t, hipStream_t s);
int launch_to_bf16(long long n, const float* x, unsigned short* y, hipStream_t s);
int launch_mc_step(int KM, const LigArgs&, const GridArgs&, const SFArgs&, const RecArgs&, const int*, int,
                   const McArgs&, const StepArgs&, hipStream_t);
int launch_mc_learned(int KM, const LigArgs&, const GridArgs&, const SFArgs&, const RecArgs&, const int*, int,
                      const McArgs&, const StepArgs&, const NetArgs&, const QueueArgs&, hipStream_t);
int launch_stage_a(const float* conf, const int* gslot, const int* na, int AM, int L, int C, const float* quats,
                   int Q, const float* trans, const int* ntrans, int NT_MAX, const GridArgs& G, float cap,
                   float* best_e, int* best_i, hipStream_t stream);
int set_waves_per_block(int w);
int get_waves_per_block();
int set_line_search_width(int ns);
int get_line_search_width();
int set_exact_staging(int on);
int get_exact_staging();
int lds_budget_ok(int KM, int P_MAX, int AM);
int mc_resident_waves(int KM, int P_MAX, int persistent, int AM);
}  // namespace adock
