// Chain kernels: launch settings and the dispatch over torsion buckets. The kernels themselves are in dock_impl.hpp,
// instantiated per bucket in dock_km<KM>.hip (see that header for the design).
#include "dock_buckets.hpp"
#include "dock_impl.hpp"

namespace adock {

// Process-wide launch defaults, atomic so that a setter racing a launch on another thread is well defined (the launch
// sees the old or the new value). The batched launch also takes its chains per workgroup per call.
std::atomic<int> g_waves_per_block{0};     // 0: adaptive (auto_waves)
std::atomic<int> g_line_search_width{0};   // 0: adaptive (line_search_width)
std::atomic<int> g_exact_staging{2};       // 0: read from memory, 1: staged, 2: staged with per-atom masks

int set_waves_per_block(int w) {
    if (w < 0 || w > 4) return -1;
    g_waves_per_block = w;
    return 0;
}
int get_waves_per_block() { return g_waves_per_block; }
int set_line_search_width(int ns) {
    if (ns != 0 && ns != 1 && ns != 2 && ns != 4) return -1;
    g_line_search_width = ns;
    return 0;
}
int get_line_search_width() { return g_line_search_width; }
int set_exact_staging(int on) {
    if (on < 0 || on > 2) return -1;
    g_exact_staging = on;
    return 0;
}
int get_exact_staging() { return g_exact_staging; }

int atoms_per_lane(int AM) { return AM <= 64 ? 1 : 2; }
int lds_budget_ok(int KM, int P_MAX, int AM) {
    return wave_lds_bytes(KM, P_MAX, atoms_per_lane(AM)) <= LDS_PER_WORKGROUP;
}

#define ADOCK_CASE_MC(KM) \
    case KM: return launch_mc_km##KM(L, G, S, R, exact, chain_lig, n_chains, M, stream);
int launch_mc(int KM, const LigArgs& L, const GridArgs& G, const SFArgs& S, const RecArgs& R, int exact,
              const int* chain_lig, int n_chains, const McArgs& M, hipStream_t stream) {
    switch (KM) {
        ADOCK_KM_LIST(ADOCK_CASE_MC)
        default: return -1;
    }
}

#define ADOCK_CASE_RW(KM) \
    case KM: return mc_resident_waves_km##KM(P_MAX, persistent, atoms_per_lane(AM));
int mc_resident_waves(int KM, int P_MAX, int persistent, int AM) {
    switch (KM) {
        ADOCK_KM_LIST(ADOCK_CASE_RW)
        default: return -1;
    }
}

#define ADOCK_CASE_PERS(KM) \
    case KM: return launch_mc_persistent_km##KM(L, G, S, R, exact, chain_lig, n_chains, M, Q, blocks, stream);
int launch_mc_persistent(int KM, const LigArgs& L, const GridArgs& G, const SFArgs& S, const RecArgs& R, int exact,
                         const int* chain_lig, int n_chains, const McArgs& M, const QueueArgs& Q, int blocks,
                         hipStream_t stream) {
    switch (KM) {
        ADOCK_KM_LIST(ADOCK_CASE_PERS)
        default: return -1;
    }
}

#define ADOCK_CASE_STEP(KM) \
    case KM: return launch_mc_step_km##KM(L, G, S, R, chain_lig, n_chains, M, P, stream);
int launch_mc_step(int KM, const LigArgs& L, const GridArgs& G, const SFArgs& S, const RecArgs& R,
                   const int* chain_lig, int n_chains, const McArgs& M, const StepArgs& P, hipStream_t stream) {
    switch (KM) {
        ADOCK_KM_LIST(ADOCK_CASE_STEP)
        default: return -1;
    }
}

#define ADOCK_CASE_LEARNED(KM) \
    case KM: return launch_mc_learned_km##KM(L, G, S, R, chain_lig, n_chains, M, P, N, Q, stream);
int launch_mc_learned(int KM, const LigArgs& L, const GridArgs& G, const SFArgs& S, const RecArgs& R,
                      const int* chain_lig, int n_chains, const McArgs& M, const StepArgs& P, const NetArgs& N,
                      const QueueArgs& Q, hipStream_t stream) {
    switch (KM) {
        ADOCK_KM_LIST(ADOCK_CASE_LEARNED)
        default: return -1;
    }
}

#define ADOCK_CASE_EG(KM) \
    case KM: return launch_eg_km##KM(L, G, S, R, exact, chain_lig, n_chains, state, e_out, g_out, stream);
int launch_eg(int KM, const LigArgs& L, const GridArgs& G, const SFArgs& S, const RecArgs& R, int exact,
              const int* chain_lig, int n_chains, const float* state, float* e_out, float* g_out,
              hipStream_t stream) {
    switch (KM) {
        ADOCK_KM_LIST(ADOCK_CASE_EG)
        default: return -1;
    }
}

__global__ void batch_reset_kernel(int* counter, int n_counters, int* done, int n_done, unsigned long long* times,
                                   int n_times) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (counter && i < n_counters) counter[i] = 0;
    if (done && i < n_done) done[i] = 0;
    if (times && i < n_times) times[i] = i == 0 ? ~0ull : 0ull;
}

int launch_batch_reset(int* counter, int n_counters, int* done, int n_done, unsigned long long* times, int n_times,
                       hipStream_t stream) {
    int n = n_done > n_times ? n_done : n_times;
    if (n < n_counters) n = n_counters;
    if (n < 1) n = 1;
    hipLaunchKernelGGL(batch_reset_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, counter, n_counters, done,
                       n_done, times, n_times);
    return (int)hipGetLastError();
}

}  // namespace adock
