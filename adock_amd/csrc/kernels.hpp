// Argument bundles and launch functions shared by the kernel translation units and the C ABI layer.
#pragma once
#include <hip/hip_runtime.h>

namespace adock {
struct LigArgs {
    const float* ref; const int* typ; const unsigned char* tm; const int* ax; const float* axp; const float* axu;
    const int* pr; const int* na; const int* nt; const int* np; const float* rg; int P_MAX; const int* gslot;
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
};
int launch_build_grid(const float*, const int*, int, const float*, const int*, const int*, const int*, const int*,
                      int, float, float, float, float, int, int, int, const float*, int, float*, hipStream_t);
int launch_mc(int KM, const LigArgs&, const GridArgs&, const SFArgs&, const RecArgs&, int, const int*, int,
              const McArgs&, hipStream_t);
int launch_eg(int KM, const LigArgs&, const GridArgs&, const SFArgs&, const RecArgs&, int, const int*, int,
              const float*, float*, float*, hipStream_t);
int set_waves_per_block(int w);
int get_waves_per_block();
int lds_budget_ok(int P_MAX);
}  // namespace adock
