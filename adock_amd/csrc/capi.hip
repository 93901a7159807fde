// extern "C" entry points of libadock_hip.so (declared in include/adock_hip.h): argument validation, bucket
// dispatch, error text. No allocation, no copies, no synchronisation.
#include <cstdio>
#include <mutex>
#include <string>

#include "../../include/adock_hip.h"
#include "common.hpp"

#include "kernels.hpp"

using namespace adock;

static thread_local std::string g_err;

static int fail(int code, const char* msg) {
    g_err = msg;
    return code;
}
static int hip_status(int st) {
    if (st != 0) g_err = std::string("HIP launch error: ") + hipGetErrorString((hipError_t)st);
    return st;
}

// AM: 64 (up to 64 heavy atoms, one per lane) or the 32-multiple buckets 96 and 128 (two atoms per lane); KM: the
// 8-multiple buckets up to 56 (state lanes 7 + KM of one wave)
static int check_buckets(int AM, int KM, int P_MAX) {
    if (AM != 64 && AM != 96 && AM != 128)
        return fail(ADOCK_E_BUCKET, "AM must be 64, 96 or 128 (at most 128 heavy atoms per ligand)");
    if (KM < 8 || KM > 56 || (KM & 7)) return fail(ADOCK_E_BUCKET, "KM must be a multiple of 8 in 8..56");
    if (P_MAX < 1 || P_MAX > 16256 || !lds_budget_ok(KM, P_MAX, AM))
        return fail(ADOCK_E_LDS, "P_MAX out of range for the per-wave LDS budget");
    // 64 heavy atoms have at most 2,016 distinct pairs; the one-atom-per-lane kernels track at most 32 rounds of 128
    if (AM == 64 && P_MAX > 4096) return fail(ADOCK_E_LDS, "P_MAX above 4096 for AM 64 (at most 2,016 distinct pairs)");
    return 0;
}

extern "C" {

const char* adock_version(void) { return "adock_hip 0.1 gfx950"; }
const char* adock_last_error(void) { return g_err.c_str(); }

int adock_build_grid(const float* rxyz, const int* rtype, int n_rec, const float* radius, const int* hyd,
                     const int* don, const int* acc, const int* lig_types, int n_lt, float ox, float oy, float oz,
                     float spacing, int nx, int ny, int nz, const float* prm, int per_term, float* out,
                     adock_stream_t stream) {
    if (n_rec < 0 || n_lt < 0 || nx < 0 || ny < 0 || nz < 0) return fail(ADOCK_E_ARG, "negative size");
    if ((long long)nx * ny * nz == 0 || n_lt == 0) return 0;
    if (!radius || !hyd || !don || !acc || !lig_types || !prm || !out || (n_rec > 0 && (!rxyz || !rtype)))
        return fail(ADOCK_E_ARG, "null pointer");
    if (n_lt > 24) return fail(ADOCK_E_TYPES, "at most 24 ligand types per call");
    if ((long long)nx * ny * nz > (1ll << 31) - 1) return fail(ADOCK_E_ARG, "grid too large");
    return hip_status(launch_build_grid(rxyz, rtype, n_rec, radius, hyd, don, acc, lig_types, n_lt, ox, oy, oz,
                                        spacing, nx, ny, nz, prm, per_term, out, (hipStream_t)stream));
}

int adock_build_grid_fit(const float* rxyz, const int* rext, int n_rec, const float* rad_rec, const float* rad_lig,
                         int n_lt, const float* W, int n_rt, float ox, float oy, float oz, float spacing, int nx,
                         int ny, int nz, const float* prm, float* out, adock_stream_t stream) {
    if (n_rec < 0 || n_lt < 0 || n_rt < 0 || nx < 0 || ny < 0 || nz < 0) return fail(ADOCK_E_ARG, "negative size");
    if ((long long)nx * ny * nz == 0 || n_lt == 0) return 0;
    if (!rad_lig || !prm || !out || (n_rec > 0 && (!rxyz || !rext || !rad_rec || !W)))
        return fail(ADOCK_E_ARG, "null pointer");
    if (n_rt > build_grid_fit_rt_max() || (n_rec > 0 && n_rt < 1))
        return fail(ADOCK_E_TYPES, "n_rt must be in 1..216 receptor types");
    if ((long long)nx * ny * nz > (1ll << 31) - 1) return fail(ADOCK_E_ARG, "grid too large");
    return hip_status(launch_build_grid_fit(rxyz, rext, n_rec, rad_rec, rad_lig, n_lt, W, n_rt, ox, oy, oz, spacing,
                                            nx, ny, nz, prm, out, (hipStream_t)stream));
}

int adock_build_grid_fit_sets(const float* rxyz, const int* rext, int n_sets, int n_rec, const float* rad_rec,
                              const float* rad_lig, int n_lt, const float* W, int n_rt, const float* origins,
                              float spacing, int nx, int ny, int nz, const float* prm, float* out,
                              adock_stream_t stream) {
    if (n_sets < 0 || n_rec < 0 || n_lt < 0 || n_rt < 0 || nx < 0 || ny < 0 || nz < 0)
        return fail(ADOCK_E_ARG, "negative size");
    if ((long long)nx * ny * nz == 0 || n_lt == 0 || n_sets == 0) return 0;
    if (n_sets > 65535) return fail(ADOCK_E_ARG, "at most 65,535 receptor sets per call");
    if (!rad_lig || !prm || !out || !origins || (n_rec > 0 && (!rxyz || !rext || !rad_rec || !W)))
        return fail(ADOCK_E_ARG, "null pointer");
    if (n_rt > build_grid_fit_rt_max() || (n_rec > 0 && n_rt < 1))
        return fail(ADOCK_E_TYPES, "n_rt must be in 1..216 receptor types");
    if ((long long)nx * ny * nz > (1ll << 31) - 1) return fail(ADOCK_E_ARG, "grid too large");
    return hip_status(launch_build_grid_fit_sets(rxyz, rext, n_sets, n_rec, rad_rec, rad_lig, n_lt, W, n_rt, origins,
                                                 spacing, nx, ny, nz, prm, out, (hipStream_t)stream));
}

int adock_inter_exact_fit(const float* xyz, const int* lig_type, int n_cx, int n_pose, int n_atom,
                          const int* rec_idx, const float* rxyz, const int* rext, int n_sets, int n_rec,
                          const float* rad_rec, const float* rad_lig, int n_lt, const float* W, int n_rt,
                          const float* prm, float* out, adock_stream_t stream) {
    if (n_cx < 0 || n_pose < 0 || n_atom < 0 || n_sets < 0 || n_rec < 0 || n_lt < 0 || n_rt < 0)
        return fail(ADOCK_E_ARG, "negative size");
    if ((long long)n_cx * n_pose == 0) return 0;
    if (!out || !rec_idx || !prm || (n_atom > 0 && (!xyz || !lig_type)))
        return fail(ADOCK_E_ARG, "null pointer");
    if (n_sets < 1) return fail(ADOCK_E_ARG, "at least one receptor set");
    if (n_rec > 0 && (!rxyz || !rext || !rad_rec || !W || !rad_lig || n_rt < 1 || n_lt < 1))
        return fail(ADOCK_E_ARG, "receptor atoms need rxyz, rext, rad_rec, rad_lig, W and n_rt, n_lt >= 1");
    if ((long long)n_cx * ((n_pose + 3) / 4) > (1ll << 31) - 1) return fail(ADOCK_E_ARG, "too many poses");
    return hip_status(launch_inter_exact_fit(xyz, lig_type, n_cx, n_pose, n_atom, rec_idx, rxyz, rext, n_sets, n_rec,
                                             rad_rec, rad_lig, n_lt, W, n_rt, prm, out, (hipStream_t)stream));
}

#define LIG_DECL                                                                                                     \
    const float *ref, const int *typ, const unsigned char *tm, const int *ax, const float *axp, const float *axu,    \
        const int *pr, const int *na, const int *nt, const int *np, const float *rg, int P_MAX, const int *gslot, \
        const unsigned *inc, const int *ist, int DEG_MAX
#define GRID_DECL                                                                                                    \
    const float *grid, const float *origin, const float *lo, const float *hi, const int *rec_idx,                    \
        const long long *goffs, const long long *gdims, int T, float spacing, float pen
#define SF_DECL const float *prm, const float *rad, const int *hyd, const int *don, const int *acc
#define REC_DECL const float *rxyz, const int *rtyp, const int *rn, int RN_MAX

static int common_checks(int AM, int KM, int P_MAX, const void* ref, const void* inc, const void* ist, int DEG_MAX,
                         const void* grid, const void* prm, int exact, const void* rxyz, const void* rn, int RN_MAX,
                         const void* chain_lig, int n_chains) {
    if (int st = check_buckets(AM, KM, P_MAX)) return st;
    if (DEG_MAX < 0 || DEG_MAX > AM || (DEG_MAX & 7))
        return fail(ADOCK_E_ARG, "DEG_MAX must be a multiple of 8 in 0..AM");
    if (n_chains < 0) return fail(ADOCK_E_ARG, "negative n_chains");
    if (!ref || !inc || !ist || !grid || !prm || (n_chains > 0 && !chain_lig)) return fail(ADOCK_E_ARG, "null pointer");
    if (exact && (!rxyz || !rn || RN_MAX <= 0)) return fail(ADOCK_E_ARG, "exact mode needs receptor atoms");
    return 0;
}

int adock_eg(int AM, int KM, LIG_DECL, GRID_DECL, SF_DECL, REC_DECL, int exact, const int* chain_lig, int n_chains,
             const float* state, float* e_out, float* g_out, adock_stream_t stream) {
    if (int st = common_checks(AM, KM, P_MAX, ref, inc, ist, DEG_MAX, grid, prm, exact, rxyz, rn, RN_MAX, chain_lig,
                               n_chains))
        return st;
    if (n_chains == 0) return 0;
    if (!state || !e_out || !g_out) return fail(ADOCK_E_ARG, "null pointer");
    LigArgs L{ref, typ, tm, ax, axp, axu, pr, na, nt, np, rg, P_MAX, gslot, inc, ist, DEG_MAX, AM};
    GridArgs G{grid, origin, lo, hi, rec_idx, goffs, gdims, T, spacing, pen};
    SFArgs S{prm, rad, hyd, don, acc};
    RecArgs R{rxyz, rtyp, rn, RN_MAX};
    return hip_status(launch_eg(KM, L, G, S, R, exact, chain_lig, n_chains, state, e_out, g_out, (hipStream_t)stream));
}

int adock_mc_counted(int AM, int KM, LIG_DECL, GRID_DECL, SF_DECL, REC_DECL, int exact, const int* chain_lig,
                     const int* chain_id, int n_chains, const int* nbr, const int* nbr_n, int NB_MAX, float* cur,
                     float* cur_e, float* best, float* best_e, int init_opt, int mc_steps, int bfgs_iters, float temp,
                     float amp_pos, float amp_rot, unsigned long long seed, int step_offset,
                     unsigned long long* eval_count, adock_stream_t stream) {
    if (int st = common_checks(AM, KM, P_MAX, ref, inc, ist, DEG_MAX, grid, prm, exact, rxyz, rn, RN_MAX, chain_lig,
                               n_chains))
        return st;
    if (n_chains == 0) return 0;
    if (!chain_id || !cur || !cur_e || !best || !best_e) return fail(ADOCK_E_ARG, "null pointer");
    if (mc_steps < 0) return fail(ADOCK_E_ARG, "negative mc_steps");
    if (nbr && (!nbr_n || NB_MAX <= 0)) return fail(ADOCK_E_ARG, "nbr needs nbr_n and NB_MAX > 0");
    LigArgs L{ref, typ, tm, ax, axp, axu, pr, na, nt, np, rg, P_MAX, gslot, inc, ist, DEG_MAX, AM};
    GridArgs G{grid, origin, lo, hi, rec_idx, goffs, gdims, T, spacing, pen};
    SFArgs S{prm, rad, hyd, don, acc};
    RecArgs R{rxyz, rtyp, rn, RN_MAX};
    McArgs M{chain_id, nbr, nbr_n, NB_MAX, cur, cur_e, best, best_e, init_opt, mc_steps, bfgs_iters, temp, amp_pos,
             amp_rot, seed, step_offset, eval_count};
    return hip_status(launch_mc(KM, L, G, S, R, exact, chain_lig, n_chains, M, (hipStream_t)stream));
}

int adock_mc(int AM, int KM, LIG_DECL, GRID_DECL, SF_DECL, REC_DECL, int exact, const int* chain_lig,
             const int* chain_id, int n_chains, const int* nbr, const int* nbr_n, int NB_MAX, float* cur, float* cur_e,
             float* best, float* best_e, int init_opt, int mc_steps, int bfgs_iters, float temp, float amp_pos,
             float amp_rot, unsigned long long seed, int step_offset, adock_stream_t stream) {
    return adock_mc_counted(AM, KM, ref, typ, tm, ax, axp, axu, pr, na, nt, np, rg, P_MAX, gslot, inc, ist, DEG_MAX,
                            grid, origin, lo, hi,
                            rec_idx, goffs, gdims, T, spacing, pen, prm, rad, hyd, don, acc, rxyz, rtyp, rn, RN_MAX,
                            exact, chain_lig, chain_id, n_chains, nbr, nbr_n, NB_MAX, cur, cur_e, best, best_e,
                            init_opt, mc_steps, bfgs_iters, temp, amp_pos, amp_rot, seed, step_offset, nullptr,
                            stream);
}

int adock_mc_batch(int AM, int KM, LIG_DECL, GRID_DECL, SF_DECL, REC_DECL, int exact, const int* chain_lig,
                   const int* chain_id, int n_chains, const int* nbr, const int* nbr_n, int NB_MAX, float* cur,
                   float* cur_e, float* best, float* best_e, int init_opt, int mc_steps, int bfgs_iters, float temp,
                   float amp_pos, float amp_rot, unsigned long long seed, int step_offset, const int* order,
                   const int* queue_start, int n_queues, int* counter, int* done, const int* lig_chains,
                   unsigned long long* times, int blocks, int waves_per_block, int lds_layout,
                   unsigned long long* eval_count, unsigned long long* lig_evals, adock_stream_t stream) {
    if (int st = common_checks(AM, KM, P_MAX, ref, inc, ist, DEG_MAX, grid, prm, exact, rxyz, rn, RN_MAX, chain_lig,
                               n_chains))
        return st;
    if (n_chains == 0) return 0;
    if (!chain_id || !cur || !cur_e || !best || !best_e || !counter) return fail(ADOCK_E_ARG, "null pointer");
    if (done && !lig_chains) return fail(ADOCK_E_ARG, "done needs lig_chains");
    if (n_queues < 0 || n_queues > 1024 || (n_queues > 1 && !queue_start))
        return fail(ADOCK_E_ARG, "n_queues must be in 0..1024, with queue_start when above 1");
    if (mc_steps < 0) return fail(ADOCK_E_ARG, "negative mc_steps");
    if (nbr && (!nbr_n || NB_MAX <= 0)) return fail(ADOCK_E_ARG, "nbr needs nbr_n and NB_MAX > 0");
    if (waves_per_block < 0 || waves_per_block > 4) return fail(ADOCK_E_ARG, "waves_per_block must be in 0..4");
    if (lds_layout != 0 && lds_layout != 1) return fail(ADOCK_E_ARG, "lds_layout must be 0 (compact) or 1 (dense)");
    LigArgs L{ref, typ, tm, ax, axp, axu, pr, na, nt, np, rg, P_MAX, gslot, inc, ist, DEG_MAX, AM};
    GridArgs G{grid, origin, lo, hi, rec_idx, goffs, gdims, T, spacing, pen};
    SFArgs S{prm, rad, hyd, don, acc};
    RecArgs R{rxyz, rtyp, rn, RN_MAX};
    McArgs M{chain_id, nbr, nbr_n, NB_MAX, cur, cur_e, best, best_e, init_opt, mc_steps, bfgs_iters, temp, amp_pos,
             amp_rot, seed, step_offset, eval_count};
    M.lig_evals = lig_evals;
    M.n_lig_evals = 0x7fffffff;
    QueueArgs Q{order, queue_start, n_queues, counter, done, lig_chains, times};
    Q.waves = waves_per_block;
    Q.dense = lds_layout;
    return hip_status(launch_mc_persistent(KM, L, G, S, R, exact, chain_lig, n_chains, M, Q, blocks,
                                           (hipStream_t)stream));
}

// ---- allocation-free batched launch ----
// Workspace of adock_mc_batch_noalloc: the work-queue counters [ADOCK_NOALLOC_MAX_QUEUES], then the finished-chain
// counts [ADOCK_NOALLOC_MAX_COMPLEXES]. One copy per GPU for the life of the process (264 KiB). Host code never
// passes the array's name to a kernel: it is a host-side stand-in, not a device address. Each device's address comes
// from hipGetSymbolAddress, taken once per device with that device current, and only that address is used.
static __device__ __align__(256) int scratch[ADOCK_NOALLOC_MAX_QUEUES + ADOCK_NOALLOC_MAX_COMPLEXES];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
int* g_scratch_dev[MAX_DEVICES] = {};

// device address of the current device's scratch array (looked up on the first call on that device)
int scratch_of_current_device(int** out) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_status((int)e);
    if (dev < 0 || dev >= MAX_DEVICES) return fail(ADOCK_E_ARG, "device ordinal out of range for the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            g_err = std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e);
            return e != hipSuccess ? (int)e : ADOCK_E_ARG;
        }
        g_scratch_dev[dev] = static_cast<int*>(p);
    }
    *out = g_scratch_dev[dev];
    return 0;
}
}  // namespace

int adock_mc_batch_noalloc(int AM, int KM, LIG_DECL, GRID_DECL, SF_DECL, REC_DECL, int exact, const int* chain_lig,
                           const int* chain_id, int n_chains, const int* nbr, const int* nbr_n, int NB_MAX, float* cur,
                           float* cur_e, float* best, float* best_e, int init_opt, int mc_steps, int bfgs_iters,
                           float temp, float amp_pos, float amp_rot, unsigned long long seed, int step_offset,
                           const int* order, const int* queue_start, int n_queues, int n_complexes,
                           const int* lig_chains, unsigned long long* times, int blocks, int waves_per_block,
                           unsigned long long* eval_count, unsigned long long* lig_evals, adock_stream_t stream) {
    if (n_queues < 0 || n_queues > ADOCK_NOALLOC_MAX_QUEUES)
        return fail(ADOCK_E_ARG, "n_queues must be in 0..1024 for the allocation-free launch");
    if (n_complexes < 0 || n_complexes > ADOCK_NOALLOC_MAX_COMPLEXES)
        return fail(ADOCK_E_ARG, "n_complexes must be in 0..65536 for the allocation-free launch");
    if (times && (!lig_chains || n_complexes < 1)) return fail(ADOCK_E_ARG, "times needs lig_chains and n_complexes");
    // the checks adock_mc_batch makes, before anything is launched
    if (int st = common_checks(AM, KM, P_MAX, ref, inc, ist, DEG_MAX, grid, prm, exact, rxyz, rn, RN_MAX, chain_lig,
                               n_chains))
        return st;
    if (n_chains == 0) return 0;
    if (!chain_id || !cur || !cur_e || !best || !best_e) return fail(ADOCK_E_ARG, "null pointer");
    if (n_queues > 1 && !queue_start) return fail(ADOCK_E_ARG, "n_queues above 1 needs queue_start");
    if (mc_steps < 0) return fail(ADOCK_E_ARG, "negative mc_steps");
    if (nbr && (!nbr_n || NB_MAX <= 0)) return fail(ADOCK_E_ARG, "nbr needs nbr_n and NB_MAX > 0");
    if (waves_per_block < 0 || waves_per_block > 4) return fail(ADOCK_E_ARG, "waves_per_block must be in 0..4");
    int* base = nullptr;
    if (int st = scratch_of_current_device(&base)) return st;
    int* counter = base;
    int* done = times ? base + ADOCK_NOALLOC_MAX_QUEUES : nullptr;
    const int nc = n_queues > 1 ? n_queues : 1;
    // the array keeps the last call's values: clear this call's counters with a kernel (no memset node in a graph)
    if (int st = hip_status(launch_batch_reset(counter, nc, done, done ? n_complexes : 0, times,
                                               times ? 1 + n_complexes : 0, (hipStream_t)stream)))
        return st;
    LigArgs L{ref, typ, tm, ax, axp, axu, pr, na, nt, np, rg, P_MAX, gslot, inc, ist, DEG_MAX, AM};
    GridArgs G{grid, origin, lo, hi, rec_idx, goffs, gdims, T, spacing, pen};
    SFArgs S{prm, rad, hyd, don, acc};
    RecArgs R{rxyz, rtyp, rn, RN_MAX};
    McArgs M{chain_id, nbr, nbr_n, NB_MAX, cur, cur_e, best, best_e, init_opt, mc_steps, bfgs_iters, temp, amp_pos,
             amp_rot, seed, step_offset, eval_count};
    QueueArgs Q{order, queue_start, n_queues, counter, done, lig_chains, times};
    Q.n_done = n_complexes;   // a chain of a complex past n_complexes is run but not counted (its slot is not ours)
    Q.waves = waves_per_block;
    M.lig_evals = lig_evals;
    M.n_lig_evals = n_complexes;
    return hip_status(launch_mc_persistent(KM, L, G, S, R, exact, chain_lig, n_chains, M, Q, blocks,
                                           (hipStream_t)stream));
}

int adock_batch_reset(int* counter, int n_counters, int* done, int n_complexes, unsigned long long* times,
                      adock_stream_t stream) {
    if (!counter || n_complexes < 0 || n_counters < 1 || n_counters > 1024)
        return fail(ADOCK_E_ARG, "null counter, n_counters outside 1..1024 or negative size");
    return hip_status(launch_batch_reset(counter, n_counters, done, done ? n_complexes : 0, times,
                                         times ? 1 + n_complexes : 0,
                                         (hipStream_t)stream));
}

int adock_mc_step(int AM, int KM, LIG_DECL, GRID_DECL, SF_DECL, REC_DECL, const int* chain_lig, const int* chain_id,
                  int n_chains, float* cur, float* cur_e, float* best, float* best_e, int init, int step, int bfgs_iters,
                  float temp, float amp_pos, float amp_rot, unsigned long long seed, const float* Y, float* feat,
                  float* log_buf, int* log_count, int log_cap, int use_net, float p_plain, float sig_t, float sig_r,
                  float sig_th, int net_mode, float tau, int log_all, unsigned long long* eval_count,
                  adock_stream_t stream) {
    if (int st = common_checks(AM, KM, P_MAX, ref, inc, ist, DEG_MAX, grid, prm, 0, rxyz, rn, RN_MAX, chain_lig,
                               n_chains))
        return st;
    if (AM != 64) return fail(ADOCK_E_BUCKET, "the learned step kernel takes ligands of at most 64 heavy atoms");
    if (n_chains == 0) return 0;
    if (!chain_id || !cur || !cur_e || !best || !best_e) return fail(ADOCK_E_ARG, "null pointer");
    if (log_buf && (!log_count || !feat || log_cap < 0)) return fail(ADOCK_E_ARG, "logging needs feat and a counter");
    if (use_net && (!Y || !feat)) return fail(ADOCK_E_ARG, "network moves need Y and feat");
    LigArgs L{ref, typ, tm, ax, axp, axu, pr, na, nt, np, rg, P_MAX, gslot, inc, ist, DEG_MAX, AM};
    GridArgs G{grid, origin, lo, hi, rec_idx, goffs, gdims, T, spacing, pen};
    SFArgs S{prm, rad, hyd, don, acc};
    RecArgs R{rxyz, rtyp, rn, RN_MAX};
    McArgs M{chain_id, nullptr, nullptr, 0, cur, cur_e, best, best_e, init, 1, bfgs_iters, temp, amp_pos, amp_rot, seed,
             step, eval_count};
    if (net_mode != 0 && net_mode != 1) return fail(ADOCK_E_ARG, "net_mode must be 0 (displacement) or 1 (entity)");
    if (net_mode == 1 && !(tau > 0.f)) return fail(ADOCK_E_ARG, "entity choice needs tau > 0");
    StepArgs P{Y, feat, log_buf, log_count, log_cap, use_net, p_plain, sig_t, sig_r, sig_th, init};
    P.net_mode = net_mode; P.tau = tau; P.log_all = log_all;
    return hip_status(launch_mc_step(KM, L, G, S, R, chain_lig, n_chains, M, P, (hipStream_t)stream));
}

int adock_mc_learned(int AM, int KM, LIG_DECL, GRID_DECL, SF_DECL, REC_DECL, const int* chain_lig, const int* chain_id,
                     int n_chains, float* cur, float* cur_e, float* best, float* best_e, int init_opt, int mc_steps,
                     int step_offset, int bfgs_iters, float temp, float amp_pos, float amp_rot,
                     unsigned long long seed, const unsigned short* W1, const float* b1, const unsigned short* W2,
                     const float* b2, const unsigned short* W3, const float* b3, float* feat, float* log_buf,
                     int* log_count, int log_cap, int use_net, float p_plain, float sig_t, float sig_r, float sig_th,
                     int net_mode, float tau, int log_all, const int* order, int* counter,
                     unsigned long long* eval_count, adock_stream_t stream) {
    if (int st = common_checks(AM, KM, P_MAX, ref, inc, ist, DEG_MAX, grid, prm, 0, rxyz, rn, RN_MAX, chain_lig,
                               n_chains))
        return st;
    if (AM != 64) return fail(ADOCK_E_BUCKET, "the learned search kernel takes ligands of at most 64 heavy atoms");
    if (mc_steps < 0) return fail(ADOCK_E_ARG, "negative mc_steps");
    if (n_chains == 0) return 0;
    if (!chain_id || !cur || !cur_e || !best || !best_e) return fail(ADOCK_E_ARG, "null pointer");
    if (log_buf && (!log_count || log_cap < 0)) return fail(ADOCK_E_ARG, "logging needs a counter");
    if (use_net && (!W1 || !b1 || !W2 || !b2 || !W3 || !b3)) return fail(ADOCK_E_ARG, "network moves need the weights");
    if (!(p_plain >= 0.f && p_plain <= 1.f)) return fail(ADOCK_E_ARG, "p_plain outside [0, 1]");
    LigArgs L{ref, typ, tm, ax, axp, axu, pr, na, nt, np, rg, P_MAX, gslot, inc, ist, DEG_MAX, AM};
    GridArgs G{grid, origin, lo, hi, rec_idx, goffs, gdims, T, spacing, pen};
    SFArgs S{prm, rad, hyd, don, acc};
    RecArgs R{rxyz, rtyp, rn, RN_MAX};
    McArgs M{chain_id, nullptr, nullptr, 0, cur, cur_e, best, best_e, init_opt, mc_steps, bfgs_iters, temp, amp_pos,
             amp_rot, seed, step_offset, eval_count};
    if (net_mode != 0 && net_mode != 1) return fail(ADOCK_E_ARG, "net_mode must be 0 (displacement) or 1 (entity)");
    if (net_mode == 1 && !(tau > 0.f)) return fail(ADOCK_E_ARG, "entity choice needs tau > 0");
    StepArgs P{nullptr, feat, log_buf, log_count, log_cap, use_net, p_plain, sig_t, sig_r, sig_th, 0};
    P.net_mode = net_mode; P.tau = tau; P.log_all = log_all;
    NetArgs N{W1, b1, W2, b2, W3, b3};
    QueueArgs Q{order, nullptr, 1, counter, nullptr, nullptr, nullptr};
    return hip_status(launch_mc_learned(KM, L, G, S, R, chain_lig, n_chains, M, P, N, Q, (hipStream_t)stream));
}

// ---- learned-proposal network ----
int adock_mlp_infer(int B, const float* X, const unsigned short* W1, const float* b1, const unsigned short* W2,
                    const float* b2, const unsigned short* W3, const float* b3, float* Y, adock_stream_t stream) {
    if (B < 0) return fail(ADOCK_E_ARG, "negative batch");
    if (B == 0) return 0;
    if (!X || !W1 || !b1 || !W2 || !b2 || !W3 || !b3 || !Y) return fail(ADOCK_E_ARG, "null pointer");
    return hip_status(launch_mlp_infer(B, X, W1, b1, W2, b2, W3, b3, Y, (hipStream_t)stream));
}
int adock_gemm_f32(int trans_a, int trans_b, int M, int N, int K, float alpha, const float* A, int lda, const float* B,
                   int ldb, float beta, float* C, int ldc, adock_stream_t stream) {
    if (M < 0 || N < 0 || K < 0) return fail(ADOCK_E_ARG, "negative size");
    if ((long long)M * N == 0) return 0;
    if (!A || !B || !C) return fail(ADOCK_E_ARG, "null pointer");
    return hip_status(launch_gemm_f32(trans_a, trans_b, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc,
                                      (hipStream_t)stream));
}
int adock_gemm_f32_splitk(int trans_a, int trans_b, int M, int N, int K, float alpha, const float* A, int lda,
                          const float* B, int ldb, float beta, float* C, int ldc, int splits, float* work,
                          adock_stream_t stream) {
    if (M < 0 || N < 0 || K < 0) return fail(ADOCK_E_ARG, "negative size");
    if ((long long)M * N == 0) return 0;
    if (splits < 1 || splits > 1024) return fail(ADOCK_E_ARG, "splits must be in 1..1024");
    if (!A || !B || !C || !work) return fail(ADOCK_E_ARG, "null pointer");
    if (K == 0) return hip_status(launch_gemm_f32(trans_a, trans_b, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc,
                                                  (hipStream_t)stream));
    return hip_status(launch_gemm_f32_splitk(trans_a, trans_b, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, splits,
                                             work, (hipStream_t)stream));
}
int adock_bias_act(int M, int N, float* Y, const float* b, int relu, adock_stream_t stream) {
    if ((long long)M * N <= 0) return M < 0 || N < 0 ? fail(ADOCK_E_ARG, "negative size") : 0;
    return hip_status(launch_bias_act(M, N, Y, b, relu, (hipStream_t)stream));
}
int adock_relu_bwd(long long n, float* dH, const float* H, adock_stream_t stream) {
    if (n <= 0) return n < 0 ? fail(ADOCK_E_ARG, "negative size") : 0;
    return hip_status(launch_relu_bwd(n, dH, H, (hipStream_t)stream));
}
int adock_colsum(int M, int N, const float* dY, float* db, adock_stream_t stream) {
    if (M < 0 || N <= 0) return N < 0 || M < 0 ? fail(ADOCK_E_ARG, "negative size") : 0;
    return hip_status(launch_colsum(M, N, dY, db, (hipStream_t)stream));
}
int adock_mse(long long n, const float* Y, const float* T, const float* mask, float inv_norm, float* dY, float* loss,
              adock_stream_t stream) {
    if (n <= 0) return n < 0 ? fail(ADOCK_E_ARG, "negative size") : 0;
    return hip_status(launch_mse(n, Y, T, mask, inv_norm, dY, loss, (hipStream_t)stream));
}
int adock_adam(long long n, float* w, const float* g, float* m, float* v, float lr, float beta1, float beta2,
               float eps, int t, adock_stream_t stream) {
    if (n <= 0 || t < 1) return n < 0 || t < 1 ? fail(ADOCK_E_ARG, "negative size or step < 1") : 0;
    return hip_status(launch_adam(n, w, g, m, v, lr, beta1, beta2, eps, t, (hipStream_t)stream));
}
int adock_to_bf16(long long n, const float* x, unsigned short* y, adock_stream_t stream) {
    if (n <= 0) return n < 0 ? fail(ADOCK_E_ARG, "negative size") : 0;
    return hip_status(launch_to_bf16(n, x, y, (hipStream_t)stream));
}

int adock_stage_a(int AM, const float* conf, const int* gslot, const int* na, int L, int C, const float* quats,
                  int Q, const float* trans, const int* ntrans, int NT_MAX, const float* grid, const float* origin,
                  const float* lo, const float* hi, const int* rec_idx, const long long* goffs,
                  const long long* gdims, float spacing, float pen, float cap, float* best_e, int* best_i,
                  adock_stream_t stream) {
    if (AM < 64 || AM > 256 || (AM & 31)) return fail(ADOCK_E_BUCKET, "AM must be a multiple of 32 in 64..256");
    if (L < 0 || C < 0 || Q < 0 || NT_MAX < 0) return fail(ADOCK_E_ARG, "negative size");
    if ((long long)L * C * Q == 0) return 0;
    if ((long long)L * C * Q > 0x7fffffffll) return fail(ADOCK_E_ARG, "L * C * Q too large");
    if (NT_MAX < 1) return fail(ADOCK_E_ARG, "NT_MAX must be at least 1");
    if (!conf || !gslot || !na || !quats || !trans || !ntrans || !grid || !origin || !lo || !hi || !rec_idx ||
        !goffs || !gdims || !best_e || !best_i)
        return fail(ADOCK_E_ARG, "null pointer");
    GridArgs G{grid, origin, lo, hi, rec_idx, goffs, gdims, 0, spacing, pen};
    return hip_status(launch_stage_a(conf, gslot, na, AM, L, C, quats, Q, trans, ntrans, NT_MAX, G, cap, best_e, best_i,
                                     (hipStream_t)stream));
}

int adock_set_chains_per_block(int w) {
    if (set_waves_per_block(w)) return fail(ADOCK_E_ARG, "chains_per_block must be 0 (adaptive) or 1..4");
    return 0;
}
int adock_get_chains_per_block(void) { return get_waves_per_block(); }

int adock_set_line_search_width(int ns) {
    if (set_line_search_width(ns)) return fail(ADOCK_E_ARG, "line_search_width must be 0 (adaptive), 1, 2 or 4");
    return 0;
}
int adock_get_line_search_width(void) { return get_line_search_width(); }
int adock_set_exact_staging(int on) {
    if (set_exact_staging(on)) return fail(ADOCK_E_ARG, "exact_staging must be 0, 1 or 2");
    return 0;
}
int adock_get_exact_staging(void) { return get_exact_staging(); }

int adock_mc_resident_waves(int AM, int KM, int P_MAX, int persistent) {
    if (int st = check_buckets(AM, KM, P_MAX)) return st;
    const int w = mc_resident_waves(KM, P_MAX, persistent ? 1 : 0, AM);
    return w < 0 ? fail(ADOCK_E_ARG, "occupancy query failed") : w;
}

}  // extern "C"
