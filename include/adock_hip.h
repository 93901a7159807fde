/* C ABI of libadock_hip.so: MI355X (gfx950) docking hot path.
 *
 * Every pointer is a device pointer (HIP / PyTorch-ROCm memory) unless stated otherwise; the library allocates
 * nothing, copies nothing, launches on `stream` (NULL = the default stream) and returns without synchronising.
 * Return value: 0 on success, a negative ADOCK_E_* code on invalid arguments, or a positive hipError_t from the
 * launch. adock_last_error() gives a text for the last non-zero status of the calling thread.
 *
 * Argument groups follow the reference's kernel parameter order (src/adock/mckernel.py:328-345, bound by
 * src/adock/cpukernel.py:97-101,263): ligand (13 + 3), grid (10), scoring function (5), receptor (4), exact, then the
 * entry's own parameters. The reference compiles one module per (A_MAX, K_MAX) bucket; here AM and KM are the first
 * two arguments and select a compiled instantiation: AM = 64 (up to 64 heavy atoms, one per lane of a wavefront) or
 * 96 / 128 (two atoms per lane; the reference's rule AM = 64 if the largest ligand has at most 64 heavy atoms, else its
 * count rounded up to 32, src/adock/mckernel.py:553), KM = the largest torsion count rounded up to 8, 8 .. 56 (the
 * state t(3) q(4) theta(KM) fits one wavefront). Larger ligands are refused with ADOCK_E_BUCKET.
 *
 * The ligand group is the reference's packed layout (src/adock/mckernel.py:547-592: ref, typ, tm, ax, axp, axu, pr,
 * na, nt, np, rg, P_MAX, gslot) plus two arrays of this library's own, built once per ligand set on the host
 * (adock_amd.ligand.slot_lists): inc [L, 2 * P128 + 8 * AM + 8] uint32 and ist [L, AM] int32, P128 = P_MAX rounded up
 * to a multiple of 128. Atom a's list, padded to a multiple of 8 entries, holds for every intra pair q of atom a,
 * ascending in q, the word 4 q | (16 p) << 16 (p: the pair's other atom): the byte offsets of the pair's slot and of
 * its partner atom in the kernels' per-wave arrays; the row ends with 8 padding entries, 4 P128 | (16 D) << 16, D the
 * dummy atom (64 for AM 64, 128 above). ist[l, lane] = s | c << 16 | h << 24 names the lane's segment inc[l, s .. s + c)
 * (s and c multiples of 8, c < 256): normally its atom's whole list, with h the lane itself. For AM 64 the lanes past
 * the ligand's atoms may each take the tail of one list: atom a keeps the head (h = the helper lane) and helper lane
 * b >= na walks the tail that starts where a's segment ends (h = a); the kernels add the helper's partial sum to the
 * atom's. DEG_MAX is the largest whole (padded) list over all atoms and ligands (at most AM). The kernels use the lists
 * to sum the pair gradient of every atom without atomics; DEG_MAX sizes the per-atom gradient rows of the
 * single-complex kernel. P_MAX is at most 16,256 (16-bit slot offsets; a 128-atom ligand has at most 8,128 pairs),
 * and at most 4,096 with AM = 64 (64 atoms have at most 2,016 pairs); ADOCK_E_LDS above.
 */
#ifndef ADOCK_HIP_H
#define ADOCK_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define ADOCK_E_ARG (-1)     /* null pointer or negative size */
#define ADOCK_E_BUCKET (-2)  /* unsupported AM / KM bucket */
#define ADOCK_E_LDS (-3)     /* per-wave pair list does not fit the LDS budget */
#define ADOCK_E_TYPES (-4)   /* too many ligand types */

typedef void* adock_stream_t; /* hipStream_t */

/* Library version string and last error text of the calling thread. */
const char* adock_version(void);
const char* adock_last_error(void);

/* Receptor scoring grids. Replaces the reference's build_grid kernel launch (src/adock/grid.py:132-166) and its host
 * twin launch_build_grid_tiled (src/adock/cpukernel.py:104-108); same arguments minus the host tile size, plus the
 * stream. prm: 15 floats in grid order (cutoff, g1_off, g1_w, g2_off, g2_w, rep_off, hyd_good, hyd_bad, hb_good,
 * hb_bad, w_g1, w_g2, w_rep, w_hyd, w_hb). out: [n_lt, nx, ny, nz] (per_term = 0) or [5, n_lt, nx, ny, nz]. */
int adock_build_grid(const float* rxyz, const int* rtype, int n_rec, const float* radius, const int* hyd,
                     const int* don, const int* acc, const int* lig_types, int n_lt, float ox, float oy, float oz,
                     float spacing, int nx, int ny, int nz, const float* prm, int per_term, float* out,
                     adock_stream_t stream);

/* Scoring grids of the fitted (linear-in-weights) function. Replaces the reference's build_grid_fit kernel launch
 * (src/adock/fit/gridfit.py:13-58, launched by build_grid_fit, gridfit.py:70-93); same arguments in the same order,
 * plus n_rt (rows of W) and the stream. rext [n_rec] receptor ext types in [0, n_rt) (atoms outside are skipped),
 * rad_rec [n_rt], rad_lig [n_lt] (compact ligand types), W [n_rt, n_lt, 6] pair weights of the basis functions
 * (gauss1, gauss2, repulsion, hydrophobic, hbond, ionic), prm: 12 floats (cutoff, g1_off, g1_w, g2_off, g2_w,
 * rep_off, hyd_good, hyd_bad, hb_good, hb_bad, ion_rmin, rep_cap). out [n_lt, nx, ny, nz], z fastest. n_rt <= 216. */
int adock_build_grid_fit(const float* rxyz, const int* rext, int n_rec, const float* rad_rec, const float* rad_lig,
                         int n_lt, const float* W, int n_rt, float ox, float oy, float oz, float spacing, int nx,
                         int ny, int nz, const float* prm, float* out, adock_stream_t stream);

/* adock_build_grid_fit for n_sets receptor sets in one launch per 8 ligand types, as make_gridset_fit needs them (the
 * reference builds them one by one, src/adock/fit/gridfit.py:96-108). rxyz [n_sets, n_rec, 3], rext [n_sets, n_rec]
 * (pad shorter sets with ext type -1: such atoms are skipped, so each set's grids equal adock_build_grid_fit's of its
 * own atoms bit for bit), origins [n_sets, 3] (device), out [n_sets, n_lt, nx, ny, nz]; the rest as
 * adock_build_grid_fit. n_sets <= 65,535. */
int adock_build_grid_fit_sets(const float* rxyz, const int* rext, int n_sets, int n_rec, const float* rad_rec,
                              const float* rad_lig, int n_lt, const float* W, int n_rt, const float* origins,
                              float spacing, int nx, int ny, int nz, const float* prm, float* out,
                              adock_stream_t stream);

/* Exact pairwise fitted inter energy of docked poses (no grid). Replaces the torch loop Problem.inter_exact
 * (src/adock/search.py:73-89) over FittedPotential.energy (src/adock/fit/terms.py:339-344) that finalize runs for the
 * fitted function (search.py:483-489). xyz [n_cx, n_pose, n_atom, 3] pose coordinates, lig_type [n_cx, n_atom]
 * compact ligand types (atoms outside [0, n_lt), e.g. -1 hydrogens and padding, are skipped), rec_idx [n_cx] receptor
 * set of each complex in [0, n_sets) (else that complex's energies are NaN), rxyz [n_sets, n_rec, 3], rext
 * [n_sets, n_rec] receptor ext types (outside [0, n_rt): padding, skipped), rad_rec [n_rt], rad_lig [n_lt],
 * W [n_rt, n_lt, 6] and prm (12 floats) as adock_build_grid_fit. out [n_cx, n_pose]. Deterministic. */
int adock_inter_exact_fit(const float* xyz, const int* lig_type, int n_cx, int n_pose, int n_atom,
                          const int* rec_idx, const float* rxyz, const int* rext, int n_sets, int n_rec,
                          const float* rad_rec, const float* rad_lig, int n_lt, const float* W, int n_rt,
                          const float* prm, float* out, adock_stream_t stream);

/* Energy + analytic gradient for n_chains poses. Replaces launch_eg_kernel (eg_kernel, src/adock/mckernel.py:411-447;
 * host binding src/adock/cpukernel.py:266-273). state [n, 7+KM], e_out [n], g_out [n, 6+KM] (entries 6+nt.. untouched).
 * prm: 15 floats in search order (g1_off, g1_w, g2_off, g2_w, rep_off, hyd_good, hyd_bad, hb_good, hb_bad, cutoff,
 * w_g1, ..., w_hb). */
int adock_eg(int AM, int KM,
             const float* ref, const int* typ, const unsigned char* tm, const int* ax, const float* axp,
             const float* axu, const int* pr, const int* na, const int* nt, const int* np, const float* rg, int P_MAX,
             const int* gslot, const unsigned* inc, const int* ist, int DEG_MAX,
             const float* grid, const float* origin, const float* lo, const float* hi, const int* rec_idx,
             const long long* goffs, const long long* gdims, int T, float spacing, float pen,
             const float* prm, const float* rad, const int* hyd, const int* don, const int* acc,
             const float* rxyz, const int* rtyp, const int* rn, int RN_MAX,
             int exact,
             const int* chain_lig, int n_chains, const float* state, float* e_out, float* g_out,
             adock_stream_t stream);

/* Monte Carlo chains with BFGS, in place on cur/cur_e/best/best_e. Replaces launch_mc_kernel (mc_kernel,
 * src/adock/mckernel.py:328-408; host binding src/adock/cpukernel.py:289-306). nbr/nbr_n may be NULL (NB_MAX 0). */
int adock_mc(int AM, int KM,
             const float* ref, const int* typ, const unsigned char* tm, const int* ax, const float* axp,
             const float* axu, const int* pr, const int* na, const int* nt, const int* np, const float* rg, int P_MAX,
             const int* gslot, const unsigned* inc, const int* ist, int DEG_MAX,
             const float* grid, const float* origin, const float* lo, const float* hi, const int* rec_idx,
             const long long* goffs, const long long* gdims, int T, float spacing, float pen,
             const float* prm, const float* rad, const int* hyd, const int* don, const int* acc,
             const float* rxyz, const int* rtyp, const int* rn, int RN_MAX,
             int exact,
             const int* chain_lig, const int* chain_id, int n_chains, const int* nbr, const int* nbr_n, int NB_MAX,
             float* cur, float* cur_e, float* best, float* best_e, int init_opt, int mc_steps, int bfgs_iters,
             float temp, float amp_pos, float amp_rot, unsigned long long seed, int step_offset,
             adock_stream_t stream);

/* adock_mc plus a device counter: *eval_count (device unsigned long long, may be NULL) is incremented by the number of
 * energy+gradient evaluations the launch performs. No reference counterpart: the reference does not count evaluations;
 * this is the instrumented variant the benchmark uses to report score+grad evals/s. */
int adock_mc_counted(int AM, int KM,
                     const float* ref, const int* typ, const unsigned char* tm, const int* ax, const float* axp,
                     const float* axu, const int* pr, const int* na, const int* nt, const int* np, const float* rg,
                     int P_MAX, const int* gslot, const unsigned* inc, const int* ist, int DEG_MAX,
                     const float* grid, const float* origin, const float* lo, const float* hi, const int* rec_idx,
                     const long long* goffs, const long long* gdims, int T, float spacing, float pen,
                     const float* prm, const float* rad, const int* hyd, const int* don, const int* acc,
                     const float* rxyz, const int* rtyp, const int* rn, int RN_MAX,
                     int exact,
                     const int* chain_lig, const int* chain_id, int n_chains, const int* nbr, const int* nbr_n,
                     int NB_MAX, float* cur, float* cur_e, float* best, float* best_e, int init_opt, int mc_steps,
                     int bfgs_iters, float temp, float amp_pos, float amp_rot, unsigned long long seed,
                     int step_offset, unsigned long long* eval_count, adock_stream_t stream);

/* Batched screening in one persistent launch: same chain arguments as adock_mc, plus work queues. The grid is sized to
 * the chip's resident capacity (blocks <= 0: CUs x occupancy). `order` (NULL = 0..n-1) holds n_queues consecutive
 * queues of chain indices, queue q = order[queue_start[q] .. queue_start[q+1]) (queue_start [n_queues + 1], from 0 to
 * n; n_queues 0 or 1: one queue over all n), each with its counter in counter[q] (counter [max(n_queues, 1)]).
 * Workgroup b pulls chains from queue b % n_queues until it is empty, then from the others in turn; the dispatcher
 * deals workgroups to the 8 XCDs in turn, so with 8 queues each queue runs on one XCD (put the complexes of a few
 * receptors in each, longest chains first, so their grids share that XCD's L2). Results do not depend on the queues
 * (every chain's computation is its own). Optional per-complex
 * bookkeeping: done [L] counts finished chains, lig_chains [L] gives chains per complex, times [1+L] receives the
 * 100 MHz real-time stamp of the first wave start (times[0]) and of each complex's last chain (times[1+l]).
 * counter, done and times must be reset before every launch (adock_batch_reset: a kernel, so the pair can be captured
 * in a HIP graph). eval_count may be NULL. lig_evals [L] (device, may be NULL) receives per complex the number of
 * energy+gradient evaluations its chains performed (added to what it holds). waves_per_block (0..4) sets the chains
 * (wavefronts) per workgroup of this launch only; 0 takes the process default of adock_set_chains_per_block.
 * lds_layout picks the wave's pair-gradient LDS tile: 0 compact (one slot per intra pair, per-atom lists read from
 * memory), 1 dense (per-atom rows of DEG_MAX entries; used in grid mode with AM 64, otherwise compact); results are
 * the same bit for bit, only speed and occupancy differ. No reference counterpart: the reference launches one block
 * per NW chains and compacts between segments on the host (src/adock/search.py:614-636). */
int adock_mc_batch(int AM, int KM,
                   const float* ref, const int* typ, const unsigned char* tm, const int* ax, const float* axp,
                   const float* axu, const int* pr, const int* na, const int* nt, const int* np, const float* rg,
                   int P_MAX, const int* gslot, const unsigned* inc, const int* ist, int DEG_MAX,
                   const float* grid, const float* origin, const float* lo, const float* hi, const int* rec_idx,
                   const long long* goffs, const long long* gdims, int T, float spacing, float pen,
                   const float* prm, const float* rad, const int* hyd, const int* don, const int* acc,
                   const float* rxyz, const int* rtyp, const int* rn, int RN_MAX,
                   int exact,
                   const int* chain_lig, const int* chain_id, int n_chains, const int* nbr, const int* nbr_n,
                   int NB_MAX, float* cur, float* cur_e, float* best, float* best_e, int init_opt, int mc_steps,
                   int bfgs_iters, float temp, float amp_pos, float amp_rot, unsigned long long seed,
                   int step_offset, const int* order, const int* queue_start, int n_queues, int* counter,
                   int* done, const int* lig_chains, unsigned long long* times, int blocks, int waves_per_block,
                   int lds_layout, unsigned long long* eval_count, unsigned long long* lig_evals,
                   adock_stream_t stream);

/* Resets the work queues of adock_mc_batch with a kernel launch: counter[0..n_counters) = 0, done[0..L) = 0,
 * times[0] = ~0, times[1..L] = 0 (done / times may be NULL). n_counters in 1..1024. */
int adock_batch_reset(int* counter, int n_counters, int* done, int n_complexes, unsigned long long* times,
                      adock_stream_t stream);

/* adock_mc_batch for callers that cannot allocate (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc). No counterpart in the reference. Same arguments as adock_mc_batch minus the workspace (counter, done)
 * and plus n_complexes (the length of lig_chains / times): the queue counters and the per-complex finished-chain counts
 * live in a static device array that the library owns, one per GPU (264 KiB, held until the process ends), cleared by
 * a kernel at the start of every call. The call allocates nothing and frees nothing. Limits: n_queues <= 1024,
 * n_complexes <= 65536 (ADOCK_E_ARG beyond; lig_evals then has n_complexes entries). Calls that overlap on one GPU (two streams, or two host threads) share
 * the array and race: serialise them. Results are those of adock_mc_batch bit for bit. */
#define ADOCK_NOALLOC_MAX_QUEUES 1024
#define ADOCK_NOALLOC_MAX_COMPLEXES 65536
int adock_mc_batch_noalloc(int AM, int KM,
                           const float* ref, const int* typ, const unsigned char* tm, const int* ax, const float* axp,
                           const float* axu, const int* pr, const int* na, const int* nt, const int* np,
                           const float* rg, int P_MAX, const int* gslot, const unsigned* inc, const int* ist,
                           int DEG_MAX,
                           const float* grid, const float* origin, const float* lo, const float* hi,
                           const int* rec_idx, const long long* goffs, const long long* gdims, int T, float spacing,
                           float pen,
                           const float* prm, const float* rad, const int* hyd, const int* don, const int* acc,
                           const float* rxyz, const int* rtyp, const int* rn, int RN_MAX,
                           int exact,
                           const int* chain_lig, const int* chain_id, int n_chains, const int* nbr, const int* nbr_n,
                           int NB_MAX, float* cur, float* cur_e, float* best, float* best_e, int init_opt,
                           int mc_steps, int bfgs_iters, float temp, float amp_pos, float amp_rot,
                           unsigned long long seed, int step_offset, const int* order, const int* queue_start,
                           int n_queues, int n_complexes, const int* lig_chains, unsigned long long* times,
                           int blocks, int waves_per_block, unsigned long long* eval_count,
                           unsigned long long* lig_evals, adock_stream_t stream);

/* ---- learned MC proposal network (128 -> 256 -> 256 -> 23, ReLU). No reference counterpart: the reference proposes
 * moves at random only (src/adock/mckernel.py:381-398). ---- */

/* One Monte Carlo step of every chain, for step-by-step search (one launch per step; capture a sequence in a HIP graph).
 * The move comes from the network output Y [n,23] (use_net = 1; with probability p_plain, or when Y is not finite, the
 * reference's random mutation is used instead) or is the reference's mutation (use_net = 0); then BFGS and Metropolis
 * with the reference's RNG slots for step `step`. feat [n,128] (may be NULL) holds the features of each chain's current
 * pose: read for logging, rewritten after the step. log_buf [log_cap, 177] (may be NULL) receives accepted moves,
 * or every move with log_all = 1 (features before the move, 23 targets, 23 masks, energy gain, entity moved: 0
 * translation, 1 rotation, 2 + k torsion k, -1 a network displacement, and the chain's best energy minus its current
 * one before the move) at *log_count (atomic). net_mode 0: the
 * network output is a displacement (translation + y_t + sig_t N, rotation normalise(y_q) q then sig_r N, torsions
 * + y_theta + sig_th N); net_mode 1: outputs 0 (translation), 3 (rotation) and 7 + k (torsion k < 16) score the
 * entities, one is drawn with probability ~ exp(score / tau) and the reference's mutation of it is applied. init = 1 runs the start
 * BFGS only (the reference's init_opt). With use_net = 0, init followed by steps 0..S-1 equals adock_mc(init_opt = 1,
 * mc_steps = S) bit for bit. */
int adock_mc_step(int AM, int KM,
                  const float* ref, const int* typ, const unsigned char* tm, const int* ax, const float* axp,
                  const float* axu, const int* pr, const int* na, const int* nt, const int* np, const float* rg,
                  int P_MAX, const int* gslot, const unsigned* inc, const int* ist, int DEG_MAX,
                  const float* grid, const float* origin, const float* lo, const float* hi, const int* rec_idx,
                  const long long* goffs, const long long* gdims, int T, float spacing, float pen,
                  const float* prm, const float* rad, const int* hyd, const int* don, const int* acc,
                  const float* rxyz, const int* rtyp, const int* rn, int RN_MAX,
                  const int* chain_lig, const int* chain_id, int n_chains, float* cur, float* cur_e, float* best,
                  float* best_e, int init, int step, int bfgs_iters, float temp, float amp_pos, float amp_rot,
                  unsigned long long seed, const float* Y, float* feat, float* log_buf, int* log_count, int log_cap,
                  int use_net, float p_plain, float sig_t, float sig_r, float sig_th, int net_mode, float tau,
                  int log_all, unsigned long long* eval_count, adock_stream_t stream);

/* The learned search in one launch: every chain's start BFGS (init_opt = 1) and mc_steps MC steps (steps step_offset
 * .. step_offset + mc_steps - 1), the chain set up once, with the proposal network run by each chain's wavefront on
 * MFMA inside its step loop (weights as for adock_mlp_infer, read in place; may be NULL when use_net = 0). Moves,
 * BFGS, Metropolis test, move log and feature definition are adock_mc_step's, so one call equals the step-by-step
 * search (init step, then adock_mlp_infer + adock_mc_step per MC step) bit for bit in cur / cur_e / best / best_e and
 * in the logged records (in another order); eval_count counts evaluations of this kernel, which takes a new pose's
 * features from the evaluation its BFGS ended on instead of evaluating it again. feat [n,128] (may be NULL) receives the features of each chain's final
 * current pose. use_net = 0 without a log is adock_mc's plain search (init_opt, mc_steps). counter NULL: one wavefront
 * per chain; else (one int, 0 at the call) persistent waves take the chains in the order `order` [n_chains] (a
 * permutation of 0 .. n_chains - 1, e.g. longest chains first; NULL: identity) from that counter, which ends at
 * n_chains plus the waves launched: reset it before every call. Results do not depend on the order. net_mode, tau
 * and log_all as for adock_mc_step. */
int adock_mc_learned(int AM, int KM,
                     const float* ref, const int* typ, const unsigned char* tm, const int* ax, const float* axp,
                     const float* axu, const int* pr, const int* na, const int* nt, const int* np, const float* rg,
                     int P_MAX, const int* gslot, const unsigned* inc, const int* ist, int DEG_MAX,
                     const float* grid, const float* origin, const float* lo, const float* hi, const int* rec_idx,
                     const long long* goffs, const long long* gdims, int T, float spacing, float pen,
                     const float* prm, const float* rad, const int* hyd, const int* don, const int* acc,
                     const float* rxyz, const int* rtyp, const int* rn, int RN_MAX,
                     const int* chain_lig, const int* chain_id, int n_chains, float* cur, float* cur_e, float* best,
                     float* best_e, int init_opt, int mc_steps, int step_offset, int bfgs_iters, float temp,
                     float amp_pos, float amp_rot, unsigned long long seed, const unsigned short* W1, const float* b1,
                     const unsigned short* W2, const float* b2, const unsigned short* W3, const float* b3,
                     float* feat, float* log_buf, int* log_count, int log_cap, int use_net, float p_plain,
                     float sig_t, float sig_r, float sig_th, int net_mode, float tau, int log_all, const int* order,
                     int* counter, unsigned long long* eval_count, adock_stream_t stream);

/* Inference on MFMA: Y[B,23] = W3 relu(W2 relu(W1 X + b1) + b2) + b3 with BF16 weights [out, in] (W1 [256,128],
 * W2 [256,256], W3 [23,256]), BF16 activations and FP32 accumulation; X [B,128] and biases FP32. */
int adock_mlp_infer(int B, const float* X, const unsigned short* W1, const float* b1, const unsigned short* W2,
                    const float* b2, const unsigned short* W3, const float* b3, float* Y, adock_stream_t stream);

/* Training primitives (FP32): C = alpha op(A) op(B) + beta C (row-major, op = transpose when trans_x != 0);
 * Y += b row-broadcast (+ ReLU); dH *= (H > 0); db = column sums of dY [M,N]; masked MSE with its gradient
 * (dY = 2 mask (Y - T) inv_norm, *loss += sum mask (Y - T)^2 inv_norm); bias-corrected Adam step t; FP32 -> BF16. */
int adock_gemm_f32(int trans_a, int trans_b, int M, int N, int K, float alpha, const float* A, int lda, const float* B,
                   int ldb, float beta, float* C, int ldc, adock_stream_t stream);
/* adock_gemm_f32 with K split over up to `splits` workgroup layers (1..1024), for products with few output tiles and a
 * long K (the weight gradients: K = the batch): work holds splits x M x N floats of partial tiles, which a second
 * kernel adds in layer order (deterministic) before alpha and beta. */
int adock_gemm_f32_splitk(int trans_a, int trans_b, int M, int N, int K, float alpha, const float* A, int lda,
                          const float* B, int ldb, float beta, float* C, int ldc, int splits, float* work,
                          adock_stream_t stream);
int adock_bias_act(int M, int N, float* Y, const float* b, int relu, adock_stream_t stream);
int adock_relu_bwd(long long n, float* dH, const float* H, adock_stream_t stream);
int adock_colsum(int M, int N, const float* dY, float* db, adock_stream_t stream);
int adock_mse(long long n, const float* Y, const float* T, const float* mask, float inv_norm, float* dY, float* loss,
              adock_stream_t stream);
int adock_adam(long long n, float* w, const float* g, float* m, float* v, float lr, float beta1, float beta2,
               float eps, int t, adock_stream_t stream);
int adock_to_bf16(long long n, const float* x, unsigned short* y, adock_stream_t stream);

/* Stage A rigid enumeration. Replaces stagea_kernel (src/adock/mckernel.py:451-521; Python seam KernelProblem.stage_a,
 * mckernel.py:690-705; host twin cpukernel.py:276-287). conf [L, C, AM, 3] heavy-atom conformer coordinates in the
 * ligand frame, gslot / na from the packed ligand, quats [Q, 4], trans [L, NT_MAX, 3] translation lattices (padded),
 * ntrans [L]; grid arguments as in the chain kernels. Out, per block b = (l * C + c) * Q + q: best_e[b] = min over
 * the lattice of sum_a min(nearest-grid energy + out-of-box penalty, cap), best_i[b] = its lattice index (ties: the
 * lowest). */
int adock_stage_a(int AM, const float* conf, const int* gslot, const int* na, int L, int C, const float* quats, int Q,
                  const float* trans, const int* ntrans, int NT_MAX,
                  const float* grid, const float* origin, const float* lo, const float* hi, const int* rec_idx,
                  const long long* goffs, const long long* gdims, float spacing, float pen, float cap,
                  float* best_e, int* best_i, adock_stream_t stream);

/* Launch shape of the chain kernels: chains (waves) per workgroup, 1..4, or 0 = adaptive (default: one chain per
 * workgroup while chains <= 8 x CUs, 2 up to 16 x CUs, else 4; the persistent batch kernel uses 4). A process-wide
 * default (atomic: a launch on another thread sees the old or the new value); the batched launches take their own per
 * call (waves_per_block). The reference's counterpart is the NW constant (src/adock/mckernel.py:525-540). Returns 0,
 * or ADOCK_E_ARG. */
int adock_set_chains_per_block(int chains_per_block);
int adock_get_chains_per_block(void);

/* Waves per chain of adock_mc / adock_mc_counted's speculative line search: 1 (off), 2 or 4, or 0 = adaptive (4 while
 * chains <= CUs, 2 up to 2 x CUs, else 1; lowered until the workgroup's LDS fits). The waves of one workgroup evaluate
 * consecutive line-search trials of one chain in parallel; results, energies and evaluation counts are the same bit
 * for bit as with 1. No counterpart in the reference (its line search is sequential). Returns 0, or ADOCK_E_ARG. */
int adock_set_line_search_width(int waves);
int adock_get_line_search_width(void);

/* Exact (pairwise) mode of adock_mc: 1 stages each chain's receptor atoms (its neighbour list, or the whole receptor)
 * in LDS once per launch when they number at most 1,024 (longer lists are read from memory), so the per-atom loop
 * reads LDS broadcasts instead of dependent global loads; 2 (default) also keeps, per atom, bit masks of the staged
 * atoms within cutoff + 1 A of where the atom was when they were built (rebuilt once any atom has moved 1 A), and each
 * atom visits only those; 0 always reads the atoms from memory. Results are the same bit for bit in all three modes.
 * No counterpart in the reference. Returns 0, or ADOCK_E_ARG. */
int adock_set_exact_staging(int on);
int adock_get_exact_staging(void);

/* Resident wavefronts per CU of the grid-mode chain kernels for buckets (AM, KM) and pair capacity P_MAX, from the
 * runtime's occupancy calculator: persistent = 0 the per-chain kernel of adock_mc (as launched for many chains), 1 the
 * persistent kernel of adock_mc_batch. Lets a caller pick the launch that keeps more chains resident. No counterpart
 * in the reference. Returns the count (> 0), or a negative ADOCK_E_* code. */
int adock_mc_resident_waves(int AM, int KM, int P_MAX, int persistent);

#ifdef __cplusplus
}
#endif
#endif /* ADOCK_HIP_H */
