"""Drivers for the batched and learned-proposal workloads (used by bench.py's opt-in modes).

screen():  one receptor grid set, many uneven ligands, all chains in one persistent HIP-graph-replayed launch.
multi():   many receptors (grid sets resident, admitted in groups under an HBM budget), per-complex latency from the
           kernel's completion stamps.
grids():   both grid builders (Vina XS and the fitted function) on the single-complex receptor, with the in-cutoff
           work counted exactly, for the builders' FP32 VALU roofline.
driver():  the full search driver (Stage A, MC segments, refine, exact polish, finalize) with a search preset on a
           one-pocket screen, grid build included, per docked ligand (Vina XS, or the fitted function's path).
wide():    the single complex widened to 4,096 chains at constant chain-steps (the reference's min_total_chains rule).
mlp():     the proposal network's kernels alone: MFMA inference on one step's batch, one FP32 training step.
eg():      the energy + gradient kernel alone on many independent poses of the single complex (throughput bound).
learned_cost(): a learned MC step's cost against a plain one (one-launch learned search and step-by-step pipeline).
learned(): trains the proposal network on accepted moves logged by the plain method on training complexes, then
           compares plain and learned search on disjoint evaluation complexes at equal MC-step budgets.
"""
import time

import numpy as np
import torch

from . import synth, workload
from .autotune import Autotuner
from .batch import XCDS, BatchRunner, GraphedBatch
from .kernel import KernelProblem
from .learned import StepSearch, entity_dataset, jump_dataset, move_weights, train_proposal
from .mlp import ProposalNet


def _bufs(st):
    N = st.shape[0]
    return st.clone(), torch.zeros(N, device=st.device), st.clone(), torch.full((N,), float("inf"), device=st.device)


def screen(n_ligands=512, chains=64, steps=500, seed=0, reps=2, tune=True, log=print, queues=XCDS, gap_sample=64):
    w, prob = workload.screen(seed=seed, n_ligands=n_ligands, chains=chains, steps=steps)
    kp = KernelProblem(prob)
    L = len(w["ligs"])
    st, cl = workload.chain_starts(seed, w["ligs"], [w["center"]] * L, [w["size"]] * L, chains, kp.KM)
    st = torch.as_tensor(st, device=kp.device)
    cl = torch.as_tensor(cl, device=kp.device)
    runner = BatchRunner(kp, cl, n_complexes=L, queues=queues)
    shape = Autotuner().tune(runner, st, probe_steps=2, reps=1) if tune else None
    bufs = _bufs(st)
    ci = torch.arange(len(cl), dtype=torch.int32, device=kp.device)
    g = GraphedBatch(runner, chain_id=ci, cur=bufs[0], cur_e=bufs[1], best=bufs[2], best_e=bufs[3], init_opt=True,
                     mc_steps=steps, bfgs_iters=0, temp=1.2, amp_pos=2.0, amp_rot=2.0, seed=seed * 1_000_003 + 12345,
                     step_offset=0)
    times = []
    for _ in range(reps):
        bufs[0].copy_(st); bufs[1].zero_(); bufs[2].copy_(st); bufs[3].fill_(float("inf"))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        g.replay()
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
    dt = min(times)
    best = bufs[3].view(L, chains).min(1).values.cpu().numpy()
    lat = runner.latencies()
    out = {"workload": f"screen: {L} ligands (10-60 atoms, 0-15 torsions) x {chains} chains x {steps} MC steps, "
                       "one grid set, one persistent graph-replayed launch",
           "docked_ligands_per_s": L / dt, "seconds": dt, "autotuned_shape": shape,
           "best_score_mean": float(best.mean()),
           "latency_p50_s": float(np.percentile(lat, 50)), "latency_p99_s": float(np.percentile(lat, 99))}
    if gap_sample:
        out["best_score_gap"] = _screen_gap(kp, w, best, chains, steps, seed, min(gap_sample, L))
    return out


def _screen_gap(kp, w, best, chains, steps, seed, n, widen=4, lengthen=2):
    """Best-score gap of the screening budget against full exhaustiveness on its first n ligands: rerun with widen x
    the chains (new start states) for lengthen x the MC steps; per ligand gap = best at the budget - min(that, best of
    the exhaustive run); mean and quantiles over the sample."""
    preps = w["ligs"][:n]
    st, c_loc = workload.chain_starts(seed + 7, preps, [w["center"]] * n, [w["size"]] * n, chains * widen, kp.KM)
    cl = torch.as_tensor(c_loc.astype(np.int32), device=kp.device)
    b = _bufs(torch.as_tensor(st, device=kp.device))
    t0 = time.perf_counter()
    BatchRunner(kp, cl, n_complexes=len(w["ligs"])).launch(torch.arange(len(cl), dtype=torch.int32, device=kp.device),
                                                          *b, True, steps * lengthen, 0, 1.2, 2.0, 2.0,
                                                          seed * 1_000_003 + 12346, 0)
    torch.cuda.synchronize()
    ex = b[3].view(n, chains * widen).min(1).values.cpu().numpy()
    gap = best[:n] - np.minimum(best[:n], ex)
    return {"mean": float(gap.mean()), "p50": float(np.percentile(gap, 50)), "p90": float(np.percentile(gap, 90)),
            "share_within_0.1": float((gap <= 0.1).mean()), "ligands": n,
            "exhaustive": f"{chains * widen} chains x {steps * lengthen} MC steps per ligand ({widen * lengthen}x the "
                          f"screen's chain-steps), {time.perf_counter() - t0:.1f} s",
            "definition": "per ligand: min(best score at the screening budget) - min(that, best score of the exhaustive "
                          "run), kcal/mol; mean over the sample"}


def multi(n_receptors=64, ligands_per=32, chains=32, steps=500, seed=0, hbm_budget=None, log=print, queues=XCDS):
    """The multi-receptor batch under HBM admission: receptors grouped so that one group's grid sets fit hbm_budget
    bytes (default: a quarter of all grid sets, so several groups), each group's grids built, its complexes docked in
    one persistent launch, and its grids released before the next group. Timed from the first build to the last
    chain; best scores are those of the all-resident batch bit for bit."""
    from .admission import AdmittedBatch, grid_bytes
    w = synth.make_multi(seed, n_receptors, ligands_per, chains, steps)
    recs = [workload.receptor_dict(r, np.zeros(3), b) for r, b in zip(w["receptors"], w["boxes"])]
    total = int(sum(grid_bytes(recs, w["ligs"], w["rec_idx"])))
    budget = float(hbm_budget) if hbm_budget else total / 4.0
    ab = AdmittedBatch(recs, w["ligs"], w["rec_idx"], budget, chains, queues=queues)
    L = len(w["ligs"])
    boxes = [w["boxes"][r] for r in w["rec_idx"]]
    st, _ = workload.chain_starts(seed, w["ligs"], [np.zeros(3)] * L, boxes, chains, ab.KM)
    bufs = _bufs(torch.as_tensor(st, device=ab.device))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res = ab.run(*bufs, True, steps, 0, 1.2, 2.0, 2.0, seed * 1_000_003 + 12345, 0, log=log)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    best = bufs[3].view(L, chains).min(1).values.cpu().numpy()
    lat = res["latency_s"]
    return {"workload": f"multi, HBM-admitted: {n_receptors} receptors (2000-6000 atoms, boxes 16-30 A) x "
                        f"{ligands_per} ligands x {chains} chains x {steps} MC steps; receptor groups whose grid sets "
                        f"fit the budget, built and docked one group at a time",
            "grid_set_bytes_total": total, "hbm_budget_bytes": int(budget), "admission_groups": len(ab.groups),
            "peak_grid_bytes": int(ab.peak_grid_bytes), "groups": res["groups"],
            "docked_ligands_per_s": L / dt, "seconds": dt, "latency_p50_s": float(np.percentile(lat, 50)),
            "latency_p99_s": float(np.percentile(lat, 99)), "best_score_mean": float(best.mean())}


def complexes(seeds, name="single"):
    """Single-workload-sized complexes (4,000-atom receptor, 35-atom ligand), one receptor + one ligand per seed,
    batched into one Problem."""
    recs, ligs = [], []
    for s in seeds:
        ws = synth.make_single(name, s)
        recs.append(workload.receptor_dict(ws["receptor"], ws["center"], ws["size"]))
        ligs.append(ws["lig"])
    prob = workload.gpu_problem(recs, ligs, np.arange(len(seeds)), with_rec=False)
    return recs, ligs, prob


# algorithmic FP32 operations of the grid builders (transcendentals counted as one): per in-cutoff (point, atom) pair
# and per (pair, ligand type) triple
GRID_OPS = {"vina": (10, 26), "fitted": (15, 38)}


def _in_cutoff_pairs(xyz, origin, dims, spacing, cutoff, dev):
    """Exact count of (grid point, receptor atom) pairs closer than the cutoff."""
    R = torch.as_tensor(np.asarray(xyz, np.float32), device=dev)
    ax = [torch.arange(n, device=dev, dtype=torch.float32) * spacing + float(o) for n, o in zip(dims, origin)]
    total = 0
    for i in range(dims[0]):
        P = torch.stack(torch.meshgrid(ax[0][i:i + 1], ax[1], ax[2], indexing="ij"), -1).reshape(-1, 3)
        total += int((torch.cdist(P, R) < cutoff).sum())
    return total


def grids(seed=0, reps=10):
    """Both grid builders on the single-complex receptor (71^3 points). kernel_ms: the C entry point on device-resident
    inputs, `reps` calls back to back between HIP events on its stream (one launch per 8 ligand types); call_ms: the
    Python facade, which also uploads the receptor and tables on every call."""
    from . import _lib
    from .config import grid_prm, sf_params
    from .fit.fdock import DATA, compact_ligands, load_fitted
    from .fit.gridfit import build_grid_fit
    from .fit.terms import FittedPotential
    from .grid import build_grid
    dev = torch.device("cuda:0")
    w = synth.make_single("single", seed)
    rec = w["receptor"]
    size = float(np.max(w["size"]))
    nx, ny, nz = dims = synth.grid_dims(size, workload.SPACING)
    origin = np.asarray(w["center"], float) - size / 2
    sfp = sf_params(workload.CFG, "vina")
    lt = np.unique(np.concatenate([[0], w["lig"].types[w["lig"].types >= 0]])).astype(np.int32)
    spec, wts = load_fitted(DATA / "adock_fitted.json", workload.CFG)
    cx = synth.fitted_complex(rec, w["lig"], w["center"], size, spec, seed=seed)
    lig_list, _ = compact_ligands([cx])
    pot = FittedPotential(spec, wts, lig_list, dev)
    D = lambda a, t=np.float32: torch.as_tensor(np.ascontiguousarray(a, t), device=dev)
    rx = D(rec["rec_xyz"])
    vina_args = [D(rec["rec_types"], np.int32), D(sfp["radius"]), D(sfp["hydrophobic"], np.int32),
                 D(sfp["donor"], np.int32), D(sfp["acceptor"], np.int32), D(lt, np.int32), D(grid_prm(sfp))]
    fit_args = [D(cx["rec_ext"], np.int32), D(spec.rec_radius), pot.radius.float().contiguous(),
                pot.W.float().contiguous(), D(spec.grid_prm())]
    out_v = torch.empty(len(lt) * nx * ny * nz, device=dev)
    out_f = torch.empty(pot.TL * nx * ny * nz, device=dev)
    L, p, o = _lib.lib(), _lib.ptr, [float(x) for x in origin]
    stream = _lib.stream_of(dev)
    a = vina_args
    raw = {"vina": lambda: L.adock_build_grid(p(rx), p(a[0]), len(rx), p(a[1]), p(a[2]), p(a[3]), p(a[4]), p(a[5]),
                                              len(lt), *o, workload.SPACING, nx, ny, nz, p(a[6]), 0, p(out_v), stream),
           "fitted": lambda: L.adock_build_grid_fit(p(rx), p(fit_args[0]), len(rx), p(fit_args[1]), p(fit_args[2]),
                                                    pot.TL, p(fit_args[3]), spec.n_rt, *o, workload.SPACING, nx, ny,
                                                    nz, p(fit_args[4]), p(out_f), stream)}
    facade = {"vina": lambda: build_grid(rec["rec_xyz"], rec["rec_types"], sfp, origin, workload.SPACING, dims,
                                         lig_types=lt, device=dev),
              "fitted": lambda: build_grid_fit(cx["rec_xyz"], cx["rec_ext"], pot, origin, workload.SPACING, dims,
                                               device=dev)}
    pairs = _in_cutoff_pairs(rec["rec_xyz"], origin, dims, workload.SPACING, float(sfp["cutoff"]), dev)
    out = {"workload": f"grid sets of the single-complex receptor: {len(rec['rec_xyz'])} atoms, "
                       f"{nx}x{ny}x{nz} points", "in_cutoff_point_atom_pairs": pairs}

    def event_ms(fn):
        """Mean time per call of `reps` back-to-back calls between one pair of HIP events (launch gaps amortised)."""
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        st = [fn() for _ in range(reps)]
        e1.record()
        torch.cuda.synchronize()
        for r in st:
            if isinstance(r, int):
                _lib.check(r)
        return e0.elapsed_time(e1) / reps

    for name, T in (("vina", len(lt)), ("fitted", pot.TL)):
        kms, cms = event_ms(raw[name]), event_ms(facade[name])
        t = kms * 1e-3
        per_pair, per_triple = GRID_OPS[name]
        flops = pairs * per_pair + pairs * T * per_triple
        out[name] = {"types": int(T), "kernel_ms": kms, "call_ms": cms, "grid_sets_per_s": 1.0 / t,
                     "point_atom_pairs_per_s": pairs / t, "triples_per_s": pairs * T / t,
                     "achieved_tflops": flops / t / 1e12, "ops_per_pair": per_pair, "ops_per_triple": per_triple}
    return out


def driver(n_ligands=256, mode="v2_default", fitted=False, seed=0, warmup=True):
    """Seconds per docked ligand through search.dock_problem_kernel (or fit.fdock.dock_fitted) on n_ligands synthetic
    ligands (10-60 heavy atoms, 0-15 torsions) in the one receptor of the screening batch, grid build included;
    ligand preparation is not part of it (the ligands are generated already prepared)."""
    from .search import dock_problem_kernel
    dev = torch.device("cuda:0")
    w = synth.make_screen(seed=seed, n_ligands=n_ligands)
    L = len(w["ligs"])
    rec = workload.receptor_dict(w["receptor"], w["center"], w["size"])
    c, s = np.asarray(w["center"], float), np.broadcast_to(np.asarray(w["size"], float), (3,))
    cfg = workload.CFG
    if fitted:
        from .fit.fdock import DATA, dock_fitted, load_fitted
        spec, wts = load_fitted(DATA / "adock_fitted.json", cfg)
        # one pocket: one annotated receptor shared by every complex (one fitted cube for the screen)
        cxs = [synth.fitted_complex(rec, lg, c, s, spec, seed=i, rec_seed=seed) for i, lg in enumerate(w["ligs"])]

        def run(n):
            return dock_fitted(cxs[:n], spec, wts, mode=mode, seed=seed, cfg=cfg, stats=stats)
    else:
        def run(n):
            prob = workload.gpu_problem([rec], w["ligs"][:n], [0] * n, device=dev)
            return dock_problem_kernel(prob, [c] * n, [s] * n, cfg["search"][mode], seed=seed, stats=stats)
    stats = {}
    if warmup:
        run(min(4, L))
    stats = {}
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res = run(L)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    best = np.array([r["total"][0] for r in res])
    return {"workload": f"search driver, preset {mode}{' (fitted function)' if fitted else ''}: {L} ligands (10-60 heavy "
                        "atoms, 0-15 torsions) in one pocket, grid build + Stage A + MC + refine + polish + finalize",
            "seconds": dt, "seconds_per_docked_ligand": dt / L, "docked_ligands_per_s": L / dt,
            "phase_seconds": {k: v for k, v in stats.items() if k.startswith("t_")}, "widen": stats.get("widen"),
            "best_total_mean": float(best.mean())}


def wide(chains=4096, base_chains=256, base_steps=1000, reps=3, seed=0):
    """The single complex with the chains widened by powers of two at constant chain-steps, as the search driver does
    for small batches (src/adock/search.py:563-569): 256 x 1,000 -> 4,096 chains x 62 steps, Vina-rule BFGS. One
    docked ligand = one launch over all chains; also the best score against the 256 x 1,000 run's."""
    dev = torch.device("cuda:0")
    w, prob = workload.single("single", seed=seed, device=dev)
    kp = KernelProblem(prob)
    f = chains // base_chains
    steps = base_steps // f
    t, q, th = synth.start_states(1, chains, w["center"], w["size"], w["lig"].n_tors, kp.KM)
    st = kp.pack(torch.as_tensor(t, device=dev), torch.as_tensor(q, device=dev), torch.as_tensor(th, device=dev))
    cl = torch.zeros(chains, dtype=torch.int32, device=dev)
    ci = torch.arange(chains, dtype=torch.int32, device=dev)
    seed64 = seed * 1_000_003 + 12345
    times, best = [], None
    for r in range(reps + 1):
        cur, ce, b, be = _bufs(st)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        kp.run_mc(cl, ci, cur, ce, b, be, True, steps, 0, 1.2, 2.0, 2.0, seed64, 0)
        e1.record()
        torch.cuda.synchronize()
        if r:
            times.append(e0.elapsed_time(e1) * 1e-3)
        best = float(be.min())
    dt = float(np.mean(times))
    return {"workload": f"single complex widened: {chains} chains x {steps} MC steps (= {base_chains} x {base_steps} "
                        "chain-steps), Vina-rule BFGS", "docked_ligands_per_s": 1.0 / dt, "seconds": dt,
            "best_score": best}


MLP_INFER_FLOP = 2 * (128 * 256 + 256 * 256 + 256 * 23)   # per row: the three matrix products (208,384)


def mlp(rows=32768, train_batch=4096, reps=20, seed=0):
    """Proposal network kernels: inference of `rows` feature rows (one MC step of the learned search's batch) on MFMA
    (BF16 in, FP32 accumulate), and one training step (forward + backward + Adam, FP32 GEMM kernels) on train_batch
    rows, each timed over `reps` back-to-back calls between HIP events on the launch stream."""
    from .mlp import IN, OUT, ProposalNet
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev).manual_seed(seed)
    net = ProposalNet(device=dev, seed=seed)
    X = torch.randn(rows, IN, generator=g, device=dev)
    Y = torch.empty(rows, OUT, device=dev)
    Xt = torch.randn(train_batch, IN, generator=g, device=dev)
    Tt = torch.randn(train_batch, OUT, generator=g, device=dev)

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps * 1e-3

    def train_step():
        net.backward(net.forward(Xt), Tt)
        net.step()

    ti = timed(lambda: net.infer(X, Y))
    tt = timed(train_step)
    return {"workload": f"proposal network 128-256-256-23: inference of {rows} rows (MFMA, BF16 / FP32 accumulate); "
                        f"training step of {train_batch} rows (FP32)",
            "infer_ms": ti * 1e3, "infer_rows_per_s": rows / ti, "infer_tflops": rows * MLP_INFER_FLOP / ti / 1e12,
            "train_step_ms": tt * 1e3, "train_tflops": 3 * train_batch * MLP_INFER_FLOP / tt / 1e12}


def eg(n_poses=65536, reps=10, seed=0, large=1 << 20):
    """adock_eg on n_poses random poses of the single-complex workload (one launch per call, `reps` calls back to back
    between HIP events on the launch stream), and again on `large` poses (a launch long enough that its start and tail
    no longer count: the figure the kernel's own throughput sets)."""
    dev = torch.device("cuda:0")
    w, prob = workload.single("single", seed=seed, device=dev)
    kp = KernelProblem(prob)

    def run(n):
        t, q, th = synth.start_states(seed + 5, n, w["center"], w["size"], w["lig"].n_tors, kp.KM)
        st = kp.pack(torch.as_tensor(t, device=dev), torch.as_tensor(q, device=dev), torch.as_tensor(th, device=dev))
        cl = torch.zeros(n, dtype=torch.int32, device=dev)
        kp.energy_grad(cl, st)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            kp.energy_grad(cl, st)
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    ms = run(n_poses)
    out = {"workload": f"energy + gradient of {n_poses} independent poses, single complex (35-atom ligand, 454 pairs)",
           "kernel_ms": ms, "evals_per_s": n_poses / (ms * 1e-3), "lig": w["lig"]}
    if large:
        ms_l = run(large)
        out["large"] = {"poses": large, "kernel_ms": ms_l, "evals_per_s": large / (ms_l * 1e-3)}
    return out


def _timed(fn):
    torch.cuda.synchronize()
    t = time.perf_counter()
    fn()
    torch.cuda.synchronize()
    return time.perf_counter() - t


def _budget_to_gap(rows, key, target):
    """Smallest step budget whose mean gap is <= target, its time, or (None, None)."""
    for r in rows:
        if r[key + "_gap"] <= target:
            return r["steps"], r[key + "_s"]
    return None, None


def learned_cost(n=512, chains=64, steps=60, seed=0, p_plain=0.25, log=print):
    """Cost of a learned MC step against a plain one on the learned-proposal workload (n single complexes x chains,
    all chains of one step in one batch): the plain fused search (adock_mc), the one-launch learned search
    (adock_mc_learned) with the network off and on (an untrained network; p_plain of the moves stay random), and the
    step-by-step pipeline (one inference + one step launch per MC step, captured in a HIP graph). Per MC step, ms."""
    _, ligs, prob = complexes(range(20_000, 20_000 + n))
    kp = KernelProblem(prob)
    E = len(ligs)
    st, cl = workload.chain_starts(seed + 2, ligs, [np.zeros(3)] * E, [np.full(3, 26.0)] * E, chains, kp.KM)
    st, cl = torch.as_tensor(st, device=kp.device), torch.as_tensor(cl, device=kp.device)
    ci = torch.arange(len(cl), dtype=torch.int32, device=kp.device)
    s64 = seed * 1_000_003 + 12345
    net = ProposalNet(kp.device, seed=seed)
    ss = StepSearch(kp, cl, net=net)
    b = _bufs(st)

    def reset():
        for x, y in zip(b, _bufs(st)):
            x.copy_(y)

    def plain():
        kp.run_mc(cl, ci, *b, True, steps, 0, 1.2, 2.0, 2.0, s64, 0)

    def evals(fn):
        kp.eval_count.zero_()
        reset()
        fn()
        return int(kp.eval_count.item())

    runs = {"plain (adock_mc)": (plain, lambda: kp.run_mc(cl, ci, *b, True, steps, 0, 1.2, 2.0, 2.0, s64, 0,
                                                           count_evals=True)),
            "learned kernel, network off": (lambda: ss.fused(*b, steps, seed=s64),
                                            lambda: ss.fused(*b, steps, seed=s64, count_evals=True)),
            "learned kernel, network on": (lambda: ss.fused(*b, steps, seed=s64, use_net=True, p_plain=p_plain),
                                           lambda: ss.fused(*b, steps, seed=s64, use_net=True, p_plain=p_plain,
                                                            count_evals=True)),
            "learned kernel, network on, entity choice": (
                lambda: ss.fused(*b, steps, seed=s64, use_net=True, p_plain=p_plain, net_mode=1, tau=0.1),
                lambda: ss.fused(*b, steps, seed=s64, use_net=True, p_plain=p_plain, net_mode=1, tau=0.1,
                                 count_evals=True))}
    out = {"workload": f"{E} complexes (4,000-atom receptor, 35-atom / 8-torsion ligand) x {chains} chains "
                       f"({E * chains} chains per step), {steps} MC steps + init BFGS", "p_plain": p_plain,
           "ms_per_step": {}, "ns_per_evaluation": {}, "evaluations": {}}
    for name, (fn, counted) in runs.items():
        reset()
        fn()   # module load, first-launch costs
        reset()
        dt = _timed(fn)
        n_ev = evals(counted)
        out["ms_per_step"][name] = 1e3 * dt / steps
        out["evaluations"][name] = n_ev
        out["ns_per_evaluation"][name] = 1e9 * dt / n_ev
        log(f"{name}: {1e3 * dt / steps:.3f} ms per MC step, {n_ev} evaluations, {1e9 * dt / n_ev:.3f} ns each")
    reset()
    g = ss.capture(*b, steps, seed=s64, use_net=True, p_plain=p_plain)
    reset()
    dt = _timed(g.replay)
    out["ms_per_step"]["step pipeline (graph), network on"] = 1e3 * dt / steps
    log(f"step pipeline: {1e3 * dt / steps:.3f} ms per MC step")
    ms, ns = out["ms_per_step"], out["ns_per_evaluation"]
    out["learned_over_plain"] = ms["learned kernel, network on"] / ms["plain (adock_mc)"]
    out["learned_over_plain_per_evaluation"] = ns["learned kernel, network on"] / ns["plain (adock_mc)"]
    out["entity_over_plain"] = ms["learned kernel, network on, entity choice"] / ms["plain (adock_mc)"]
    out["entity_over_plain_per_evaluation"] = (ns["learned kernel, network on, entity choice"] /
                                               ns["plain (adock_mc)"])
    out["step_pipeline_over_plain"] = ms["step pipeline (graph), network on"] / ms["plain (adock_mc)"]
    return out


def _time_to_gap(rows, key, target):
    """Search time at which the method's mean gap reaches `target`: log-log interpolation between the last budget
    above it and the first at or below it; None if no budget reaches it or the first one already does."""
    prev = None
    for r in rows:
        g, t = r[key + "_gap"], r[key + "_s"]
        if g <= target:
            if prev is None or prev[0] <= 0 or g <= 0:
                return None if prev is None else t
            f = (np.log(prev[0]) - np.log(target)) / (np.log(prev[0]) - np.log(g))
            return float(np.exp(np.log(prev[1]) + f * (np.log(t) - np.log(prev[1]))))
        prev = (g, t)
    return None


def learned(n_train=512, train_chains=64, train_steps=400, n_eval=512, eval_chains=64,
            budgets=(10, 15, 20, 30, 40, 60, 80, 120, 160, 240, 320, 480), full_steps=1000, epochs=5, p_plain=0.5,
            gap_target=0.1, seed=0, target="move", every=2, weighting="improving", tau=0.1, log=print):
    """Learned proposal: the proposal network trained on the plain method's runs over n_train complexes, then plain (fused
    kernel, one launch) and learned (adock_mc_learned, one launch with MFMA inference inside the chain loop) searches
    on n_eval disjoint complexes at a range of MC-step budgets. Gap = mean over complexes of
    best(budget) - best(plain, full_steps). Reports, per method, the smallest budget reaching gap_target and the
    docked ligands/s there. target "move": regress the accepted moves (features before the move -> the move);
    "jump": regress, from the pose every `every` steps, the jump to the best pose its chain ends with; "entity": learn
    each entity's expected energy gain from every logged move (accepted or not) and let the network pick the entity
    the reference's mutation perturbs (probability ~ exp(score / tau)); "entity_best": the same with the improvement
    of the chain's best energy as the gain. weighting (move target): "accepted" weighs every accepted move (downhill ones by their gain), "improving" only downhill moves, by
    their gain. The speed-up at a matched gap compares the search times at which the two methods' mean gaps reach
    the same value (interpolated between budgets), at gap_target and at a few other gaps."""
    t_all = time.perf_counter()
    # 1. accepted moves of the plain method on the training complexes
    _, tligs, tprob = complexes(range(10_000, 10_000 + n_train))
    tkp = KernelProblem(tprob)
    L = len(tligs)
    st, cl = workload.chain_starts(seed + 1, tligs, [np.zeros(3)] * L, [np.full(3, 26.0)] * L, train_chains, tkp.KM)
    st, cl = torch.as_tensor(st, device=tkp.device), torch.as_tensor(cl, device=tkp.device)
    holder = {}
    if target in ("entity", "entity_best"):
        cap = len(cl) * train_steps + 1024
        ss = StepSearch(tkp, cl, log_cap=cap)
        log(f"training set: {L} complexes x {train_chains} chains x {train_steps} steps, every move logged")
        t_log = _timed(lambda: ss.fused(*_bufs(st), train_steps, seed=seed + 777, log=True, log_all=True))
        Xa, _, _, ga, ea, da = ss.logged(entity=True)
        X, T, W = entity_dataset(Xa, ga, ea, best_gap=da if target == "entity_best" else None)
        M = gain = None
        del Xa, ga, ea, da
    elif target == "move":
        cap = int(len(cl) * train_steps * 0.85) + 1024
        ss = StepSearch(tkp, cl, log_cap=cap)
        log(f"training set: {L} complexes x {train_chains} chains x {train_steps} steps (log capacity {cap} moves)")
        t_log = _timed(lambda: ss.fused(*_bufs(st), train_steps, seed=seed + 777, log=True))
        X, T, M, gain = ss.logged()
        W = move_weights(M, gain, mode=weighting)
    else:
        ss = StepSearch(tkp, cl)
        log(f"training set: {L} complexes x {train_chains} chains x {train_steps} steps, jumps every {every}")
        t_log = _timed(lambda: holder.update(d=jump_dataset(ss, st, train_steps, seed=seed + 777, every=every)))
        X, T, W = holder["d"]
        M = gain = None
    n_moves = int(X.shape[0])
    log(f"{n_moves} training samples in {t_log:.1f} s")
    t_train = _timed(lambda: holder.update(r=train_proposal(X, T, M, gain, epochs=epochs, batch=4096, lr=1e-3,
                                                            seed=seed, log=log, weights=W)))
    net, losses = holder["r"]
    del ss, X, T, M, gain, W, tkp, tprob, holder
    torch.cuda.empty_cache()
    # 2. disjoint evaluation complexes
    _, eligs, eprob = complexes(range(20_000, 20_000 + n_eval))
    ekp = KernelProblem(eprob)
    E = len(eligs)
    st, cl = workload.chain_starts(seed + 2, eligs, [np.zeros(3)] * E, [np.full(3, 26.0)] * E, eval_chains, ekp.KM)
    st, cl = torch.as_tensor(st, device=ekp.device), torch.as_tensor(cl, device=ekp.device)
    ci = torch.arange(len(cl), dtype=torch.int32, device=ekp.device)
    s64 = seed * 1_000_003 + 12345

    def plain(steps):
        b = _bufs(st)
        dt = _timed(lambda: ekp.run_mc(cl, ci, *b, True, steps, 0, 1.2, 2.0, 2.0, s64, 0))
        return b[3].view(E, eval_chains).min(1).values.cpu().numpy(), dt

    plain(10)   # warm-up (module load, first-launch costs)
    ref, t_full = plain(full_steps)
    log(f"plain reference: {full_steps} steps in {t_full:.2f} s, mean best {ref.mean():.3f}")
    ss = StepSearch(ekp, cl, net=net)
    # control: the same search with a network whose output is zero (every network move is the current pose plus the
    # Gaussian noise alone), which separates what the learned move mean adds from what the move shape adds
    zero = ProposalNet(ekp.device, seed=seed)
    zero.p["W3"].zero_()
    zero.p["b3"].zero_()
    zero.refresh_bf16()
    sz = StepSearch(ekp, cl, net=zero)
    mode = 1 if target.startswith("entity") else 0
    rows = []
    for S in budgets:
        bp, tp = plain(S)
        b = _bufs(st)
        tl = _timed(lambda: ss.fused(*b, S, seed=s64, use_net=True, p_plain=p_plain, net_mode=mode, tau=tau))
        bl = b[3].view(E, eval_chains).min(1).values.cpu().numpy()
        b = _bufs(st)
        tz = _timed(lambda: sz.fused(*b, S, seed=s64, use_net=True, p_plain=p_plain, net_mode=mode, tau=tau))
        bz = b[3].view(E, eval_chains).min(1).values.cpu().numpy()
        rows.append({"steps": S, "plain_gap": float((bp - ref).mean()), "learned_gap": float((bl - ref).mean()),
                     "control_gap": float((bz - ref).mean()), "plain_s": tp, "learned_s": tl, "control_s": tz})
        log(f"S={S}: plain gap {rows[-1]['plain_gap']:.3f} ({tp:.2f} s), learned gap {rows[-1]['learned_gap']:.3f} "
            f"({tl:.2f} s), zero-output control gap {rows[-1]['control_gap']:.3f}")
    sp, tp = _budget_to_gap(rows, "plain", gap_target)
    sl, tl = _budget_to_gap(rows, "learned", gap_target)
    matched = {}
    for g in sorted({0.3, 0.2, gap_target, 0.05, 0.03, 0.02}, reverse=True):
        a, b, z = (_time_to_gap(rows, k, g) for k in ("plain", "learned", "control"))
        matched[f"{g:g}"] = {"plain_s": a, "learned_s": b, "control_s": z, "speedup": (a / b) if (a and b) else None,
                             "control_speedup": (a / z) if (a and z) else None}
        log(f"gap {g:g}: plain {a} s, learned {b} s, speed-up {matched[f'{g:g}']['speedup']}, "
            f"zero-output control {matched[f'{g:g}']['control_speedup']}")
    out = {"workload": f"learned proposal ({target} target): trained on {n_moves} samples from {L} single complexes, "
                       f"evaluated on {E} disjoint complexes x {eval_chains} chains ({E * eval_chains} chains per step)",
           "gap_target_kcal": gap_target, "full_plain_steps": full_steps, "full_plain_s": t_full, "log_s": t_log,
           "train_s": t_train, "train_loss": losses, "rows": rows, "total_s": time.perf_counter() - t_all,
           "plain_steps_to_gap": sp, "learned_steps_to_gap": sl,
           "plain_docked_ligands_per_s": E / tp if tp else None,
           "learned_docked_ligands_per_s": E / tl if tl else None}
    out["speedup_at_matched_gap"] = matched[f"{gap_target:g}"]["speedup"]
    out["matched_gaps"] = matched
    out["weighting"] = weighting
    out["learned_over_plain_time_at_equal_steps"] = (sum(r["learned_s"] for r in rows) /
                                                     sum(r["plain_s"] for r in rows))
    return out
