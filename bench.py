"""Benchmark: docked ligands/s on one MI355X for the multi-receptor batch (the largest single-GPU workload): 64 seeded
receptors of 2,000-6,000 heavy atoms with box edges 16-30 A at 0.375 A (one grid set each, 0.8 GB in all, resident),
32 ligands per receptor (10-60 heavy atoms, 0-15 torsions), 32 chains x 500 MC steps per ligand, each step followed by
Vina-rule BFGS, FP32.

One step = rebuild all 64 receptors' grid sets in place + run every chain of the 2,048 complexes in one persistent
work-queue launch = 2,048 docked ligands. Inputs (receptor atoms, packed ligands, start states) are resident on the GPU
before the timed region; each step replays two HIP graphs (grid builds + buffer resets; queue reset + persistent
search kernel). Prints one JSON line with the roofline of the search kernel, the best-score gap against a run at full
exhaustiveness, per-complex latency and the CPU baseline (the reference's own host kernels where they are built).

python bench.py [--gpus 1] [--steps 3] [--warmup 1] [--no-cpu] [--workload multi|single|screen|...]
"""
import argparse
import json
import os
import platform
import sys
import time

import numpy as np
import torch

PEAK_FP32_TFLOPS = 157.3      # MI355X published FP32 vector peak
PEAK_BF16_TFLOPS = 2500.0     # dense MFMA
HBM_PEAK_GBS = 8000.0
FP32_VALU = ("FP32 vector ALU (no matrix instructions); peak = MI355X published FP32 157.3 TFLOP/s (packed FP32; "
             "this build issues unpacked FP32, whose ceiling is 78.6 TFLOP/s)")
HERE = os.path.dirname(os.path.abspath(__file__))


def flops_per_eval(lig):
    """Algorithmic FP32 operations of one energy+gradient evaluation: 98 A + 40 M + 47 P + nt sincos + P sqrt + 2P exp
    (A heavy atoms, M = sum over atoms of the torsions moving them, P intra pairs)."""
    A = int(lig.heavy.sum())
    M = int(lig.tors_mask[:, lig.heavy].sum()) if lig.n_tors else 0
    P = len(lig.pairs)
    return 98 * A + 40 * M + 47 * P + lig.n_tors + 3 * P


def cpu_info():
    model = platform.processor() or "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"model": model, "os_cpu_count": os.cpu_count(), "affinity": len(os.sched_getaffinity(0))}


def host_threads():
    return int(os.environ.get("OMP_NUM_THREADS") or len(os.sched_getaffinity(0)))


def measured_traffic(name):
    """HBM bytes per launch from a committed counter summary (profiles/<name>.json), or None."""
    try:
        with open(os.path.join(HERE, "profiles", name + ".json")) as fh:
            d = json.load(fh)
        return float(d["traffic_bytes"]), d.get("kernel")
    except (OSError, ValueError, KeyError):
        return None, None


# ------------------------------------------------------------------------------------------- multi-receptor batch
def multi_headline(args):
    from adock_amd import workload
    from adock_amd.batch import BatchRunner
    from adock_amd.kernel import KernelProblem
    dev = torch.device("cuda:0")
    n_rec, per, chains, steps = 64, 32, 32, 500
    t_setup = time.perf_counter()
    w, prob = workload.multi(seed=0, n_receptors=n_rec, ligands_per=per, chains=chains, steps=steps, device=dev,
                             keep_builder=True)
    kp = KernelProblem(prob)
    L = len(w["ligs"])
    boxes = [w["boxes"][r] for r in w["rec_idx"]]
    st_h, cl_h = workload.chain_starts(0, w["ligs"], [np.zeros(3)] * L, boxes, chains, kp.KM)
    st0 = torch.as_tensor(st_h, device=dev)
    cl = torch.as_tensor(cl_h, device=dev)
    N = len(cl_h)
    ci = torch.arange(N, dtype=torch.int32, device=dev)
    runner = BatchRunner(kp, cl, n_complexes=L, queues=args.queues)
    runner.lds_layout = args.lds_layout
    cur, best = st0.clone(), st0.clone()
    ce, be = torch.zeros(N, device=dev), torch.full((N,), float("inf"), device=dev)
    seed64 = 0 * 1_000_003 + 12345
    gs = prob.gs

    def prep():
        gs.rebuild()
        cur.copy_(st0)
        best.copy_(st0)
        ce.fill_(0.0)
        be.fill_(float("inf"))

    def search():
        runner.launch(ci, cur, ce, best, be, True, steps, 0, 1.2, 2.0, 2.0, seed64, 0, count_evals=True)

    # capture: warm-up launches outside capture first (the launch shape is looked up once per kernel and size)
    s = torch.cuda.Stream(device=dev)
    s.wait_stream(torch.cuda.current_stream(dev))
    with torch.cuda.stream(s):
        prep()
        search()
    torch.cuda.current_stream(dev).wait_stream(s)
    torch.cuda.synchronize()
    g_prep, g_search = torch.cuda.CUDAGraph(), torch.cuda.CUDAGraph()
    with torch.cuda.graph(g_prep):
        prep()
    with torch.cuda.graph(g_search):
        search()
    t_setup = time.perf_counter() - t_setup

    def one_step(evs=None):
        if evs is not None:
            evs[0].record()
        g_prep.replay()
        if evs is not None:
            evs[1].record()
        g_search.replay()
        if evs is not None:
            evs[2].record()

    for _ in range(args.warmup):
        one_step()
    torch.cuda.synchronize()
    kp.eval_count.zero_()
    runner.lig_evals.zero_()
    events = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        one_step(events[k])
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    prep_ms = float(np.mean([e[0].elapsed_time(e[1]) for e in events]))
    search_ms = float(np.mean([e[1].elapsed_time(e[2]) for e in events]))
    lig_evals = runner.lig_evals.cpu().numpy().astype(np.float64) / args.steps
    n_eval = int(kp.eval_count.item())
    fpe = np.array([flops_per_eval(p) for p in w["ligs"]], np.float64)
    flop_per_launch = float((lig_evals * fpe).sum())
    achieved = flop_per_launch / (search_ms * 1e-3) / 1e12
    lat = prep_ms * 1e-3 + runner.latencies()
    best_l = be.view(L, chains).min(1).values.cpu().numpy()
    traffic, tkernel = measured_traffic("mc_persistent_kernel_multi_receptor_traffic")
    kname = f"mc_persistent_kernel<{kp.KM}, {2 if kp.AM > 64 else 1}, false>"
    out = {
        "metric": "docked ligands/s/GPU at matched best-score gap; score+grad evals/s",
        "value": L * args.steps / dt, "unit": "docked ligands/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": dt * 1e3 / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded receptors, ligands and start states; no downloads)",
        "config": {"workload": f"multi-receptor batch: {n_rec} receptors (2,000-6,000 heavy atoms, box edges 16-30 A at "
                               f"0.375 A, one grid set each) x {per} ligands (10-60 heavy atoms, 0-15 torsions) x "
                               f"{chains} chains x {steps} MC steps + Vina-rule BFGS; one step = {n_rec} grid-set "
                               f"builds + {L} complexes in one persistent launch (2 HIP graphs)",
                   "receptors": n_rec, "complexes": L, "chains": N, "mc_steps": steps,
                   "grid_types_per_receptor": [int(x) for x in np.unique((gs.slot >= 0).sum(1).cpu().numpy())[[0, -1]]],
                   "grid_bytes": int(gs.flat.numel() * 4), "KM": kp.KM,
                   "lds_layout": "dense" if args.lds_layout else "compact", "work_queues": args.queues},
        "score_grad_evals_per_s": n_eval / dt,
        "evals_per_step": n_eval / args.steps,
        "grid_build_and_reset_ms": prep_ms, "search_kernel_ms": search_ms,
        "latency_s": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                      "definition": "per complex, from the start of the step (grid builds) to its last chain's end "
                                    "(in-kernel 100 MHz stamps), last timed step"},
        "best_score_mean": float(best_l.mean()),
        "roofline": {"bound": "valu", "compute_unit": FP32_VALU, "achieved": achieved, "peak": PEAK_FP32_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_TFLOPS,
                     "traffic": traffic if tkernel == kname else None, "kernel": kname,
                     "flop_per_launch": flop_per_launch,
                     "flop_definition": "sum over complexes of its evaluations in the launch (in-kernel per-complex "
                                        "counters) x (98 A + 40 M + 47 P + nt + 3 P) of its ligand",
                     "traffic_source": "profiles/mc_persistent_kernel_multi_receptor_traffic.json (rocprofv3 "
                                       "FETCH_SIZE + WRITE_SIZE per dispatch, copy-kernel calibrated)"},
        "setup_s": t_setup,
    }
    if not args.no_gap:
        out["best_score_gap"] = exhaustive_gap(kp, w, prob, st0, cl_h, chains, steps, be, seed64, dev)
    if not args.no_single:
        out["single_complex"] = single_headline(args, extra=True)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline_multi(w, kp, chains, steps, seed64)
    return out


def exhaustive_gap(kp, w, prob, st0, cl_h, chains, steps, be, seed64, dev, n_rec_sample=8, widen=4, lengthen=2):
    """Best-score gap of the benchmarked budget against full exhaustiveness on a sample: the complexes of the first
    n_rec_sample receptors rerun with widen x the chains (new start states) for lengthen x the MC steps; per complex
    gap = best score at the benchmark budget - best score of the exhaustive run (>= 0 when the budget finds the same
    minimum); reported as the mean and quantiles over the sample."""
    from adock_amd import workload
    from adock_amd.batch import BatchRunner
    L = len(w["ligs"])
    ligs = np.nonzero(np.asarray(w["rec_idx"]) < n_rec_sample)[0]
    n_ch = chains * widen
    preps = [w["ligs"][i] for i in ligs]
    boxes = [w["boxes"][w["rec_idx"][i]] for i in ligs]
    st_h, c_loc = workload.chain_starts(7, preps, [np.zeros(3)] * len(ligs), boxes, n_ch, kp.KM)
    cl = torch.as_tensor(ligs[c_loc].astype(np.int32), device=dev)
    N = len(c_loc)
    st = torch.as_tensor(st_h, device=dev)
    cur, best = st.clone(), st.clone()
    ce, bex = torch.zeros(N, device=dev), torch.full((N,), float("inf"), device=dev)
    t0 = time.perf_counter()
    BatchRunner(kp, cl, n_complexes=L).launch(torch.arange(N, dtype=torch.int32, device=dev), cur, ce, best, bex, True,
                                              steps * lengthen, 0, 1.2, 2.0, 2.0, seed64 + 1, 0)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    full = bex.view(len(ligs), n_ch).min(1).values.cpu().numpy()
    bench = be.view(L, chains).min(1).values.cpu().numpy()[ligs]
    gap = bench - np.minimum(full, bench)
    return {"mean": float(gap.mean()), "p50": float(np.percentile(gap, 50)), "p90": float(np.percentile(gap, 90)),
            "share_within_0.1": float((gap <= 0.1).mean()), "complexes": int(len(ligs)),
            "exhaustive": f"{n_ch} chains x {steps * lengthen} MC steps per complex ({widen * lengthen}x the "
                          f"benchmark's chain-steps), {dt:.1f} s",
            "definition": "per complex: min(best score at the benchmark budget) - min(that, best score of the "
                          "exhaustive run), kcal/mol; mean over the sample"}


def cpu_baseline_multi(w, kp, chains, steps, seed64, rec_sample=0, single_thread_complexes=2, single_thread_steps=100):
    """CPU baseline on a bounded sample of the same workload: receptor `rec_sample`'s grid set (the reference's tiled
    host builder, cpukernel.py:104-137) and its 32 complexes x 32 chains x 500 MC steps (the reference's mc_kernel, one
    OpenMP task per chain), all host threads; then the MC of 2 complexes x 32 chains x 100 steps on one thread. Uses
    the reference's own host kernels (oracle/_ref) when they are built, else the FP32 restatement of the oracle."""
    sys.path.insert(0, HERE)
    import oracle
    from oracle import ref as oref
    from adock_amd import workload
    from adock_amd.config import grid_prm, search_prm, sf_params
    from adock_amd.ligand import pack_ligands
    cfg = workload.CFG
    sfp = sf_params(cfg, "vina")
    ligs = np.nonzero(np.asarray(w["rec_idx"]) == rec_sample)[0]
    preps = [w["ligs"][i] for i in ligs]
    rec = w["receptor_dicts"][rec_sample]
    types = np.union1d([0], np.concatenate([p.types[p.types >= 0] for p in preps])).astype(np.int32)
    n_all = len(sfp["radius"])
    slot = -np.ones((1, n_all), np.int64)
    slot[0, types] = np.arange(len(types))
    size = np.broadcast_to(np.asarray(rec["size"], float), (3,))
    origin = np.asarray(rec["center"], float) - size / 2
    dims = tuple(int(np.ceil(s / workload.SPACING)) + 1 for s in size)
    tabs = (np.asarray(sfp["radius"], np.float32), np.asarray(sfp["hydrophobic"], np.int32),
            np.asarray(sfp["donor"], np.int32), np.asarray(sfp["acceptor"], np.int32))
    threads = host_threads()
    kind = "reference" if oref.available(64, kp.KM) else "port"
    t0 = time.perf_counter()
    if kind == "reference":
        oref._load(64, kp.KM).set_threads(threads)
        G = oref.build_grid_tiled(rec["rec_xyz"], rec["rec_types"], tabs, grid_prm(sfp), origin, workload.SPACING,
                                  dims, types, threads=threads)
    else:
        oracle.set_threads(threads)
        G = oracle.build_grid(rec["rec_xyz"], rec["rec_types"], tabs, grid_prm(sfp), origin, workload.SPACING, dims,
                              types, f64=False)
    t_grid = time.perf_counter() - t0
    AM, KM, lig = pack_ligands(preps, slot, np.zeros(len(preps), np.int64))
    o32 = origin.astype(np.float32)[None]
    grid = dict(flat=np.asarray(G, np.float32).reshape(-1), origin=o32, lo=o32,
                hi=(origin + size).astype(np.float32)[None], rec_idx=np.zeros(len(preps), np.int32),
                offs=np.zeros(1, np.int64), dims=np.array([dims], np.int64), spacing=workload.SPACING,
                penalty=workload.PENALTY)
    sf = dict(prm=search_prm(sfp), rad=tabs[0], hyd=tabs[1], don=tabs[2], acc=tabs[3])
    st, cl = workload.chain_starts(0, preps, [np.zeros(3)] * len(preps), [rec["size"]] * len(preps), chains, KM)
    if kind == "reference":
        eng = oref.RefKernels(AM, KM, lig, grid, sf)
        run = lambda *a: eng.run_mc(*a)
        set_thr = eng.set_threads
    else:
        eng = oracle.OracleProblem(AM, KM, lig, grid, sf)
        run = lambda *a: eng.run_mc(*a, f64=False, libname="libcpubase.so")
        set_thr = lambda n: oracle.set_threads(n, "libcpubase.so")

    def mc(n_cx, n_steps, n_thr):
        n = n_cx * chains
        cur, best = st[:n].copy(), st[:n].copy()
        ce, be = np.zeros(n, np.float32), np.full(n, np.inf, np.float32)
        set_thr(n_thr)
        t = time.perf_counter()
        run(cl[:n], np.arange(n), cur, ce, best, be, True, n_steps, 0, 1.2, 2.0, 2.0, seed64, 0)
        return time.perf_counter() - t, be

    t_mc, be = mc(len(preps), steps, threads)
    t1, _ = mc(single_thread_complexes, single_thread_steps, 1)
    chain_steps_1 = single_thread_complexes * chains * single_thread_steps
    return {"value": len(preps) / (t_grid + t_mc), "unit": "docked ligands/s", "cores": threads, "kind": kind,
            "sample": f"receptor {rec_sample} ({len(rec['rec_xyz'])} atoms, {dims[0]}x{dims[1]}x{dims[2]} x "
                      f"{len(types)} grids) built on the host + its {len(preps)} complexes x {chains} chains x {steps} "
                      f"MC steps, {threads} threads: grid {t_grid:.2f} s + search {t_mc:.1f} s",
            "cpu": cpu_info(), "grid_build_s": t_grid, "search_s": t_mc,
            "one_thread": {"chain_steps_per_s": chain_steps_1 / t1,
                           "cpu_s_per_docked_ligand": t1 / chain_steps_1 * chains * steps + t_grid * threads / len(
                               preps),
                           "sample": f"{single_thread_complexes} complexes x {chains} chains x {single_thread_steps} "
                                     f"MC steps on 1 thread ({t1:.1f} s); grid build per ligand taken as the "
                                     f"all-thread grid time x threads / {len(preps)}"},
            "best_score_mean_sample": float(be.reshape(len(preps), chains).min(1).mean())}


# ------------------------------------------------------------------------------------------------- single complex
def single_headline(args, extra=False):
    """The single-complex workload: 4,000-atom receptor, 35-atom / 8-torsion ligand, 71^3 grids, 256 chains x 1,000
    MC steps + Vina-rule BFGS; one step = grid build + every chain's search."""
    from adock_amd import _lib, synth, workload
    from adock_amd.config import sf_params
    from adock_amd.grid import make_gridset
    from adock_amd.kernel import KernelProblem
    dev = torch.device("cuda:0")
    w, prob = workload.single("single", seed=0, device=dev)
    kp = KernelProblem(prob)
    if args.chains_per_block:
        _lib.check(_lib.lib().adock_set_chains_per_block(args.chains_per_block))
    _lib.check(_lib.lib().adock_set_line_search_width(args.line_search_width))
    sfp = sf_params(workload.CFG, "vina")
    rec = workload.receptor_dict(w["receptor"], w["center"], w["size"])
    lt = [np.unique(w["lig"].types[w["lig"].types >= 0])]
    N = w["chains"]
    t, q, th = synth.start_states(1, N, w["center"], w["size"], w["lig"].n_tors, kp.KM)
    st0 = kp.pack(torch.as_tensor(t, device=dev), torch.as_tensor(q, device=dev), torch.as_tensor(th, device=dev))
    cl = torch.zeros(N, dtype=torch.int32, device=dev)
    ci = torch.arange(N, dtype=torch.int32, device=dev)
    gs = make_gridset([rec], sfp, workload.SPACING, workload.PENALTY, device=dev, lig_types=lt, keep_builder=True)
    kp.grid["flat"] = gs.flat
    cur, best = st0.clone(), st0.clone()
    ce, be = torch.zeros(N, device=dev), torch.full((N,), float("inf"), device=dev)

    def dock_once(count):
        gs.rebuild()
        cur.copy_(st0)
        best.copy_(st0)
        ce.zero_()
        be.fill_(float("inf"))
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        kp.run_mc(cl, ci, cur, ce, best, be, True, w["steps"], w["bfgs_iters"], 1.2, 2.0, 2.0, 12345, 0,
                  count_evals=count)
        e1.record()
        return e0, e1

    steps = 3 if extra else args.steps
    for _ in range(1 if extra else args.warmup):
        dock_once(False)
    torch.cuda.synchronize()
    kp.eval_count.zero_()
    t0 = time.perf_counter()
    events = [dock_once(True) for _ in range(steps)]
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    mc_ms = float(np.mean([a.elapsed_time(b) for a, b in events]))
    n_eval = int(kp.eval_count.item())
    fl = flops_per_eval(w["lig"])
    achieved = n_eval / steps * fl / (mc_ms * 1e-3) / 1e12
    kname = mc_kernel_name(kp.KM, N)
    traffic, tkernel = measured_traffic("mc_spec_kernel_single_complex_traffic")
    out = {"workload": "single complex: 4,000-atom receptor, 35-atom ligand / 8 torsions, 71^3 grids of its "
                       f"{int((gs.slot[0] >= 0).sum())} types, 256 chains x 1,000 MC steps + Vina-rule BFGS",
           "docked_ligands_per_s": steps / dt, "ms_per_docked_ligand": dt * 1e3 / steps,
           "score_grad_evals_per_s": n_eval / dt, "search_kernel_ms": mc_ms, "best_score": float(be.min()),
           "roofline": {"bound": "valu", "compute_unit": FP32_VALU, "achieved": achieved, "peak": PEAK_FP32_TFLOPS,
                        "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_TFLOPS, "kernel": kname,
                        "flops_per_eval": fl, "traffic": traffic if tkernel == kname else None}}
    if extra:
        return out
    line = {"metric": "docked ligands/s/GPU at matched best-score gap; score+grad evals/s",
            "value": steps / dt, "unit": "docked ligands/s", "n_gpus": 1, "steps": steps, "warmup": args.warmup,
            "ms_per_step": dt * 1e3 / steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f32", "data": "synthetic", "config": {"workload": out.pop("workload")}}
    line.update(out)
    return line


def mc_kernel_name(KM, n_chains):
    """Kernel adock_mc launches for this many chains with the adaptive settings: the speculative line search on 4
    waves per chain while the chains fit one per CU, 2 up to two per CU, else one wave per chain."""
    from adock_amd import _lib
    ns = int(_lib.lib().adock_get_line_search_width())
    if ns == 0:
        cus = torch.cuda.get_device_properties(0).multi_processor_count
        ns = 4 if n_chains <= cus else 2 if n_chains <= 2 * cus else 1
    return f"mc_spec_kernel<{KM}, {ns}, false>" if ns > 1 else f"mc_kernel<{KM}, false>"


def other_workload(args):
    from adock_amd import experiments as ex
    wl = args.workload
    if wl == "screen":
        return ex.screen(n_ligands=64 if args.small else 512, chains=64, steps=50 if args.small else 500,
                         queues=args.queues)
    if wl in ("driver", "driver_fitted"):
        return ex.driver(n_ligands=32 if args.small else args.n_ligands, fitted=wl == "driver_fitted")
    if wl == "mlp":
        out = ex.mlp()
        out["roofline"] = {"infer": {"bound": "mfma", "achieved": out["infer_tflops"], "peak": PEAK_BF16_TFLOPS,
                                     "unit": "TFLOP/s", "frac": out["infer_tflops"] / PEAK_BF16_TFLOPS},
                           "train": {"bound": "valu", "compute_unit": FP32_VALU, "achieved": out["train_tflops"],
                                     "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                                     "frac": out["train_tflops"] / PEAK_FP32_TFLOPS}}
        return out
    if wl == "wide":
        return ex.wide()
    if wl == "eg":
        out = ex.eg()
        fl = flops_per_eval(out.pop("lig"))
        tf = out["evals_per_s"] * fl / 1e12
        out["roofline"] = {"bound": "valu", "compute_unit": FP32_VALU, "achieved": tf, "peak": PEAK_FP32_TFLOPS,
                           "unit": "TFLOP/s", "frac": tf / PEAK_FP32_TFLOPS, "flops_per_eval": fl}
        if "large" in out:
            tl = out["large"]["evals_per_s"] * fl / 1e12
            out["large"]["roofline"] = {"bound": "valu", "achieved": tl, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                                        "frac": tl / PEAK_FP32_TFLOPS}
        return out
    if wl == "grids":
        out = ex.grids()
        for k in ("vina", "fitted"):
            out[k]["roofline"] = {"bound": "valu", "compute_unit": FP32_VALU, "achieved": out[k]["achieved_tflops"],
                                  "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                                  "frac": out[k]["achieved_tflops"] / PEAK_FP32_TFLOPS}
        return out
    if wl == "multi_groups":
        return ex.multi(n_receptors=8 if args.small else 64, ligands_per=8 if args.small else 32, chains=32,
                        steps=50 if args.small else 500, queues=args.queues,
                        hbm_budget=args.hbm_budget_mb * 1e6 if args.hbm_budget_mb else None)
    if wl == "learned_cost":
        return ex.learned_cost(n=16 if args.small else 512, steps=10 if args.small else 60)
    return (ex.learned(n_train=8, train_steps=40, n_eval=8, budgets=(10, 20, 40), full_steps=100, epochs=2,
                       target=args.target, weighting=args.weighting, p_plain=args.p_plain, tau=args.tau)
            if args.small else ex.learned(target=args.target, weighting=args.weighting, p_plain=args.p_plain,
                                          tau=args.tau, seed=args.seed))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-gap", action="store_true", help="skip the best-score gap against full exhaustiveness")
    ap.add_argument("--no-single", action="store_true", help="skip the single-complex extra keys")
    ap.add_argument("--chains-per-block", type=int, default=0)
    ap.add_argument("--line-search-width", type=int, default=0, help="trials per speculative batch (0 = adaptive)")
    ap.add_argument("--workload", default="multi",
                    choices=["multi", "single", "screen", "multi_groups", "learned", "learned_cost", "grids", "eg",
                             "wide", "driver", "driver_fitted", "mlp"],
                    help="multi = the headline line (default); the others print their own JSON line")
    ap.add_argument("--small", action="store_true", help="scaled-down sizes for the opt-in workloads")
    ap.add_argument("--n-ligands", type=int, default=256,
                    help="driver / driver_fitted: ligands docked in the one pocket (the reference's scale report "
                         "uses 2,000)")
    ap.add_argument("--queues", type=int, default=8,
                    help="multi / screen / multi_groups: work queues of the persistent launch (8: one per XCD)")
    ap.add_argument("--hbm-budget-mb", type=float, default=0.0,
                    help="multi_groups: device memory one admission group's grid sets may take (0: a quarter of all)")
    ap.add_argument("--target", choices=["move", "jump", "entity", "entity_best"], default="move",
                    help="training target of the learned proposal (--workload learned)")
    ap.add_argument("--weighting", choices=["accepted", "improving"], default="improving",
                    help="regression weights of logged moves (--workload learned, move target)")
    ap.add_argument("--tau", type=float, default=0.1,
                    help="entity target: temperature of the network's entity choice, kcal/mol (--workload learned)")
    ap.add_argument("--seed", type=int, default=0,
                    help="learned: seed of the training and evaluation chains, the MC draws and the network's init")
    ap.add_argument("--p-plain", type=float, default=0.5,
                    help="share of random (plain) moves in the learned search (--workload learned)")
    ap.add_argument("--lds-layout", type=int, choices=[0, 1], default=0,
                    help="multi: pair-gradient LDS tile of the persistent launch (0 compact slots, 1 dense rows)")
    ap.add_argument("--copy-calib-mib", type=int, default=0,
                    help="after the timed steps, one device-to-device copy of this many MiB (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build runs the docking path on a single MI355X "
              "(the path shards by complex; no multi-GPU run is part of this benchmark)", file=sys.stderr)
        sys.exit(2)
    if args.workload == "multi":
        out = multi_headline(args)
    elif args.workload == "single":
        out = single_headline(args)
    else:
        out = other_workload(args)
    if args.copy_calib_mib > 0:
        src = torch.empty(args.copy_calib_mib << 18, device="cuda:0")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        out["copy_calib_bytes"] = 2 * src.numel() * 4
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
