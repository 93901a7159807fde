"""Benchmark: docked ligands/s on one MI355X for the single-complex GPU workload (4,000-atom receptor in a 26 A sphere
with an 8 A cavity, 35-atom ligand with 8 torsions, 26 A box at 0.375 A -> 71^3 x 8 FP32 grids, 256 chains x 1,000
MC steps, each followed by Vina-rule BFGS, FP32).

One step = build the receptor grid set + run every chain's MC/BFGS search = one docked ligand. Inputs (receptor atoms,
packed ligand, start states) are resident on the GPU before the timed region. Prints one JSON line.

python bench.py [--gpus 1] [--steps 3] [--warmup 1] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

import numpy as np
import torch

from adock_amd import synth, workload
from adock_amd.config import sf_params
from adock_amd.grid import make_gridset
from adock_amd.kernel import KernelProblem

PEAK_FP32_TFLOPS = 157.3
PEAK_BF16_TFLOPS = 2500.0   # dense MFMA
# compute-bound kernels are reported under the contract's compute category ("mfma") against the FP32 peak; these
# kernels issue FP32 vector instructions, no matrix instructions, and say so in "compute_unit"
FP32_VALU = "fp32 VALU (no matrix instructions); peak = MI355X published FP32 157.3 TFLOP/s"
HBM_PEAK_GBS = 8000.0


def flops_per_eval(lig):
    """Algorithmic FP32 operations of one energy+gradient evaluation: 98 A + 40 M + 47 P + nt sincos + P sqrt + 2P exp
    (A heavy atoms, M = sum over atoms of the torsions moving them, P intra pairs)."""
    A = int(lig.heavy.sum())
    M = int(lig.tors_mask[:, lig.heavy].sum()) if lig.n_tors else 0
    P = len(lig.pairs)
    return 98 * A + 40 * M + 47 * P + lig.n_tors + 3 * P


def cpu_baseline(w, prob, threads, steps=300):
    """FP32 restatement (-O3 -ffast-math, OpenMP) on a bounded sample: all 256 chains, `steps` MC steps (about 25
    core-seconds), scaled to docked ligands/s; the grid build (0.3 ms of a docked ligand on the GPU) is left out."""
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    import oracle
    oracle.build()
    AM, KM, lig, grid, sf, rec = workload.host_arrays(prob)
    op = oracle.OracleProblem(AM, KM, lig, grid, sf, rec)
    oracle.set_threads(threads, "libcpubase.so")
    N = w["chains"]
    t, q, th = synth.start_states(1, N, w["center"], w["size"], w["lig"].n_tors, KM)
    st = np.zeros((N, 7 + KM), np.float32)
    st[:, :3], st[:, 3:7], st[:, 7:] = t, q, th
    cur, best = st.copy(), st.copy()
    ce, be = np.zeros(N, np.float32), np.full(N, np.inf, np.float32)
    t0 = time.perf_counter()
    op.run_mc(np.zeros(N), np.arange(N), cur, ce, best, be, True, steps, w["bfgs_iters"], 1.2, 2.0, 2.0, 12345, 0,
              f64=False, libname="libcpubase.so")
    dt = time.perf_counter() - t0
    chain_steps_per_s = N * steps / dt
    return {"value": chain_steps_per_s / (w["chains"] * w["steps"]), "unit": "docked ligands/s", "cores": threads,
            "kind": "port",
            "sample": f"{N} chains x {steps} MC steps of the same complex (FP32 restatement, {threads} OpenMP threads), "
                      f"{dt:.1f} s wall; scaled by {w['chains']}x{w['steps']} chain-steps per docked ligand",
            "chain_steps_per_s": chain_steps_per_s}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--chains-per-block", type=int, default=0)
    ap.add_argument("--line-search-width", type=int, default=0, help="trials per speculative batch (0 = adaptive)")
    ap.add_argument("--workload", choices=["single", "screen", "multi", "learned", "grids", "eg", "wide", "driver", "driver_fitted", "mlp"], default="single",
                    help="single = the headline line (default); the others print their own JSON line")
    ap.add_argument("--small", action="store_true", help="scaled-down sizes for the opt-in workloads")
    ap.add_argument("--queues", type=int, default=8,
                    help="screen / multi: work queues of the persistent launch (8: one per XCD; 1: one global queue)")
    ap.add_argument("--target", choices=["move", "jump"], default="move",
                    help="training target of the learned proposal (--workload learned)")
    ap.add_argument("--copy-calib-mib", type=int, default=0,
                    help="after the timed steps, one device-to-device copy of this many MiB (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build runs the docking path on a single MI355X "
              "(the path shards by complex; no multi-GPU run is part of this benchmark)", file=sys.stderr)
        sys.exit(2)
    if args.workload != "single":
        from adock_amd import experiments as ex
        if args.workload == "screen":
            out = ex.screen(n_ligands=64 if args.small else 512, chains=64, steps=50 if args.small else 500,
                            queues=args.queues)
        elif args.workload in ("driver", "driver_fitted"):
            out = ex.driver(n_ligands=32 if args.small else 256, fitted=args.workload == "driver_fitted")
        elif args.workload == "mlp":
            out = ex.mlp()
            out["roofline"] = {"infer": {"bound": "mfma", "achieved": out["infer_tflops"], "peak": PEAK_BF16_TFLOPS,
                                         "unit": "TFLOP/s", "frac": out["infer_tflops"] / PEAK_BF16_TFLOPS},
                               "train": {"bound": "mfma", "compute_unit": FP32_VALU, "achieved": out["train_tflops"], "peak": PEAK_FP32_TFLOPS,
                                         "unit": "TFLOP/s", "frac": out["train_tflops"] / PEAK_FP32_TFLOPS}}
        elif args.workload == "wide":
            out = ex.wide()
        elif args.workload == "eg":
            out = ex.eg()
            fl = flops_per_eval(out.pop("lig"))
            tf = out["evals_per_s"] * fl / 1e12
            out["roofline"] = {"bound": "mfma", "compute_unit": FP32_VALU, "achieved": tf, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                               "frac": tf / PEAK_FP32_TFLOPS, "flops_per_eval": fl}
        elif args.workload == "grids":
            out = ex.grids()
            for k in ("vina", "fitted"):
                out[k]["roofline"] = {"bound": "mfma", "compute_unit": FP32_VALU, "achieved": out[k]["achieved_tflops"], "peak": PEAK_FP32_TFLOPS,
                                      "unit": "TFLOP/s", "frac": out[k]["achieved_tflops"] / PEAK_FP32_TFLOPS}
        elif args.workload == "multi":
            out = ex.multi(n_receptors=8 if args.small else 64, ligands_per=8 if args.small else 32, chains=32,
                           steps=50 if args.small else 500, queues=args.queues)
        else:
            out = (ex.learned(n_train=8, train_steps=40, n_eval=8, budgets=(10, 20, 40), full_steps=100, epochs=2,
                              target=args.target)
                   if args.small else ex.learned(target=args.target))
        print(json.dumps(out), flush=True)
        return
    dev = torch.device("cuda:0")
    w, prob = workload.single("single", seed=0, device=dev)
    kp = KernelProblem(prob)
    from adock_amd import _lib
    if args.chains_per_block:
        _lib.check(_lib.lib().adock_set_chains_per_block(args.chains_per_block))
    _lib.check(_lib.lib().adock_set_line_search_width(args.line_search_width))
    sfp = sf_params(workload.CFG, "vina")
    rec = workload.receptor_dict(w["receptor"], w["center"], w["size"])
    lt = [np.unique(w["lig"].types)]
    N = w["chains"]
    t, q, th = synth.start_states(1, N, w["center"], w["size"], w["lig"].n_tors, kp.KM)
    st0 = kp.pack(torch.as_tensor(t, device=dev), torch.as_tensor(q, device=dev), torch.as_tensor(th, device=dev))
    cl = torch.zeros(N, dtype=torch.int32, device=dev)
    ci = torch.arange(N, dtype=torch.int32, device=dev)
    seed64 = 0 * 1_000_003 + 12345

    def dock_once(count):
        gs = make_gridset([rec], sfp, workload.SPACING, workload.PENALTY, device=dev, lig_types=lt)
        kp.grid["flat"] = gs.flat
        cur, best = st0.clone(), st0.clone()
        ce = torch.zeros(N, device=dev)
        be = torch.full((N,), float("inf"), device=dev)
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        kp.run_mc(cl, ci, cur, ce, best, be, True, w["steps"], w["bfgs_iters"], 1.2, 2.0, 2.0, seed64, 0,
                  count_evals=count)
        e1.record()
        return be, (e0, e1)

    for _ in range(args.warmup):
        dock_once(False)
    torch.cuda.synchronize()
    kp.eval_count.zero_()
    events, bests = [], []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        be, ev = dock_once(True)
        events.append(ev)
        bests.append(be)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    mc_ms = float(np.mean([a.elapsed_time(b) for a, b in events]))
    n_eval = int(kp.eval_count.item())
    fl = flops_per_eval(w["lig"])
    evals_per_launch = n_eval / args.steps
    achieved = evals_per_launch * fl / (mc_ms * 1e-3) / 1e12
    best_e = torch.stack(bests).min(1).values.cpu().numpy()
    out = {
        "metric": "docked ligands/s/GPU at matched best-score gap; score+grad evals/s",
        "value": args.steps / dt, "unit": "docked ligands/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": dt * 1e3 / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded receptor, ligand and start states)",
        "config": {"workload": "single complex: 4000-atom receptor, 35-atom ligand / 8 torsions, 71^3 x 8 grids, "
                               "256 chains x 1000 MC steps + Vina-rule BFGS (20 iterations)",
                   "chains": N, "mc_steps": w["steps"], "grid": list(map(int, prob.gs.dims[0].tolist())),
                   "types": int(prob.gs.slot[0].ge(0).sum()), "chains_per_block": kp_chains_per_block()},
        "score_grad_evals_per_s": n_eval / (dt),
        "evals_per_step": evals_per_launch,
        "mc_kernel_ms": mc_ms,
        "best_score": {"min": float(best_e.min()), "per_step": [float(x) for x in best_e]},
        "roofline": {"bound": "mfma", "compute_unit": FP32_VALU, "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_TFLOPS, "traffic": measured_traffic(mc_kernel_name(kp.KM, N)),
                     "kernel": mc_kernel_name(kp.KM, N), "flops_per_eval": fl,
                     "traffic_source": "profiles/mc_spec_kernel_single_complex_traffic.json (rocprofv3 FETCH_SIZE / WRITE_SIZE, "
                                       "calibrated on a 1 GiB device copy)"},
    }
    if args.copy_calib_mib > 0:
        src = torch.empty(args.copy_calib_mib << 18, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        out["copy_calib_bytes"] = 2 * src.numel() * 4
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        out["cpu_baseline"] = cpu_baseline(w, prob, threads)
    print(json.dumps(out), flush=True)


def measured_traffic(kernel):
    """HBM bytes per launch of this workload's search kernel from the committed counter summary, or None when the
    summary is absent or was measured on another kernel."""
    f = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                     "mc_spec_kernel_single_complex_traffic.json")
    try:
        with open(f) as fh:
            d = json.load(fh)
    except (OSError, ValueError):
        return None
    return float(d["traffic_bytes"]) if d.get("kernel") == kernel else None


def mc_kernel_name(KM, n_chains):
    """Kernel adock_mc launches for this many chains with the adaptive settings: the speculative line search on 4
    waves per chain while the chains fit one per CU, 2 up to two per CU, else one wave per chain."""
    from adock_amd import _lib
    ns = int(_lib.lib().adock_get_line_search_width())
    if ns == 0:
        cus = torch.cuda.get_device_properties(0).multi_processor_count
        ns = 4 if n_chains <= cus else 2 if n_chains <= 2 * cus else 1
    return f"mc_spec_kernel<{KM}, {ns}, false>" if ns > 1 else f"mc_kernel<{KM}, false>"


def kp_chains_per_block():
    from adock_amd import _lib
    return int(_lib.lib().adock_get_chains_per_block())


if __name__ == "__main__":
    main()
