"""Drivers for the batched and learned-proposal workloads (used by bench.py's opt-in modes).

screen():  one receptor grid set, many uneven ligands, all chains in one persistent HIP-graph-replayed launch.
multi():   many receptors (grid sets resident, admitted in groups under an HBM budget), per-complex latency from the
           kernel's completion stamps.
learned(): trains the proposal network on accepted moves logged by the plain method on training complexes, then
           compares plain and learned search on disjoint evaluation complexes at equal MC-step budgets.
"""
import time

import numpy as np
import torch

from . import synth, workload
from .autotune import Autotuner
from .batch import BatchRunner, GraphedBatch, admit
from .kernel import KernelProblem
from .learned import StepSearch, train_proposal


def _bufs(st):
    N = st.shape[0]
    return st.clone(), torch.zeros(N, device=st.device), st.clone(), torch.full((N,), float("inf"), device=st.device)


def screen(n_ligands=512, chains=64, steps=500, seed=0, reps=2, tune=True, log=print):
    w, prob = workload.screen(seed=seed, n_ligands=n_ligands, chains=chains, steps=steps)
    kp = KernelProblem(prob)
    L = len(w["ligs"])
    st, cl = workload.chain_starts(seed, w["ligs"], [w["center"]] * L, [w["size"]] * L, chains, kp.KM)
    st = torch.as_tensor(st, device=kp.device)
    cl = torch.as_tensor(cl, device=kp.device)
    runner = BatchRunner(kp, cl, n_complexes=L)
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
    return {"workload": f"screen: {L} ligands (10-60 atoms, 0-15 torsions) x {chains} chains x {steps} MC steps, "
                        "one grid set, one persistent graph-replayed launch",
            "docked_ligands_per_s": L / dt, "seconds": dt, "autotuned_shape": shape, "best_score_mean": float(best.mean()),
            "latency_p50_s": float(np.percentile(lat, 50)), "latency_p99_s": float(np.percentile(lat, 99))}


def multi(n_receptors=64, ligands_per=32, chains=32, steps=500, seed=0, hbm_budget=64e9, log=print):
    w, prob = workload.multi(seed=seed, n_receptors=n_receptors, ligands_per=ligands_per, chains=chains, steps=steps)
    grid_bytes = [int(prob.gs.dims[r].prod()) * int((prob.gs.slot[r] >= 0).sum()) * 4 for r in range(n_receptors)]
    groups = admit(grid_bytes, hbm_budget)
    kp = KernelProblem(prob)
    L = len(w["ligs"])
    boxes = [w["boxes"][r] for r in w["rec_idx"]]
    st, cl = workload.chain_starts(seed, w["ligs"], [np.zeros(3)] * L, boxes, chains, kp.KM)
    st = torch.as_tensor(st, device=kp.device)
    cl_all = torch.as_tensor(cl, device=kp.device)
    bufs = _bufs(st)
    ci = torch.arange(len(cl), dtype=torch.int32, device=kp.device)
    lat = np.zeros(L)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for grp in groups:
        ligs = np.nonzero(np.isin(w["rec_idx"], grp))[0]
        sel = torch.as_tensor(np.concatenate([np.arange(l * chains, (l + 1) * chains) for l in ligs]),
                              device=kp.device)
        c_, ce_, b_, be_ = [x[sel].contiguous() for x in bufs]
        runner = BatchRunner(kp, cl_all[sel], n_complexes=L)
        t1 = time.perf_counter()
        runner.launch(ci[sel].contiguous(), c_, ce_, b_, be_, True, steps, 0, 1.2, 2.0, 2.0, seed * 1_000_003 + 12345, 0)
        torch.cuda.synchronize()
        lat[ligs] = runner.latencies()[ligs] + (t1 - t0)
        bufs[3][sel] = be_
    dt = time.perf_counter() - t0
    best = bufs[3].view(L, chains).min(1).values.cpu().numpy()
    return {"workload": f"multi: {n_receptors} receptors (2000-6000 atoms, boxes 16-30 A) x {ligands_per} ligands x "
                        f"{chains} chains x {steps} MC steps, all grid sets resident",
            "grid_set_bytes_total": int(sum(grid_bytes)), "admission_groups": len(groups),
            "docked_ligands_per_s": L / dt, "seconds": dt, "latency_p50_s": float(np.percentile(lat, 50)),
            "latency_p99_s": float(np.percentile(lat, 99)), "best_score_mean": float(best.mean())}


def complexes(seeds, config=2):
    """Config-2-sized complexes, one receptor + one ligand per seed, batched into one Problem."""
    recs, ligs = [], []
    for s in seeds:
        ws = synth.make_single(config, s)
        recs.append(workload.receptor_dict(ws["receptor"], ws["center"], ws["size"]))
        ligs.append(ws["lig"])
    prob = workload.gpu_problem(recs, ligs, np.arange(len(seeds)), with_rec=False)
    return recs, ligs, prob


def learned(train_seeds=range(1000, 1032), eval_seeds=range(8), chains=256, train_steps=200, budgets=(50, 100, 200),
            full_steps=1000, epochs=6, p_plain=0.25, seed=0, log=print):
    # 1. training data from the plain method on disjoint seeds
    t_all = time.perf_counter()
    _, tligs, tprob = complexes(list(train_seeds))
    tkp = KernelProblem(tprob)
    L = len(tligs)
    st, cl = workload.chain_starts(seed + 1, tligs, [np.zeros(3)] * L, [np.full(3, 26.0)] * L, chains, tkp.KM)
    st, cl = torch.as_tensor(st, device=tkp.device), torch.as_tensor(cl, device=tkp.device)
    cap = int(len(cl) * train_steps * 0.6) + 1024
    ss = StepSearch(tkp, cl, log_cap=cap)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ss.run(*_bufs(st), train_steps, seed=seed + 777, log=True)
    torch.cuda.synchronize()
    t_log = time.perf_counter() - t0
    X, T, M, gain = ss.logged()
    n_moves = int(X.shape[0])
    log(f"logged {n_moves} accepted moves in {t_log:.1f} s")
    t0 = time.perf_counter()
    net, losses = train_proposal(X, T, M, gain, epochs=epochs, batch=4096, lr=1e-3, seed=seed, log=log)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    del ss, X, T, M, gain
    # 2. evaluation on disjoint complexes
    _, eligs, eprob = complexes(list(eval_seeds))
    ekp = KernelProblem(eprob)
    E = len(eligs)
    st, cl = workload.chain_starts(seed + 2, eligs, [np.zeros(3)] * E, [np.full(3, 26.0)] * E, chains, ekp.KM)
    st, cl = torch.as_tensor(st, device=ekp.device), torch.as_tensor(cl, device=ekp.device)
    ci = torch.arange(len(cl), dtype=torch.int32, device=ekp.device)
    s64 = seed * 1_000_003 + 12345

    def plain(steps):
        b = _bufs(st)
        torch.cuda.synchronize()
        t = time.perf_counter()
        ekp.run_mc(cl, ci, *b, True, steps, 0, 1.2, 2.0, 2.0, s64, 0)
        torch.cuda.synchronize()
        return b[3].view(E, chains).min(1).values.cpu().numpy(), time.perf_counter() - t

    ref, t_full = plain(full_steps)
    rows = []
    ss = StepSearch(ekp, cl, net=net)
    for S in budgets:
        bp, tp = plain(S)
        b = _bufs(st)
        torch.cuda.synchronize()
        t = time.perf_counter()
        ss.run(*b, S, seed=s64, use_net=True, p_plain=p_plain)
        torch.cuda.synchronize()
        tl = time.perf_counter() - t
        bl = b[3].view(E, chains).min(1).values.cpu().numpy()
        rows.append({"steps": S, "plain_gap": float((bp - ref).mean()), "learned_gap": float((bl - ref).mean()),
                     "plain_s": tp, "learned_s": tl})
        log(f"S={S}: plain gap {rows[-1]['plain_gap']:.3f} ({tp:.2f} s), learned gap {rows[-1]['learned_gap']:.3f} "
            f"({tl:.2f} s)")
    return {"workload": f"learned proposal: trained on {n_moves} accepted moves from {L} complexes (seeds "
                        f"{min(train_seeds)}-{max(train_seeds)}), evaluated on {E} disjoint complexes x {chains} chains",
            "full_plain_steps": full_steps, "full_plain_s": t_full, "log_s": t_log, "train_s": t_train,
            "train_loss": losses, "rows": rows, "total_s": time.perf_counter() - t_all}
