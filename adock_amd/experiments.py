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
from .learned import StepSearch, jump_dataset, train_proposal


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


def learned(n_train=512, train_chains=64, train_steps=400, n_eval=512, eval_chains=64,
            budgets=(10, 15, 20, 30, 40, 60, 80, 120, 160, 240, 320, 480), full_steps=1000, epochs=5, p_plain=0.25,
            gap_target=0.1, seed=0, target="move", every=2, log=print):
    """Config 5: the proposal network trained on the plain method's runs over n_train complexes, then plain (fused
    kernel, one launch) and learned (step kernel + MFMA inference, the whole search captured in one HIP graph)
    searches on n_eval disjoint complexes at a range of MC-step budgets. Gap = mean over complexes of
    best(budget) - best(plain, full_steps). Reports, per method, the smallest budget reaching gap_target and the
    docked ligands/s there. target "move": regress the accepted moves (features before the move -> the move);
    "jump": regress, from the pose every `every` steps, the jump to the best pose its chain ends with."""
    t_all = time.perf_counter()
    # 1. accepted moves of the plain method on the training complexes
    _, tligs, tprob = complexes(range(10_000, 10_000 + n_train))
    tkp = KernelProblem(tprob)
    L = len(tligs)
    st, cl = workload.chain_starts(seed + 1, tligs, [np.zeros(3)] * L, [np.full(3, 26.0)] * L, train_chains, tkp.KM)
    st, cl = torch.as_tensor(st, device=tkp.device), torch.as_tensor(cl, device=tkp.device)
    holder = {}
    if target == "move":
        cap = int(len(cl) * train_steps * 0.85) + 1024
        ss = StepSearch(tkp, cl, log_cap=cap)
        log(f"training set: {L} complexes x {train_chains} chains x {train_steps} steps (log capacity {cap} moves)")
        t_log = _timed(lambda: ss.run(*_bufs(st), train_steps, seed=seed + 777, log=True))
        X, T, M, gain = ss.logged()
        W = None
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
    rows = []
    for S in budgets:
        bp, tp = plain(S)
        b = _bufs(st)
        g = ss.capture(*b, S, seed=s64, use_net=True, p_plain=p_plain)
        for x, y in zip(b, _bufs(st)):
            x.copy_(y)
        tl = _timed(g.replay)
        bl = b[3].view(E, eval_chains).min(1).values.cpu().numpy()
        del g
        rows.append({"steps": S, "plain_gap": float((bp - ref).mean()), "learned_gap": float((bl - ref).mean()),
                     "plain_s": tp, "learned_s": tl})
        log(f"S={S}: plain gap {rows[-1]['plain_gap']:.3f} ({tp:.2f} s), learned gap {rows[-1]['learned_gap']:.3f} "
            f"({tl:.2f} s)")
    sp, tp = _budget_to_gap(rows, "plain", gap_target)
    sl, tl = _budget_to_gap(rows, "learned", gap_target)
    out = {"workload": f"learned proposal ({target} target): trained on {n_moves} samples from {L} config-2 complexes, "
                       f"evaluated on {E} disjoint complexes x {eval_chains} chains ({E * eval_chains} chains per step)",
           "gap_target_kcal": gap_target, "full_plain_steps": full_steps, "full_plain_s": t_full, "log_s": t_log,
           "train_s": t_train, "train_loss": losses, "rows": rows, "total_s": time.perf_counter() - t_all,
           "plain_steps_to_gap": sp, "learned_steps_to_gap": sl,
           "plain_docked_ligands_per_s": E / tp if tp else None,
           "learned_docked_ligands_per_s": E / tl if tl else None}
    out["speedup_at_matched_gap"] = (tp / tl) if (tp and tl) else None
    return out
