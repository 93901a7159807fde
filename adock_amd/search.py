"""The search driver around the kernels: Stage A start poses, Monte Carlo + BFGS in segments, refinement, exact
polish and finalize (exact rescoring, clustering, affinities).

dock_problem_kernel reproduces the reference's engine-v2 control flow (src/adock/search.py:544-671): chain widening
for small batches, Stage A (stage_a_kernel, src/adock/search.py:322-366), MC segments of mc_segment steps with
init_opt on segment 0 and step offsets sg * mc_segment (so a split run equals an unsplit one), compaction of
finished / early-stopped complexes, the size-scaled segment budget (mc_size_ref), the longer BFGS refinement of every
chain's best state, the exact-pairwise polish of the polish_top best chains per complex with receptor neighbour lists,
and finalize_kernel (src/adock/search.py:506-537). The kernels come from a KernelProblem (libadock_hip.so, MI355X);
tests may pass another object with the same methods as `kernel_problem`.
"""
import math
import time

import numpy as np
import torch

from .pose import internal_coords, pose_coords, random_quat


def _translation_lattices(prob, boxes_c, boxes_s, cfg):
    """(trans [L, maxT, 3] f32, nT [L]): each complex's translation lattice (step stageA_translation_step inside its
    box shrunk by 1 A), points whose nearest-grid C_H probe energy exceeds stageA_probe_max dropped (the box centre
    alone when none is left). All complexes' lattices go through one probe lookup and one host round trip."""
    gs = prob.gs
    dev = prob.lb.ref.device
    L = len(boxes_c)
    step = cfg["stageA_translation_step"]
    centres = torch.as_tensor(np.stack([np.asarray(c, np.float64) for c in boxes_c]), device=dev)
    Tall, sizes = [], []
    for l in range(L):
        s = boxes_s[l]
        axes = [torch.arange(-float(s[d]) / 2 + 1.0, float(s[d]) / 2 - 1.0 + 1e-6, step, device=dev) for d in range(3)]
        T = torch.stack(torch.meshgrid(*axes, indexing="ij"), -1).reshape(-1, 3) + centres[l]
        Tall.append(T)
        sizes.append(T.shape[0])
    owner = torch.repeat_interleave(torch.arange(L, device=dev), torch.as_tensor(sizes, device=dev))
    T = torch.cat(Tall)
    ep = gs.lookup_nearest(T, torch.zeros(len(T), dtype=torch.long, device=dev), prob.rec_idx.to(dev)[owner])
    keep = ep <= cfg.get("stageA_probe_max", 0.0)
    T, owner = T[keep], owner[keep]
    counts = torch.bincount(owner, minlength=L)
    first = torch.cumsum(counts, 0) - counts
    nT_h = counts.cpu().numpy()
    empty = np.nonzero(nT_h == 0)[0]
    nT_h = np.maximum(nT_h, 1)
    trans = torch.zeros(L, int(nT_h.max()), 3, device=dev)
    trans[owner, torch.arange(len(owner), device=dev) - first[owner]] = T.to(trans.dtype)
    for l in empty:
        trans[l, 0] = centres[l].to(trans.dtype)
    return trans, torch.as_tensor(nT_h, device=dev)


@torch.no_grad()
def stage_a_kernel(prob, kp, boxes_c, boxes_s, cfg, gen):
    """Rigid enumeration: NT random torsion draws per ligand, the C lowest-intra-energy conformers x Q random rotations
    x the complex's translation lattice (step stageA_translation_step inside the box shrunk by 1 A, points whose
    nearest-grid C_H probe energy exceeds stageA_probe_max dropped), scored by kp.stage_a with the per-atom cap
    stageA_atom_cap; the n_chains best (conformer, rotation, translation) triples per complex become the chains' start
    poses. Returns (t [L,M,3], q [L,M,4], theta [L,M,K], e [L,M])."""
    lb, gs = prob.lb, prob.gs
    dev = lb.ref.device
    L, K = lb.L, lb.K
    C, NT, Q, M = cfg["stageA_conformers"], cfg["stageA_conformer_trials"], cfg["stageA_rotations"], cfg["n_chains"]
    th = (torch.rand((L, NT, K), generator=gen, device=dev) * 2 - 1) * math.pi
    th[:, 0] = 0.0
    th = th * lb.tvalid[:, None, :]
    e_intra = prob.intra(internal_coords(lb, th))
    ci = torch.topk(-e_intra, min(C, NT), dim=1).indices
    th_c = torch.gather(th, 1, ci[..., None].expand(L, ci.shape[1], K))
    e_c = torch.gather(e_intra, 1, ci)
    Xc = internal_coords(lb, th_c)
    C = Xc.shape[1]
    hidx_h = np.zeros((L, kp.AM), np.int64)
    for l, p in enumerate(lb.preps):
        hv = np.where(p.heavy)[0]
        hidx_h[l, :len(hv)] = hv
    hidx = torch.as_tensor(hidx_h, device=dev)
    conf_h = torch.gather(Xc, 2, hidx[:, None, :, None].expand(L, C, kp.AM, 3))
    quats = random_quat((Q,), gen, dev)
    trans, nT = _translation_lattices(prob, boxes_c, boxes_s, cfg)
    best_e, best_i = kp.stage_a(conf_h, quats, trans, nT, cfg.get("stageA_atom_cap", 1e9))
    best_e = best_e + e_c.repeat_interleave(Q, dim=1)
    k = min(M, C * Q)
    top = torch.topk(-best_e, k, dim=1).indices
    if k < M:
        top = top[:, torch.arange(M, device=dev) % k]
    li = torch.arange(L, device=dev)[:, None]
    t0 = trans[li, best_i[li, top]]
    return t0, quats[top % Q], th_c[li, top // Q], best_e[li, top]


@torch.no_grad()
def finalize_kernel(prob, kp, states, chain_lig, L, P, cfg):
    """Exact total energies (pairwise inter + penalty + intra, grid total without receptor atoms) of the L x P final
    states, intra from the torch pair potential, heavy-atom RMSD, greedy clustering at cluster_rmsd (lowest energy
    first, at most n_out representatives) and Vina-style affinities (src/adock/search.py:506-537)."""
    lb = prob.lb
    bt, bq, bth = kp.unpack(states, lb.K)
    X = pose_coords(lb, bt.view(L, P, 3), bq.view(L, P, 4), torch.zeros_like(bt).view(L, P, 3),
                    bth.view(L, P, lb.K))
    etot = kp.energy_grad(chain_lig, states, exact=prob.rec is not None)[0].view(L, P)
    eintra = prob.intra(X)
    D = _rmsd(X, lb.heavy)
    return _cluster_all(X, etot, eintra, D, lb.n_tors_vina, prob.pot.p["rot_weight"], cfg["cluster_rmsd"],
                        cfg["n_out"])


def _rmsd(X, heavy, chunk=50_000_000):
    """Heavy-atom RMSD between every two poses of each complex: X [L, P, A, 3] -> [L, P, P], over groups of complexes
    so the [l, P, P, A] difference stays near `chunk` elements."""
    L, P, A = X.shape[:3]
    out = torch.empty(L, P, P, device=X.device, dtype=X.dtype)
    step = max(1, chunk // max(1, P * P * A))
    for l0 in range(0, L, step):
        Xs, hv = X[l0:l0 + step], heavy[l0:l0 + step, None, None, :].to(X.dtype)
        d2 = ((Xs[:, :, None] - Xs[:, None, :]) ** 2).sum(-1)
        out[l0:l0 + step] = ((d2 * hv).sum(-1) / hv.sum(-1).clamp(min=1)).sqrt()
    return out


def _cluster_all(X, E, intra, D, n_tors_vina, rot_weight, cut, n_out, inter=None):
    """Greedy clustering of every complex's final poses at once (lowest energy first; a pose becomes a representative
    when its RMSD D to every earlier one is at least `cut`; at most n_out) and the result dicts with Vina-style
    affinities (src/adock/search.py:389-396, 520-537). X [L, M, A, 3], E / intra / inter [L, M] (inter None: E -
    intra), D [L, M, M], n_tors_vina [L], on any device. Round k takes, in every complex, the first pose in energy order
    (stable) that is at least `cut` from all k earlier representatives: the sequential rule's picks, since a pose it
    rejected was within `cut` of an earlier representative and stays so. Only the representatives leave the device."""
    L, M = E.shape
    dev = E.device
    order = torch.sort(E, dim=1, stable=True).indices
    rank = torch.empty_like(order)
    rank.scatter_(1, order, torch.arange(M, device=dev).expand(L, M).contiguous())
    eligible = torch.ones(L, M, dtype=torch.bool, device=dev)
    reps = torch.full((L, n_out), -1, dtype=torch.long, device=dev)
    li = torch.arange(L, device=dev)
    for k in range(n_out):
        first = torch.where(eligible, rank, torch.full_like(rank, M)).min(1)
        has = first.values < M
        pick = first.indices
        reps[:, k] = torch.where(has, pick, torch.full_like(pick, -1))
        ok = D[li, pick] >= cut
        ok[li, pick] = False
        eligible &= ok | ~has[:, None]
    take = reps.clamp(min=0)
    Xr = X[li[:, None], take].cpu().numpy()
    Er, Ir = E.gather(1, take).cpu().numpy(), intra.gather(1, take).cpu().numpy()
    Fr = inter.gather(1, take).cpu().numpy() if inter is not None else None
    Nt, rh = n_tors_vina.cpu().numpy(), reps.cpu().numpy()
    out = []
    for l in range(L):
        k = int((rh[l] >= 0).sum())
        e, ia = Er[l, :k], Ir[l, :k]
        fi = Fr[l, :k] if Fr is not None else e - ia
        aff = (fi + ia - ia[0]) / (1 + rot_weight * Nt[l])
        out.append({"coords": Xr[l, :k], "affinity": aff, "inter": fi, "intra": ia, "total": e})
    return out


@torch.no_grad()
def finalize(prob, bt, bq, bth, cfg):
    """Torch finalize of poses [L, M] (the finalize_torch path, src/adock/search.py:476-503): grid energy, then,
    when receptor atoms are present, the inter part replaced by the exact pairwise energy through prob.pot (any
    potential, e.g. a fitted one) plus the grid's out-of-box penalty; clustering and affinities as finalize_kernel."""
    lb = prob.lb
    etot, einter, eintra, X = prob.energy(bt, bq, torch.zeros_like(bt), bth)
    if prob.rec is not None:
        pen = einter - prob.inter_nopen(X)
        einter = prob.inter_exact(X) + pen
        etot = einter + eintra
    D = _rmsd(X, lb.heavy)
    return _cluster_all(X, etot, eintra, D, lb.n_tors_vina, prob.pot.p["rot_weight"], cfg["cluster_rmsd"],
                        cfg["n_out"], inter=einter)


def dock_problem_kernel(prob, boxes_c, boxes_s, cfg, seed=0, log=None, stats=None, kernel_problem=None):
    """Full search of every complex of `prob` (boxes_c / boxes_s: per-complex box centre and edge lengths), settings
    from a search preset (config.load_config()["search"]["v2_default"] etc.). Returns one dict per complex (coords
    [n_out, A, 3], affinity, inter, intra, total). `stats` (dict) receives the widening factor, chain-steps and phase
    times."""
    lb = prob.lb
    dev = lb.ref.device
    if kernel_problem is None:
        from .kernel import KernelProblem as kernel_problem
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    L = lb.L
    # small batches leave the GPU idle and latency is set by chain depth: widen (more, shorter chains) at constant
    # chain-steps until L * n_chains >= min_total_chains, keeping at least min_chain_steps steps per chain
    f = 1
    while (L * cfg["n_chains"] * f * 2 <= cfg.get("min_total_chains", 0)
           and cfg["mc_steps"] // (f * 2) >= cfg.get("min_chain_steps", 32)):
        f *= 2
    if f > 1:
        cfg = dict(cfg, n_chains=cfg["n_chains"] * f, mc_steps=cfg["mc_steps"] // f,
                   mc_segment=max(1, cfg.get("mc_segment", cfg["mc_steps"]) // f))
    M = cfg["n_chains"]
    tick = {}

    def mark(k):
        if dev.type == "cuda":
            torch.cuda.synchronize()
        tick[k] = time.perf_counter()

    mark("t0")
    kp = kernel_problem(prob)
    t, q0, th, _ = stage_a_kernel(prob, kp, boxes_c, boxes_s, cfg, gen)
    mark("stageA")
    N = L * M
    cur = kp.pack(t.reshape(N, 3), q0.reshape(N, 4), th.reshape(N, -1))
    best = cur.clone()
    cur_e = torch.zeros(N, device=dev)
    best_e = torch.full((N,), float("inf"), device=dev)
    chain_lig = torch.arange(L, device=dev, dtype=torch.int32).repeat_interleave(M)
    chain_id = torch.arange(N, device=dev, dtype=torch.int32)
    seg = max(1, cfg.get("mc_segment", cfg["mc_steps"]))
    n_seg = max(1, math.ceil(cfg["mc_steps"] / seg)) if cfg["mc_steps"] > 0 else 0
    active = torch.ones(L, dtype=torch.bool, device=dev)
    stall = torch.zeros(L, dtype=torch.long, device=dev)
    glob = torch.full((L,), float("inf"), device=dev)
    es_rounds, es_tol = cfg.get("early_stop_rounds", 0), cfg.get("early_stop_tol", 0.01)
    # size-scaled budget (Vina scales its MC steps with 50 + heavy atoms): complex l runs
    # n_seg * (50 + nh_l) / (50 + mc_size_ref) segments; the loop runs to the largest of them
    if cfg.get("mc_size_ref", 0) > 0:
        nh = lb.heavy.sum(1).float()
        seg_lim = torch.ceil(n_seg * (50 + nh) / (50 + cfg["mc_size_ref"])).long().clamp(min=1)
        n_seg = int(seg_lim.max()) if n_seg > 0 else 0
    else:
        seg_lim = torch.full((L,), n_seg, dtype=torch.long, device=dev)
    chain_steps = 0
    seed64 = int(seed) * 1_000_003 + 12345
    mc_args = (cfg["bfgs_iters"], cfg["mc_temperature"], cfg["mc_amp_pos"], cfg["mc_amp_rot"], seed64)
    if n_seg == 0:
        kp.run_mc(chain_lig, chain_id, cur, cur_e, best, best_e, True, 0, *mc_args, 0)
    for sg in range(n_seg):
        active = active & (seg_lim > sg)
        idx = active.repeat_interleave(M).nonzero().squeeze(1)
        if len(idx) == 0:
            break
        if len(idx) == N:
            kp.run_mc(chain_lig, chain_id, cur, cur_e, best, best_e, sg == 0, seg, *mc_args, sg * seg)
        else:   # compaction: only the chains of still-active complexes are launched
            c_, ce_ = cur[idx].contiguous(), cur_e[idx].contiguous()
            b_, be_ = best[idx].contiguous(), best_e[idx].contiguous()
            kp.run_mc(chain_lig[idx], chain_id[idx], c_, ce_, b_, be_, sg == 0, seg, *mc_args, sg * seg)
            cur[idx], cur_e[idx], best[idx], best_e[idx] = c_, ce_, b_, be_
        chain_steps += len(idx) * seg
        gb = best_e.view(L, M).min(1).values
        improved = gb < glob - es_tol
        glob = torch.minimum(glob, gb)
        stall = torch.where(improved, torch.zeros_like(stall), stall + 1)
        if es_rounds > 0:
            active = active & (stall < es_rounds)
        if log is not None:
            log(f"segment {sg}: active {int(active.sum())}/{L}, best E mean {glob.mean().item():.3f}")
    mark("mc")
    # refine every chain's best state with a longer BFGS
    if cfg.get("final_bfgs_iters", 0) > 0:
        fe = torch.zeros(N, device=dev)
        bb = best.clone()
        kp.run_mc(chain_lig, chain_id, best, fe, bb, best_e, True, 0, cfg["final_bfgs_iters"], *mc_args[1:], 0)
        best = bb
    mark("refine")
    # exact polish: the polish_top lowest-grid-energy chain bests per complex get polish_iters BFGS iterations on the
    # exact pairwise energy, each over its receptor neighbour list
    P = min(cfg.get("polish_top", 0), M)
    if P > 0 and cfg.get("polish_iters", 0) > 0 and prob.rec is not None:
        top = torch.topk(-best_e.view(L, M), P, dim=1).indices
        sel = (top + torch.arange(L, device=dev)[:, None] * M).reshape(-1)
        pc, pb = best[sel].contiguous(), best[sel].clone()
        pe, pbe = torch.zeros(L * P, device=dev), torch.full((L * P,), float("inf"), device=dev)
        margin = cfg.get("polish_nbr_margin", 2.0)
        nb = kp.receptor_neighbours(prob, chain_lig[sel], pc, lb.K, margin) if margin > 0 else (None, None)
        mark("neighbours")
        kp.run_mc(chain_lig[sel], chain_id[sel], pc, pe, pb, pbe, True, 0, cfg["polish_iters"], *mc_args[1:], 0,
                  exact=True, nbr=nb[0], nbr_n=nb[1])
        best, M, chain_lig = pb, P, chain_lig[sel]
    mark("polish")
    if cfg.get("finalize_torch", False):
        bt, bq, bth = kp.unpack(best, lb.K)
        res = finalize(prob, bt.view(L, M, 3), bq.view(L, M, 4), bth.view(L, M, lb.K), cfg)
    else:
        res = finalize_kernel(prob, kp, best, chain_lig, L, M, cfg)
    mark("final")
    if stats is not None:
        ks = list(tick)
        stats.update(widen=f, chain_steps=chain_steps, chain_steps_full=int(seg_lim.sum()) * cfg["n_chains"] * seg,
                     **{f"t_{b}": round(tick[b] - tick[a], 3) for a, b in zip(ks[:-1], ks[1:])})
    return res
