"""The search driver around the kernels: Stage A start poses, Monte Carlo + BFGS in segments, refinement of every
chain's best, exact-pairwise polish and finalize (exact rescoring, clustering, affinities).

Contract kept from the reference's engine-v2 driver (src/adock/search.py:322-366, 506-671), because it fixes the
results at a given seed: the widening rule for small batches, Stage A's draws from the torch Generator (torsion sets,
then rotations) and its selection of start poses, MC segments of `mc_segment` steps whose step offsets are
segment * mc_segment (so a split run equals an unsplit one), init_opt on the first segment only, the size-scaled
budget (`mc_size_ref`), optional early stop, the refine and polish passes and their RNG seeds.

How it runs here: the whole MC schedule (how many segments each complex runs, and so which chains each launch holds) is
planned on the host from ligand sizes known before the search starts, so the segment loop launches back to back
without reading anything from the device; only early stop (off in the presets) and the log callback need the
per-segment best scores on the host. The kernels come from a KernelProblem (libadock_hip.so, MI355X); tests may pass
another object with the same methods as `kernel_problem`.
"""
import math
import time
from dataclasses import dataclass

import numpy as np
import torch

from .pose import internal_coords, pose_coords, random_quat


# ------------------------------------------------------------------------------------------------------ Stage A
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


def _conformer_pool(prob, gen, trials, keep):
    """Torsion sets for Stage A: `trials` draws per ligand, uniform in [-pi, pi) on its valid torsions, with draw 0 the
    input conformation (all zero); the `keep` lowest in intra-ligand energy survive.
    Returns (theta [L, C, K], intra energy [L, C], coordinates [L, C, A, 3])."""
    lb = prob.lb
    u = torch.rand((lb.L, trials, lb.K), generator=gen, device=lb.ref.device)
    theta = (2.0 * u - 1.0) * math.pi
    theta[:, 0] = 0.0
    theta = theta * lb.tvalid[:, None, :]
    e = prob.intra(internal_coords(lb, theta))
    pick = torch.topk(-e, min(keep, trials), dim=1).indices
    theta = theta.gather(1, pick[..., None].expand(-1, -1, lb.K))
    return theta, e.gather(1, pick), internal_coords(lb, theta)


def _heavy_slots(lb, AM):
    """[L, AM] indices of each ligand's heavy atoms in kernel order (padding slots repeat atom 0)."""
    idx = np.zeros((lb.L, AM), np.int64)
    for l, p in enumerate(lb.preps):
        hv = np.flatnonzero(p.heavy)
        idx[l, :len(hv)] = hv
    return torch.as_tensor(idx, device=lb.ref.device)


@torch.no_grad()
def stage_a_kernel(prob, kp, boxes_c, boxes_s, cfg, gen):
    """Rigid enumeration for the chains' start poses: stageA_conformer_trials torsion draws per ligand, the
    stageA_conformers lowest in intra energy, stageA_rotations random rotations, each complex's probe-filtered
    translation lattice; kp.stage_a scores every (conformer, rotation) over the lattice (per-atom cap stageA_atom_cap)
    and the n_chains lowest (rigid + intra) energies become the start poses (cycled when there are fewer candidates).
    Returns (t [L, M, 3], q [L, M, 4], theta [L, M, K], e [L, M])."""
    lb = prob.lb
    dev = lb.ref.device
    M, Q = cfg["n_chains"], cfg["stageA_rotations"]
    theta_c, e_conf, X = _conformer_pool(prob, gen, cfg["stageA_conformer_trials"], cfg["stageA_conformers"])
    C = theta_c.shape[1]
    hs = _heavy_slots(lb, kp.AM)
    conf = X.gather(2, hs[:, None, :, None].expand(-1, C, -1, 3))
    rot = random_quat((Q,), gen, dev)
    trans, n_trans = _translation_lattices(prob, boxes_c, boxes_s, cfg)
    rigid_e, rigid_i = kp.stage_a(conf, rot, trans, n_trans, cfg.get("stageA_atom_cap", 1e9))
    # candidate c * Q + r = conformer c under rotation r
    score = rigid_e + e_conf.repeat_interleave(Q, dim=1)
    n_best = min(M, C * Q)
    order = torch.topk(-score, n_best, dim=1).indices
    if n_best < M:
        order = order[:, torch.arange(M, device=dev) % n_best]
    rows = torch.arange(lb.L, device=dev)[:, None]
    return (trans[rows, rigid_i[rows, order]], rot[order % Q], theta_c[rows, order // Q], score[rows, order])


# ----------------------------------------------------------------------------------------------------- finalize
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


# ------------------------------------------------------------------------------------------------- MC schedule
@dataclass
class MCSchedule:
    """How the MC phase runs: chains per complex, steps per chain, steps per launch (segment) and, per complex, the
    number of segments it runs (all equal unless the budget scales with ligand size)."""
    widen: int
    chains: int
    steps: int
    segment: int
    seg_limit: np.ndarray

    @property
    def n_seg(self):
        return int(self.seg_limit.max()) if self.steps > 0 and len(self.seg_limit) else 0


def widen_factor(n_complexes, cfg):
    """The reference's widening for small batches: the largest power of two f with
    n_complexes * n_chains * f <= min_total_chains and mc_steps // f >= min_chain_steps (f = 1 when none)."""
    target, floor = cfg.get("min_total_chains", 0), cfg.get("min_chain_steps", 32)
    f = 1
    while n_complexes * cfg["n_chains"] * 2 * f <= target and cfg["mc_steps"] // (2 * f) >= floor:
        f *= 2
    return f


def plan_schedule(cfg, heavy_counts):
    """MCSchedule of a search over complexes with these heavy-atom counts, and the cfg with the widened chain and step
    counts. Size-scaled budget (mc_size_ref > 0, after Vina's 50 + heavy atoms): complex l runs
    ceil(n_seg * (50 + nh_l) / (50 + mc_size_ref)) segments (at least one), computed in float32."""
    L = len(heavy_counts)
    f = widen_factor(L, cfg)
    if f > 1:
        cfg = dict(cfg, n_chains=cfg["n_chains"] * f, mc_steps=cfg["mc_steps"] // f,
                   mc_segment=max(1, cfg.get("mc_segment", cfg["mc_steps"]) // f))
    seg = max(1, cfg.get("mc_segment", cfg["mc_steps"]))
    base = max(1, math.ceil(cfg["mc_steps"] / seg)) if cfg["mc_steps"] > 0 else 0
    ref = cfg.get("mc_size_ref", 0)
    if ref > 0 and base > 0:
        nh = np.asarray(heavy_counts, np.float32)
        lim = np.ceil(np.float32(base) * (np.float32(50.0) + nh) / np.float32(50 + ref)).astype(np.int64)
        lim = np.maximum(lim, 1)
    else:
        lim = np.full(L, base, np.int64)
    return MCSchedule(f, cfg["n_chains"], cfg["mc_steps"], seg, lim), cfg


class _Chains:
    """The chains of a search: state buffers [N, SD], their complex and global id, and launches over all of them or
    over a subset (gathered into contiguous buffers and scattered back)."""

    def __init__(self, kp, t, q, theta, L, M, dev):
        N = L * M
        self.kp, self.L, self.M, self.N = kp, L, M, N
        self.cur = kp.pack(t.reshape(N, 3), q.reshape(N, 4), theta.reshape(N, -1))
        self.best = self.cur.clone()
        self.cur_e = torch.zeros(N, device=dev)
        self.best_e = torch.full((N,), float("inf"), device=dev)
        self.lig = torch.arange(L, device=dev, dtype=torch.int32).repeat_interleave(M)
        self.gid = torch.arange(N, device=dev, dtype=torch.int32)

    def launch(self, sel, init_opt, steps, mc, step_offset):
        bufs = (self.cur, self.cur_e, self.best, self.best_e)
        if sel is None:
            self.kp.run_mc(self.lig, self.gid, *bufs, init_opt, steps, *mc, step_offset)
            return
        part = [b[sel].contiguous() for b in bufs]
        self.kp.run_mc(self.lig[sel], self.gid[sel], *part, init_opt, steps, *mc, step_offset)
        for b, p in zip(bufs, part):
            b[sel] = p

    def complex_best(self):
        return self.best_e.view(self.L, self.M).min(1).values


def _run_segments(ch, plan, cfg, seed64, dev, log):
    """The MC phase: segment sg launches the chains of every complex still within its segment budget (and not stopped
    early) for plan.segment steps at step offset sg * segment, init_opt on segment 0. Returns the chain-steps run."""
    mc = (cfg["bfgs_iters"], cfg["mc_temperature"], cfg["mc_amp_pos"], cfg["mc_amp_rot"], seed64)
    if plan.n_seg == 0:
        ch.launch(None, True, 0, mc, 0)
        return 0
    es_rounds, es_tol = cfg.get("early_stop_rounds", 0), cfg.get("early_stop_tol", 0.01)
    stopped = np.zeros(ch.L, bool)
    record = torch.full((ch.L,), float("inf"), device=dev)
    stall = torch.zeros(ch.L, dtype=torch.long, device=dev)
    done_steps = 0
    for sg in range(plan.n_seg):
        run = (plan.seg_limit > sg) & ~stopped
        n_run = int(run.sum())
        if n_run == 0:
            break
        sel = None if n_run == ch.L else torch.as_tensor(
            (np.flatnonzero(run)[:, None] * ch.M + np.arange(ch.M)).reshape(-1), device=dev)
        ch.launch(sel, sg == 0, plan.segment, mc, sg * plan.segment)
        done_steps += n_run * ch.M * plan.segment
        # per-complex progress, kept on the device unless early stop or the log needs it on the host
        cb = ch.complex_best()
        stall = torch.where(cb < record - es_tol, torch.zeros_like(stall), stall + 1)
        record = torch.minimum(record, cb)
        if es_rounds > 0:
            stopped |= run & (stall.cpu().numpy() >= es_rounds)
        if log is not None:
            log(f"segment {sg}: active {int(((plan.seg_limit > sg + 1) & ~stopped).sum())}/{ch.L}, "
                f"best E mean {record.mean().item():.3f}")
    return done_steps


def _polish(prob, kp, ch, cfg, seed64, dev, mark):
    """Exact-pairwise polish of each complex's polish_top lowest-grid-energy chain bests (polish_iters BFGS
    iterations over receptor neighbour lists). Returns (states, chains per complex, complex of each state)."""
    P = min(cfg.get("polish_top", 0), ch.M)
    if P <= 0 or cfg.get("polish_iters", 0) <= 0 or prob.rec is None:
        return ch.best, ch.M, ch.lig
    top = torch.topk(-ch.best_e.view(ch.L, ch.M), P, dim=1).indices
    sel = (top + torch.arange(ch.L, device=dev)[:, None] * ch.M).reshape(-1)
    start = ch.best[sel].contiguous()
    out = start.clone()
    e_cur = torch.zeros(ch.L * P, device=dev)
    e_best = torch.full((ch.L * P,), float("inf"), device=dev)
    margin = cfg.get("polish_nbr_margin", 2.0)
    nbr, nbr_n = kp.receptor_neighbours(prob, ch.lig[sel], start, prob.lb.K, margin) if margin > 0 else (None, None)
    mark("neighbours")
    kp.run_mc(ch.lig[sel], ch.gid[sel], start, e_cur, out, e_best, True, 0, cfg["polish_iters"],
              cfg["mc_temperature"], cfg["mc_amp_pos"], cfg["mc_amp_rot"], seed64, 0, exact=True, nbr=nbr,
              nbr_n=nbr_n)
    return out, P, ch.lig[sel]


def dock_problem_kernel(prob, boxes_c, boxes_s, cfg, seed=0, log=None, stats=None, kernel_problem=None, starts=None):
    """Full search of every complex of `prob` (boxes_c / boxes_s: per-complex box centre and edge lengths), settings
    from a search preset (config.load_config()["search"]["v2_default"] etc.). Returns one dict per complex (coords
    [n_out, A, 3], affinity, inter, intra, total). `stats` (dict) receives the widening factor, chain-steps and phase
    times. starts = (t [L, M, 3], q [L, M, 4], theta [L, M, K]), M = n_chains: start poses to use instead of Stage A's
    (e.g. one Stage A run shared by two engines, so that they search from identical inputs)."""
    lb = prob.lb
    dev = lb.ref.device
    if kernel_problem is None:
        from .kernel import KernelProblem as kernel_problem
    plan, cfg = plan_schedule(cfg, [int(p.heavy.sum()) for p in lb.preps])
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    seed64 = int(seed) * 1_000_003 + 12345
    tick = {}

    def mark(k):
        if dev.type == "cuda":
            torch.cuda.synchronize()
        tick[k] = time.perf_counter()

    mark("t0")
    kp = kernel_problem(prob)
    if starts is None:
        t, q, theta, _ = stage_a_kernel(prob, kp, boxes_c, boxes_s, cfg, gen)
    else:
        t, q, theta = (torch.as_tensor(a, dtype=torch.float32, device=dev) for a in starts)
        if t.shape[:2] != (lb.L, cfg["n_chains"]) or q.shape[:2] != t.shape[:2] or theta.shape[:2] != t.shape[:2]:
            raise ValueError("starts must hold n_chains poses per complex")
    mark("stageA")
    ch = _Chains(kp, t, q, theta, lb.L, plan.chains, dev)
    chain_steps = _run_segments(ch, plan, cfg, seed64, dev, log)
    mark("mc")
    if cfg.get("final_bfgs_iters", 0) > 0:
        # each chain's best state gets a longer BFGS (its MC state is left as it was)
        e0 = torch.zeros(ch.N, device=dev)
        refined = ch.best.clone()
        kp.run_mc(ch.lig, ch.gid, ch.best, e0, refined, ch.best_e, True, 0, cfg["final_bfgs_iters"],
                  cfg["mc_temperature"], cfg["mc_amp_pos"], cfg["mc_amp_rot"], seed64, 0)
        ch.best = refined
    mark("refine")
    states, P, lig_of = _polish(prob, kp, ch, cfg, seed64, dev, mark)
    mark("polish")
    if cfg.get("finalize_torch", False):
        bt, bq, bth = kp.unpack(states, lb.K)
        res = finalize(prob, bt.view(lb.L, P, 3), bq.view(lb.L, P, 4), bth.view(lb.L, P, lb.K), cfg)
    else:
        res = finalize_kernel(prob, kp, states, lig_of, lb.L, P, cfg)
    mark("final")
    if stats is not None:
        ks = list(tick)
        stats.update(widen=plan.widen, chain_steps=chain_steps,
                     chain_steps_full=int(plan.seg_limit.sum()) * plan.chains * plan.segment if plan.n_seg else 0,
                     **{f"t_{b}": round(tick[b] - tick[a], 3) for a, b in zip(ks[:-1], ks[1:])})
    return res
