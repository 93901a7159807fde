"""Learned Monte Carlo proposal: the search with the proposal network inside the chain loop, and step by step.

StepSearch.fused runs the whole search in one launch (adock_mc_learned): every chain is set up once and its wavefront
runs the network on MFMA before each network move. StepSearch.run is the step-by-step form: one adock_mc_step launch
per MC step over all chains, with adock_mlp_infer (BF16 MFMA) on the features of every chain's current pose right
before it; the two give the same results bit for bit. The step sequence can be captured once in a HIP graph
(torch.cuda.CUDAGraph) and replayed. The exact Vina score and the reference's Metropolis test decide every move; the
reference's random mutation is the fallback (probability p_plain, or whenever the network output is not finite).
The network proposal is not symmetric, so the chain no longer satisfies detailed balance: it is an optimiser, not a
sampler, which is all the docking search needs (the best score is what is reported).

Training data: accepted moves logged by the plain method (features before the move, the move as target, a mask of the
components that exist, the energy gain), on generator seeds disjoint from the evaluation seeds.
"""
import math

import numpy as np
import torch

from . import _lib
from .mlp import IN, OUT, ProposalNet

LOG_W = IN + 2 * OUT + 3   # features, target, mask, energy gain, entity moved, best - current energy before
ENTITY_OUT = (0, 3)        # output scoring the translation / rotation (torsion k: 7 + k)


class StepSearch:
    def __init__(self, kp, chain_lig, chain_id=None, net=None, log_cap=0):
        self.kp = kp
        dev = kp.device
        self.N = int(chain_lig.numel())
        self.chain_lig = chain_lig.to(device=dev, dtype=torch.int32).contiguous()
        self.chain_id = (torch.arange(self.N, dtype=torch.int32, device=dev) if chain_id is None
                         else chain_id.to(device=dev, dtype=torch.int32).contiguous())
        self.net = net
        self.feat = torch.zeros(self.N, IN, device=dev)
        self.Y = torch.zeros(self.N, OUT, device=dev)
        self.log_cap = int(log_cap)
        self.log_buf = torch.zeros(max(self.log_cap, 1), LOG_W, device=dev) if self.log_cap else None
        self.log_count = torch.zeros(1, dtype=torch.int32, device=dev)
        self.graphs = {}
        # the one-launch search's chain queue: longest chains first (their complexes' cost), pulled from a counter
        self.order = torch.sort(-kp.cost_dev[self.chain_lig.long()], stable=True).indices.to(torch.int32)
        self.counter = torch.zeros(1, dtype=torch.int32, device=dev)

    def _step(self, cur, cur_e, best, best_e, init, step, bfgs_iters, temp, amp_pos, amp_rot, seed, use_net, p_plain,
              sig, log, feat_on, count_evals, net_mode=0, tau=0.1, log_all=False):
        kp = self.kp
        kp.check_chain_buffers(self.N, cur, cur_e, best, best_e)
        if use_net and not init:
            self.net.infer(self.feat, self.Y)
        args = (kp._args(False)[:-1] +
                [_lib.ptr(self.chain_lig), _lib.ptr(self.chain_id), self.N, _lib.ptr(cur), _lib.ptr(cur_e),
                 _lib.ptr(best), _lib.ptr(best_e), int(init), int(step), int(bfgs_iters), float(temp), float(amp_pos),
                 float(amp_rot), int(seed) & (2 ** 64 - 1), _lib.ptr(self.Y) if use_net else None,
                 _lib.ptr(self.feat) if feat_on else None,
                 _lib.ptr(self.log_buf) if log and self.log_buf is not None else None, _lib.ptr(self.log_count),
                 self.log_cap, int(bool(use_net)), float(p_plain), float(sig[0]), float(sig[1]), float(sig[2]),
                 int(net_mode), float(tau), int(bool(log_all)),
                 _lib.ptr(kp.eval_count) if count_evals else None, _lib.stream_of(kp.device)])
        _lib.check(_lib.lib().adock_mc_step(*args))

    def run(self, cur, cur_e, best, best_e, steps, bfgs_iters=0, temp=1.2, amp_pos=2.0, amp_rot=2.0, seed=0,
            step_offset=0, init_opt=True, use_net=False, p_plain=0.25, sig=(0.5, 0.15, 0.3), log=False,
            count_evals=False, features=False, net_mode=0, tau=0.1, log_all=False):
        """In place on cur / cur_e / best / best_e (float32 contiguous device tensors), like KernelProblem.run_mc.
        Features (self.feat) are kept current whenever the network, the move log or `features` needs them; every step
        reads the features of its current pose, so a search continued with init_opt=False relies on self.feat being
        the one left by the previous call on the same chains. net_mode 0: the network's output is a displacement;
        1: it scores the entities (translation, rotation, torsions) and the reference's mutation of the one drawn
        with probability ~ exp(score / tau) is applied. log_all: log rejected moves too."""
        if use_net and self.net is None:
            raise _lib.AdockError("use_net needs a ProposalNet")
        feat_on = bool(use_net or log or features)
        if init_opt:
            self._step(cur, cur_e, best, best_e, 1, step_offset, bfgs_iters, temp, amp_pos, amp_rot, seed, False,
                       p_plain, sig, log, feat_on, count_evals, net_mode, tau, log_all)
        for st in range(steps):
            self._step(cur, cur_e, best, best_e, 0, step_offset + st, bfgs_iters, temp, amp_pos, amp_rot, seed,
                       use_net, p_plain, sig, log, feat_on, count_evals, net_mode, tau, log_all)

    def fused(self, cur, cur_e, best, best_e, steps, bfgs_iters=0, temp=1.2, amp_pos=2.0, amp_rot=2.0, seed=0,
              step_offset=0, init_opt=True, use_net=False, p_plain=0.25, sig=(0.5, 0.15, 0.3), log=False,
              count_evals=False, features=False, queued=True, net_mode=0, tau=0.1, log_all=False):
        """run(...) in one launch (adock_mc_learned): each chain is set up once and its wavefront runs the network on
        MFMA inside its step loop, so there is no launch, chain set-up or feature evaluation per MC step. Same poses,
        energies and logged moves (in another order) as run(...) bit for bit; self.feat holds the features of the
        final current poses when the network, the log or `features` is on (a later run(..., init_opt=False) can
        continue from them). queued: persistent waves take the chains longest first from a queue (else one wave per
        chain in chain order); the results are the same."""
        if use_net and self.net is None:
            raise _lib.AdockError("use_net needs a ProposalNet")
        kp = self.kp
        kp.check_chain_buffers(self.N, cur, cur_e, best, best_e)
        net = self.net if use_net else None
        w = ([_lib.ptr(net.bf["W1"]), _lib.ptr(net.p["b1"]), _lib.ptr(net.bf["W2"]), _lib.ptr(net.p["b2"]),
              _lib.ptr(net.bf["W3"]), _lib.ptr(net.p["b3"])] if net is not None else [None] * 6)
        feat_on = bool(use_net or log or features)
        args = (kp._args(False)[:-1] +
                [_lib.ptr(self.chain_lig), _lib.ptr(self.chain_id), self.N, _lib.ptr(cur), _lib.ptr(cur_e),
                 _lib.ptr(best), _lib.ptr(best_e), int(bool(init_opt)), int(steps), int(step_offset), int(bfgs_iters),
                 float(temp), float(amp_pos), float(amp_rot), int(seed) & (2 ** 64 - 1)] + w +
                [_lib.ptr(self.feat) if feat_on else None,
                 _lib.ptr(self.log_buf) if log and self.log_buf is not None else None, _lib.ptr(self.log_count),
                 self.log_cap, int(bool(use_net)), float(p_plain), float(sig[0]), float(sig[1]), float(sig[2]),
                 int(net_mode), float(tau), int(bool(log_all)),
                 _lib.ptr(self.order) if queued else None, _lib.ptr(self.counter) if queued else None,
                 _lib.ptr(kp.eval_count) if count_evals else None, _lib.stream_of(kp.device)])
        if queued:
            self.counter.zero_()
        _lib.check(_lib.lib().adock_mc_learned(*args))

    def capture(self, cur, cur_e, best, best_e, steps, **kw):
        """Captures run(...) on these buffers in a HIP graph; returns the graph (replay() re-runs the search; reset the
        buffers between replays)."""
        dev = self.kp.device
        s = torch.cuda.Stream(device=dev)
        s.wait_stream(torch.cuda.current_stream(dev))
        with torch.cuda.stream(s):
            self._step(cur, cur_e, best, best_e, 1, 0, kw.get("bfgs_iters", 0), kw.get("temp", 1.2),
                       kw.get("amp_pos", 2.0), kw.get("amp_rot", 2.0), kw.get("seed", 0), False, 0.0,
                       kw.get("sig", (0.5, 0.15, 0.3)), False, False, False)
        torch.cuda.current_stream(dev).wait_stream(s)
        torch.cuda.synchronize(dev)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self.run(cur, cur_e, best, best_e, steps, **kw)
        return g

    def logged(self, entity=False):
        """(features [n,128], targets [n,23], masks [n,23], energy gain [n]) of the moves logged so far, and with
        entity=True also the entity each moved ([n] int64: 0 translation, 1 rotation, 2 + k torsion k, -1 a network
        displacement) and the chain's best energy minus its current one before the move ([n], <= 0)."""
        n = min(int(self.log_count.item()), self.log_cap)
        L = self.log_buf[:n]
        out = (L[:, :IN], L[:, IN:IN + OUT], L[:, IN + OUT:IN + 2 * OUT], L[:, IN + 2 * OUT])
        return out + (L[:, IN + 2 * OUT + 1].round().long(), L[:, IN + 2 * OUT + 2]) if entity else out


def entity_dataset(X, gain, entity, max_gain=5.0, best_gap=None):
    """Training set of the entity-choice target from moves logged with log_all (accepted and rejected): each move's
    energy gain, clipped to [0, max_gain], is the target of the output that scores the entity it moved (translation
    0, rotation 3, torsion k < 16: 7 + k), with weight 1 there and 0 elsewhere, so the network learns each entity's
    expected gain from the features of the pose. With best_gap (the chain's best minus its current energy before each
    move) the gain counted is the improvement of the chain's best, gain + best_gap, instead. Network displacement
    moves and torsions past 16 are dropped. Returns (X, T, W)."""
    ok = (entity >= 0) & (entity < 2 + (OUT - 7))
    if best_gap is not None:
        gain = gain + best_gap
    X, gain, entity = X[ok], gain[ok], entity[ok]
    col = torch.where(entity == 0, 0, torch.where(entity == 1, 3, entity - 2 + 7))
    rows = torch.arange(len(col), device=X.device)
    T = torch.zeros(len(col), OUT, device=X.device)
    W = torch.zeros(len(col), OUT, device=X.device)
    T[rows, col] = gain.clamp(min=0.0, max=max_gain)
    W[rows, col] = 1.0
    return X.contiguous(), T.contiguous(), W.contiguous()


def pose_delta(best, cur, n_tors):
    """Jump from the poses cur to the poses best [N, 7+KM], in the network's output layout: translation difference,
    rotation q_best conj(q_cur) with w >= 0, wrapped torsion differences; mask 1 on components that exist."""
    N, SD = cur.shape
    out = torch.zeros(N, OUT, device=cur.device)
    mask = torch.zeros(N, OUT, device=cur.device)
    out[:, :3] = best[:, :3] - cur[:, :3]
    a = best[:, 3:7] / best[:, 3:7].norm(dim=1, keepdim=True)
    b = cur[:, 3:7] / cur[:, 3:7].norm(dim=1, keepdim=True)
    bw, bx, by, bz = b[:, 0], -b[:, 1], -b[:, 2], -b[:, 3]
    aw, ax, ay, az = a.unbind(1)
    q = torch.stack([aw * bw - ax * bx - ay * by - az * bz, aw * bx + ax * bw + ay * bz - az * by,
                     aw * by - ax * bz + ay * bw + az * bx, aw * bz + ax * by - ay * bx + az * bw], 1)
    out[:, 3:7] = q * torch.where(q[:, :1] < 0, -1.0, 1.0)
    mask[:, :7] = 1.0
    K = min(SD - 7, OUT - 7)
    d = best[:, 7:7 + K] - cur[:, 7:7 + K]
    d = d - 2 * math.pi * torch.round(d / (2 * math.pi))
    valid = (torch.arange(K, device=cur.device)[None, :] < n_tors[:, None]).float()
    out[:, 7:7 + K] = d * valid
    mask[:, 7:7 + K] = valid
    return out, mask


def jump_dataset(search, start, steps, seed=0, every=4, max_gain=5.0):
    """Training set of the 'jump' target: run the plain method step by step from `start`, keep the features and the
    current pose of every chain every `every` steps, and after the run take as target the jump from each kept pose to
    the best pose its chain found (pose_delta), weighted 1 + min(E(kept) - E(best), max_gain).
    Returns (X [n,128], T [n,23], W [n,23])."""
    kp = search.kp
    cur, best = start.clone(), start.clone()
    cur_e = torch.zeros(start.shape[0], device=start.device)
    best_e = torch.full((start.shape[0],), float("inf"), device=start.device)
    search.run(cur, cur_e, best, best_e, 0, seed=seed, init_opt=True, features=True)
    feats, poses, energies = [], [], []
    for s in range(steps):
        search.run(cur, cur_e, best, best_e, 1, seed=seed, step_offset=s, init_opt=False, features=True)
        if s % every == 0:
            feats.append(search.feat.clone())
            poses.append(cur.clone())
            energies.append(cur_e.clone())
    nt = torch.as_tensor([kp.preps[l].n_tors for l in search.chain_lig.tolist()], device=start.device)
    X, T, W = [], [], []
    for f, c, e in zip(feats, poses, energies):
        t, m = pose_delta(best, c, nt)
        gap = (e - best_e).clamp(min=0.0, max=max_gain)
        X.append(f)
        T.append(t)
        W.append(m * (1.0 + gap)[:, None])
    return torch.cat(X), torch.cat(T), torch.cat(W)


def move_weights(mask, gain, mode="accepted"):
    """Per-component regression weights, 0 on missing components. mode "accepted": moves that lowered the energy count
    1 + min(gain, 5), uphill (Metropolis) moves 0.25; "improving": only moves that lowered the energy, weighted
    min(gain, 5)."""
    if mode == "accepted":
        w = torch.where(gain > 0, 1.0 + gain.clamp(max=5.0), torch.full_like(gain, 0.25))
    elif mode == "improving":
        w = gain.clamp(min=0.0, max=5.0)
    else:
        raise ValueError(f"unknown weighting {mode!r}")
    return (mask * w[:, None]).contiguous()


def train_proposal(X, T, M, gain, epochs=8, batch=4096, lr=1e-3, seed=0, log=None, device="cuda", weights=None):
    """Trains a ProposalNet on (features, targets) with move_weights(M, gain), or with explicit per-component
    weights. Returns (net, per-epoch losses)."""
    net = ProposalNet(device, seed=seed)
    W = move_weights(M, gain) if weights is None else weights.contiguous()
    losses = net.fit(X.contiguous(), T.contiguous(), W, epochs=epochs, batch=batch, lr=lr, seed=seed, log=log)
    return net, losses
