"""KernelProblem: the drop-in facade over libadock_hip.so for the docking search kernels.

Same constructor input, attributes (AM, KM, L, SD, DM) and methods (pack, unpack, energy_grad, run_mc) with the same
argument meaning and in-place conventions as the reference's KernelProblem (src/adock/mckernel.py:602-726), so its
caller (src/adock/search.py:544-671) runs unchanged. Tensors are torch ROCm device tensors; launches go to the current
stream and are asynchronous, like the reference's.
"""
import math

import numpy as np
import torch

from . import _lib
from .config import search_prm
from .ligand import pack_ligands


def ligand_arrays(prob, bucket=None):
    """Packed ligand arrays (numpy) of a Problem: (AM, KM, dict); bucket (AM, KM): smallest buckets to pack to."""
    slot = prob.gs.slot.cpu().numpy()
    return pack_ligands(prob.lb.preps, slot, prob.rec_idx.cpu().numpy(), floor=bucket or (0, 0))


# from this many chains in grid mode, over complexes whose chain costs differ by more than UNEVEN_COST, and where the
# persistent kernel keeps as many waves resident as the per-chain one, run_mc uses the persistent queued launch: its
# longest-first queues keep every CU busy to the end of a launch of uneven chains (8 % off the search driver's MC phase
# at 65,536 chains); equal chains (one complex, widened) are faster in adock_mc
QUEUE_MIN_CHAINS = 4096
UNEVEN_COST = 1.1


class KernelProblem:
    """State layout per chain: [t(3), q(4) (w,x,y,z), theta(KM)]; gradient [d/dt(3), torque(3), d/dtheta(KM)]."""

    use_queues = True
    _queue_occupancy = None

    def __init__(self, prob, bucket=None):
        """bucket (AM, KM): smallest ligand buckets to pack to (default: those the ligands need), so that problems
        run one after another share one state layout [N, 7 + KM] and one kernel bucket."""
        _lib.require_gpu()
        self.AM, self.KM, lig = ligand_arrays(prob, bucket)
        if self.AM > 128 or self.KM > 56:
            raise _lib.AdockError(f"ligand bucket AM={self.AM}, KM={self.KM}: the HIP kernels take at most 128 heavy "
                                  "atoms and 56 torsions per ligand")
        gs = prob.gs
        dev = gs.flat.device
        if dev.type != "cuda":
            raise _lib.AdockError("the grid set must live on the GPU")
        self.device = dev
        if int(gs.dims.min()) < 2:
            raise _lib.AdockError("every grid axis needs at least 2 points")
        T = lambda a: torch.as_tensor(np.ascontiguousarray(a), device=dev)
        self.lig = {k: (T(v) if isinstance(v, np.ndarray) else v) for k, v in lig.items()}
        self.P_MAX = int(lig["P"])
        self.DEG_MAX = int(lig["DEG"])
        self.grid = dict(flat=gs.flat.float().contiguous(), origin=gs.origin.float().contiguous(),
                         lo=gs.lo.float().contiguous(), hi=gs.hi.float().contiguous(),
                         rec_idx=prob.rec_idx.to(device=dev, dtype=torch.int32).contiguous(),
                         offs=gs.offs.to(device=dev, dtype=torch.int64).contiguous(),
                         dims=gs.dims.to(device=dev, dtype=torch.int64).contiguous(),
                         T=int(gs.T), spacing=float(gs.spacing), penalty=float(gs.penalty))
        sp = prob.pot.p
        self.sf = dict(prm=T(search_prm(sp)), rad=T(np.asarray(sp["radius"], np.float32)),
                       hyd=T(np.asarray(sp["hydrophobic"], np.int32)), don=T(np.asarray(sp["donor"], np.int32)),
                       acc=T(np.asarray(sp["acceptor"], np.int32)))
        if prob.rec is not None:
            rx, rt, rm = prob.rec
            self.rec = dict(rxyz=rx.to(dev).float().contiguous(), rtyp=rt.to(device=dev, dtype=torch.int32).contiguous(),
                            rn=rm.to(dev).sum(1).to(torch.int32).contiguous())
            self.RN_MAX = int(rx.shape[1])
        else:
            self.rec = dict(rxyz=torch.zeros(3, device=dev), rtyp=torch.zeros(1, dtype=torch.int32, device=dev),
                            rn=torch.zeros(1, dtype=torch.int32, device=dev))
            self.RN_MAX = 0
        self.L, self.SD, self.DM = prob.lb.L, 7 + self.KM, 6 + self.KM
        from .batch import chain_cost
        self.preps = prob.lb.preps
        self.cost = chain_cost(self.preps)
        self.cost_dev = torch.as_tensor(self.cost, dtype=torch.float64, device=dev)
        # routing of run_mc to the queued launch: decided from all complexes of the problem (host data), so run_mc
        # never reads chain_lig back from the device
        self._uneven = len(self.cost) > 1 and float(np.max(self.cost)) > UNEVEN_COST * float(np.min(self.cost))
        self.eval_count = torch.zeros(1, dtype=torch.int64, device=dev)

    # ---- state packing ----
    def pack(self, t, q, theta):
        """t [N,3], q [N,4], theta [N,K<=KM] -> state [N, SD] float32 contiguous."""
        N, K = theta.shape
        s = torch.zeros(N, self.SD, device=t.device, dtype=torch.float32)
        s[:, :3], s[:, 3:7] = t, q
        s[:, 7:7 + min(K, self.KM)] = theta[:, :self.KM]
        return s.contiguous()

    def unpack(self, s, K):
        th = torch.zeros(s.shape[0], K, device=s.device, dtype=s.dtype)
        th[:, :min(K, self.KM)] = s[:, 7:7 + min(K, self.KM)]
        th = torch.remainder(th + math.pi, 2 * math.pi) - math.pi
        return s[:, :3].clone(), s[:, 3:7] / s[:, 3:7].norm(dim=-1, keepdim=True), th

    # ---- argument groups ----
    def _args(self, exact):
        L, G, S, R = self.lig, self.grid, self.sf, self.rec
        p = _lib.ptr
        return ([self.AM, self.KM] + [p(L[k]) for k in ("ref", "typ", "tm", "ax", "axp", "axu", "pr", "na", "nt",
                                                         "np", "rg")]
                + [self.P_MAX, p(L["gslot"]), p(L["inc"]), p(L["ist"]), self.DEG_MAX]
                + [p(G[k]) for k in ("flat", "origin", "lo", "hi", "rec_idx", "offs", "dims")]
                + [G["T"], G["spacing"], G["penalty"]]
                + [p(S[k]) for k in ("prm", "rad", "hyd", "don", "acc")]
                + [p(R["rxyz"]), p(R["rtyp"]), p(R["rn"]), self.RN_MAX, int(exact)])

    def energy_grad(self, chain_lig, state, exact=False):
        """chain_lig int32 [N], state [N, SD] (device) -> (e [N], g [N, DM])."""
        if exact and self.RN_MAX <= 0:
            raise _lib.AdockError("exact mode needs receptor atoms (Problem.rec)")
        N = state.shape[0]
        if state.dim() != 2 or state.shape[1] != self.SD or chain_lig.shape != (N,):
            raise _lib.AdockError(f"state must be [N, {self.SD}] and chain_lig [N]")
        e = torch.zeros(N, device=self.device)
        g = torch.zeros(N, self.DM, device=self.device)
        cl = chain_lig.to(device=self.device, dtype=torch.int32).contiguous()
        st = state.to(self.device).float().contiguous()
        self._keep = (cl, st)
        _lib.check(_lib.lib().adock_eg(*self._args(exact), _lib.ptr(cl), N, _lib.ptr(st), _lib.ptr(e), _lib.ptr(g),
                                       _lib.stream_of(self.device)))
        return e, g

    def stage_a(self, conf_h, quats, trans, ntrans, cap):
        """conf_h [L, C, AM, 3] heavy-atom conformer coordinates (ligand frame), quats [Q, 4], trans [L, NT, 3]
        (padded lattices), ntrans [L] -> (best_e [L, C*Q], best_i [L, C*Q] int64 index into trans)
        (src/adock/mckernel.py:690-705)."""
        L, C = conf_h.shape[:2]
        Q, NT = quats.shape[0], trans.shape[1]
        if L != self.L:
            raise _lib.AdockError("conf_h must hold one row per complex")
        be = torch.empty(L * C * Q, device=self.device)
        bi = torch.empty(L * C * Q, dtype=torch.int32, device=self.device)
        cf = conf_h.to(self.device).float().contiguous()
        qs = quats.to(self.device).float().contiguous()
        tr = trans.to(self.device).float().contiguous()
        nt = ntrans.to(device=self.device, dtype=torch.int32).contiguous()
        if cf.shape[2] != self.AM:
            raise _lib.AdockError(f"conf_h must have AM = {self.AM} atom slots")
        G, L_ = self.grid, self.lig
        self._keep = (cf, qs, tr, nt)
        _lib.check(_lib.lib().adock_stage_a(
            self.AM, _lib.ptr(cf), _lib.ptr(L_["gslot"]), _lib.ptr(L_["na"]), L, C, _lib.ptr(qs), Q, _lib.ptr(tr),
            _lib.ptr(nt), max(int(NT), 1), *[_lib.ptr(G[k]) for k in ("flat", "origin", "lo", "hi", "rec_idx", "offs",
                                                                       "dims")],
            G["spacing"], G["penalty"], float(cap), _lib.ptr(be), _lib.ptr(bi), _lib.stream_of(self.device)))
        return be.view(L, C * Q), bi.view(L, C * Q).long()

    def receptor_neighbours(self, prob, chain_lig, states, K, margin, chunk=100_000_000):
        """Per chain: the receptor atoms within cutoff + margin of any heavy atom of the chain's pose, in receptor order
        -> (nbr [N, NB] int32, nbr_n [N] int32), NB the longest list (src/adock/mckernel.py:662-688). All chains at
        once, in groups sized so the [chains, atoms, receptor atoms] distance block stays near `chunk` elements."""
        from .pose import pose_coords
        rx, rt, rm = prob.rec
        lb = prob.lb
        dev = states.device
        cut = float(prob.pot.p["cutoff"]) + margin
        N = states.shape[0]
        cl = chain_lig.to(device=dev).long()
        t, q, th = self.unpack(states, K)
        X = pose_coords(lb.subset(cl), t[:, None], q[:, None], torch.zeros_like(t)[:, None], th[:, None])[:, 0]
        heavy = lb.heavy[cl]
        ri = prob.rec_idx.to(dev)[cl]
        rxd, rmd = rx.to(dev).float(), rm.to(dev)
        near = torch.zeros(N, rxd.shape[1], dtype=torch.bool, device=dev)
        step = max(1, chunk // max(1, X.shape[1] * rxd.shape[1]))
        for c0 in range(0, N, step):
            d = torch.cdist(X[c0:c0 + step].float(), rxd[ri[c0:c0 + step]])
            d = d.masked_fill(~heavy[c0:c0 + step, :, None], float("inf"))
            near[c0:c0 + step] = (d < cut).any(1) & rmd[ri[c0:c0 + step]]
        cnt = near.sum(1)
        NB = max(1, int(cnt.max())) if N else 1
        order = torch.argsort((~near).to(torch.int8), dim=1, stable=True)[:, :NB]
        nbr = torch.zeros(N, NB, dtype=torch.int32, device=dev)
        nbr[:, :order.shape[1]] = order.to(torch.int32)
        return nbr, cnt.to(torch.int32)

    def check_chain_buffers(self, N, cur, cur_e, best, best_e):
        """cur / best [N, SD] and cur_e / best_e [N]: contiguous float32 tensors on this problem's device (the kernels
        read and write exactly these extents)."""
        for x in (cur, cur_e, best, best_e):
            if not (x.is_contiguous() and x.dtype == torch.float32 and x.device == self.device):
                raise _lib.AdockError("cur/cur_e/best/best_e must be contiguous float32 tensors on the problem's device")
        if cur.shape != (N, self.SD) or best.shape != (N, self.SD):
            raise _lib.AdockError(f"cur and best must be [{N}, {self.SD}] (chains x 7 + KM)")
        if cur_e.shape != (N,) or best_e.shape != (N,):
            raise _lib.AdockError(f"cur_e and best_e must be [{N}]")

    def queues_keep_occupancy(self):
        """True when the persistent kernel keeps at least as many chains resident per CU as the per-chain kernel at
        this problem's bucket and pair capacity (the runtime's occupancy calculator, asked once): the queued launch
        only pays where it does not lose a wave per SIMD to its extra registers."""
        if self._queue_occupancy is None:
            L = _lib.lib()
            a = L.adock_mc_resident_waves(self.AM, self.KM, self.P_MAX, 1)
            b = L.adock_mc_resident_waves(self.AM, self.KM, self.P_MAX, 0)
            self._queue_occupancy = (a, b)
        a, b = self._queue_occupancy
        return a > 0 and b > 0 and a >= b

    def run_mc(self, chain_lig, chain_id, cur, cur_e, best, best_e, init_opt, mc_steps, bfgs_iters, temp,
               amp_pos, amp_rot, seed, step_offset, exact=False, nbr=None, nbr_n=None, count_evals=False):
        """In place on cur/cur_e/best/best_e (float32 contiguous device tensors). exact=True with nbr [N, NB] /
        nbr_n [N]: each chain's exact inter energy loops only over its receptor neighbour list. From QUEUE_MIN_CHAINS
        chains in grid mode over complexes of uneven cost (and use_queues) the chains run in the persistent queued
        launch (adock_mc_batch), otherwise in adock_mc (speculative line search for few chains); the results are the
        same bit for bit."""
        N = cur.shape[0]
        if N == 0:
            return
        if exact and self.RN_MAX <= 0:
            raise _lib.AdockError("exact mode needs receptor atoms (Problem.rec)")
        self.check_chain_buffers(N, cur, cur_e, best, best_e)
        cl = chain_lig.to(device=self.device, dtype=torch.int32).contiguous()
        ci = chain_id.to(device=self.device, dtype=torch.int32).contiguous()
        if cl.shape != (N,) or ci.shape != (N,):
            raise _lib.AdockError("chain_lig and chain_id must hold one entry per chain")
        if nbr is not None:
            nbr = nbr.to(device=self.device, dtype=torch.int32).contiguous()
            nbr_n = nbr_n.to(device=self.device, dtype=torch.int32).contiguous()
            if nbr.dim() != 2 or nbr.shape[0] != N or nbr_n.shape != (N,):
                raise _lib.AdockError("nbr must be [N, NB] and nbr_n [N]")
            nb_args = [_lib.ptr(nbr), _lib.ptr(nbr_n), int(nbr.shape[1])]
        else:
            nb_args = [None, None, 0]
        if (self.use_queues and not exact and nbr is None and N >= QUEUE_MIN_CHAINS and self._uneven
                and self.queues_keep_occupancy()):
            # many chains over complexes of uneven cost in grid mode: the persistent launch with one longest-first
            # queue, built on the device (same results bit for bit; the tail of long chains no longer idles CUs; no
            # host synchronisation, so the call stays asynchronous and capturable). Problems whose complexes all
            # cost the same stay on adock_mc.
            from .batch import BatchRunner
            runner = BatchRunner.on_device(self, cl, self.L)
            runner.launch(ci, cur, cur_e, best, best_e, init_opt, mc_steps, bfgs_iters, temp, amp_pos, amp_rot,
                          seed, step_offset, count_evals=count_evals)
            self._keep = (cl, ci, runner)
            return
        self._keep = (cl, ci, nbr, nbr_n)
        args = (self._args(exact) + [_lib.ptr(cl), _lib.ptr(ci), N] + nb_args +
                [_lib.ptr(cur), _lib.ptr(cur_e), _lib.ptr(best), _lib.ptr(best_e), int(bool(init_opt)), int(mc_steps),
                 int(bfgs_iters), float(temp), float(amp_pos), float(amp_rot), int(seed) & (2 ** 64 - 1),
                 int(step_offset)])
        if count_evals:
            _lib.check(_lib.lib().adock_mc_counted(*args, _lib.ptr(self.eval_count), _lib.stream_of(self.device)))
        else:
            _lib.check(_lib.lib().adock_mc(*args, _lib.stream_of(self.device)))
