"""Batched screening: many ligands (and receptor grid sets) in one persistent launch.

BatchRunner drives adock_mc_batch: all chains of all complexes in a single launch whose wavefronts pull chains from a
device work queue, longest chains first (cost model below), so uneven ligands (10-60 atoms, 0-15 torsions) balance
without host round trips. The queue reset and the launch are both kernels on the current stream, so the pair can be
captured once in a HIP graph (torch.cuda.CUDAGraph) and replayed. Per-complex completion stamps (100 MHz real-time
clock) give per-complex latency from inside the kernel.

The reference has no counterpart: it runs L x M chains per launch with one block per NW chains and compacts finished
complexes on the host between segments (src/adock/search.py:614-636).
"""
import numpy as np
import torch

from . import _lib

RT_CLOCK_HZ = 100e6   # s_memrealtime frequency


def chain_cost(preps, iters_rule=True):
    """Relative cost of one MC step of a chain on each ligand: BFGS iterations x FLOPs per evaluation."""
    out = []
    for p in preps:
        A = int(p.heavy.sum())
        M = int(p.tors_mask[:, p.heavy].sum()) if p.n_tors else 0
        P = len(p.pairs)
        it = (25 + A) // 3 if iters_rule else 1
        out.append(it * (98 * A + 40 * M + 50 * P + 400))
    return np.asarray(out, np.float64)


class BatchRunner:
    def __init__(self, kp, chain_lig, n_complexes=None, cost=None):
        """kp: KernelProblem; chain_lig [N] (complex of each chain, host or device); cost: per-complex cost (default
        chain_cost of kp's ligands)."""
        self.kp = kp
        dev = kp.device
        cl = torch.as_tensor(chain_lig).to(dtype=torch.int32)
        self.N = int(cl.numel())
        self.L = int(n_complexes if n_complexes is not None else int(cl.max()) + 1)
        self.chain_lig = cl.to(dev).contiguous()
        cost = np.asarray(cost if cost is not None else kp.cost, np.float64)
        c = cost[cl.cpu().numpy()]
        order = np.argsort(-c, kind="stable").astype(np.int32)
        self.order = torch.as_tensor(order, device=dev)
        self.lig_chains = torch.as_tensor(np.bincount(cl.cpu().numpy(), minlength=self.L).astype(np.int32), device=dev)
        self.counter = torch.zeros(1, dtype=torch.int32, device=dev)
        self.done = torch.zeros(self.L, dtype=torch.int32, device=dev)
        self.times = torch.zeros(1 + self.L, dtype=torch.int64, device=dev)
        self.blocks = 0

    def launch(self, chain_id, cur, cur_e, best, best_e, init_opt, mc_steps, bfgs_iters, temp, amp_pos, amp_rot, seed,
               step_offset, count_evals=False):
        """Queue reset + one persistent launch on the current stream (asynchronous, capturable)."""
        kp = self.kp
        L = _lib.lib()
        st = _lib.stream_of(kp.device)
        kp.check_chain_buffers(self.N, cur, cur_e, best, best_e)
        _lib.check(L.adock_batch_reset(_lib.ptr(self.counter), _lib.ptr(self.done), self.L, _lib.ptr(self.times), st))
        ci = chain_id if chain_id.dtype == torch.int32 else chain_id.to(torch.int32)
        self._ci = ci.contiguous()
        if self._ci.shape != (self.N,):
            raise _lib.AdockError("chain_id must hold one entry per chain")
        args = (kp._args(False) + [_lib.ptr(self.chain_lig), _lib.ptr(self._ci), self.N, None, None, 0,
                                   _lib.ptr(cur), _lib.ptr(cur_e), _lib.ptr(best), _lib.ptr(best_e), int(bool(init_opt)),
                                   int(mc_steps), int(bfgs_iters), float(temp), float(amp_pos), float(amp_rot),
                                   int(seed) & (2 ** 64 - 1), int(step_offset), _lib.ptr(self.order),
                                   _lib.ptr(self.counter), _lib.ptr(self.done), _lib.ptr(self.lig_chains),
                                   _lib.ptr(self.times), int(self.blocks),
                                   _lib.ptr(kp.eval_count) if count_evals else None, st])
        _lib.check(L.adock_mc_batch(*args))

    def latencies(self):
        """Per-complex completion time (s) after the first wavefront started, from the in-kernel stamps."""
        t = self.times.cpu().numpy().astype(np.int64)
        return (t[1:] - t[0]) / RT_CLOCK_HZ


class GraphedBatch:
    """The reset + persistent launch captured once in a HIP graph; replay() re-runs it on the same buffers."""

    def __init__(self, runner, **launch_kw):
        self.runner = runner
        self.kw = launch_kw
        s = torch.cuda.Stream(device=runner.kp.device)
        s.wait_stream(torch.cuda.current_stream(runner.kp.device))
        with torch.cuda.stream(s):       # warm-up launch outside capture
            runner.launch(**launch_kw)
        torch.cuda.current_stream(runner.kp.device).wait_stream(s)
        torch.cuda.synchronize(runner.kp.device)
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            runner.launch(**launch_kw)

    def replay(self):
        self.graph.replay()


def admit(grid_bytes, budget_bytes):
    """HBM admission: split receptors into groups whose grid sets fit the budget (first-fit decreasing). Returns a
    list of index lists; a receptor larger than the budget gets a group of its own."""
    order = np.argsort(-np.asarray(grid_bytes, np.float64), kind="stable")
    groups, free = [], []
    for i in order:
        b = float(grid_bytes[i])
        for g, f in enumerate(free):
            if b <= f:
                groups[g].append(int(i))
                free[g] -= b
                break
        else:
            groups.append([int(i)])
            free.append(budget_bytes - b)
    return [sorted(g) for g in groups]
