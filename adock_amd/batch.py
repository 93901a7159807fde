"""Batched screening: many ligands (and receptor grid sets) in one persistent launch.

BatchRunner drives adock_mc_batch: all chains of all complexes in a single launch whose wavefronts pull chains from
device work queues, longest chains first (cost model below), so uneven ligands (10-60 atoms, 0-15 torsions) balance
without host round trips. There is one queue per XCD holding the complexes of a few receptors (queue_partition), so
the grids a die's chains read stay in that die's L2; a wave whose queue is drained serves the others. The queue reset and the launch are both kernels on the current stream, so the pair can be
captured once in a HIP graph (torch.cuda.CUDAGraph) and replayed. Per-complex completion stamps (100 MHz real-time
clock) give per-complex latency from inside the kernel.

The reference has no counterpart: it runs L x M chains per launch with one block per NW chains and compacts finished
complexes on the host between segments (src/adock/search.py:614-636).
"""
import numpy as np
import torch

from . import _lib

RT_CLOCK_HZ = 100e6   # s_memrealtime frequency
XCDS = 8              # MI355X: 8 accelerator dies (XCDs), each with its own L2; workgroups are dealt to them in turn


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


def queue_partition(cost, group, n_queues):
    """(order [N] int32, queue_start [n_queues + 1] int32) for adock_mc_batch: the N chains (per-chain cost, group =
    receptor set of each chain's complex) in n_queues queues of about equal cost. Groups are kept whole where they fit
    one queue's share and otherwise dealt over ceil(cost / share) units; units go to the least-loaded queue, largest
    first. Inside a queue the units run one after another, largest first, each with its longest chains first: a
    queue (one XCD) works on one receptor's grids at a time and ends on short work. n_queues <= 1: one queue, longest
    chains first."""
    cost = np.asarray(cost, np.float64)
    group = np.asarray(group)
    N = len(cost)
    if n_queues <= 1 or N == 0:
        return np.argsort(-cost, kind="stable").astype(np.int32), np.array([0, N], np.int32)
    share = cost.sum() / n_queues
    units = []
    for g in np.unique(group):
        idx = np.nonzero(group == g)[0]
        gc = float(cost[idx].sum())
        k = min(len(idx), max(1, int(np.ceil(gc / share - 1e-9)))) if share > 0 else 1
        if k == 1:
            units.append((gc, idx))
            continue
        srt = idx[np.argsort(-cost[idx], kind="stable")]
        for j in range(k):
            part = srt[j::k]
            units.append((float(cost[part].sum()), part))
    units.sort(key=lambda u: -u[0])
    load = np.zeros(n_queues)
    members = [[] for _ in range(n_queues)]
    for c, idx in units:
        q = int(np.argmin(load))
        load[q] += c
        members[q].append(idx)
    order, start = [], [0]
    for m in members:
        # members arrive largest unit first (LPT order); each unit longest chains first
        idx = (np.concatenate([u[np.argsort(-cost[u], kind="stable")] for u in m]) if m
               else np.zeros(0, np.int64))
        order.append(idx)
        start.append(start[-1] + len(idx))
    return np.concatenate(order).astype(np.int32), np.array(start, np.int32)


class BatchRunner:
    def __init__(self, kp, chain_lig, n_complexes=None, cost=None, groups=None, queues=XCDS):
        """kp: KernelProblem; chain_lig [N] (complex of each chain, host or device); cost: per-complex cost (default
        chain_cost of kp's ligands); groups: per-complex receptor set (default kp's rec_idx); queues: work queues, one
        per XCD (queue_partition)."""
        self.kp = kp
        dev = kp.device
        cl = torch.as_tensor(chain_lig).to(dtype=torch.int32)
        self.N = int(cl.numel())
        self.L = int(n_complexes if n_complexes is not None else int(cl.max()) + 1)
        self.chain_lig = cl.to(dev).contiguous()
        cost = np.asarray(cost if cost is not None else kp.cost, np.float64)
        cl_h = cl.cpu().numpy()
        grp = np.asarray(groups if groups is not None else kp.grid["rec_idx"].cpu().numpy())
        order, qstart = queue_partition(cost[cl_h], grp[cl_h], int(queues))
        self.order = torch.as_tensor(order, device=dev)
        self.qstart = torch.as_tensor(qstart, device=dev)
        self.nq = len(qstart) - 1
        self.lig_chains = torch.as_tensor(np.bincount(cl_h, minlength=self.L).astype(np.int32), device=dev)
        self.counter = torch.zeros(max(self.nq, 1), dtype=torch.int32, device=dev)
        self.done = torch.zeros(self.L, dtype=torch.int32, device=dev)
        self.times = torch.zeros(1 + self.L, dtype=torch.int64, device=dev)
        self.lig_evals = torch.zeros(self.L, dtype=torch.int64, device=dev)   # evaluations per complex (count_evals)
        self.blocks = 0
        self.waves = 0   # chains per workgroup of this runner's launches (0: the library's default)
        self.lds_layout = 0   # pair-gradient LDS tile: 0 compact slots, 1 dense per-atom rows (same results)

    @classmethod
    def on_device(cls, kp, chain_lig, n_complexes):
        """A runner built without reading chain_lig on the host (no synchronisation, so it can be made inside a
        stream's work or a HIP-graph capture): one queue over all chains, longest first by their complexes' cost (a
        stable sort on the device: the order queue_partition gives one queue), chains per complex by a scatter-add."""
        self = cls.__new__(cls)
        self.kp = kp
        dev = kp.device
        cl = chain_lig.to(device=dev, dtype=torch.int32).contiguous()
        self.N = int(cl.numel())
        self.L = int(n_complexes)
        self.chain_lig = cl
        idx = cl.long()
        self.order = torch.sort(-kp.cost_dev[idx], stable=True).indices.to(torch.int32)
        self.qstart = None
        self.nq = 1
        self.lig_chains = torch.zeros(self.L, dtype=torch.int32, device=dev).scatter_add_(
            0, idx, torch.ones_like(cl))
        self.counter = torch.zeros(1, dtype=torch.int32, device=dev)
        self.done = torch.zeros(self.L, dtype=torch.int32, device=dev)
        self.times = torch.zeros(1 + self.L, dtype=torch.int64, device=dev)
        self.lig_evals = torch.zeros(self.L, dtype=torch.int64, device=dev)
        self.blocks = 0
        self.waves = 0
        self.lds_layout = 0
        return self

    def launch(self, chain_id, cur, cur_e, best, best_e, init_opt, mc_steps, bfgs_iters, temp, amp_pos, amp_rot, seed,
               step_offset, count_evals=False):
        """Queue reset + one persistent launch on the current stream (asynchronous, capturable)."""
        kp = self.kp
        L = _lib.lib()
        st = _lib.stream_of(kp.device)
        kp.check_chain_buffers(self.N, cur, cur_e, best, best_e)
        _lib.check(L.adock_batch_reset(_lib.ptr(self.counter), int(self.counter.numel()), _lib.ptr(self.done), self.L,
                                       _lib.ptr(self.times), st))
        ci = chain_id if chain_id.dtype == torch.int32 else chain_id.to(torch.int32)
        self._ci = ci.contiguous()
        if self._ci.shape != (self.N,):
            raise _lib.AdockError("chain_id must hold one entry per chain")
        args = (kp._args(False) + [_lib.ptr(self.chain_lig), _lib.ptr(self._ci), self.N, None, None, 0,
                                   _lib.ptr(cur), _lib.ptr(cur_e), _lib.ptr(best), _lib.ptr(best_e), int(bool(init_opt)),
                                   int(mc_steps), int(bfgs_iters), float(temp), float(amp_pos), float(amp_rot),
                                   int(seed) & (2 ** 64 - 1), int(step_offset), _lib.ptr(self.order),
                                   _lib.ptr(self.qstart), self.nq, _lib.ptr(self.counter), _lib.ptr(self.done), _lib.ptr(self.lig_chains),
                                   _lib.ptr(self.times), int(self.blocks), int(self.waves), int(self.lds_layout),
                                   _lib.ptr(kp.eval_count) if count_evals else None,
                                   _lib.ptr(self.lig_evals) if count_evals else None, st])
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
