"""HBM admission for multi-receptor batches.

Every receptor needs a grid set of (box points) x (its ligands' XS types + C_H) float32 values, known from its box and
its ligands before anything is built (grid_bytes). Receptors are put in groups whose grid sets fit a device-memory
budget, first-fit decreasing (batch.admit); then, one group at a time, the group's grid sets are built on the GPU, all
chains of its complexes run in one persistent queued launch, and the grid sets are released before the next group is
built, so at most one group's grids are resident. Every group is packed to the same ligand buckets (AM, KM), so the
chains of all groups share one state layout and run the same kernel: a complex's result does not depend on its group
(the same bits as with every grid set resident at once).

The reference has no counterpart: dock_batch builds the grid sets of every complex it is given at once and keeps no
device budget, so its callers bound memory through the batch they pass (src/adock/dock.py:69-110).
"""
import time

import numpy as np
import torch

from . import workload
from .batch import XCDS, BatchRunner, admit
from .grid import C_H, box_points
from .kernel import KernelProblem
from .ligand import bucket as _bucket


def grid_bytes(receptors, preps, rec_idx, spacing=workload.SPACING):
    """Bytes of each receptor's grid set as workload.gpu_problem builds it: its box points x (the XS types of its
    ligands + C_H) x 4."""
    rec_idx = np.asarray(rec_idx)
    out = []
    for r, rec in enumerate(receptors):
        ts = [p.types[p.types >= 0] for p, ri in zip(preps, rec_idx) if ri == r]
        types = np.union1d([C_H], np.concatenate(ts)) if ts else np.array([C_H])
        size = np.broadcast_to(np.asarray(rec["size"], float), (3,))
        out.append(int(np.prod(box_points(size, spacing))) * len(types) * 4)
    return out


def buckets_of(preps):
    """(AM, KM) that every ligand of the batch fits (the packing rule of ligand.pack_ligands)."""
    nh = max(int(p.heavy.sum()) for p in preps)
    return (64 if nh <= 64 else _bucket(nh, 32)), _bucket(max(max(p.n_tors for p in preps), 1), 8)


class AdmittedBatch:
    """Complexes (receptor dicts, one LigandPrep per complex, rec_idx) docked in HBM-admitted receptor groups.

    budget_bytes: device memory the grid sets of one group may take (a receptor larger than the budget gets a group of
    its own). Chain states are the caller's [L * chains, 7 + KM] (chains of complex l at rows l * chains ...), with KM
    from self.KM; chain ids are the row indices, so every chain's random stream is the same whatever the grouping."""

    def __init__(self, receptors, preps, rec_idx, budget_bytes, chains, sf="vina", device="cuda", queues=XCDS):
        self.receptors, self.preps = receptors, preps
        self.rec_idx = np.asarray(rec_idx)
        self.chains, self.sf, self.device, self.queues = int(chains), sf, torch.device(device), queues
        self.bytes = grid_bytes(receptors, preps, self.rec_idx)
        self.groups = admit(self.bytes, float(budget_bytes))
        self.AM, self.KM = buckets_of(preps)
        self.L = len(preps)
        self.peak_grid_bytes = 0

    def group_complexes(self, g):
        return np.flatnonzero(np.isin(self.rec_idx, self.groups[g]))

    def run(self, cur, cur_e, best, best_e, init_opt, mc_steps, bfgs_iters, temp, amp_pos, amp_rot, seed, step_offset,
            log=None):
        """In place on the caller's chain buffers (device, float32). Returns per-group timings and, per complex, the
        latency (s) from the call's start to its last chain's end."""
        dev = self.device
        M = self.chains
        latency = np.zeros(self.L)
        stats = []
        t0 = time.perf_counter()
        for g, grp in enumerate(self.groups):
            cx = self.group_complexes(g)
            if len(cx) == 0:
                continue
            t_build = time.perf_counter()
            local = {int(r): i for i, r in enumerate(grp)}
            recs = [self.receptors[r] for r in grp]
            prob = workload.gpu_problem(recs, [self.preps[l] for l in cx], [local[int(self.rec_idx[l])] for l in cx],
                                        sf=self.sf, device=dev, with_rec=False)
            self.peak_grid_bytes = max(self.peak_grid_bytes, int(prob.gs.flat.numel()) * 4)
            kp = KernelProblem(prob, bucket=(self.AM, self.KM))
            rows = torch.as_tensor((cx[:, None] * M + np.arange(M)[None, :]).ravel(), device=dev)
            cl = torch.as_tensor(np.repeat(np.arange(len(cx), dtype=np.int32), M), device=dev)
            bufs = [x[rows].contiguous() for x in (cur, cur_e, best, best_e)]
            runner = BatchRunner(kp, cl, n_complexes=len(cx), queues=self.queues)
            torch.cuda.synchronize(dev)
            t_launch = time.perf_counter()
            runner.launch(rows.to(torch.int32), *bufs, init_opt, mc_steps, bfgs_iters, temp, amp_pos, amp_rot, seed,
                          step_offset)
            torch.cuda.synchronize(dev)
            t_end = time.perf_counter()
            for x, b in zip((cur, cur_e, best, best_e), bufs):
                x[rows] = b
            latency[cx] = (t_launch - t0) + runner.latencies()
            stats.append({"receptors": len(grp), "complexes": int(len(cx)), "grid_bytes": int(prob.gs.flat.numel()) * 4,
                          "build_s": t_launch - t_build, "search_s": t_end - t_launch})
            if log:
                log(f"group {g}: {len(grp)} receptors, {len(cx)} complexes, "
                    f"{stats[-1]['grid_bytes'] / 1e6:.1f} MB grids, build {stats[-1]['build_s']:.2f} s, "
                    f"search {stats[-1]['search_s']:.2f} s")
            del runner, kp, prob, bufs
        return {"groups": stats, "latency_s": latency}
