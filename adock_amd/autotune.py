"""Launch-shape autotuner for the chain kernels.

Chooses the wave's pair-gradient LDS tile (compact: one slot per intra pair with the per-atom lists read from memory;
dense: per-atom rows of DEG_MAX float4 entries with no list indirection, more LDS per wave), chains
(wavefronts) per workgroup W in {1, 2, 4} and the persistent grid size (resident capacity x {1, 2}, or one workgroup
per W chains) for a given batch by timing short probes of the real workload (the same chains, a few MC steps, on
scratch copies of the states) and keeps the fastest. Every candidate gives the same results bit for bit. Results are
cached per (KM, chain-count bucket, pair-list size bucket, slot-list length bucket, device name) for the process and
optionally in a JSON file.
"""
import json
import math
import os
import time

import torch

from . import _lib
from .batch import BatchRunner


def _key(kp, n):
    name = torch.cuda.get_device_name(kp.device).replace(" ", "_")
    return (f"{name}|KM{kp.KM}|N{2 ** int(math.log2(max(n, 1)))}|P{(kp.P_MAX + 255) // 256}"
            f"|D{(kp.DEG_MAX + 7) // 8}|A{kp.AM}")


class Autotuner:
    def __init__(self, cache_file=None):
        self.cache_file = cache_file or os.environ.get("ADOCK_TUNE_CACHE")
        self.cache = {}
        if self.cache_file and os.path.exists(self.cache_file):
            self.cache = json.loads(open(self.cache_file).read())

    def tune(self, runner, state, probe_steps=4, reps=2, candidates=None):
        """runner: BatchRunner; state: start states [N, SD] (device). Returns (W, blocks, lds_layout) and applies
        it."""
        kp = runner.kp
        key = _key(kp, runner.N)
        if key in self.cache:
            W, blocks, layout = self.cache[key]
            self.apply(runner, W, blocks, layout=layout)
            return W, blocks, layout
        cus = torch.cuda.get_device_properties(kp.device).multi_processor_count
        layouts = (0, 1) if kp.AM == 64 else (0,)   # the dense tile holds one atom per lane
        candidates = candidates or [(W, b, d) for d in layouts for W in (4, 2, 1) for b in (0, -2, -1)]
        best, best_t = None, float("inf")
        dev = kp.device
        ci = torch.arange(runner.N, dtype=torch.int32, device=dev)
        for W, b, d in candidates:
            self.apply(runner, W, b, cus, layout=d)
            times = []
            for r in range(reps + 1):
                cur, bst = state.clone(), state.clone()
                ce = torch.zeros(runner.N, device=dev)
                be = torch.full((runner.N,), float("inf"), device=dev)
                torch.cuda.synchronize(dev)
                t0 = time.perf_counter()
                runner.launch(ci, cur, ce, bst, be, True, probe_steps, 0, 1.2, 2.0, 2.0, 1234, 0)
                torch.cuda.synchronize(dev)
                if r:
                    times.append(time.perf_counter() - t0)
            t = min(times)
            if t < best_t:
                best, best_t = (W, b, d), t
        self.cache[key] = list(best)
        if self.cache_file:
            with open(self.cache_file, "w") as fh:
                json.dump(self.cache, fh, indent=1)
        self.apply(runner, best[0], best[1], cus, layout=best[2])
        return tuple(best)

    @staticmethod
    def apply(runner, W, blocks, cus=None, layout=0):
        """W: chains per workgroup of this runner's launches (the library's process-wide default is left alone);
        blocks: 0 = resident capacity (CUs x occupancy); -2 = twice that; -1 = one workgroup per W chains; layout: the
        pair-gradient LDS tile (0 compact, 1 dense)."""
        runner.waves = int(W)
        runner.lds_layout = int(layout)
        if blocks == -1:
            runner.blocks = (runner.N + W - 1) // W
        elif blocks == -2:
            cus = cus or torch.cuda.get_device_properties(runner.kp.device).multi_processor_count
            runner.blocks = 2 * _resident_blocks(runner, W, cus)
        else:
            runner.blocks = int(blocks)


def _resident_blocks(runner, W, cus):
    """Resident-capacity estimate: the library's own occupancy query runs when blocks == 0; here CUs x (8 waves per
    SIMD x 4 SIMDs / W) bounded by the chain count, a safe upper bound that the hardware trims."""
    return min(cus * max(1, 16 // W), (runner.N + W - 1) // W)
