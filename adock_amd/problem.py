"""Containers the kernel facade reads: ligand batch, pair potential, problem.

Attribute names match the reference's LigandBatch / PairPotential / Problem (src/adock/ligand.py:125-182,
src/adock/scoring.py:34-61, src/adock/search.py:65-71) for everything KernelProblem touches (`prob.lb.preps`, `prob.gs`,
`prob.pot.p`, `prob.rec`, `prob.rec_idx`), so a reference Problem object can be handed to KernelProblem unchanged.
"""
import math

import numpy as np
import torch

K_BUCKET = 8


class LigandBatch:
    """L ligand preps (one per complex) with the torsion count bucketed to a multiple of 8."""

    def __init__(self, preps, device):
        self.preps = preps
        self.L = len(preps)
        K = max(max(p.n_tors for p in preps), 1)
        self.K = K_BUCKET * math.ceil(K / K_BUCKET)
        tv = np.zeros((self.L, self.K), bool)
        for l, p in enumerate(preps):
            tv[l, :p.n_tors] = True
        self.tvalid = torch.as_tensor(tv, device=device)
        self.n_tors_vina = torch.tensor([float(p.n_tors_vina) for p in preps], device=device)
        self.heavy_count = torch.tensor([int(p.heavy.sum()) for p in preps], device=device)


class PairPotential:
    def __init__(self, sfp, device="cpu"):
        self.p = sfp
        self.device = device


class Problem:
    def __init__(self, gridset, lb, rec_idx, pot, rec=None):
        self.gs, self.lb, self.rec_idx, self.pot = gridset, lb, rec_idx, pot
        self.rec = rec   # optional (xyz [R,N,3], types [R,N], mask [R,N]) for the exact (grid-free) energy


def pad_receptors(receptors, device):
    """(xyz [R,N,3], types [R,N], mask [R,N]) padded to the largest receptor (src/adock/search.py:369-379)."""
    N = max(len(r["rec_xyz"]) for r in receptors)
    R = len(receptors)
    xyz = np.zeros((R, N, 3), np.float32)
    typ = np.zeros((R, N), np.int64)
    msk = np.zeros((R, N), bool)
    for i, r in enumerate(receptors):
        n = len(r["rec_xyz"])
        xyz[i, :n], typ[i, :n], msk[i, :n] = r["rec_xyz"], r["rec_types"], True
    T = lambda a: torch.as_tensor(a, device=device)
    return T(xyz), T(typ), T(msk)
