"""Containers the kernel facade reads: ligand batch, pair potential, problem.

Attribute names match the reference's LigandBatch / PairPotential / Problem (src/adock/ligand.py:125-182,
src/adock/scoring.py:34-61, src/adock/search.py:65-71) for everything KernelProblem touches (`prob.lb.preps`, `prob.gs`,
`prob.pot.p`, `prob.rec`, `prob.rec_idx`), so a reference Problem object can be handed to KernelProblem unchanged.
"""
import math

import numpy as np
import torch

from .grid import trilinear
from .scoring import pair_terms

K_BUCKET = 8


class LigandBatch:
    """L ligands (one per complex) as padded tensors (the attribute set of the reference's LigandBatch,
    src/adock/ligand.py:125-182): ref [L,A,3], types [L,A] (-1 = hydrogen or padding), heavy, amask, axis [L,K,2],
    tmask [L,K,A], tvalid [L,K], pairs [L,P,2], pmask [L,P], rg, n_tors_vina, heavy_count, axis_p / axis_u [L,K,3]
    (torsion axis point and unit vector in the ligand frame). K is the largest torsion count rounded up to a multiple
    of 8; A and P are the largest atom and pair counts."""

    # per-ligand tensors (first axis = ligand): what subset() slices
    FIELDS = ("ref", "types", "heavy", "amask", "axis", "tmask", "tvalid", "pairs", "pmask", "rg", "n_tors_vina",
              "heavy_count", "axis_p", "axis_u")

    def __init__(self, preps, device):
        self.preps = list(preps)
        self.L = len(self.preps)
        self.A = max(p.n for p in self.preps)
        self.K = K_BUCKET * math.ceil(max(1, max(p.n_tors for p in self.preps)) / K_BUCKET)
        self.P = max(1, max(len(p.pairs) for p in self.preps))
        dev = torch.device(device)

        def stack(get, tail, fill, dtype):
            out = np.full((self.L,) + tail, fill, dtype)
            for l, p in enumerate(self.preps):
                v = np.asarray(get(p))
                out[(l,) + tuple(slice(0, n) for n in v.shape)] = v
            return torch.as_tensor(out, device=dev)

        A, K, P = self.A, self.K, self.P
        self.ref = stack(lambda p: p.ref, (A, 3), 0.0, np.float32)
        self.types = stack(lambda p: p.types, (A,), -1, np.int64)
        self.amask = stack(lambda p: np.ones(p.n, bool), (A,), False, bool)
        self.axis = stack(lambda p: p.tors_axis.reshape(-1, 2), (K, 2), 0, np.int64)
        self.tmask = stack(lambda p: p.tors_mask.reshape(-1, p.n), (K, A), False, bool)
        self.tvalid = stack(lambda p: np.ones(p.n_tors, bool), (K,), False, bool)
        self.pairs = stack(lambda p: p.pairs.reshape(-1, 2), (P, 2), 0, np.int64)
        self.pmask = stack(lambda p: np.ones(len(p.pairs), bool), (P,), False, bool)
        self.rg = torch.as_tensor(np.array([p.rg for p in self.preps], np.float32), device=dev)
        self.n_tors_vina = torch.as_tensor(np.array([p.n_tors_vina for p in self.preps], np.float32), device=dev)
        self.heavy = self.types >= 0
        self.heavy_count = self.heavy.sum(1)
        ends = self.ref.gather(1, self.axis.reshape(self.L, 2 * K, 1).expand(-1, -1, 3)).view(self.L, K, 2, 3)
        self.axis_p = ends[:, :, 0]
        axis = ends[:, :, 1] - ends[:, :, 0]
        self.axis_u = axis / axis.norm(dim=-1, keepdim=True).clamp(min=1e-6)

    def subset(self, idx):
        """The ligands idx (a long tensor) as a new batch with the same padded sizes A, K, P."""
        sub = LigandBatch.__new__(LigandBatch)
        sub.preps = [self.preps[i] for i in idx.tolist()]
        sub.L, sub.A, sub.K, sub.P = len(sub.preps), self.A, self.K, self.P
        for name in self.FIELDS:
            setattr(sub, name, getattr(self, name)[idx])
        return sub


class PairPotential:
    """Scoring-function parameters (sf_params dict) and the per-type tables as tensors
    (src/adock/scoring.py:34-61)."""

    def __init__(self, sfp, device="cpu"):
        self.p = sfp
        self.device = device
        self.radius = torch.tensor(sfp["radius"], dtype=torch.float32, device=device)
        self.hyd = torch.tensor(sfp["hydrophobic"], device=device).bool()
        self.don = torch.tensor(sfp["donor"], device=device).bool()
        self.acc = torch.tensor(sfp["acceptor"], device=device).bool()

    def prm(self):
        p = self.p
        return (p["g1_off"], p["g1_w"], p["g2_off"], p["g2_w"], p["rep_off"], p["hyd_good"], p["hyd_bad"],
                p["hb_good"], p["hb_bad"], p["cutoff"])

    def terms(self, r, ti, tj):
        return pair_terms(r, ti, tj, self.radius, self.hyd, self.don, self.acc, self.prm())

    def weights(self, device=None):
        p = self.p
        return torch.tensor([p["w_g1"], p["w_g2"], p["w_rep"], p["w_hyd"], p["w_hb"]],
                            device=device or self.radius.device)

    def energy(self, r, ti, tj):
        return (self.terms(r, ti, tj) * self.weights(r.device)).sum(-1)


class Problem:
    """A docking batch: grid set `gs`, ligand batch `lb`, receptor index of each complex `rec_idx`, pair potential
    `pot`, and optionally the padded receptor atoms `rec` = (xyz [R,N,3], types [R,N], mask [R,N]) for exact
    (grid-free) energies. Attribute names are the reference's (src/adock/search.py:65-71)."""

    def __init__(self, gridset, lb, rec_idx, pot, rec=None):
        self.gs, self.lb, self.rec_idx, self.pot, self.rec = gridset, lb, rec_idx, pot, rec

    def subset(self, idx):
        """The complexes idx (long tensor), sharing grids, potential and receptors."""
        return Problem(self.gs, self.lb.subset(idx), self.rec_idx[idx], self.pot, self.rec)

    def _atom_grid_energy(self, X, penalty):
        """Per heavy atom of poses X [L, M, A, 3]: trilinear grid energy (+ out-of-box penalty when `penalty`), zero on
        hydrogens and padding."""
        lb, gs = self.lb, self.gs
        L, M, A = X.shape[:3]
        rows = (L, M, A)
        t = lb.types[:, None, :].expand(rows)
        ri = self.rec_idx[:, None, None].expand(rows)
        return trilinear(gs.flat, gs.origin, gs.lo, gs.hi, X, t, ri, lb.heavy[:, None, :].expand(rows), gs.dims,
                         gs.offs, gs.slot, gs.spacing, gs.penalty if penalty else 0.0)

    def inter(self, X):
        """Grid inter energy + out-of-box penalty [L, M] of poses X [L, M, A, 3] (src/adock/search.py:91-97)."""
        return self._atom_grid_energy(X, True).sum(-1)

    def inter_nopen(self, X):
        """Grid inter energy without the out-of-box penalty [L, M] (src/adock/search.py:99-105)."""
        return self._atom_grid_energy(X, False).sum(-1)

    def inter_exact(self, X, chunk=20_000_000):
        """Direct pairwise ligand-receptor energy [L, M], no grid (src/adock/search.py:73-89). For cuda poses and a
        potential with its own exact kernel (pot.inter_exact: the fitted function) that kernel; otherwise through
        pot.energy in torch, over groups of complexes (or of poses of one complex) at a time, so the [complexes, poses,
        atoms, receptor atoms] distance block stays near `chunk` elements."""
        if self.rec is None:
            raise ValueError("inter_exact needs receptor atoms (Problem.rec)")
        kernel = getattr(self.pot, "inter_exact", None)
        if kernel is not None and X.is_cuda:
            # a potential with its own exact kernel (FittedPotential: adock_inter_exact_fit)
            return kernel(X, self.lb.types, self.rec_idx, self.rec).to(X.dtype)
        rx, rt, rm = self.rec
        lb = self.lb
        L, M, A = X.shape[:3]
        dev = X.device
        rx, rt, rm = rx.to(dev), rt.to(dev), rm.to(dev)
        # padding slots get a real receptor type (they are masked out; the potential only sees valid pair types)
        rt = torch.where(rm, rt, rt[rm].max() if bool(rm.any()) else rt.new_zeros(()))
        NR = rx.shape[1]
        out = torch.zeros(L, M, device=dev, dtype=X.dtype)
        per_l = M * A * NR
        lstep = max(1, chunk // max(1, per_l))
        mstep = M if per_l <= chunk else max(1, chunk // max(1, A * NR))
        for l0 in range(0, L, lstep):
            ls = slice(l0, min(L, l0 + lstep))
            ri = self.rec_idx[ls].to(dev)
            xr = rx[ri].to(X.dtype)
            n = xr.shape[0]
            mask = lb.heavy[ls][:, None, :, None] & rm[ri][:, None, None, :]
            tl = lb.types[ls].clamp(min=0)[:, None, :, None]
            tr = rt[ri][:, None, None, :]
            for m0 in range(0, M, mstep):
                ms = slice(m0, min(M, m0 + mstep))
                Xs = X[ls, ms]
                m = Xs.shape[1]
                r = torch.cdist(Xs.reshape(n, m * A, 3), xr).view(n, m, A, NR)
                e = self.pot.energy(r, tl.expand(n, m, A, NR), tr.expand(n, m, A, NR))
                out[ls, ms] = torch.where(mask, e, torch.zeros_like(e)).sum((-1, -2))
        return out

    def energy(self, t, q0, r, theta):
        """(total, inter, intra, X) of poses [L, M] from the grid and the pair potential (the uncompiled branch of
        src/adock/search.py:123-133)."""
        from .pose import pose_coords
        X = pose_coords(self.lb, t, q0, r, theta)
        ei, ea = self.inter(X), self.intra(X)
        return ei + ea, ei, ea, X

    def intra(self, X):
        """Intra-ligand energy [L, M] of poses X [L, M, A, 3]: the pair potential over each ligand's pair list
        (src/adock/search.py:108-119)."""
        lb = self.lb
        L, M = X.shape[:2]
        P = lb.pairs.shape[1]
        ends = lb.pairs.reshape(L, 1, 2 * P, 1).expand(L, M, 2 * P, 3)
        Y = X.gather(2, ends).view(L, M, P, 2, 3)
        r = (Y[..., 0, :] - Y[..., 1, :]).norm(dim=-1)
        tt = lb.types.gather(1, lb.pairs.reshape(L, 2 * P)).view(L, 1, P, 2).clamp(min=0).expand(L, M, P, 2)
        e = self.pot.energy(r, tt[..., 0], tt[..., 1])
        return torch.where(lb.pmask[:, None, :], e, torch.zeros_like(e)).sum(-1)


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
