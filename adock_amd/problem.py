"""Containers the kernel facade reads: ligand batch, pair potential, problem.

Attribute names match the reference's LigandBatch / PairPotential / Problem (src/adock/ligand.py:125-182,
src/adock/scoring.py:34-61, src/adock/search.py:65-71) for everything KernelProblem touches (`prob.lb.preps`, `prob.gs`,
`prob.pot.p`, `prob.rec`, `prob.rec_idx`), so a reference Problem object can be handed to KernelProblem unchanged.
"""
import math

import numpy as np
import torch

from .scoring import pair_terms

K_BUCKET = 8


class LigandBatch:
    """Padded tensors of L ligand preps (one per complex), the torsion count bucketed to a multiple of 8
    (src/adock/ligand.py:125-182): ref [L,A,3], types [L,A] (-1 = hydrogen / padding), heavy, amask, axis [L,K,2],
    tmask [L,K,A], tvalid [L,K], pairs [L,P,2], pmask [L,P], rg, n_tors_vina, axis_p / axis_u [L,K,3]."""

    def __init__(self, preps, device):
        self.preps = preps
        L = len(preps)
        A = max(p.n for p in preps)
        K = max(max(p.n_tors for p in preps), 1)
        K = K_BUCKET * math.ceil(K / K_BUCKET)
        P = max(max(len(p.pairs) for p in preps), 1)
        ref = np.zeros((L, A, 3), np.float32)
        types = -np.ones((L, A), np.int64)
        amask = np.zeros((L, A), bool)
        axis = np.zeros((L, K, 2), np.int64)
        tmask = np.zeros((L, K, A), bool)
        tvalid = np.zeros((L, K), bool)
        pairs = np.zeros((L, P, 2), np.int64)
        pmask = np.zeros((L, P), bool)
        for l, p in enumerate(preps):
            ref[l, :p.n] = p.ref
            types[l, :p.n] = p.types
            amask[l, :p.n] = True
            axis[l, :p.n_tors] = p.tors_axis
            tmask[l, :p.n_tors, :p.n] = p.tors_mask
            tvalid[l, :p.n_tors] = True
            pairs[l, :len(p.pairs)] = p.pairs
            pmask[l, :len(p.pairs)] = True
        T = lambda a: torch.as_tensor(a, device=device)
        self.L, self.A, self.K, self.P = L, A, K, P
        self.ref, self.types, self.amask = T(ref), T(types), T(amask)
        self.heavy = self.types >= 0
        self.axis, self.tmask, self.tvalid = T(axis), T(tmask), T(tvalid)
        self.pairs, self.pmask = T(pairs), T(pmask)
        self.rg = T(np.array([p.rg for p in preps], np.float32))
        self.n_tors_vina = T(np.array([p.n_tors_vina for p in preps], np.float32))
        self.heavy_count = self.heavy.sum(1)
        li = torch.arange(L, device=device)[:, None]
        pa, pb = self.ref[li, self.axis[..., 0]], self.ref[li, self.axis[..., 1]]
        u = pb - pa
        self.axis_p = pa
        self.axis_u = u / u.norm(dim=-1, keepdim=True).clamp(min=1e-6)

    def subset(self, idx):
        new = LigandBatch.__new__(LigandBatch)
        new.preps = [self.preps[i] for i in idx.tolist()]
        new.A, new.K, new.P = self.A, self.K, self.P
        new.L = len(idx)
        for k in ("ref", "types", "heavy", "amask", "axis", "tmask", "tvalid", "pairs", "pmask", "rg", "n_tors_vina",
                  "heavy_count", "axis_p", "axis_u"):
            setattr(new, k, getattr(self, k)[idx])
        return new


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
    def __init__(self, gridset, lb, rec_idx, pot, rec=None):
        self.gs, self.lb, self.rec_idx, self.pot = gridset, lb, rec_idx, pot
        self.rec = rec   # optional (xyz [R,N,3], types [R,N], mask [R,N]) for the exact (grid-free) energy

    def subset(self, idx):
        return Problem(self.gs, self.lb.subset(idx), self.rec_idx[idx], self.pot, self.rec)

    def inter(self, X):
        """Grid inter energy + out-of-box penalty [L, M] of heavy atoms X [L, M, A, 3] (src/adock/search.py:91-97)."""
        lb = self.lb
        L, M, A, _ = X.shape
        t = lb.types[:, None, :].expand(L, M, A)
        ri = self.rec_idx[:, None, None].expand(L, M, A)
        return self.gs.lookup(X, t, ri, mask=lb.heavy[:, None, :].expand(L, M, A)).sum(-1)

    def inter_nopen(self, X):
        """Grid inter energy without the out-of-box penalty (src/adock/search.py:99-105)."""
        pen, self.gs.penalty = self.gs.penalty, 0.0
        try:
            return self.inter(X)
        finally:
            self.gs.penalty = pen

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
        """Intra-ligand energy [L, M] of coordinates X [L, M, A, 3] (src/adock/search.py:108-119)."""
        lb = self.lb
        L, M, A, _ = X.shape
        pi = lb.pairs[:, None, :, :].expand(L, M, lb.P, 2)
        xi = torch.gather(X, 2, pi[..., 0:1].expand(L, M, lb.P, 3))
        xj = torch.gather(X, 2, pi[..., 1:2].expand(L, M, lb.P, 3))
        r = (xi - xj).norm(dim=-1)
        ti = torch.gather(lb.types, 1, lb.pairs[..., 0])[:, None].expand(L, M, lb.P)
        tj = torch.gather(lb.types, 1, lb.pairs[..., 1])[:, None].expand(L, M, lb.P)
        e = self.pot.energy(r, ti.clamp(min=0), tj.clamp(min=0))
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
