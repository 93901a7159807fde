"""Receptor scoring grids on the GPU: build_grid (HIP kernel), GridSet (flat multi-receptor layout), make_gridset, and
the torch trilinear / nearest lookups.

Drop-in for the reference's src/adock/grid.py: same function names, argument meaning, layouts ([T, nx, ny, nz] per
receptor, z fastest; flat concatenation with offs / dims / slot) and box rule (origin = centre - size / 2, n_d =
ceil(size_d / spacing) + 1, types = ligand types + C_H).
"""
import math

import numpy as np
import torch

from . import _lib
from .config import grid_prm

MAX_TYPES = 24


def build_grid(rec_xyz, rec_types, sfp, origin, spacing, n, lig_types=None, per_term=False, device="cuda"):
    """Tensor [T, nx, ny, nz] (or [5, T, nx, ny, nz] if per_term) on `device`, built by the HIP grid kernel
    (reference: src/adock/grid.py:132-166). n: int or (nx, ny, nz); lig_types: XS indices (default all types)."""
    if str(device).startswith("cpu"):
        raise _lib.AdockError("adock_amd builds grids on the GPU only (device='cuda')")
    _lib.require_gpu()
    if np.isscalar(n):
        n = (int(n),) * 3
    nx, ny, nz = map(int, n)
    T_all = len(sfp["radius"])
    lig_types = np.arange(T_all, dtype=np.int32) if lig_types is None else np.asarray(lig_types, dtype=np.int32)
    if len(lig_types) > MAX_TYPES:
        raise _lib.AdockError(f"at most {MAX_TYPES} ligand types per grid build")
    dev = torch.device(device)
    f32 = lambda a: torch.as_tensor(np.ascontiguousarray(a, np.float32), device=dev)
    i32 = lambda a: torch.as_tensor(np.ascontiguousarray(a, np.int32), device=dev)
    rx = f32(np.asarray(rec_xyz, np.float32).reshape(-1, 3))
    rt = i32(rec_types)
    keep = [rx, rt, f32(sfp["radius"]), i32(sfp["hydrophobic"]), i32(sfp["donor"]), i32(sfp["acceptor"]),
            i32(lig_types), f32(grid_prm(sfp))]
    nlt = len(lig_types)
    out = torch.empty(((5 if per_term else 1) * nlt * nx * ny * nz,), dtype=torch.float32, device=dev)
    with torch.cuda.device(dev):
        st = _lib.lib().adock_build_grid(
            _lib.ptr(rx), _lib.ptr(rt), len(rx), _lib.ptr(keep[2]), _lib.ptr(keep[3]), _lib.ptr(keep[4]),
            _lib.ptr(keep[5]), _lib.ptr(keep[6]), nlt, float(origin[0]), float(origin[1]), float(origin[2]),
            float(spacing), nx, ny, nz, _lib.ptr(keep[7]), int(per_term), _lib.ptr(out), _lib.stream_of(dev))
    _lib.check(st)   # argument tensors may be freed now: their memory is reused only by later work on this stream
    shape = ((5, nlt) if per_term else (nlt,)) + (nx, ny, nz)
    return out.view(shape)


def trilinear(flat, origin, lo, hi, x, t, rec_idx, mask, dims, offs, slot, spacing: float, penalty: float):
    """Trilinear lookup into the flat per-receptor grids + quadratic out-of-box penalty (torch; differentiable in x).
    Same clamping as the kernels (src/adock/grid.py:169-194)."""
    o = origin[rec_idx]
    dm = dims[rec_idx]
    f = (x - o) / spacing
    fc = torch.minimum(f.clamp(min=0.0), (dm - 1).to(f.dtype) - 1e-4)
    i0 = fc.detach().floor().long()
    w = fc - i0
    ny, nz = dm[..., 1], dm[..., 2]
    base = offs[rec_idx] + slot[rec_idx, t.clamp(min=0)].clamp(min=0) * (dm[..., 0] * ny * nz)
    e = 0.0
    for dx in (0, 1):
        wx = w[..., 0] if dx else 1 - w[..., 0]
        for dy in (0, 1):
            wy = w[..., 1] if dy else 1 - w[..., 1]
            for dz in (0, 1):
                wz = w[..., 2] if dz else 1 - w[..., 2]
                idx = base + ((i0[..., 0] + dx) * ny + (i0[..., 1] + dy)) * nz + (i0[..., 2] + dz)
                e = e + wx * wy * wz * flat[idx]
    out = (lo[rec_idx] - x).clamp(min=0) ** 2 + (x - hi[rec_idx]).clamp(min=0) ** 2
    e = e + penalty * out.sum(-1)
    if mask is not None:
        e = torch.where(mask, e, torch.zeros_like(e))
    return e


class GridSet:
    """Receptor grids with per-receptor origin and box (src/adock/grid.py:197-247).

    grids: a stacked tensor [R, T, n, n, n] (common cube) or a list of [T_r, nx, ny, nz] tensors (per-receptor size).
    slot: optional [R, n_types] (global XS type -> grid slot of that receptor, -1 = not built)."""

    def __init__(self, grids, origins, box_lo, box_hi, spacing, penalty, slot=None, n_types=None):
        dev = origins.device
        if torch.is_tensor(grids):
            self.G = grids
            self.R, self.T, self.n = grids.shape[0], grids.shape[1], grids.shape[2]
            self.flat = grids.reshape(-1)
            sz = self.T * self.n ** 3
            self.dims = torch.full((self.R, 3), self.n, dtype=torch.long, device=dev)
            self.offs = torch.arange(self.R, device=dev, dtype=torch.long) * sz
        else:
            self.G = None
            self.R, self.T = len(grids), grids[0].shape[0]
            self.flat = torch.cat([g.reshape(-1) for g in grids])
            self.dims = torch.tensor([list(g.shape[1:]) for g in grids], dtype=torch.long, device=dev)
            sizes = torch.tensor([g.numel() for g in grids], dtype=torch.long, device=dev)
            self.offs = torch.cumsum(sizes, 0) - sizes
            self.n = int(self.dims.max())
        if slot is None:
            slot = torch.arange(self.T, device=dev).expand(self.R, self.T).contiguous()
        else:
            self.T = n_types
        self.slot = slot
        self.origin = origins
        self.lo, self.hi = box_lo, box_hi
        self.spacing = spacing
        self.penalty = penalty

    def lookup(self, x, t, rec_idx, mask=None):
        return trilinear(self.flat, self.origin, self.lo, self.hi, x, t, rec_idx, mask, self.dims, self.offs, self.slot,
                         self.spacing, self.penalty)

    def lookup_nearest(self, x, t, rec_idx):
        dm = self.dims[rec_idx]
        f = torch.minimum(((x - self.origin[rec_idx]) / self.spacing).round().clamp(min=0).long(), dm - 1)
        ny, nz = dm[..., 1], dm[..., 2]
        base = self.offs[rec_idx] + self.slot[rec_idx, t.clamp(min=0)].clamp(min=0) * (dm[..., 0] * ny * nz)
        e = self.flat[base + (f[..., 0] * ny + f[..., 1]) * nz + f[..., 2]]
        lo, hi = self.lo[rec_idx], self.hi[rec_idx]
        out = (lo - x).clamp(min=0) ** 2 + (x - hi).clamp(min=0) ** 2
        return e + self.penalty * out.sum(-1)


def make_gridset(receptors, sfp, spacing, penalty, device="cuda", per_term=False, cubic=None, lig_types=None):
    """receptors: dicts with rec_xyz, rec_types, center, size (src/adock/grid.py:250-282).
    cubic (default: per_term): common cube for all receptors; otherwise per-receptor n_d = ceil(size_d / h) + 1.
    lig_types (non-cubic): per receptor the XS types to build; C_H (type 0) is always added."""
    cubic = per_term if cubic is None else cubic
    T_all = len(sfp["radius"])
    if lig_types is not None and not cubic:
        sel = [np.unique(np.concatenate([[0], np.asarray(lt, dtype=np.int64)])) for lt in lig_types]
    else:
        sel = [np.arange(T_all)] * len(receptors)
    slot = -np.ones((len(receptors), T_all), np.int64)
    for i, st in enumerate(sel):
        slot[i, st] = np.arange(len(st))
    n = int(math.ceil(max(max(np.broadcast_to(r["size"], (3,))) for r in receptors) / spacing)) + 1
    grids, origins, lo, hi = [], [], [], []
    for r, st in zip(receptors, sel):
        c, s = np.asarray(r["center"], float), np.broadcast_to(np.asarray(r["size"], float), (3,))
        origin = c - s / 2.0
        nn = n if cubic else tuple(int(math.ceil(x / spacing)) + 1 for x in s)
        grids.append(build_grid(r["rec_xyz"], r["rec_types"], sfp, origin, spacing, nn, per_term=per_term,
                                device=device, lig_types=None if cubic else st))
        origins.append(origin)
        lo.append(c - s / 2.0)
        hi.append(c + s / 2.0)
    tt = lambda a: torch.tensor(np.array(a), dtype=torch.float32, device=device)
    if cubic:
        return GridSet(torch.stack(grids), tt(origins), tt(lo), tt(hi), spacing, penalty)
    return GridSet(grids, tt(origins), tt(lo), tt(hi), spacing, penalty, slot=torch.as_tensor(slot, device=device),
                   n_types=T_all)
