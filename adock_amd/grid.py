"""Receptor scoring grids on the GPU.

Public names and data contract follow the reference's grid module (src/adock/grid.py:132-282): build_grid,
GridSet (attributes flat / origin / lo / hi / offs / dims / slot / T / spacing / penalty, and G for a common cube),
make_gridset and trilinear, with the same grid layout ([T_r, nx, ny, nz] per receptor, z fastest, receptors
concatenated in one flat buffer), box rule (origin = centre - size / 2, n_d = ceil(size_d / spacing) + 1) and type
rule (a receptor's grids cover its ligands' types plus C_H).

How it is built here: make_gridset first plans the whole set (per receptor: type slots, point counts, offset into
the flat buffer), allocates the flat buffer once, and a GridSetBuilder - which keeps every receptor's atoms and the
scoring tables resident on the device - launches the HIP builder (adock_build_grid) straight into each receptor's
slice. Rebuilding a set (a new batch of the same receptors, or every step of a benchmark) is then launches only: no
host-to-device copies and no allocation, so it can be captured in a HIP graph.
"""
import math

import numpy as np
import torch

from . import _lib
from .config import grid_prm

MAX_TYPES = 24          # ligand types per adock_build_grid launch
C_H = 0                 # XS index of C_H: always built (the Stage A probe type)


def box_points(size, spacing):
    """Grid points per axis of a box of edge `size` (scalar or 3 edges): ceil(edge / spacing) + 1."""
    return tuple(int(math.ceil(float(e) / spacing)) + 1 for e in np.broadcast_to(np.asarray(size, float), (3,)))


class _Tables:
    """Scoring-function tables of one sfp dict, on one device (radius, type flags, the 15-float grid-order vector)."""

    def __init__(self, sfp, dev):
        f32 = lambda a: torch.as_tensor(np.ascontiguousarray(a, np.float32), device=dev)
        i32 = lambda a: torch.as_tensor(np.ascontiguousarray(a, np.int32), device=dev)
        self.radius, self.prm = f32(sfp["radius"]), f32(grid_prm(sfp))
        self.hyd, self.don, self.acc = i32(sfp["hydrophobic"]), i32(sfp["donor"]), i32(sfp["acceptor"])
        self.n_types = len(sfp["radius"])


def _launch(rx, rt, tab, types, origin, spacing, dims, per_term, out_ptr, dev):
    """One adock_build_grid call: receptor rx [N,3] / rt [N] and ligand types (device int32) -> out_ptr."""
    nx, ny, nz = dims
    st = _lib.lib().adock_build_grid(
        _lib.ptr(rx), _lib.ptr(rt), int(rx.shape[0]), _lib.ptr(tab.radius), _lib.ptr(tab.hyd), _lib.ptr(tab.don),
        _lib.ptr(tab.acc), _lib.ptr(types), int(types.numel()), float(origin[0]), float(origin[1]), float(origin[2]),
        float(spacing), int(nx), int(ny), int(nz), _lib.ptr(tab.prm), int(per_term), out_ptr, _lib.stream_of(dev))
    _lib.check(st)


def build_grid(rec_xyz, rec_types, sfp, origin, spacing, n, lig_types=None, per_term=False, device="cuda"):
    """Tensor [T, nx, ny, nz] (or [5, T, nx, ny, nz] with per_term: the five unweighted terms) of receptor potentials
    for the ligand types `lig_types` (default: every XS type), built by the HIP kernel (reference:
    src/adock/grid.py:132-166). n: points per axis, an int or (nx, ny, nz)."""
    if str(device).startswith("cpu"):
        raise _lib.AdockError("adock_amd builds grids on the GPU only (device='cuda')")
    _lib.require_gpu()
    dims = (int(n),) * 3 if np.isscalar(n) else tuple(int(v) for v in n)
    dev = torch.device(device)
    tab = _Tables(sfp, dev)
    types = np.arange(tab.n_types) if lig_types is None else np.asarray(lig_types)
    if len(types) > MAX_TYPES:
        raise _lib.AdockError(f"at most {MAX_TYPES} ligand types per grid build")
    rx = torch.as_tensor(np.ascontiguousarray(np.asarray(rec_xyz, np.float32).reshape(-1, 3)), device=dev)
    rt = torch.as_tensor(np.ascontiguousarray(rec_types, np.int32), device=dev)
    ty = torch.as_tensor(np.ascontiguousarray(types, np.int32), device=dev)
    lead = (5, len(types)) if per_term else (len(types),)
    out = torch.empty(lead + dims, dtype=torch.float32, device=dev)
    with torch.cuda.device(dev):
        _launch(rx, rt, tab, ty, origin, spacing, dims, per_term, _lib.ptr(out), dev)
    # the argument tensors may be freed now: their memory is reused only by later work on this stream
    return out


# ------------------------------------------------------------------------------------------------ lookups (torch)
# the eight corners of a cell as (dx, dy, dz), in the order the weights below are formed
_CORNERS = np.array([(dx, dy, dz) for dx in (0, 1) for dy in (0, 1) for dz in (0, 1)], np.int64)


def _box_penalty(x, lo, hi, penalty):
    """penalty * sum_d (lo_d - x_d)_+^2 + (x_d - hi_d)_+^2 (the reference's quadratic out-of-box term)."""
    return penalty * (torch.relu(lo - x).square() + torch.relu(x - hi).square()).sum(-1)


def _type_base(offs, dims, slot, t, rec_idx):
    """Flat index of grid point (0, 0, 0) of type t's grid in receptor rec_idx's block (unbuilt types map to slot 0;
    callers mask those atoms out)."""
    d = dims[rec_idx]
    n3 = d[..., 0] * d[..., 1] * d[..., 2]
    return offs[rec_idx] + slot[rec_idx, t.clamp(min=0)].clamp(min=0) * n3, d


def trilinear(flat, origin, lo, hi, x, t, rec_idx, mask, dims, offs, slot, spacing: float, penalty: float):
    """Per-atom grid energy by trilinear interpolation plus the out-of-box penalty (torch, differentiable in x), with
    the kernels' clamping: cell coordinates are clamped to [0, n - 1 - 1e-4] (reference: src/adock/grid.py:169-194).
    x [..., 3], t / rec_idx [...]; atoms where mask is False get 0."""
    base, d = _type_base(offs, dims, slot, t, rec_idx)
    f = (x - origin[rec_idx]) / spacing
    f = torch.minimum(f.clamp(min=0.0), (d - 1).to(f.dtype) - 1e-4)
    cell = f.detach().floor()
    w1 = f - cell                                                     # weight of the upper corner per axis
    corners = torch.as_tensor(_CORNERS, device=x.device)             # [8, 3]
    idx3 = cell.long().unsqueeze(-2) + corners                        # [..., 8, 3]
    strides = torch.stack([d[..., 1] * d[..., 2], d[..., 2], torch.ones_like(d[..., 2])], -1).unsqueeze(-2)
    idx = base.unsqueeze(-1) + (idx3 * strides).sum(-1)               # [..., 8]
    cw = torch.where(corners.bool(), w1.unsqueeze(-2), 1.0 - w1.unsqueeze(-2)).prod(-1)   # [..., 8]
    e = (flat[idx] * cw).sum(-1) + _box_penalty(x, lo[rec_idx], hi[rec_idx], penalty)
    return e if mask is None else torch.where(mask, e, torch.zeros_like(e))


class GridSet:
    """The grids of R receptors with their boxes: one flat buffer of per-receptor blocks [T_r, nx, ny, nz]
    (src/adock/grid.py:197-247).

    grids: a stacked tensor [R, T, n, n, n] (common cube; kept as .G) or a list of per-receptor [T_r, nx, ny, nz]
    tensors. slot: optional [R, n_types] map from XS type to the receptor's grid slot (-1 = not built); without it
    every receptor holds every type in slot order."""

    def __init__(self, grids, origins, box_lo, box_hi, spacing, penalty, slot=None, n_types=None):
        dev = origins.device
        if torch.is_tensor(grids):
            self.G = grids
            blocks = [grids[r] for r in range(grids.shape[0])]
            self.flat = grids.reshape(-1)
        else:
            self.G = None
            blocks = list(grids)
            self.flat = torch.cat([b.reshape(-1) for b in blocks]) if len(blocks) > 1 else blocks[0].reshape(-1)
        shapes = np.array([tuple(b.shape[-3:]) for b in blocks], np.int64).reshape(-1, 3)
        sizes = np.array([b.numel() for b in blocks], np.int64)
        self._init_layout(shapes, np.concatenate([[0], np.cumsum(sizes)[:-1]]), blocks[0].shape[0], dev)
        self.origin, self.lo, self.hi = origins, box_lo, box_hi
        self.spacing, self.penalty = spacing, penalty
        if slot is None:
            self.slot = torch.arange(self.T, device=dev).expand(self.R, self.T).contiguous()
        else:
            self.slot, self.T = slot, n_types

    def _init_layout(self, shapes, offsets, n_lead, dev):
        self.R = len(shapes)
        self.T = int(n_lead)
        self.dims = torch.as_tensor(shapes, device=dev)
        self.offs = torch.as_tensor(offsets, device=dev)
        self.n = int(shapes.max()) if len(shapes) else 0
        self.host_dims, self.host_offs = shapes, offsets     # host copies: launches need no device round trip

    @classmethod
    def allocate(cls, shapes, type_counts, origins, box_lo, box_hi, spacing, penalty, slot, n_types):
        """An unfilled set: per receptor type_counts[r] grids of shapes[r] points in one flat buffer (builders write
        into it). origins / box_lo / box_hi [R, 3] device tensors, slot [R, n_types] device tensor."""
        self = cls.__new__(cls)
        shapes = np.asarray(shapes, np.int64).reshape(-1, 3)
        sizes = np.asarray(type_counts, np.int64) * shapes.prod(1)
        offsets = np.concatenate([[0], np.cumsum(sizes)[:-1]])
        dev = origins.device
        self.G = None
        self.flat = torch.empty(int(sizes.sum()), dtype=torch.float32, device=dev)
        self._init_layout(shapes, offsets, int(max(type_counts)) if len(type_counts) else 0, dev)
        self.origin, self.lo, self.hi = origins, box_lo, box_hi
        self.spacing, self.penalty = spacing, penalty
        self.slot, self.T = slot, n_types
        return self

    def lookup(self, x, t, rec_idx, mask=None):
        """Trilinear energy + out-of-box penalty per atom (x [..., 3], XS type t and receptor rec_idx [...])."""
        return trilinear(self.flat, self.origin, self.lo, self.hi, x, t, rec_idx, mask, self.dims, self.offs,
                         self.slot, self.spacing, self.penalty)

    def lookup_nearest(self, x, t, rec_idx):
        """Value at the nearest grid point (clamped into the grid) + out-of-box penalty, per atom (Stage A's probe)."""
        base, d = _type_base(self.offs, self.dims, self.slot, t, rec_idx)
        p = torch.minimum(((x - self.origin[rec_idx]) / self.spacing).round().clamp(min=0).long(), d - 1)
        idx = base + (p[..., 0] * d[..., 1] + p[..., 1]) * d[..., 2] + p[..., 2]
        return self.flat[idx] + _box_penalty(x, self.lo[rec_idx], self.hi[rec_idx], self.penalty)


class GridSetBuilder:
    """Keeps receptors' atoms and the scoring tables on the device and (re)builds a planned GridSet's blocks in place:
    one adock_build_grid launch per receptor and MAX_TYPES ligand types, each writing into its slice of gs.flat."""

    def __init__(self, receptors, sfp, spacing, types_per_receptor, device="cuda", per_term=False):
        dev = torch.device(device)
        self.device, self.spacing, self.per_term = dev, spacing, per_term
        self.tables = _Tables(sfp, dev)
        self.rec = []
        for r, types in zip(receptors, types_per_receptor):
            rx = torch.as_tensor(np.ascontiguousarray(np.asarray(r["rec_xyz"], np.float32).reshape(-1, 3)), device=dev)
            rt = torch.as_tensor(np.ascontiguousarray(r["rec_types"], np.int32), device=dev)
            chunks = [torch.as_tensor(np.ascontiguousarray(types[i:i + MAX_TYPES], np.int32), device=dev)
                      for i in range(0, len(types), MAX_TYPES)]
            self.rec.append((rx, rt, chunks))

    def build_into(self, gs, origins):
        """Launch every receptor's builds into gs.flat on the current stream (asynchronous, capturable).
        origins: host [R, 3] grid origins (the same values gs.origin holds)."""
        _lib.require_gpu()
        base = gs.flat.data_ptr()
        with torch.cuda.device(self.device):
            for r, (rx, rt, chunks) in enumerate(self.rec):
                dims = tuple(int(v) for v in gs.host_dims[r])
                n3 = int(np.prod(dims))
                first = 0
                for ty in chunks:
                    if self.per_term:
                        raise _lib.AdockError("per-term grids are built whole by build_grid, not in place")
                    _launch(rx, rt, self.tables, ty, origins[r], self.spacing, dims, False,
                            base + 4 * (int(gs.host_offs[r]) + first * n3), self.device)
                    first += int(ty.numel())


def make_gridset(receptors, sfp, spacing, penalty, device="cuda", per_term=False, cubic=None, lig_types=None,
                 keep_builder=False):
    """GridSet of `receptors` (dicts with rec_xyz, rec_types, center, size; src/adock/grid.py:250-282).
    cubic (default: per_term): one common cube edge for all receptors, every XS type, stacked as .G; otherwise each
    receptor gets its own box_points and, when lig_types gives its ligands' XS types, only those plus C_H.
    keep_builder: attach the GridSetBuilder as gs.builder (gs.rebuild() then refills the set in place)."""
    cubic = per_term if cubic is None else cubic
    dev = torch.device(device)
    n_all = len(sfp["radius"])
    R = len(receptors)
    if cubic or lig_types is None:
        types = [np.arange(n_all)] * R
    else:
        types = [np.union1d([C_H], np.asarray(lt, np.int64)) for lt in lig_types]
    centres = np.array([np.asarray(r["center"], float) for r in receptors]).reshape(R, 3)
    sizes = np.array([np.broadcast_to(np.asarray(r["size"], float), (3,)) for r in receptors]).reshape(R, 3)
    origins = centres - sizes / 2.0
    f32 = lambda a: torch.as_tensor(np.asarray(a, np.float32), device=dev)
    if cubic:
        edge = box_points(sizes.max(), spacing)[0]
        grids = [build_grid(r["rec_xyz"], r["rec_types"], sfp, o, spacing, edge, per_term=per_term, device=dev)
                 for r, o in zip(receptors, origins)]
        return GridSet(torch.stack(grids), f32(origins), f32(origins), f32(centres + sizes / 2.0), spacing, penalty)
    slot = -np.ones((R, n_all), np.int64)
    for r, ts in enumerate(types):
        slot[r, ts] = np.arange(len(ts))
    shapes = [box_points(s, spacing) for s in sizes]
    gs = GridSet.allocate(shapes, [len(ts) for ts in types], f32(origins), f32(origins), f32(centres + sizes / 2.0),
                          spacing, penalty, torch.as_tensor(slot, device=dev), n_all)
    builder = GridSetBuilder(receptors, sfp, spacing, types, device=dev)
    builder.build_into(gs, origins)
    if keep_builder:
        gs.builder = builder
        gs.rebuild = lambda: builder.build_into(gs, origins)
    return gs
