"""Per-ligand-type scoring grids of the fitted function, built by the HIP kernel adock_build_grid_fit.

Drop-in for the reference's src/adock/fit/gridfit.py: build_grid_fit (one receptor, one cube of n^3 points per
compact ligand type) and make_gridset_fit (a GridSet over several complexes, common cube edge). Grid values are the
same basis-times-pair-weight sums as FittedPotential.energy at the grid points.
"""
import math

import numpy as np
import torch

from .. import _lib
from ..grid import GridSet

MAX_RT = 216      # receptor ext types the kernel's LDS table holds


def build_grid_fit(rec_xyz, rec_ext, pot, origin, spacing, n, device="cuda"):
    """Tensor [TL, nx, ny, nz] of the fitted inter energy for the compact ligand types of `pot` (a FittedPotential)
    (src/adock/fit/gridfit.py:70-93). n: int (cube) or (nx, ny, nz)."""
    if str(device).startswith("cpu"):
        raise _lib.AdockError("adock_amd builds grids on the GPU only (device='cuda')")
    _lib.require_gpu()
    nx, ny, nz = (int(n),) * 3 if np.isscalar(n) else map(int, n)
    spec = pot.spec
    rext = np.ascontiguousarray(rec_ext, np.int32)
    rx = np.ascontiguousarray(rec_xyz, np.float32).reshape(-1, 3)
    if len(rext) != len(rx):
        raise _lib.AdockError("rec_xyz and rec_ext differ in length")
    if tuple(pot.W.shape) != (spec.n_rt, pot.TL, 6):
        raise _lib.AdockError("pot.W must be [n_rt, TL, 6]")
    if spec.n_rt > MAX_RT:
        raise _lib.AdockError(f"at most {MAX_RT} receptor ext types")
    if len(rext) and (rext.min() < 0 or rext.max() >= spec.n_rt):
        raise _lib.AdockError(f"receptor ext types must lie in [0, {spec.n_rt})")
    dev = torch.device(device)
    f32 = lambda a: torch.as_tensor(np.ascontiguousarray(a, np.float32), device=dev)
    keep = [f32(rx), torch.as_tensor(rext, device=dev), f32(spec.rec_radius),
            pot.radius.to(dev, torch.float32).contiguous(), pot.W.to(dev, torch.float32).contiguous(),
            f32(spec.grid_prm())]
    TL = pot.TL
    out = torch.empty((TL * nx * ny * nz,), dtype=torch.float32, device=dev)
    with torch.cuda.device(dev):
        st = _lib.lib().adock_build_grid_fit(
            _lib.ptr(keep[0]), _lib.ptr(keep[1]), len(rx), _lib.ptr(keep[2]), _lib.ptr(keep[3]), TL,
            _lib.ptr(keep[4]), spec.n_rt, float(origin[0]), float(origin[1]), float(origin[2]), float(spacing), nx, ny,
            nz, _lib.ptr(keep[5]), _lib.ptr(out), _lib.stream_of(dev))
    _lib.check(st)
    return out.view(TL, nx, ny, nz)


def make_gridset_fit(preps, pot, spacing, penalty, device="cuda"):
    """GridSet of the complexes `preps` (dicts with rec_xyz, rec_ext, center, size): one [TL, n, n, n] cube each with
    the common edge n = ceil(max box edge / spacing) + 1, origin = centre - size / 2, box = centre -+ size / 2
    (src/adock/fit/gridfit.py:96-108)."""
    n = int(math.ceil(max(float(np.max(p["size"])) for p in preps) / spacing)) + 1
    grids, origins, lo, hi = [], [], [], []
    for p in preps:
        c = np.asarray(p["center"], np.float64)
        s = np.broadcast_to(np.asarray(p["size"], np.float64), (3,))
        grids.append(build_grid_fit(p["rec_xyz"], p["rec_ext"], pot, c - s / 2, spacing, n, device=device))
        origins.append(c - s / 2)
        lo.append(c - s / 2)
        hi.append(c + s / 2)
    T = lambda a: torch.as_tensor(np.array(a), dtype=torch.float32, device=device)
    return GridSet(torch.stack(grids), T(origins), T(lo), T(hi), spacing, penalty)
