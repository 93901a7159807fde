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
    (src/adock/fit/gridfit.py:96-108). All complexes in one adock_build_grid_fit_sets launch per 8 ligand types
    (receptors padded with ext type -1, which the kernel skips: each cube equals build_grid_fit's bit for bit)."""
    if str(device).startswith("cpu"):
        raise _lib.AdockError("adock_amd builds grids on the GPU only (device='cuda')")
    _lib.require_gpu()
    spec = pot.spec
    n = int(math.ceil(max(float(np.max(p["size"])) for p in preps) / spacing)) + 1
    R = len(preps)
    NR = max(max(len(p["rec_ext"]) for p in preps), 1)
    rx = np.zeros((R, NR, 3), np.float32)
    re = -np.ones((R, NR), np.int32)
    origins, lo, hi = [], [], []
    for i, p in enumerate(preps):
        e = np.asarray(p["rec_ext"], np.int32)
        x = np.asarray(p["rec_xyz"], np.float32).reshape(-1, 3)
        if len(e) != len(x):
            raise _lib.AdockError("rec_xyz and rec_ext differ in length")
        if len(e) and (e.min() < 0 or e.max() >= spec.n_rt):
            raise _lib.AdockError(f"receptor ext types must lie in [0, {spec.n_rt})")
        rx[i, :len(x)], re[i, :len(e)] = x, e
        c = np.asarray(p["center"], np.float64)
        sz = np.broadcast_to(np.asarray(p["size"], np.float64), (3,))
        origins.append(c - sz / 2)
        lo.append(c - sz / 2)
        hi.append(c + sz / 2)
    if tuple(pot.W.shape) != (spec.n_rt, pot.TL, 6):
        raise _lib.AdockError("pot.W must be [n_rt, TL, 6]")
    if spec.n_rt > MAX_RT:
        raise _lib.AdockError(f"at most {MAX_RT} receptor ext types")
    dev = torch.device(device)
    f32 = lambda a: torch.as_tensor(np.ascontiguousarray(a, np.float32), device=dev)
    keep = [f32(rx), torch.as_tensor(re, device=dev), f32(spec.rec_radius),
            pot.radius.to(dev, torch.float32).contiguous(), pot.W.to(dev, torch.float32).contiguous(),
            f32(np.array(origins)), f32(spec.grid_prm())]
    TL = pot.TL
    out = torch.empty((R, TL, n, n, n), dtype=torch.float32, device=dev)
    with torch.cuda.device(dev):
        st = _lib.lib().adock_build_grid_fit_sets(
            _lib.ptr(keep[0]), _lib.ptr(keep[1]), R, NR, _lib.ptr(keep[2]), _lib.ptr(keep[3]), TL, _lib.ptr(keep[4]),
            spec.n_rt, _lib.ptr(keep[5]), float(spacing), n, n, n, _lib.ptr(keep[6]), _lib.ptr(out),
            _lib.stream_of(dev))
    _lib.check(st)
    T = lambda a: torch.as_tensor(np.array(a), dtype=torch.float32, device=device)
    return GridSet(out, T(origins), T(lo), T(hi), spacing, penalty)


def inter_exact_fit(X, lig_types, rec_idx, rec_xyz, rec_ext, pot):
    """Exact pairwise fitted inter energy [L, M] of poses X [L, M, A, 3] by the HIP kernel adock_inter_exact_fit:
    the reference's Problem.inter_exact over FittedPotential.energy (src/adock/search.py:73-89,
    src/adock/fit/terms.py:339-344). lig_types [L, A] compact ligand types (-1: hydrogen / padding, skipped), rec_idx
    [L] receptor set of each complex, rec_xyz [R, N, 3], rec_ext [R, N] receptor ext types (-1: padding, skipped)."""
    _lib.require_gpu()
    if not X.is_cuda:
        raise _lib.AdockError("inter_exact_fit runs on the GPU only (X must be a cuda tensor)")
    dev = X.device
    L, M, A = X.shape[:3]
    spec = pot.spec
    xs = X.detach().to(torch.float32).contiguous()
    lt = torch.as_tensor(lig_types, device=dev).to(torch.int32).contiguous()
    ri = torch.as_tensor(rec_idx, device=dev).to(torch.int32).contiguous()
    rx = torch.as_tensor(rec_xyz, device=dev).to(torch.float32).contiguous()
    re = torch.as_tensor(rec_ext, device=dev).to(torch.int32).contiguous()
    if tuple(lt.shape) != (L, A) or tuple(ri.shape) != (L,):
        raise _lib.AdockError("lig_types must be [L, A] and rec_idx [L] for X [L, M, A, 3]")
    if rx.dim() != 3 or rx.shape[2] != 3 or tuple(re.shape) != tuple(rx.shape[:2]):
        raise _lib.AdockError("rec_xyz must be [R, N, 3] and rec_ext [R, N]")
    R, N = re.shape
    if L and R == 0:
        raise _lib.AdockError("no receptor sets")
    if L and (int(ri.min()) < 0 or int(ri.max()) >= R):
        raise _lib.AdockError(f"rec_idx must lie in [0, {R})")
    if tuple(pot.W.shape) != (spec.n_rt, pot.TL, 6):
        raise _lib.AdockError("pot.W must be [n_rt, TL, 6]")
    keep = [pot.W.to(dev, torch.float32).contiguous(), torch.as_tensor(spec.rec_radius, dtype=torch.float32,
            device=dev), pot.radius.to(dev, torch.float32).contiguous(),
            torch.as_tensor(np.asarray(spec.grid_prm(), np.float32), device=dev)]
    out = torch.empty((L, M), dtype=torch.float32, device=dev)
    with torch.cuda.device(dev):
        st = _lib.lib().adock_inter_exact_fit(
            _lib.ptr(xs), _lib.ptr(lt), L, M, A, _lib.ptr(ri), _lib.ptr(rx), _lib.ptr(re), R, N, _lib.ptr(keep[1]),
            _lib.ptr(keep[2]), pot.TL, _lib.ptr(keep[0]), spec.n_rt, _lib.ptr(keep[3]), _lib.ptr(out),
            _lib.stream_of(dev))
    _lib.check(st)
    return out
