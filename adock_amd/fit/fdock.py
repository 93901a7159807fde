"""Docking with a fitted scoring function on the unchanged kernel search (src/adock/fit/fdock.py).

    from adock_amd.fit.fdock import load_fitted, dock_fitted
    spec, w = load_fitted("adock_amd/data/adock_fitted.json", cfg)
    res = dock_fitted(preps, spec, w, mode="v2_default", seed=0)

MC chains and BFGS run in the HIP kernels on the fitted grids (adock_build_grid_fit) with the intra-ligand Vina form
at the fitted intra weights. The chain kernels' exact pairwise mode implements only the stock function, so the exact
polish is off and finalize rescores the final poses with the exact pairwise fitted energy from the HIP kernel
adock_inter_exact_fit (kernel_cfg; Problem.inter_exact dispatches to FittedPotential.inter_exact).

A complex here is a dict with rec_xyz [N,3], rec_ext [N] (receptor ext types, terms.receptor_ext_types), center,
size and lig (a LigandPrep with XS types) plus lig_ext (its ext types, terms.ligand_ext_types). The reference builds
these from PDB / SDF files with RDKit (prepare_complex_fit, src/adock/fit/fdock.py:45-53).
"""
import copy
import hashlib
import json
import time
from pathlib import Path

import numpy as np
import torch

from ..config import DEFAULT_CONFIG, load_config
from ..problem import LigandBatch, Problem, pad_receptors
from .gridfit import make_gridset_fit
from .terms import TERM_NAMES, FittedPotential, Spec

DATA = Path(__file__).resolve().parents[1] / "data"


def load_fitted(src, cfg):
    """(Spec, weights [18]) from a fitted-function JSON file or dict (src/adock/fit/fdock.py:25-32)."""
    d = src if isinstance(src, dict) else json.loads(Path(src).read_text())
    nl = d["nonlinear"]
    spec = Spec(cfg, base=nl["base"], buried_threshold=nl["buried_threshold"], buried_radius=nl["buried_radius"],
                ion_rmin=nl["ion_rmin"], inter_prm=nl.get("inter_prm"), rep_cap=nl.get("rep_cap"),
                sc_radius_scale=nl.get("sc_radius_scale"), sc_radius_delta=nl.get("sc_radius_delta"))
    return spec, np.array([d["weights"][k] for k in TERM_NAMES], np.float32)


def save_fitted(path, spec, w, extra=None):
    """Write a fitted-function JSON file that load_fitted reads back (src/adock/fit/fdock.py:35-42)."""
    d = {"_comment": "fitted scoring function: basis parameters and term weights (adock_amd.fit.terms)",
         "nonlinear": spec.nonlinear(), "weights": {k: float(v) for k, v in zip(TERM_NAMES, w)}}
    d.update(extra or {})
    Path(path).write_text(json.dumps(d, indent=1))


def kernel_cfg(scfg):
    """A kernel search preset adapted to a fitted function: no exact polish, the torch-orchestrated finalize with the
    exact pairwise fitted energy (src/adock/fit/fdock.py:109-116)."""
    return dict(scfg, polish_top=0, polish_iters=0, finalize_torch=True)


def compact_ligands(preps):
    """(lig_list, ligands): the sorted ligand ext types present plus ext type 0 (neutral C_H, always slot 0: Stage A
    probes the translation lattice with type 0), and copies of the complexes' LigandPreps whose types are slots in
    lig_list (-1 kept for hydrogens). The caller's LigandPrep objects are not modified."""
    exts = [np.asarray(p["lig_ext"], np.int64) for p in preps]
    lig_list = np.unique(np.concatenate([[0]] + [e[e >= 0] for e in exts]))
    slot = {int(e): i for i, e in enumerate(lig_list)}
    ligs = []
    for p, e in zip(preps, exts):
        if len(e) != p["lig"].n or not np.array_equal(e >= 0, p["lig"].types >= 0):
            raise ValueError("lig_ext must give every heavy atom of lig an ext type (-1 for hydrogens)")
        lp = copy.copy(p["lig"])
        lp.types = np.array([slot[int(x)] if x >= 0 else -1 for x in e], np.int64)
        ligs.append(lp)
    return lig_list, ligs


def unique_receptors(preps):
    """(index of the first complex of each distinct receptor, receptor of every complex): complexes with the same
    receptor atoms, ext types, box centre and box size share one grid cube and one padded receptor."""
    first, idx, seen = [], [], {}
    for i, p in enumerate(preps):
        h = hashlib.sha1()
        for a, dt in ((p["rec_xyz"], np.float32), (p["rec_ext"], np.int64), (p["center"], np.float64),
                      (np.broadcast_to(np.asarray(p["size"], np.float64), (3,)), np.float64)):
            h.update(np.ascontiguousarray(np.asarray(a, dt)).tobytes())
            h.update(b"|")
        k = h.digest()
        if k not in seen:
            seen[k] = len(first)
            first.append(i)
        idx.append(seen[k])
    return first, np.asarray(idx, np.int64)


def build_problem(preps, spec, w, cfg, device="cuda"):
    """Problem of fitted complexes: compact ligand types (compact_ligands), a FittedPotential, the fitted grid set and
    the receptors with pair types TL + ext type (src/adock/fit/fdock.py:56-70). The reference builds a cube per
    complex; here each distinct receptor (and box) gets one cube, which its complexes share through rec_idx (the same
    values: a screen of many ligands against one pocket builds one cube instead of one per ligand)."""
    lig_list, ligs = compact_ligands(preps)
    pot = FittedPotential(spec, w, lig_list, device)
    first, ridx = unique_receptors(preps)
    upreps = [preps[i] for i in first]
    gs = make_gridset_fit(upreps, pot, cfg["grid"]["spacing"], cfg["grid"]["out_of_box_penalty"], device=device)
    lb = LigandBatch(ligs, device)
    recs = [{"rec_xyz": p["rec_xyz"], "rec_types": np.asarray(p["rec_ext"], np.int64) + pot.TL} for p in upreps]
    return Problem(gs, lb, torch.as_tensor(ridx, device=device), pot, rec=pad_receptors(recs, device))


def dock_fitted(preps, spec, w, mode="v2_default", seed=0, device="cuda", cfg=None, log=None, timings=None,
                stats=None):
    """Dock fitted complexes with a kernel search preset (src/adock/fit/fdock.py:119-143). Returns one dict per complex
    (coords [n_out, n, 3], affinity, inter, intra, total, n_tors_vina); `timings` receives grid_s / search_s."""
    from ..search import dock_problem_kernel
    cfg = cfg or load_config()
    mode = cfg.get("mode_aliases", {}).get(mode, mode)
    scfg = kernel_cfg(cfg["search"][mode])
    t0 = time.perf_counter()
    prob = build_problem(preps, spec, w, cfg, device)
    if str(device).startswith("cuda"):
        torch.cuda.synchronize()
    t1 = time.perf_counter()
    res = dock_problem_kernel(prob, [np.asarray(p["center"], np.float64) for p in preps],
                              [np.broadcast_to(np.asarray(p["size"], np.float64), (3,)) for p in preps], scfg,
                              seed=seed, log=log, stats=stats)
    if str(device).startswith("cuda"):
        torch.cuda.synchronize()
    t2 = time.perf_counter()
    if timings is not None:
        timings.update(grid_s=t1 - t0, search_s=t2 - t1, n=len(preps))
    for p, r in zip(preps, res):
        r["n_tors_vina"] = p["lig"].n_tors_vina
        r["coords"] = r["coords"][:, :p["lig"].n]
    return res


def dock_preset(preps, mode="v2_fitted", seed=0, device="cuda", cfg=None, **kw):
    """Dock with a preset that names a fitted function (fitted_sf, base_preset), as the reference's dock_batch does for
    v2_fitted (src/adock/dock.py:86-90); fitted_sf resolves against the package data folder, then the config's."""
    cfg = cfg or load_config()
    scfg = cfg["search"][mode]
    if not scfg.get("fitted_sf"):
        raise ValueError(f"preset {mode!r} names no fitted scoring function")
    path = Path(scfg["fitted_sf"])
    for base in (DATA, Path(DEFAULT_CONFIG).parent):
        if (base / path).exists():
            path = base / path
            break
    spec, w = load_fitted(path, cfg)
    return dock_fitted(preps, spec, w, mode=scfg["base_preset"], seed=seed, device=device, cfg=cfg, **kw)
