"""Synthetic single-complex and batched workloads assembled into the containers the kernels read.

`host_arrays` gives the same inputs as plain numpy dicts (packed ligands, flat grids, scoring tables, receptors) so a CPU
implementation can be driven with identical data.
"""
import numpy as np
import torch

from . import synth
from .config import load_config, search_prm, sf_params
from .grid import GridSet, make_gridset
from .ligand import pack_ligands
from .problem import LigandBatch, PairPotential, Problem, pad_receptors

CFG = load_config()
SPACING = CFG["grid"]["spacing"]
PENALTY = CFG["grid"]["out_of_box_penalty"]


def receptor_dict(rec, center, size):
    return {"rec_xyz": rec["rec_xyz"], "rec_types": rec["rec_types"], "center": np.asarray(center, float),
            "size": np.broadcast_to(np.asarray(size, float), (3,)).copy()}


def grid_types(preps):
    """Per receptor: ligand XS types present (C_H is added by make_gridset)."""
    return [np.unique(p.types[p.types >= 0]) for p in preps]


def gpu_problem(receptors, preps, rec_idx, sf="vina", device="cuda", with_rec=True, keep_builder=False):
    """receptors: list of dicts (rec_xyz, rec_types, center, size); preps: one LigandPrep per complex; rec_idx [L].
    Grids hold, per receptor, the union of the types of its ligands (+ C_H). keep_builder: prob.gs.rebuild() refills
    the grids in place from device-resident receptor atoms."""
    sfp = sf_params(CFG, sf)
    rec_idx = np.asarray(rec_idx, np.int64)
    lt = []
    for r in range(len(receptors)):
        ts = [p.types[p.types >= 0] for p, ri in zip(preps, rec_idx) if ri == r]
        lt.append(np.unique(np.concatenate(ts)) if ts else np.zeros(0, np.int64))
    gs = make_gridset(receptors, sfp, SPACING, PENALTY, device=device, lig_types=lt, keep_builder=keep_builder)
    lb = LigandBatch(preps, device)
    rec = pad_receptors(receptors, device) if with_rec else None
    return Problem(gs, lb, torch.as_tensor(rec_idx, device=device), PairPotential(sfp, device), rec=rec)


def host_arrays(prob):
    """Numpy copies of everything the kernels read: (AM, KM, lig, grid, sf, rec)."""
    gs = prob.gs
    AM, KM, lig = pack_ligands(prob.lb.preps, gs.slot.cpu().numpy(), prob.rec_idx.cpu().numpy())
    grid = dict(flat=gs.flat.float().cpu().numpy(), origin=gs.origin.float().cpu().numpy(),
                lo=gs.lo.float().cpu().numpy(), hi=gs.hi.float().cpu().numpy(),
                rec_idx=prob.rec_idx.cpu().numpy().astype(np.int32), offs=gs.offs.cpu().numpy().astype(np.int64),
                dims=gs.dims.cpu().numpy().astype(np.int64), spacing=float(gs.spacing), penalty=float(gs.penalty))
    sp = prob.pot.p
    sf = dict(prm=search_prm(sp), rad=np.asarray(sp["radius"], np.float32), hyd=np.asarray(sp["hydrophobic"], np.int32),
              don=np.asarray(sp["donor"], np.int32), acc=np.asarray(sp["acceptor"], np.int32))
    rec = None
    if prob.rec is not None:
        rx, rt, rm = prob.rec
        rec = dict(rxyz=rx.float().cpu().numpy(), rtyp=rt.cpu().numpy().astype(np.int32),
                   rn=rm.sum(1).cpu().numpy().astype(np.int32))
    return AM, KM, lig, grid, sf, rec


def single(name, seed=0, sf="vina", device="cuda"):
    """A single-complex workload ("small" or "single", synth.make_single): one receptor, one ligand. Returns (workload
    dict, Problem)."""
    w = synth.make_single(name, seed)
    rec = receptor_dict(w["receptor"], w["center"], w["size"])
    prob = gpu_problem([rec], [w["lig"]], [0], sf=sf, device=device)
    return w, prob


def screen(seed=0, n_ligands=512, chains=64, steps=500, sf="vina", device="cuda"):
    """Screening batch: one grid set (union of ligand types + C_H) shared by n_ligands ligands."""
    w = synth.make_screen(seed, n_ligands, chains, steps)
    rec = receptor_dict(w["receptor"], w["center"], w["size"])
    prob = gpu_problem([rec], w["ligs"], np.zeros(len(w["ligs"]), np.int64), sf=sf, device=device, with_rec=False)
    return w, prob


def multi(seed=0, n_receptors=64, ligands_per=32, chains=32, steps=500, sf="vina", device="cuda", keep_builder=False):
    """Multi-receptor batch: one grid set per receptor (its ligands' types + C_H), all resident."""
    w = synth.make_multi(seed, n_receptors, ligands_per, chains, steps)
    recs = [receptor_dict(r, np.zeros(3), b) for r, b in zip(w["receptors"], w["boxes"])]
    w["receptor_dicts"] = recs
    prob = gpu_problem(recs, w["ligs"], w["rec_idx"], sf=sf, device=device, with_rec=False, keep_builder=keep_builder)
    return w, prob


def chain_starts(seed, preps, centers, sizes, chains, KM):
    """Start states for `chains` chains per complex (translation uniform in the shrunk box, uniform quaternion,
    theta ~ U(-pi, pi)); returns (state [L*chains, 7+KM] f32, chain_lig [L*chains] i32)."""
    L = len(preps)
    st = np.zeros((L * chains, 7 + KM), np.float32)
    for l, p in enumerate(preps):
        t, q, th = synth.start_states(seed * 7 + l, chains, centers[l], sizes[l], p.n_tors, KM)
        st[l * chains:(l + 1) * chains, :3] = t
        st[l * chains:(l + 1) * chains, 3:7] = q
        st[l * chains:(l + 1) * chains, 7:] = th
    return st, np.repeat(np.arange(L, dtype=np.int32), chains)
