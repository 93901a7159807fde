"""Shared test helpers: independent closed forms, small synthetic problems, CPU oracle problems."""
import math

import numpy as np

import oracle
from adock_amd import synth
from adock_amd.config import grid_prm, load_config, search_prm, sf_params
from adock_amd.ligand import pack_ligands

CFG = load_config()
IDX = CFG["_xs_index"]
SPACING = CFG["grid"]["spacing"]
PENALTY = CFG["grid"]["out_of_box_penalty"]

# tolerances of the parity contract
GRID_REL = 1e-5          # grids vs the FP64 oracle, relative with an absolute floor
GRID_FLOOR = 1e-5        # kcal/mol
EG_REL = 1e-4            # per-pose energies and gradients vs the FP64 oracle
EG_FLOOR = 1e-4


def analytic_pair(sf, r, a, b):
    """Two-atom energy from the published Vina / Vinardo form (written independently of the kernels and oracle)."""
    p = CFG["scoring"][sf]
    tab = dict(zip(CFG["xs_types"]["_columns"], zip(*[CFG["xs_types"][n] for n in CFG["_xs_names"]])))
    R = lambda t: np.float32(tab[p["radius_column"]][t])
    if r >= p["cutoff"]:
        return 0.0
    d = r - float(R(a)) - float(R(b))
    e = p["gauss1"]["weight"] * math.exp(-((d - p["gauss1"]["offset"]) / p["gauss1"]["width"]) ** 2)
    e += p["gauss2"]["weight"] * math.exp(-((d - p["gauss2"]["offset"]) / p["gauss2"]["width"]) ** 2)
    e += p["repulsion"]["weight"] * (d * d if d < 0 else 0.0)
    if tab["hydrophobic"][a] and tab["hydrophobic"][b]:
        g, bd = p["hydrophobic"]["good"], p["hydrophobic"]["bad"]
        e += p["hydrophobic"]["weight"] * (1.0 if d < g else 0.0 if d > bd else (bd - d) / (bd - g))
    if (tab["donor"][a] and tab["acceptor"][b]) or (tab["acceptor"][a] and tab["donor"][b]):
        g, bd = p["hbond"]["good"], p["hbond"]["bad"]
        e += p["hbond"]["weight"] * (1.0 if d < g else 0.0 if d > bd else (bd - d) / (bd - g))
    return e


def sf_tables(sfp):
    return (np.asarray(sfp["radius"], np.float32), np.asarray(sfp["hydrophobic"], np.int32),
            np.asarray(sfp["donor"], np.int32), np.asarray(sfp["acceptor"], np.int32))


def random_receptor(n=300, seed=0, extent=14.0):
    """The reference's known-answer receptor recipe (tests/test_scoring_grid.py:54-58)."""
    rng = np.random.default_rng(seed)
    xyz = rng.uniform(0, extent, (n, 3)).astype(np.float32)
    types = rng.choice([IDX[t] for t in ("C_H", "C_P", "N_D", "N_A", "O_A", "O_DA", "S_P", "Met_D")], n)
    return xyz, types.astype(np.int32)


def oracle_grid(rec, sfp, origin, n, lig_types, f64=True, per_term=False):
    return oracle.build_grid(rec["rec_xyz"], rec["rec_types"], sf_tables(sfp), grid_prm(sfp), origin, SPACING, n,
                             lig_types, per_term=per_term, f64=f64)


def cpu_problem(rec, preps, center, size, sf="vina", grid=None, rec_idx=None):
    """OracleProblem over one receptor with grids built by the FP64 oracle (rounded to float32, as the kernels store
    them) for the ligands' types + C_H. grid: optional callable(origin, n, types) -> [T, nx, ny, nz] array."""
    sfp = sf_params(CFG, sf)
    types = np.unique(np.concatenate([[0]] + [p.types[p.types >= 0] for p in preps]))
    slot = -np.ones((1, len(sfp["radius"])), np.int64)
    slot[0, types] = np.arange(len(types))
    center = np.asarray(center, float)
    size = np.broadcast_to(np.asarray(size, float), (3,))
    origin = center - size / 2
    n = tuple(int(math.ceil(s / SPACING)) + 1 for s in size)
    G = grid(origin, n, types) if grid is not None else oracle_grid(rec, sfp, origin, n, types)
    G = np.asarray(G, np.float32)
    rec_idx = np.zeros(len(preps), np.int64) if rec_idx is None else rec_idx
    AM, KM, lig = pack_ligands(preps, slot, rec_idx)
    gd = dict(flat=G.reshape(-1), origin=origin.astype(np.float32)[None], lo=origin.astype(np.float32)[None],
              hi=(center + size / 2).astype(np.float32)[None], rec_idx=rec_idx.astype(np.int32),
              offs=np.zeros(1, np.int64), dims=np.array([n], np.int64), spacing=SPACING, penalty=PENALTY)
    sft = dict(prm=search_prm(sfp), rad=sf_tables(sfp)[0], hyd=sf_tables(sfp)[1], don=sf_tables(sfp)[2],
               acc=sf_tables(sfp)[3])
    rd = dict(rxyz=np.asarray(rec["rec_xyz"], np.float32)[None], rtyp=np.asarray(rec["rec_types"], np.int32)[None],
              rn=np.array([len(rec["rec_xyz"])], np.int32))
    return oracle.OracleProblem(AM, KM, lig, gd, sft, rd), dict(G=G, origin=origin, n=n, types=types, slot=slot)


def states(seed, n, center, size, prep, KM, sigma=None):
    """Random poses: translation uniform in the shrunk box (or N(center, sigma)), uniform quaternion, theta U(-pi, pi)."""
    t, q, th = synth.start_states(seed, n, center, size, prep.n_tors, KM)
    if sigma is not None:
        rng = np.random.default_rng(seed + 7)
        t = (np.asarray(center) + rng.standard_normal((n, 3)) * sigma).astype(np.float32)
    s = np.zeros((n, 7 + KM), np.float32)
    s[:, :3], s[:, 3:7], s[:, 7:] = t, q, th
    return s


def rel_err(a, b, floor, rel):
    """max |a - b| / max(|b|, floor / rel) (the reference's _rel_close, tests/test_cpukernel.py:210-213)."""
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float((np.abs(a - b) / np.maximum(np.abs(b), floor / rel)).max()) if a.size else 0.0


def vec_rel(a, b, floor):
    """max over rows of ||a - b|| / max(||b||, floor) (tests/test_cpukernel.py:203-207)."""
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float((np.linalg.norm(a - b, axis=1) / np.maximum(np.linalg.norm(b, axis=1), floor)).max()) if a.size else 0.0
