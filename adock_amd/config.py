"""Engine parameters: XS type table, scoring-function constants, grid and search presets.

Mirrors the reference's `load_config` / `sf_params` (src/adock/config.py:34-62) so that a reference
config/adock.json loads unchanged and yields the same flat scoring dict and per-type arrays.
"""
import json
import os
from pathlib import Path

import numpy as np

DEFAULT_CONFIG = Path(os.environ.get("ADOCK_CONFIG", Path(__file__).resolve().parent / "data" / "adock.json"))


def load_config(path=None):
    """Read the JSON config and derive `_xs_names`, `_xs_index`, `_xs_table` (src/adock/config.py:34-42)."""
    cfg = json.loads(Path(path or DEFAULT_CONFIG).read_text())
    cols = cfg["xs_types"]["_columns"]
    names = [k for k in cfg["xs_types"] if not k.startswith("_")]
    table = np.array([cfg["xs_types"][k] for k in names], dtype=np.float64)
    cfg["_xs_names"] = names
    cfg["_xs_index"] = {n: i for i, n in enumerate(names)}
    cfg["_xs_table"] = {c: table[:, i] for i, c in enumerate(cols)}
    return cfg


def sf_params(cfg, sf="vina"):
    """Flat scoring parameters + per-type radius / flag arrays (src/adock/config.py:45-62)."""
    p = cfg["scoring"][sf]
    t = cfg["_xs_table"]
    return {
        "name": sf,
        "cutoff": p["cutoff"],
        "g1_off": p["gauss1"]["offset"], "g1_w": p["gauss1"]["width"], "w_g1": p["gauss1"]["weight"],
        "g2_off": p["gauss2"]["offset"], "g2_w": p["gauss2"]["width"], "w_g2": p["gauss2"]["weight"],
        "rep_off": p["repulsion"]["offset"], "w_rep": p["repulsion"]["weight"],
        "hyd_good": p["hydrophobic"]["good"], "hyd_bad": p["hydrophobic"]["bad"], "w_hyd": p["hydrophobic"]["weight"],
        "hb_good": p["hbond"]["good"], "hb_bad": p["hbond"]["bad"], "w_hb": p["hbond"]["weight"],
        "rot_weight": p["rot_weight"],
        "radius": t[p["radius_column"]].astype(np.float32),
        "hydrophobic": t["hydrophobic"].astype(np.int32),
        "donor": t["donor"].astype(np.int32),
        "acceptor": t["acceptor"].astype(np.int32),
    }


# Order of the 15-float parameter vector taken by the grid builder (src/adock/grid.py:84-87): cutoff first.
GRID_PRM_KEYS = ("cutoff", "g1_off", "g1_w", "g2_off", "g2_w", "rep_off", "hyd_good", "hyd_bad", "hb_good", "hb_bad",
                 "w_g1", "w_g2", "w_rep", "w_hyd", "w_hb")
# Order taken by the search kernels (src/adock/mckernel.py:594-596): cutoff tenth. The two orders differ on purpose:
# each follows the reference's own binding.
SEARCH_PRM_KEYS = ("g1_off", "g1_w", "g2_off", "g2_w", "rep_off", "hyd_good", "hyd_bad", "hb_good", "hb_bad", "cutoff",
                   "w_g1", "w_g2", "w_rep", "w_hyd", "w_hb")


def grid_prm(sfp):
    return np.array([sfp[k] for k in GRID_PRM_KEYS], dtype=np.float32)


def search_prm(sfp):
    return np.array([sfp[k] for k in SEARCH_PRM_KEYS], dtype=np.float32)
