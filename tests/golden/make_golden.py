"""Writes the committed fixtures of tests/golden/.

two_atom_terms.json: the two-atom energy at the reference's known-answer distances (tests/test_scoring_grid.py:40-51)
for 8 type pairs x {vina, vinardo} (each distance rounded to float32, as the atom coordinate is stored), from the published Vina / Vinardo functional form (tests/helpers.analytic_pair,
written independently of the kernels and of the oracle). splitmix64.json holds the published first outputs of
SplitMix64 seeded with 0 (Vigna's reference generator: state += 0x9E3779B97F4A7C15, then the finaliser), i.e. smix of
k * 0x9E3779B97F4A7C15 for k = 0, 1, 2; they are data from the literature, not computed here.
"""
import json
import sys

import numpy as np
from pathlib import Path

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))
sys.path.insert(0, str(HERE.parent.parent))

from helpers import IDX, analytic_pair  # noqa: E402

PAIRS = [("C_H", "C_H"), ("C_H", "C_P"), ("N_D", "O_A"), ("O_DA", "O_DA"), ("N_P", "O_A"), ("Cl_H", "C_H"),
         ("Met_D", "O_A"), ("S_P", "N_DA")]
RS = [1.0, 2.0, 2.9, 3.3, 3.6, 4.0, 4.5, 5.5, 7.99, 8.01]

if __name__ == "__main__":
    out = {sf: {f"{a}-{b}": {str(r): analytic_pair(sf, float(np.float32(r)), IDX[a], IDX[b]) for r in RS} for a, b in PAIRS}
           for sf in ("vina", "vinardo")}
    (HERE / "two_atom_terms.json").write_text(json.dumps(out, indent=1) + "\n")
    print("wrote two_atom_terms.json")
