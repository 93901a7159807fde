"""Host-side measurement and problem-building logic (no GPU): the FLOP count behind the roofline figures and the
fitted driver's receptor deduplication."""
import numpy as np

from adock_amd import synth
from adock_amd.fit.fdock import unique_receptors


def test_flops_per_eval_of_the_single_complex_ligand():
    """98 A + 40 M + 47 P + nt + 3 P with A heavy atoms, M atom-torsion memberships, P intra pairs: the single-complex
    ligand (35 atoms, 8 torsions) has M = 50 and P = 454, 28,138 operations; and the count follows each term."""
    import bench
    lig = synth.make_single("single", seed=0)["lig"]
    A = int(lig.heavy.sum())
    M = int(lig.tors_mask[:, lig.heavy].sum())
    P = len(lig.pairs)
    assert (A, lig.n_tors, M, P) == (35, 8, 50, 454)
    assert bench.flops_per_eval(lig) == 98 * 35 + 40 * 50 + 47 * 454 + 8 + 3 * 454 == 28138
    rigid = synth.make_ligand(3, 12, 0)
    assert bench.flops_per_eval(rigid) == 98 * int(rigid.heavy.sum()) + 50 * len(rigid.pairs)


def _prep(seed, center=(0.0, 0.0, 0.0), size=20.0):
    rng = np.random.default_rng(seed)
    return {"rec_xyz": rng.normal(size=(50, 3)).astype(np.float32), "rec_ext": rng.integers(0, 10, 50),
            "center": np.asarray(center, np.float64), "size": size}


def test_unique_receptors_shares_identical_receptors_only():
    a, b = _prep(1), _prep(2)
    same = dict(a, rec_xyz=a["rec_xyz"].copy())                 # equal content, another array
    moved = dict(a, center=np.array([0.0, 0.0, 0.5]))           # same atoms, another box centre
    bigger = dict(a, size=np.array([20.0, 20.0, 22.0]))         # same atoms, another box size
    retyped = dict(a, rec_ext=a["rec_ext"] + 1)                 # same atoms, other ext types
    first, idx = unique_receptors([a, b, same, moved, a, bigger, retyped])
    assert first == [0, 1, 3, 5, 6]
    assert idx.tolist() == [0, 1, 0, 2, 0, 3, 4]
    # a scalar box edge and the same edge on every axis are one box
    first, idx = unique_receptors([a, dict(a, size=np.full(3, 20.0))])
    assert idx.tolist() == [0, 0]
