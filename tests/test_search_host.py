"""Host side of the search (CPU): the torch pair potential against the reference's two-atom table, intra energy and
pose geometry against the FP64 oracle, and the search driver's control flow run end to end on the oracle twin
(Stage A, segments, refine, exact polish, finalize)."""
import json
from pathlib import Path

import numpy as np
import pytest
import torch

from adock_amd import synth
from adock_amd.config import sf_params
from adock_amd.pose import internal_coords, pose_coords
from adock_amd.search import dock_problem_kernel
from helpers import CFG, IDX, OracleKernel, analytic_pair, cpu_problem, cpu_search_problem, states

GOLD = Path(__file__).resolve().parent / "golden"


@pytest.mark.parametrize("sf", ["vina", "vinardo"])
def test_torch_pair_terms_match_two_atom_table(sf):
    from adock_amd.problem import PairPotential
    pot = PairPotential(sf_params(CFG, sf))
    gold = json.loads((GOLD / "two_atom_terms.json").read_text())[sf]
    for key, tab in gold.items():
        a, b = key.split("-")
        for r, want in tab.items():
            e = pot.energy(torch.tensor([float(np.float32(float(r)))], dtype=torch.float64),
                           torch.tensor([IDX[a]]), torch.tensor([IDX[b]]))
            # radii are held as float32: <= 1e-7 relative
            assert abs(float(e) - want) <= 2e-6 * abs(want) + 1e-7, (key, r)


def test_intra_and_pose_coords_match_oracle():
    from adock_amd import workload
    rec = workload.receptor_dict(synth.make_receptor(3, 800, 16.0, 5.0), np.zeros(3), 18.0)
    lig = synth.make_ligand(11, 24, 5)
    prob = cpu_search_problem(rec, [lig], np.zeros(3), 18.0)
    op, _ = cpu_problem(rec, [lig], np.zeros(3), 18.0)
    st = states(5, 16, np.zeros(3), 18.0, lig, op.KM)
    e, g, X, m = op.energy_grad(np.zeros(16, np.int32), st)
    from adock_amd.kernel import KernelProblem
    tw = OracleKernel(prob)
    t, q, th = KernelProblem.unpack(tw, torch.as_tensor(st), prob.lb.K)
    Xt = pose_coords(prob.lb, t[None], q[None], torch.zeros_like(t)[None], th[None])[0]
    hv = np.where(lig.heavy)[0]
    np.testing.assert_allclose(Xt[:, hv].numpy(), X[:, :len(hv)], atol=2e-4)
    # intra energy: pairwise closed form at the oracle's coordinates
    ei = prob.intra(torch.as_tensor(X[:, :len(hv)][None], dtype=torch.float64))[0].numpy()
    typ = lig.types[hv]
    pl = lig.pairs
    for i in range(4):
        want = sum(analytic_pair("vina", float(np.linalg.norm(X[i, p] - X[i, q])), int(typ[p]), int(typ[q]))
                   for p, q in pl)
        assert abs(ei[i] - want) <= 1e-6 * (1 + abs(want)), i
    # torsion rotations preserve bonded distances
    Xi = internal_coords(prob.lb, th[None].float())[0]
    for i, j in lig.bonds:
        d0 = np.linalg.norm(lig.ref[i] - lig.ref[j])
        np.testing.assert_allclose(torch.linalg.norm(Xi[:, i] - Xi[:, j], dim=-1).numpy(), d0, atol=1e-4)


def _search_cfg(**kw):
    cfg = dict(CFG["search"]["v2_default"])
    cfg.update(n_chains=16, stageA_conformers=4, stageA_conformer_trials=16, stageA_rotations=24,
               stageA_translation_step=2.0, mc_steps=8, mc_segment=8, mc_size_ref=0, final_bfgs_iters=10,
               polish_top=4, polish_iters=8, min_total_chains=0, n_out=3)
    cfg.update(kw)
    return cfg


def test_search_driver_on_oracle_twin():
    """Stage A -> MC segments -> refine -> exact polish -> finalize on the oracle twin: clusters are cluster_rmsd
    apart, totals ascend, affinities are finite; a run split into segments equals the unsplit run."""
    from adock_amd import workload
    rec = workload.receptor_dict(synth.make_receptor(4, 700, 16.0, 5.0), np.zeros(3), 16.0)
    preps = [synth.make_ligand(21, 14, 3), synth.make_ligand(22, 18, 2)]
    prob = cpu_search_problem(rec, preps, np.zeros(3), 16.0)
    boxes_c = [np.zeros(3)] * 2
    boxes_s = [np.full(3, 16.0)] * 2
    stats = {}
    res = dock_problem_kernel(prob, boxes_c, boxes_s, _search_cfg(), seed=1, stats=stats, kernel_problem=OracleKernel)
    assert len(res) == 2 and stats["widen"] == 1 and stats["chain_steps"] == 2 * 16 * 8
    for r, p in zip(res, preps):
        n = len(r["total"])
        assert 1 <= n <= 3 and r["coords"].shape == (n, prob.lb.A, 3)
        assert np.all(np.diff(r["total"]) >= 0) and np.isfinite(r["affinity"]).all()
        hv = np.zeros(prob.lb.A, bool)
        hv[:p.n] = p.heavy
        for i in range(n):
            for j in range(i):
                d = np.sqrt(((r["coords"][i, hv] - r["coords"][j, hv]) ** 2).sum(-1).mean())
                assert d >= 2.0 - 1e-6
    res2 = dock_problem_kernel(prob, boxes_c, boxes_s, _search_cfg(mc_segment=4), seed=1, kernel_problem=OracleKernel)
    for a, b in zip(res, res2):
        np.testing.assert_array_equal(a["total"], b["total"])
        np.testing.assert_array_equal(a["coords"], b["coords"])


def test_widening_rule():
    """Few complexes: chains widen by powers of two at constant chain-steps (mc_steps shrinks), bounded below by
    min_chain_steps (src/adock/search.py:563-569)."""
    from adock_amd import workload
    rec = workload.receptor_dict(synth.make_receptor(4, 500, 14.0, 4.0), np.zeros(3), 14.0)
    preps = [synth.make_ligand(23, 10, 1)]
    prob = cpu_search_problem(rec, preps, np.zeros(3), 14.0)
    stats = {}
    cfg = _search_cfg(n_chains=4, mc_steps=16, mc_segment=16, min_total_chains=16, min_chain_steps=4, polish_top=0)
    dock_problem_kernel(prob, [np.zeros(3)], [np.full(3, 14.0)], cfg, seed=0, stats=stats, kernel_problem=OracleKernel)
    assert stats["widen"] == 4 and stats["chain_steps"] == 16 * 4


def test_receptor_neighbour_lists():
    """Neighbour lists (receptor atoms within cutoff + margin of any heavy atom of each chain's pose, receptor order)
    against a per-chain brute force, for chains of two ligands."""
    from adock_amd import workload
    rec = workload.receptor_dict(synth.make_receptor(4, 700, 16.0, 5.0), np.zeros(3), 16.0)
    preps = [synth.make_ligand(21, 14, 3), synth.make_ligand(22, 18, 2)]
    prob = cpu_search_problem(rec, preps, np.zeros(3), 16.0)
    tw = OracleKernel(prob)
    cl = torch.tensor([0, 1, 1, 0, 1], dtype=torch.int32)
    st = torch.as_tensor(np.concatenate([states(40 + i, 1, np.zeros(3), 10.0, preps[int(c)], tw.KM)
                                         for i, c in enumerate(cl)]))
    nbr, nbr_n = OracleKernel.receptor_neighbours(tw, prob, cl, st, prob.lb.K, 2.0)
    rx = np.asarray(rec["rec_xyz"], np.float64)
    cut = float(prob.pot.p["cutoff"]) + 2.0
    for i, c in enumerate(cl.tolist()):
        t, q, th = tw.unpack(st[i:i + 1], prob.lb.K)
        sub = prob.lb.subset(torch.tensor([c]))
        X = pose_coords(sub, t[None], q[None], torch.zeros_like(t)[None], th[None])[0, 0][prob.lb.heavy[c]].numpy()
        d = np.sqrt(((X[:, None, :] - rx[None]) ** 2).sum(-1)).min(0)
        want = np.nonzero(d < cut)[0]
        assert int(nbr_n[i]) == len(want)
        np.testing.assert_array_equal(nbr[i, :len(want)].numpy(), want)


def _greedy_reference(X, E, inter, intra, D, nt, rw, cut, n_out):
    """The sequential greedy rule of one complex, as the reference states it (src/adock/search.py:389-396, 520-537)."""
    reps = []
    for i in np.argsort(E, kind="stable"):
        if all(D[i, j] >= cut for j in reps):
            reps.append(i)
            if len(reps) == n_out:
                break
    ri = np.array(reps)
    aff = (inter[ri] + intra[ri] - intra[ri[0]]) / (1 + rw * nt)
    return {"coords": X[ri], "affinity": aff, "inter": inter[ri], "intra": intra[ri], "total": E[ri]}


@pytest.mark.parametrize("cut", [2.0, 0.0, 5.5])
def test_cluster_all_equals_sequential_greedy(cut):
    """The all-complex clustering picks the sequential rule's representatives, in its order, with the same values
    bit for bit: ties in energy (stable order), complexes with fewer than n_out clusters, a zero cutoff."""
    from adock_amd.search import _cluster_all
    rng = np.random.default_rng(int(cut * 10))
    L, M, A, n_out = 7, 40, 5, 9
    X = rng.normal(0, 3, (L, M, A, 3)).astype(np.float32)
    E = np.round(rng.normal(-5, 2, (L, M)), 1).astype(np.float32)         # many ties
    intra = rng.normal(0, 1, (L, M)).astype(np.float32)
    inter = (E - intra).astype(np.float32)
    D = rng.uniform(0, 8, (L, M, M)).astype(np.float32)
    D = (D + D.transpose(0, 2, 1)) / 2
    D[:, np.arange(M), np.arange(M)] = 0
    D[2] *= 0.05                                                          # one complex with a single cluster
    nt = rng.integers(0, 8, L).astype(np.float32)
    T = torch.as_tensor
    for inter_arg in (None, T(inter)):
        got = _cluster_all(T(X), T(E), T(intra), T(D), T(nt), 0.05846, cut, n_out, inter=inter_arg)
        for l in range(L):
            want = _greedy_reference(X[l], E[l], E[l] - intra[l] if inter_arg is None else inter[l], intra[l], D[l],
                                     nt[l], 0.05846, cut, n_out)
            for key in want:
                np.testing.assert_array_equal(got[l][key], want[key])
