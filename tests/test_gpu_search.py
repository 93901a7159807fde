"""Stage A and the full search driver on the GPU.

Stage A (adock_stage_a) against the FP64 oracle on the same conformers, rotations and lattices; the search driver
(Stage A -> MC segments -> refine -> exact polish -> finalize) on libadock_hip.so: deterministic, a run split into
segments equals the unsplit run bit for bit, clusters ascend in energy, and the best score matches the same driver
run on the oracle twin (independent Stage A draws) within 1.5 kcal/mol; from identical inputs (one Stage A run's start
poses handed to both engines, the same counter-RNG seed) the top-1 best scores of 6 complexes agree in distribution
(median difference <= 0.1 kcal/mol, mean <= 0.25, each within 1.0)."""
import math

import numpy as np
import pytest
import torch

from adock_amd import synth, workload
from adock_amd.kernel import KernelProblem
from adock_amd.pose import internal_coords, random_quat
from adock_amd.search import dock_problem_kernel, stage_a_kernel
from helpers import CFG, OracleKernel, cpu_search_problem

pytestmark = pytest.mark.gpu


def _cfg(**kw):
    cfg = dict(CFG["search"]["v2_default"])
    cfg.update(n_chains=32, stageA_conformers=4, stageA_conformer_trials=32, stageA_rotations=64,
               stageA_translation_step=1.5, mc_steps=32, mc_segment=32, mc_size_ref=0, final_bfgs_iters=30,
               polish_top=6, polish_iters=20, min_total_chains=0, n_out=4)
    cfg.update(kw)
    return cfg


def _complexes():
    w1 = synth.make_single("small", 0)
    rec = workload.receptor_dict(w1["receptor"], w1["center"], w1["size"])
    preps = [w1["lig"], synth.make_ligand(31, 16, 3)]
    return rec, preps, np.asarray(w1["center"], float), float(np.max(w1["size"]))


def _stage_a_vs_oracle(prob, boxes, n_quats, step, seed):
    """adock_stage_a and the oracle on the same conformers (6 per ligand), rotations and per-ligand translation
    lattices (boxes[l] = (centre, edge); the lattices differ in length): argmin translations agree on >= 99 % of
    (ligand, conformer, rotation) rows, energies within 1e-4 relative where they agree and within 0.05 everywhere."""
    kp = KernelProblem(prob)
    ok = OracleKernel(prob)
    lb = prob.lb
    gen = torch.Generator(device="cuda").manual_seed(seed)
    th = (torch.rand((lb.L, 6, lb.K), generator=gen, device="cuda") * 2 - 1) * math.pi * lb.tvalid[:, None, :]
    Xc = internal_coords(lb, th)
    hidx = torch.zeros(lb.L, kp.AM, dtype=torch.long, device="cuda")
    for l, p in enumerate(lb.preps):
        hv = torch.as_tensor(np.where(p.heavy)[0], device="cuda")
        hidx[l, :len(hv)] = hv
    conf = torch.gather(Xc, 2, hidx[:, None, :, None].expand(lb.L, 6, kp.AM, 3)).contiguous()
    quats = random_quat((n_quats,), gen, "cuda")
    lat = []
    for c, s in boxes:
        axes = [torch.arange(-s / 2 + 1.0, s / 2 - 1.0 + 1e-6, step, device="cuda") for _ in range(3)]
        T = torch.stack(torch.meshgrid(*axes, indexing="ij"), -1).reshape(-1, 3) + torch.as_tensor(c, device="cuda")
        lat.append(T[torch.randperm(len(T), generator=gen, device="cuda")].float())
    nT = torch.tensor([len(T) for T in lat], device="cuda")
    trans = torch.zeros(lb.L, int(nT.max()), 3, device="cuda")
    for l, T in enumerate(lat):
        trans[l, :len(T)] = T
    be, bi = kp.stage_a(conf, quats, trans, nT, 0.3)
    bo, io = ok.stage_a(conf, quats, trans, nT, 0.3)
    be, bi = be.cpu().numpy(), bi.cpu().numpy()
    bo, io = bo.numpy(), io.numpy()
    assert (bi < nT.cpu().numpy()[:, None]).all()
    same = bi == io
    print(f"  argmin agrees on {same.mean():.4f} of {same.size} rows", flush=True)
    assert same.mean() >= 0.99, same.mean()
    assert np.abs(be - bo)[same].max() <= 1e-4 * (1 + np.abs(bo[same]).max())
    # where the argmin differs (grid-cell rounding of FP32 vs FP64 at half-cell ties), the energies still agree
    assert np.abs(be - bo).max() <= 0.05


def test_stage_a_matches_oracle():
    print("[gpu] stage A vs oracle", flush=True)
    rec, preps, c, s = _complexes()
    prob = workload.gpu_problem([rec], preps, [0, 0])
    _stage_a_vs_oracle(prob, [(c, s), (c, s - 3.0)], 40, 1.5, 3)


def test_stage_a_two_receptors_uneven_ligands():
    """Two receptor grid sets (rec_idx indirection) and ligands of 10 to 64 heavy atoms (every lane of the wavefront
    holding an atom) with 0 to 15 torsions."""
    print("[gpu] stage A vs oracle: two receptors, 10-64 atoms", flush=True)
    edges = (20.0, 24.0)
    recs = [workload.receptor_dict(synth.make_receptor(r, 1500, 20.0, 6.0), np.zeros(3), edges[r]) for r in (0, 1)]
    shapes = [(10, 0), (64, 15), (35, 8), (23, 2), (64, 5)]
    preps = [synth.make_ligand(500 + i, a, t) for i, (a, t) in enumerate(shapes)]
    rec_idx = [0, 1, 1, 0, 1]
    prob = workload.gpu_problem(recs, preps, rec_idx)
    boxes = [(np.zeros(3), edges[r] - 0.5 * (i % 2)) for i, r in enumerate(rec_idx)]
    _stage_a_vs_oracle(prob, boxes, 32, 2.0, 7)


def test_search_driver_gpu():
    print("[gpu] search driver", flush=True)
    rec, preps, c, s = _complexes()
    prob = workload.gpu_problem([rec], preps, [0, 0])
    bc, bs = [c, c], [np.full(3, s)] * 2
    stats = {}
    res = dock_problem_kernel(prob, bc, bs, _cfg(), seed=2, stats=stats)
    res_again = dock_problem_kernel(prob, bc, bs, _cfg(), seed=2)
    res_split = dock_problem_kernel(prob, bc, bs, _cfg(mc_segment=8), seed=2)
    for a, b, d in zip(res, res_again, res_split):
        np.testing.assert_array_equal(a["total"], b["total"])
        np.testing.assert_array_equal(a["coords"], b["coords"])
        np.testing.assert_array_equal(a["total"], d["total"])
    assert stats["chain_steps"] == 2 * 32 * 32
    for r in res:
        assert np.all(np.diff(r["total"]) >= 0) and np.isfinite(r["affinity"]).all()
    # same driver on the oracle twin (host tensors, its own Stage A draws): comparable best scores
    cprob = cpu_search_problem(rec, preps, c, s)
    cres = dock_problem_kernel(cprob, bc, bs, _cfg(), seed=2, kernel_problem=OracleKernel)
    for r, cr in zip(res, cres):
        assert r["total"][0] < 0 and cr["total"][0] < 0
        assert abs(r["total"][0] - cr["total"][0]) <= 1.5, (r["total"][0], cr["total"][0])


def test_search_driver_matches_oracle_twin_from_identical_starts():
    """Six complexes (the small receptor, ligands of 12-30 heavy atoms) searched by the GPU driver and by the same
    driver on the oracle twin from identical start poses (one GPU Stage A run) and the same seed: the MC moves come
    from the counter RNG, so the engines differ only by float rounding (FP32 kernels vs the FP64 oracle). Rounding
    parts the chains' paths after a few BFGS line searches (chaotic, as between any two float builds), so the bar is
    on the top-1 distribution at a fixed seed: median |difference| <= 0.1 kcal/mol, mean <= 0.25, every complex
    within 1.0 (measured: 0.001-0.064 for five complexes, 0.63 for the 30-atom ligand whose shallow landscape sends
    the two engines' best chains to different minima)."""
    print("[gpu] search driver vs oracle twin from identical Stage A starts", flush=True)
    w1 = synth.make_single("small", 0)
    rec = workload.receptor_dict(w1["receptor"], w1["center"], w1["size"])
    shapes = [(20, 4), (16, 3), (12, 2), (24, 5), (30, 6), (18, 0)]
    preps = [synth.make_ligand(900 + i, a, t) for i, (a, t) in enumerate(shapes)]
    L = len(preps)
    c, s = np.asarray(w1["center"], float), float(np.max(w1["size"]))
    bc, bs = [c] * L, [np.full(3, s)] * L
    cfg = _cfg()
    prob = workload.gpu_problem([rec], preps, [0] * L)
    gen = torch.Generator(device="cuda").manual_seed(11)
    t, q, th, _ = stage_a_kernel(prob, KernelProblem(prob), bc, bs, cfg, gen)
    starts = (t.cpu(), q.cpu(), th.cpu())
    res = dock_problem_kernel(prob, bc, bs, cfg, seed=4, starts=starts)
    cprob = cpu_search_problem(rec, preps, c, s)
    cres = dock_problem_kernel(cprob, bc, bs, cfg, seed=4, kernel_problem=OracleKernel, starts=starts)
    g = np.array([r["total"][0] for r in res])
    o = np.array([r["total"][0] for r in cres])
    print(f"  top-1 GPU {np.round(g, 3)}", flush=True)
    print(f"  top-1 oracle twin {np.round(o, 3)}", flush=True)
    d = np.abs(g - o)
    print(f"  |difference| {np.round(d, 3)}: median {np.median(d):.3f}, mean {d.mean():.3f}", flush=True)
    assert np.all(g < 0) and np.all(o < 0)
    assert np.median(d) <= 0.1 and d.mean() <= 0.25 and d.max() <= 1.0, d
