"""GPU parity through the C ABI: HIP kernels vs the FP64 CPU oracle on identical float32 inputs.

Bars (written into each test): grids 1e-5 relative with a 1e-5 kcal/mol floor; per-pose energies 1e-4 relative with a
1e-4 floor and gradients 1e-4 in vector norm (poses closer than 2e-5 A to a discontinuity of E or dE are excluded and
counted: a cell face of the trilinear lookup, the 8 A cutoff, a slope-term kink: within 2e-5 A the float32 and
float64 poses can fall on different sides); one BFGS iteration: >= 99% of
chains within 1e-3; longer BFGS / MC runs: distributions (median, min, mean) within 0.5 kcal/mol, because float
rounding flips Armijo branches (the reference's own CPU-vs-GPU bars, tests/test_cpukernel.py:118-148)."""
import math

import numpy as np
import pytest
import torch

import oracle
from adock_amd import synth, workload
from adock_amd.config import load_config, sf_params
from adock_amd.grid import build_grid
from adock_amd.kernel import KernelProblem
from helpers import (EG_FLOOR, EG_REL, GRID_FLOOR, GRID_REL, IDX, SPACING, oracle_grid, random_receptor, rel_err,
                     sf_tables, states, vec_rel)

pytestmark = pytest.mark.gpu
CFG = load_config()
MARGIN = 2e-5    # A: about 10x the float32 position error of a forward-kinematics result near 13 A


def _oracle_of(prob):
    AM, KM, lig, grid, sf, rec = workload.host_arrays(prob)
    return oracle.OracleProblem(AM, KM, lig, grid, sf, rec)


@pytest.fixture(scope="module")
def small_cx():
    w, prob = workload.single("small", seed=0)
    return w, prob, KernelProblem(prob), _oracle_of(prob)


@pytest.fixture(scope="module")
def single_cx():
    w, prob = workload.single("single", seed=0)
    return w, prob, KernelProblem(prob), _oracle_of(prob)


# ------------------------------------------------------------------------------------------------ grids
@pytest.mark.parametrize("sf", ["vina", "vinardo"])
def test_grid_random_receptor(sf):
    """Reference recipe (tests/test_scoring_grid.py:54-75): 300 atoms in [0,14]^3, origin (3,3,3), n 21."""
    print(f"[gpu] grid random receptor sf={sf}", flush=True)
    sfp = sf_params(CFG, sf)
    xyz, types = random_receptor()
    origin, n = np.array([3.0, 3.0, 3.0]), (21, 21, 21)
    lt = np.arange(18)
    g = build_grid(xyz, types, sfp, origin, SPACING, n, lig_types=lt).cpu().numpy()
    o = oracle_grid({"rec_xyz": xyz, "rec_types": types}, sfp, origin, n, lt)
    assert rel_err(g, o, GRID_FLOOR, GRID_REL) < GRID_REL
    gp = build_grid(xyz, types, sfp, origin, SPACING, n, lig_types=[0, 3, 7, 12], per_term=True).cpu().numpy()
    op = oracle_grid({"rec_xyz": xyz, "rec_types": types}, sfp, origin, n, [0, 3, 7, 12], per_term=True)
    assert rel_err(gp, op, GRID_FLOOR, GRID_REL) < GRID_REL


def test_grid_small_full(small_cx):
    """Whole grid set of the small workload (55^3 x types): 1,500 receptor atoms."""
    print("[gpu] grid small complex", flush=True)
    w, prob, kp, op = small_cx
    sfp = sf_params(CFG, "vina")
    rec = w["receptor"]
    types = np.nonzero(prob.gs.slot[0].cpu().numpy() >= 0)[0]
    o = oracle_grid(rec, sfp, -w["size"] / 2, (55, 55, 55), types)
    g = prob.gs.flat.cpu().numpy().reshape(o.shape)
    assert rel_err(g, o, GRID_FLOOR, GRID_REL) < GRID_REL


def test_grid_single_subbox(single_cx):
    """Single-complex workload (71^3, 4,000 atoms): a 14^3 sub-box of the GPU grid vs the oracle built on that sub-box. The origin
    -13 A and h = 0.375 make every point an exact multiple of 1/8, so both sides place the points identically."""
    print("[gpu] grid single complex sub-box", flush=True)
    w, prob, kp, op = single_cx
    sfp = sf_params(CFG, "vina")
    types = np.nonzero(prob.gs.slot[0].cpu().numpy() >= 0)[0]
    g = prob.gs.flat.cpu().numpy().reshape(len(types), 71, 71, 71)
    for s0 in (0, 28, 57):
        origin = -w["size"] / 2 + s0 * SPACING
        o = oracle_grid(w["receptor"], sfp, origin, (14, 14, 14), types)
        assert rel_err(g[:, s0:s0 + 14, s0:s0 + 14, s0:s0 + 14], o, GRID_FLOOR, GRID_REL) < GRID_REL


# ------------------------------------------------------------------------------------------------ energy + gradient
def _eg_parity(kp, op, w, st, exact=False):
    N = len(st)
    cl = np.zeros(N, np.int32)
    e, g = kp.energy_grad(torch.as_tensor(cl).cuda(), torch.as_tensor(st).cuda(), exact=exact)
    e, g = e.cpu().numpy(), g.cpu().numpy()
    eo, go, _, m = op.energy_grad(cl, st, exact=exact)
    keep = m >= MARGIN
    D = 6 + w["lig"].n_tors
    assert keep.mean() > 0.9, f"too many poses near a discontinuity: {(~keep).sum()}"
    print(f"  excluded {(~keep).sum()} of {len(keep)} poses within {MARGIN} A of a discontinuity", flush=True)
    er = rel_err(e[keep], eo[keep], EG_FLOOR, EG_REL)
    gr = vec_rel(g[keep, :D], go[keep, :D], EG_FLOOR)
    assert er < EG_REL, er
    assert gr < EG_REL, gr
    return er, gr


@pytest.mark.parametrize("which", ["near", "box"])
def test_eg_small(small_cx, which):
    print(f"[gpu] eg small complex {which}", flush=True)
    w, prob, kp, op = small_cx
    st = states(5, 256, w["center"], w["size"], w["lig"], kp.KM, sigma=1.5 if which == "near" else None)
    _eg_parity(kp, op, w, st)


def test_eg_single(single_cx):
    print("[gpu] eg single complex", flush=True)
    w, prob, kp, op = single_cx
    st = states(6, 256, w["center"], w["size"], w["lig"], kp.KM, sigma=2.0)
    _eg_parity(kp, op, w, st)


def test_eg_vinardo():
    print("[gpu] eg vinardo", flush=True)
    w, prob = workload.single("small", seed=3, sf="vinardo")
    kp, op = KernelProblem(prob), _oracle_of(prob)
    st = states(7, 256, w["center"], w["size"], w["lig"], kp.KM, sigma=1.5)
    _eg_parity(kp, op, w, st)


def test_eg_exact_mode(small_cx):
    print("[gpu] eg exact", flush=True)
    w, prob, kp, op = small_cx
    st = states(8, 128, w["center"], w["size"], w["lig"], kp.KM, sigma=1.5)
    _eg_parity(kp, op, w, st, exact=True)


def test_eg_many_ligands_two_receptors():
    """Multi-receptor flat grid set with uneven ligands (10-60 atoms, 0-15 torsions: KM bucket 16): chain_lig and
    rec_idx indirection."""
    print("[gpu] eg batch", flush=True)
    rng = np.random.default_rng(0)
    recs = [workload.receptor_dict(synth.make_receptor(s, 1500, 20.0, 6.0), np.zeros(3), 20.0 + 4 * s) for s in (0, 1)]
    preps = [synth.make_ligand(100 + i, int(rng.integers(10, 61)), int(rng.integers(0, 16))) for i in range(6)]
    prob = workload.gpu_problem(recs, preps, [0, 1, 0, 1, 1, 0])
    kp, op = KernelProblem(prob), _oracle_of(prob)
    assert kp.KM == 16
    N = 240
    cl = rng.integers(0, len(preps), N).astype(np.int32)
    st = np.zeros((N, kp.SD), np.float32)
    for i in range(N):
        p = preps[cl[i]]
        st[i] = states(1000 + i, 1, np.zeros(3), 18.0, p, kp.KM, sigma=2.0)[0]
    e, g = kp.energy_grad(torch.as_tensor(cl).cuda(), torch.as_tensor(st).cuda())
    e, g = e.cpu().numpy(), g.cpu().numpy()
    eo, go, _, m = op.energy_grad(cl, st)
    keep = m >= MARGIN
    assert rel_err(e[keep], eo[keep], EG_FLOOR, EG_REL) < EG_REL
    for i in np.nonzero(keep)[0]:
        D = 6 + preps[cl[i]].n_tors
        assert vec_rel(g[i:i + 1, :D], go[i:i + 1, :D], EG_FLOOR) < EG_REL, i


def test_eg_edge_ligands():
    """A rigid ligand (no torsions, no intra pairs), a 64-atom ligand with 15 torsions (the largest pair list, one
    chain per workgroup by LDS) and a one-torsion ligand in one batch."""
    print("[gpu] eg edge ligands", flush=True)
    rec = workload.receptor_dict(synth.make_receptor(3, 1500, 20.0, 6.0), np.zeros(3), 24.0)
    preps = [synth.make_ligand(7, 12, 0), synth.make_ligand(8, 64, 15), synth.make_ligand(9, 33, 1)]
    assert len(preps[0].pairs) == 0 and int(preps[1].heavy.sum()) == 64
    prob = workload.gpu_problem([rec], preps, [0, 0, 0])
    kp, op = KernelProblem(prob), _oracle_of(prob)
    N = 96
    cl = (np.arange(N) % 3).astype(np.int32)
    st = np.zeros((N, kp.SD), np.float32)
    for i in range(N):
        st[i] = states(2000 + i, 1, np.zeros(3), 20.0, preps[cl[i]], kp.KM, sigma=2.0)[0]
    e, g = kp.energy_grad(torch.as_tensor(cl).cuda(), torch.as_tensor(st).cuda())
    e, g = e.cpu().numpy(), g.cpu().numpy()
    eo, go, _, m = op.energy_grad(cl, st)
    keep = m >= MARGIN
    assert keep.sum() > N // 2
    assert rel_err(e[keep], eo[keep], EG_FLOOR, EG_REL) < EG_REL
    for i in np.nonzero(keep)[0]:
        D = 6 + preps[cl[i]].n_tors
        assert vec_rel(g[i:i + 1, :D], go[i:i + 1, :D], EG_FLOOR) < EG_REL, i


# ------------------------------------------------------------------------------------------------ BFGS / MC
def _run_both(kp, op, st, init_opt, mc_steps, iters, seed=0, step_offset=0, f64=True, cl=None, nbr=None):
    """The same chains through the GPU and the oracle; nbr = (lists [N, NB], counts [N]) runs both in exact mode over
    those receptor neighbour lists."""
    N = len(st)
    cl = np.zeros(N, np.int32) if cl is None else cl
    ci = np.arange(N, dtype=np.int32)
    ex = {} if nbr is None else dict(exact=True, nbr=torch.as_tensor(nbr[0]).cuda(), nbr_n=torch.as_tensor(nbr[1]).cuda())
    out = {}
    cur, best = torch.as_tensor(st).cuda(), torch.as_tensor(st).cuda()
    ce, be = torch.zeros(N).cuda(), torch.full((N,), float("inf")).cuda()
    kp.run_mc(torch.as_tensor(cl).cuda(), torch.as_tensor(ci).cuda(), cur, ce, best, be, init_opt, mc_steps, iters,
              1.2, 2.0, 2.0, seed, step_offset, **ex)
    out["gpu"] = (cur.cpu().numpy(), ce.cpu().numpy(), best.cpu().numpy(), be.cpu().numpy())
    cur, best = st.copy(), st.copy()
    ce, be = np.zeros(N, np.float32), np.full(N, np.inf, np.float32)
    exo = {} if nbr is None else dict(exact=True, nbr=nbr[0], nbr_n=nbr[1])
    op.run_mc(cl, ci, cur, ce, best, be, init_opt, mc_steps, iters, 1.2, 2.0, 2.0, seed, step_offset, f64=f64, **exo)
    out["cpu"] = (cur, ce, best, be)
    return out


def test_bfgs_one_iteration(small_cx):
    print("[gpu] bfgs 1 iteration", flush=True)
    w, prob, kp, op = small_cx
    st = states(9, 512, w["center"], w["size"], w["lig"], kp.KM, sigma=1.0)
    st[:, 7:] = 0
    out = _run_both(kp, op, st, True, 0, 1)
    eg, ec = out["gpu"][3], out["cpu"][3]
    ok = np.abs(eg - ec) <= 1e-3 * np.maximum(np.abs(ec), 1)
    assert ok.mean() >= 0.99, ok.mean()


@pytest.mark.parametrize("N", [256, 1024])
def test_exact_bfgs_over_neighbour_lists(small_cx, N):
    """Exact mode with receptor neighbour lists (the polish) against the FP64 oracle on the same lists: one BFGS
    iteration from poses near the pocket, 256 chains (4-wave speculative kernel) and 1,024 (one wave per chain), the
    lists staged in LDS and walked through per-atom masks; >= 99 % of chains within 1e-3, as in grid mode."""
    print(f"[gpu] exact bfgs over neighbour lists, {N} chains", flush=True)
    w, prob, kp, op = small_cx
    st = states(21, N, w["center"], w["size"], w["lig"], kp.KM, sigma=1.0)
    cl = np.zeros(N, np.int32)
    nbr, nbr_n = kp.receptor_neighbours(prob, torch.as_tensor(cl).cuda(), torch.as_tensor(st).cuda(), prob.lb.K, 2.0)
    assert 0 < nbr.shape[1] <= 1024
    out = _run_both(kp, op, st, True, 0, 1, cl=cl, nbr=(nbr.cpu().numpy(), nbr_n.cpu().numpy()))
    eg, ec = out["gpu"][3], out["cpu"][3]
    ok = np.abs(eg - ec) <= 1e-3 * np.maximum(np.abs(ec), 1)
    print(f"  within 1e-3: {ok.mean():.4f}", flush=True)
    assert ok.mean() >= 0.99, ok.mean()


@pytest.mark.parametrize("n_heavy,n_tors,KM", [(40, 12, 16), (50, 20, 24), (64, 31, 32)])
def test_bfgs_one_iteration_large_buckets(n_heavy, n_tors, KM):
    """One BFGS iteration (direction from the inverse-Hessian rows, Armijo line search, rank-2 update) at the larger
    torsion buckets against the FP64 oracle: >= 99 % of 256 chains within 1e-3, as at KM 8."""
    print(f"[gpu] bfgs 1 iteration, {n_heavy} atoms / {n_tors} torsions (KM {KM})", flush=True)
    rec = workload.receptor_dict(synth.make_receptor(3, 1500, 20.0, 6.0), np.zeros(3), 24.0)
    lig = synth.make_ligand(101 + KM, n_heavy, n_tors)
    assert lig.n_tors == n_tors
    prob = workload.gpu_problem([rec], [lig], [0])
    kp, op = KernelProblem(prob), _oracle_of(prob)
    assert kp.KM == KM
    st = states(30 + KM, 256, np.zeros(3), 20.0, lig, kp.KM, sigma=1.5)
    out = _run_both(kp, op, st, True, 0, 1)
    eg, ec = out["gpu"][3], out["cpu"][3]
    ok = np.abs(eg - ec) <= 1e-3 * np.maximum(np.abs(ec), 1)
    print(f"  within 1e-3: {ok.mean():.4f}", flush=True)
    assert ok.mean() >= 0.99, ok.mean()


def test_bfgs_one_iteration_vinardo():
    """One BFGS iteration under the Vinardo function (its own radii, Gaussian and repulsion terms, no h-bond
    distance offset) against the FP64 oracle: >= 99 % of 512 chains within 1e-3."""
    print("[gpu] bfgs 1 iteration vinardo", flush=True)
    w, prob = workload.single("small", seed=3, sf="vinardo")
    kp, op = KernelProblem(prob), _oracle_of(prob)
    st = states(19, 512, w["center"], w["size"], w["lig"], kp.KM, sigma=1.0)
    out = _run_both(kp, op, st, True, 0, 1)
    eg, ec = out["gpu"][3], out["cpu"][3]
    ok = np.abs(eg - ec) <= 1e-3 * np.maximum(np.abs(ec), 1)
    print(f"  within 1e-3: {ok.mean():.4f}", flush=True)
    assert ok.mean() >= 0.99, ok.mean()


def test_exact_bfgs_large_bucket():
    """Exact mode at the KM 32 bucket (64 heavy atoms, 31 torsions): one BFGS iteration over receptor neighbour lists
    against the FP64 oracle on the same lists; lists longer than the 1,024 staged atoms read their atoms from memory
    in the same launch. >= 99 % of 256 chains within 1e-3."""
    print("[gpu] exact bfgs, 64 atoms / 31 torsions (KM 32)", flush=True)
    # 2,600 receptor atoms at >= 2.3 A: the chains' lists straddle the 1,024-atom staging area
    rec = workload.receptor_dict(synth.make_receptor(3, 2600, 20.0, 6.0, min_sep=2.3), np.zeros(3), 24.0)
    lig = synth.make_ligand(133, 64, 31)
    assert lig.n_tors == 31
    prob = workload.gpu_problem([rec], [lig], [0])
    kp, op = KernelProblem(prob), _oracle_of(prob)
    assert kp.KM == 32
    N = 256
    st = states(50, N, np.zeros(3), 20.0, lig, kp.KM, sigma=1.5)
    cl = np.zeros(N, np.int32)
    nbr, nbr_n = kp.receptor_neighbours(prob, torch.as_tensor(cl).cuda(), torch.as_tensor(st).cuda(), prob.lb.K, 2.0)
    nn = nbr_n.cpu().numpy()
    print(f"  neighbour lists: {nn.min()}-{nn.max()} atoms, {(nn > 1024).sum()} of {N} longer than 1,024", flush=True)
    assert 0 < (nn > 1024).sum() < N
    out = _run_both(kp, op, st, True, 0, 1, cl=cl, nbr=(nbr.cpu().numpy(), nn))
    eg, ec = out["gpu"][3], out["cpu"][3]
    ok = np.abs(eg - ec) <= 1e-3 * np.maximum(np.abs(ec), 1)
    print(f"  within 1e-3: {ok.mean():.4f}", flush=True)
    assert ok.mean() >= 0.99, ok.mean()


@pytest.mark.parametrize("n_heavy,n_tors,KM", [(50, 20, 24), (64, 31, 32)])
def test_mc_step_large_buckets(n_heavy, n_tors, KM):
    """One MC step (the counter-RNG mutation: translation, rotation or one of up to 31 torsions) plus one BFGS
    iteration at the larger buckets against the FP64 oracle: >= 97 % of 256 chains within 1e-3, as at KM 8."""
    print(f"[gpu] mc one step, {n_heavy} atoms / {n_tors} torsions (KM {KM})", flush=True)
    rec = workload.receptor_dict(synth.make_receptor(3, 1500, 20.0, 6.0), np.zeros(3), 24.0)
    lig = synth.make_ligand(101 + KM, n_heavy, n_tors)
    prob = workload.gpu_problem([rec], [lig], [0])
    kp, op = KernelProblem(prob), _oracle_of(prob)
    assert kp.KM == KM
    st = states(40 + KM, 256, np.zeros(3), 20.0, lig, kp.KM, sigma=1.5)
    out = _run_both(kp, op, st, False, 1, 1, seed=4242, step_offset=3)
    eg, ec = out["gpu"][1], out["cpu"][1]
    ok = np.abs(eg - ec) <= 1e-3 * np.maximum(np.abs(ec), 1)
    print(f"  within 1e-3: {ok.mean():.4f}", flush=True)
    assert ok.mean() >= 0.97, ok.mean()


def test_bfgs_100_iterations(small_cx):
    print("[gpu] bfgs 100 iterations", flush=True)
    w, prob, kp, op = small_cx
    st = states(10, 512, w["center"], w["size"], w["lig"], kp.KM, sigma=1.0)
    out = _run_both(kp, op, st, True, 0, 100)
    eg, ec = out["gpu"][3], out["cpu"][3]
    assert abs(np.median(eg) - np.median(ec)) < 0.5
    assert abs(eg.min() - ec.min()) < 0.5
    assert (np.abs(eg - ec) <= 1e-3 * np.maximum(np.abs(ec), 1)).mean() > 0.2


def test_mc_rng_and_mutation_match(small_cx):
    """init_opt off, one MC step, one BFGS iteration: the move (entity, ball vector, torsion draw) comes from the
    counter-based RNG, so matching energies show the GPU draws the same numbers as the oracle."""
    print("[gpu] mc one step", flush=True)
    w, prob, kp, op = small_cx
    st = states(11, 512, w["center"], w["size"], w["lig"], kp.KM, sigma=1.0)
    out = _run_both(kp, op, st, False, 1, 1, seed=777, step_offset=5)
    xg, xc = out["gpu"][0], out["cpu"][0]
    eg, ec = out["gpu"][1], out["cpu"][1]
    ok = np.abs(eg - ec) <= 1e-3 * np.maximum(np.abs(ec), 1)
    assert ok.mean() >= 0.97, ok.mean()


def test_mc_small_properties(small_cx):
    """The small workload itself: 8 chains x 200 MC steps, Vina-rule BFGS, seed 0."""
    print("[gpu] mc small complex (8 x 200)", flush=True)
    w, prob, kp, op = small_cx
    N = w["chains"]
    st = states(0, N, w["center"], w["size"], w["lig"], kp.KM)
    cl, ci = torch.zeros(N, dtype=torch.int32).cuda(), torch.arange(N, dtype=torch.int32).cuda()
    e0, _ = kp.energy_grad(cl, torch.as_tensor(st).cuda())
    runs = []
    for split in (False, True, False):
        cur, best = torch.as_tensor(st).cuda(), torch.as_tensor(st).cuda()
        ce, be = torch.zeros(N).cuda(), torch.full((N,), float("inf")).cuda()
        seed64 = 0 * 1_000_003 + 12345
        if split:
            kp.run_mc(cl, ci, cur, ce, best, be, True, 100, 0, 1.2, 2.0, 2.0, seed64, 0)
            kp.run_mc(cl, ci, cur, ce, best, be, False, 100, 0, 1.2, 2.0, 2.0, seed64, 100)
        else:
            kp.run_mc(cl, ci, cur, ce, best, be, True, 200, 0, 1.2, 2.0, 2.0, seed64, 0)
        runs.append([x.cpu().numpy() for x in (cur, ce, best, be)])
    be = runs[0][3]
    assert (be <= e0.cpu().numpy() + 1e-4).all()
    e1, _ = kp.energy_grad(cl, torch.as_tensor(runs[0][2]).cuda())
    np.testing.assert_allclose(e1.cpu().numpy(), be, rtol=1e-4, atol=1e-4)
    for a, b in zip(runs[0], runs[2]):            # deterministic: same launch twice, same bits
        np.testing.assert_array_equal(a, b)
    for a, b in zip(runs[0], runs[1]):            # segment split == one launch
        np.testing.assert_array_equal(a, b)
    assert be.min() < 0.0


def test_mc_best_score_distribution(small_cx):
    """Same seeded starts, same seed: the per-chain best-score distributions of GPU and FP64 oracle agree."""
    print("[gpu] mc distribution (64 x 30)", flush=True)
    w, prob, kp, op = small_cx
    st = states(12, 64, w["center"], w["size"], w["lig"], kp.KM)
    out = _run_both(kp, op, st, True, 30, 0, seed=12345)
    eg, ec = out["gpu"][3], out["cpu"][3]
    assert abs(np.median(eg) - np.median(ec)) < 0.5, (np.median(eg), np.median(ec))
    assert abs(eg.min() - ec.min()) < 0.5, (eg.min(), ec.min())
    assert abs(eg.mean() - ec.mean()) < 0.5
