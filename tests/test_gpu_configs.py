"""The batched and learned workloads through their own launch paths, at reduced size against the FP64 oracle and at full
size through properties that hold at any size (the oracle would need minutes there).

- multi-receptor batch (64 receptors of 2,000-6,000 atoms in 16-30 A boxes x 32 ligands x 32 chains x 500 steps):
  reduced to 4 receptors x 4 ligands, grids (sub-boxes), per-pose energies / gradients, one BFGS iteration and one MC
  step through the persistent queued launch (adock_mc_batch) against the oracle; at full size, best_e re-evaluates to
  itself, best_e <= cur_e and <= the optimised start, two launches are bit-identical, and chains re-run through the
  per-chain launch equal the batched ones bit for bit;
- screening batch (512 ligands of 10-60 atoms / 0-15 torsions x 64 chains x 500 steps, one grid set): one BFGS
  iteration through the persistent launch against the oracle on two chains of every ligand, and the same full-size
  properties;
- learned proposal on the screening batch (512 x 64 chains per step): network-driven steps keep the exact-score
  invariants, are deterministic, replay from a HIP graph bit for bit, and the 128 features the kernel writes match a
  numpy restatement from the oracle's FP64 pose (atom coordinates, energy, gradient) and a numpy trilinear lookup.
Bars as in the rest of the GPU parity tests: per-pose energies 1e-4 relative (poses within 2e-5 A of a
discontinuity excluded and counted), one BFGS iteration >= 99 % of chains within 1e-3, one MC step >= 97 %."""
import numpy as np
import pytest
import torch

from adock_amd import workload
from adock_amd.batch import BatchRunner
from adock_amd.config import load_config, sf_params
from adock_amd.kernel import KernelProblem
from adock_amd.learned import StepSearch
from adock_amd.mlp import ProposalNet
from helpers import EG_FLOOR, EG_REL, GRID_FLOOR, GRID_REL, oracle_grid, rel_err, states, vec_rel
from test_gpu_parity import MARGIN, _oracle_of

pytestmark = pytest.mark.gpu
CFG = load_config()
SEED64 = 12345


def _bufs(st):
    N = st.shape[0]
    return st.clone(), torch.zeros(N, device=st.device), st.clone(), torch.full((N,), float("inf"), device=st.device)


def _batched(kp, runner, st, init_opt, mc_steps, iters, seed, step_offset=0, count=False):
    N = st.shape[0]
    b = _bufs(st)
    if count:
        kp.eval_count.zero_()
    runner.launch(torch.arange(N, dtype=torch.int32, device=st.device), *b, init_opt, mc_steps, iters, 1.2, 2.0, 2.0,
                  seed, step_offset, count_evals=count)
    torch.cuda.synchronize()
    out = [x.cpu().numpy() for x in b]
    return out + [int(kp.eval_count.item())] if count else out


def _oracle_run(op, cl, ci, st, init_opt, mc_steps, iters, seed, step_offset=0):
    cur, best = st.copy(), st.copy()
    ce, be = np.zeros(len(cl), np.float32), np.full(len(cl), np.inf, np.float32)
    op.run_mc(cl, ci, cur, ce, best, be, init_opt, mc_steps, iters, 1.2, 2.0, 2.0, seed, step_offset)
    return cur, ce, best, be


def _within(a, b, tol=1e-3):
    return np.abs(a - b) <= tol * np.maximum(np.abs(b), 1)


def _full_size_properties(kp, st, cl_h, steps, subset):
    """best_e re-evaluates, best_e <= cur_e and <= the optimised start, two launches equal, per-chain == batched."""
    dev = st.device
    N = st.shape[0]
    cl = torch.as_tensor(cl_h, device=dev)
    runner = BatchRunner(kp, cl, n_complexes=int(cl_h.max()) + 1)
    cur, ce, best, be, nev = _batched(kp, runner, st, True, steps, 0, SEED64, count=True)
    print(f"  best_e mean {be.mean():.3f} min {be.min():.3f}; evaluations {nev}", flush=True)
    eb = kp.energy_grad(cl, torch.as_tensor(best, device=dev))[0].cpu().numpy()
    ec = kp.energy_grad(cl, torch.as_tensor(cur, device=dev))[0].cpu().numpy()
    np.testing.assert_allclose(eb, be, rtol=1e-5, atol=1e-4)
    np.testing.assert_allclose(ec, ce, rtol=1e-5, atol=1e-4)
    assert np.isfinite(be).all() and np.all(be <= ce + 1e-6)
    e_init = _batched(kp, runner, st, True, 0, 0, SEED64)[1]
    assert np.all(be <= e_init + 1e-6)
    again = _batched(kp, runner, st, True, steps, 0, SEED64, count=True)
    for a, b in zip((cur, ce, best, be, nev), again):
        np.testing.assert_array_equal(a, b)
    # chains of a few complexes through the per-chain launch (same chain ids, so the same random streams)
    idx = np.flatnonzero(np.isin(cl_h, subset))
    sub = _bufs(st[torch.as_tensor(idx, device=dev)].contiguous())
    kp.run_mc(torch.as_tensor(cl_h[idx], device=dev), torch.as_tensor(idx.astype(np.int32), device=dev), *sub, True,
              steps, 0, 1.2, 2.0, 2.0, SEED64, 0)
    torch.cuda.synchronize()
    for a, b in zip((cur, ce, best, be), sub):
        np.testing.assert_array_equal(a[idx], b.cpu().numpy())
    return be


# ------------------------------------------------------------------------------------------------ multi-receptor
@pytest.fixture(scope="module")
def multi_small():
    w, prob = workload.multi(seed=3, n_receptors=4, ligands_per=4, chains=8, steps=4)
    kp = KernelProblem(prob)
    return w, prob, kp, _oracle_of(prob)


def test_multi_receptor_grids(multi_small):
    """Every receptor's grid set (2,000-6,000 atoms, its own box and types): a corner and a central 12^3 sub-box."""
    print("[gpu] multi-receptor grids", flush=True)
    w, prob, kp, op = multi_small
    gs = prob.gs
    sfp = sf_params(CFG, "vina")
    slot, offs, dims = gs.slot.cpu().numpy(), gs.offs.cpu().numpy(), gs.dims.cpu().numpy()
    origin = gs.origin.cpu().numpy().astype(np.float64)
    flat = gs.flat.cpu().numpy()
    assert len(w["receptors"]) == 4
    for r, rec in enumerate(w["receptors"]):
        assert 2000 <= len(rec["rec_xyz"]) <= 6000 and 16.0 <= w["boxes"][r] <= 30.0
        types = np.flatnonzero(slot[r] >= 0)
        nx, ny, nz = (int(v) for v in dims[r])
        n3 = nx * ny * nz
        for s0 in (0, min(nx, ny, nz) // 2 - 6):
            o = oracle_grid(rec, sfp, origin[r] + s0 * gs.spacing, (12, 12, 12), types)
            for q, t in enumerate(types):
                g = flat[offs[r] + slot[r, t] * n3: offs[r] + (slot[r, t] + 1) * n3].reshape(nx, ny, nz)
                assert rel_err(g[s0:s0 + 12, s0:s0 + 12, s0:s0 + 12], o[q], GRID_FLOOR, GRID_REL) < GRID_REL, (r, t)


def test_multi_receptor_energy_gradient(multi_small):
    print("[gpu] multi-receptor energy / gradient", flush=True)
    w, prob, kp, op = multi_small
    L = len(w["ligs"])
    sts, cls = [], []
    for l, p in enumerate(w["ligs"]):
        b = w["boxes"][w["rec_idx"][l]]
        sts.append(states(40 + l, 32, np.zeros(3), b, p, kp.KM, sigma=0.2 * b))
        cls.append(np.full(32, l, np.int32))
    st, cl = np.concatenate(sts), np.concatenate(cls)
    e, g = kp.energy_grad(torch.as_tensor(cl).cuda(), torch.as_tensor(st).cuda())
    e, g = e.cpu().numpy(), g.cpu().numpy()
    eo, go, _, m = op.energy_grad(cl, st)
    keep = m >= MARGIN
    assert keep.mean() > 0.9
    assert rel_err(e[keep], eo[keep], EG_FLOOR, EG_REL) < EG_REL
    for l in range(L):
        rows = keep & (cl == l)
        D = 6 + w["ligs"][l].n_tors
        assert vec_rel(g[rows, :D], go[rows, :D], EG_FLOOR) < EG_REL, l


def _multi_starts(w, kp, chains, seed):
    L = len(w["ligs"])
    boxes = [w["boxes"][r] for r in w["rec_idx"]]
    return workload.chain_starts(seed, w["ligs"], [np.zeros(3)] * L, boxes, chains, kp.KM)


def test_multi_receptor_batched_bfgs_and_mc_step(multi_small):
    """Through the persistent queued launch: one BFGS iteration from the start states, and one MC step (init_opt off,
    one BFGS iteration after the move) from the same states."""
    print("[gpu] multi-receptor batched BFGS iteration and MC step vs oracle", flush=True)
    w, prob, kp, op = multi_small
    st_h, cl_h = _multi_starts(w, kp, 16, 5)
    st = torch.as_tensor(st_h).cuda()
    ci = np.arange(len(cl_h), dtype=np.int32)
    runner = BatchRunner(kp, torch.as_tensor(cl_h).cuda(), n_complexes=len(w["ligs"]))
    got = _batched(kp, runner, st, True, 0, 1, 7)
    ref = _oracle_run(op, cl_h, ci, st_h, True, 0, 1, 7)
    ok = _within(got[3], ref[3])
    print(f"  BFGS iteration within 1e-3: {ok.mean():.4f}", flush=True)
    assert ok.mean() >= 0.99
    got = _batched(kp, runner, st, False, 1, 1, 777, step_offset=5)
    ref = _oracle_run(op, cl_h, ci, st_h, False, 1, 1, 777, step_offset=5)
    ok = _within(got[1], ref[1])
    print(f"  MC step within 1e-3: {ok.mean():.4f}", flush=True)
    assert ok.mean() >= 0.97


def test_multi_receptor_full_size_properties():
    print("[gpu] multi-receptor batch at full size: 64 receptors x 32 ligands x 32 chains x 500 steps", flush=True)
    w, prob = workload.multi(seed=0)
    kp = KernelProblem(prob)
    assert len(w["ligs"]) == 2048 and w["chains"] == 32 and w["steps"] == 500
    st_h, cl_h = _multi_starts(w, kp, w["chains"], 0)
    _full_size_properties(kp, torch.as_tensor(st_h).cuda(), cl_h, w["steps"], subset=[0, 777, 2047])


def test_hbm_admission_groups_equal_resident_batch():
    """HBM admission with a budget of a third of all grid bytes: several receptor groups, each group's grid sets within
    the budget (the predicted bytes are those the build allocates), and every chain ends as in the all-resident batch
    bit for bit (one state layout and kernel bucket for all groups; chain ids are global)."""
    print("[gpu] HBM admission: groups under a budget vs all grid sets resident", flush=True)
    from adock_amd.admission import AdmittedBatch
    M, steps = 8, 3
    w, prob = workload.multi(seed=5, n_receptors=6, ligands_per=3, chains=M, steps=steps)
    kp = KernelProblem(prob)
    probe = AdmittedBatch(w["receptor_dicts"], w["ligs"], w["rec_idx"], 1e18, M)
    total = sum(probe.bytes)
    assert total == int(prob.gs.flat.numel()) * 4
    budget = total / 3
    ab = AdmittedBatch(w["receptor_dicts"], w["ligs"], w["rec_idx"], budget, M)
    assert len(ab.groups) >= 3 and (ab.AM, ab.KM) == (kp.AM, kp.KM)
    st_h, cl_h = _multi_starts(w, kp, M, 9)
    st = torch.as_tensor(st_h).cuda()
    runner = BatchRunner(kp, torch.as_tensor(cl_h).cuda(), n_complexes=len(w["ligs"]))
    ref = _batched(kp, runner, st, True, steps, 0, SEED64)
    b = _bufs(st)
    res = ab.run(*b, True, steps, 0, 1.2, 2.0, 2.0, SEED64, 0, log=print)
    torch.cuda.synchronize()
    for x, y in zip(ref, b):
        np.testing.assert_array_equal(x, y.cpu().numpy())
    if max(ab.bytes) <= budget:
        assert ab.peak_grid_bytes <= budget
    assert len(res["groups"]) == len(ab.groups) and np.all(res["latency_s"] > 0)


# ------------------------------------------------------------------------------------------------ screening batch
@pytest.fixture(scope="module")
def screen_full():
    w, prob = workload.screen(seed=0)
    kp = KernelProblem(prob)
    L = len(w["ligs"])
    st_h, cl_h = workload.chain_starts(0, w["ligs"], [w["center"]] * L, [w["size"]] * L, w["chains"], kp.KM)
    return w, prob, kp, st_h, cl_h


def test_screening_batch_bfgs_iteration_vs_oracle(screen_full):
    """All 512 x 64 chains through the persistent launch, one BFGS iteration; the oracle re-runs two chains of every
    ligand (chains are independent, so a subset with the same chain ids is the same computation)."""
    print("[gpu] screening batch: one BFGS iteration, persistent launch vs oracle", flush=True)
    w, prob, kp, st_h, cl_h = screen_full
    assert len(w["ligs"]) == 512 and w["chains"] == 64
    runner = BatchRunner(kp, torch.as_tensor(cl_h).cuda(), n_complexes=512)
    got = _batched(kp, runner, torch.as_tensor(st_h).cuda(), True, 0, 1, 11)
    idx = np.concatenate([np.flatnonzero(cl_h == l)[:2] for l in range(512)])
    ref = _oracle_run(_oracle_of(prob), cl_h[idx], idx.astype(np.int32), st_h[idx], True, 0, 1, 11)
    ok = _within(got[3][idx], ref[3])
    print(f"  within 1e-3: {ok.mean():.4f} of {len(idx)}", flush=True)
    assert ok.mean() >= 0.99


def test_screening_batch_full_size_properties(screen_full):
    print("[gpu] screening batch at full size: 512 ligands x 64 chains x 500 steps", flush=True)
    w, prob, kp, st_h, cl_h = screen_full
    be = _full_size_properties(kp, torch.as_tensor(st_h).cuda(), cl_h, w["steps"], subset=[0, 255, 511])
    # the 26 A box around the 8 A cavity docks the mix: most ligands reach a negative best score
    best_l = be.reshape(512, -1).min(1)
    print(f"  per-ligand best: mean {best_l.mean():.3f}, negative for {(best_l < 0).mean():.3f}", flush=True)
    assert (best_l < 0).mean() > 0.5


# ------------------------------------------------------------------------------------------------ learned proposal
EDGES = np.array([-1.2, -0.9, -0.7, -0.5, -0.35, -0.2, -0.1, 0.0, 0.1, 0.25, 0.5, 1.0, 2.0, 5.0, 20.0])


def _atom_energies(grid, lig, l, X):
    """Per-atom grid energy (trilinear, clamped at the grid edge) + out-of-box penalty at world coordinates X [na, 3],
    FP64 numpy."""
    r = int(grid["rec_idx"][l])
    nx, ny, nz = (int(v) for v in grid["dims"][r])
    o, lo, hi = grid["origin"][r].astype(np.float64), grid["lo"][r], grid["hi"][r]
    G = grid["flat"][grid["offs"][r]:].astype(np.float64)
    na = int(lig["na"][l])
    fmx = np.array([nx, ny, nz], np.float64) - 1 - 1e-4
    f = np.clip((X - o) / grid["spacing"], 0, fmx)
    i = np.floor(f).astype(np.int64)
    w_ = f - i
    slot = lig["gslot"][l, :na].astype(np.int64)
    e = np.zeros(na)
    for dx in (0, 1):
        for dy in (0, 1):
            for dz in (0, 1):
                wt = ((w_[:, 0] if dx else 1 - w_[:, 0]) * (w_[:, 1] if dy else 1 - w_[:, 1]) *
                      (w_[:, 2] if dz else 1 - w_[:, 2]))
                e += wt * G[slot * nx * ny * nz + ((i[:, 0] + dx) * ny + i[:, 1] + dy) * nz + i[:, 2] + dz]
    pen = np.maximum(lo - X, 0) ** 2 + np.maximum(X - hi, 0) ** 2
    return e + grid["penalty"] * pen.sum(1)


def _features(x, E, g, ea, lig, l, lo, hi):
    """The 128 proposal-network features of one pose, restated in numpy."""
    f = np.zeros(128)
    na, nt = int(lig["na"][l]), int(lig["nt"][l])
    cen, hal = 0.5 * (lo + hi), np.maximum(0.5 * (hi - lo), 1e-3)
    f[0:3] = (x[:3] - cen) / hal
    q = x[3:7] / np.linalg.norm(x[3:7])
    f[3:7] = q * (-1.0 if x[3] < 0 else 1.0)
    k = min(nt, 16)
    f[7:7 + k] = np.sin(x[7:7 + k])
    f[23:23 + k] = np.cos(x[7:7 + k])
    f[39:45] = np.tanh(0.1 * g[:6])
    f[45:45 + k] = np.tanh(0.1 * g[6:6 + k])
    f[61] = np.tanh(0.1 * E)
    f[62] = np.tanh(0.1 * ea.sum())
    f[63] = (ea > 1.0).sum() / na
    h = f[64:]
    h[:16] = np.bincount((ea[:, None] > EDGES[None, :]).sum(1), minlength=16) / na
    ref = lig["ref"][l, :na]
    octant = (ref[:, 0] > 0) * 1 + (ref[:, 1] > 0) * 2 + (ref[:, 2] > 0) * 4
    for o in range(8):
        sel = octant == o
        if sel.any():
            h[16 + o] = np.tanh(ea[sel].mean())
            h[24 + o] = np.tanh(ea[sel].max())
    h[32], h[33], h[34], h[35] = na / 64.0, nt / 16.0, lig["rg"][l] / 5.0, lig["np"][l] / 1024.0
    for t in range(k):
        sel = lig["tm"][l, t, :na] != 0
        if sel.any():
            h[48 + t] = np.tanh(ea[sel].mean())
    return f


@pytest.fixture(scope="module")
def learned_full(screen_full):
    w, prob, kp, st_h, cl_h = screen_full
    st = torch.as_tensor(st_h).cuda()
    cl = torch.as_tensor(cl_h).cuda()
    assert st.shape[0] == 512 * 64
    ss = StepSearch(kp, cl, net=ProposalNet("cuda", seed=1))
    return w, prob, kp, st, cl, ss


def test_learned_steps_full_batch_invariants_and_graph(learned_full):
    """3 network-driven steps (p_plain 0.25) on 32,768 chains: best_e re-evaluates to itself and is <= cur_e, the
    run is deterministic, and a HIP-graph capture of the same steps replays it bit for bit."""
    print("[gpu] learned proposal, 512 x 64 chains per step: invariants and graph replay", flush=True)
    w, prob, kp, st, cl, ss = learned_full
    kw = dict(seed=21, use_net=True, p_plain=0.25, features=True)
    eager = _bufs(st)
    ss.run(*eager, 3, **kw)
    torch.cuda.synchronize()
    cur, ce, best, be = eager
    assert bool(torch.isfinite(be).all()) and bool((be <= ce + 1e-6).all())
    torch.testing.assert_close(kp.energy_grad(cl, best)[0], be, rtol=1e-4, atol=1e-4)
    again = _bufs(st)
    ss.run(*again, 3, **kw)
    torch.cuda.synchronize()
    for a, b in zip(eager, again):
        assert torch.equal(a, b)
    bufs = _bufs(st)
    g = ss.capture(*bufs, 3, **kw)
    bufs[0].copy_(st); bufs[1].zero_(); bufs[2].copy_(st); bufs[3].fill_(float("inf"))
    torch.cuda.synchronize()
    g.replay()
    torch.cuda.synchronize()
    for a, b in zip(eager, bufs):
        assert torch.equal(a, b)


def test_learned_features_match_numpy(learned_full):
    """After two steps, the features of 512 sampled chains' current poses (one per ligand) against the numpy
    restatement; poses with an atom energy within 2e-5 kcal/mol of a histogram edge (but not on it: the float32 and
    FP64 energies differ by ~1e-6) or whose FP64 pose is within 2e-5 A of a discontinuity are excluded and counted."""
    print("[gpu] learned proposal features vs numpy", flush=True)
    w, prob, kp, st, cl, ss = learned_full
    b = _bufs(st)
    ss.run(*b, 2, seed=33, use_net=True, p_plain=0.25, features=True)
    torch.cuda.synchronize()
    cl_h = cl.cpu().numpy()
    idx = np.array([np.flatnonzero(cl_h == l)[5] for l in range(512)])
    cur = b[0].cpu().numpy()[idx]
    feat = ss.feat.cpu().numpy()[idx]
    AM, KM, lig, grid, sf, _ = workload.host_arrays(prob)
    op = _oracle_of(prob)
    e, g, X, m = op.energy_grad(cl_h[idx], cur)
    lo, hi = grid["lo"][0].astype(np.float64), grid["hi"][0].astype(np.float64)
    checked = 0
    for r, (i, l) in enumerate(zip(idx, cl_h[idx])):
        na = int(lig["na"][l])
        ea = _atom_energies(grid, lig, l, X[r, :na])
        d = np.abs(ea[:, None] - EDGES[None, :])
        # an atom exactly on an edge (zero energy far from the receptor) falls in the same bin on both sides
        if m[r] < MARGIN or ((d < 2e-5) & (d > 0)).any():
            continue
        want = _features(cur[r].astype(np.float64), e[r], g[r], ea, lig, l, lo, hi)
        np.testing.assert_allclose(feat[r], want, rtol=0, atol=2e-3, err_msg=f"chain {i} ligand {l}")
        checked += 1
    print(f"  checked {checked} of {len(idx)} chains", flush=True)
    assert checked > 0.8 * len(idx)
