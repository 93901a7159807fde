"""Fitted scoring function on the GPU: adock_build_grid_fit against the FP64 oracle (random weights, every receptor
ext-type block, several passes of ligand types, non-cubic boxes), against the exact pairwise energy at grid points
(the reference's test_grid_kernel_matches_exact, tests/test_fit_features.py:103-123), against the Vina grid kernel at
the base weights, and the fitted docking driver end to end (deterministic, segment split = unsplit, outputs rescored
independently, best score close to the oracle twin's)."""
import numpy as np
import pytest
import torch

import oracle
from adock_amd import synth, workload
from adock_amd.config import sf_params
from adock_amd.fit.fdock import DATA, build_problem, dock_fitted, dock_preset, kernel_cfg, load_fitted
from adock_amd.fit.gridfit import build_grid_fit, make_gridset_fit
from adock_amd.fit.terms import (N_INTER, TERM_NAMES, FittedPotential, Spec, default_weights, inter_features,
                                 intra_features)
from adock_amd.grid import build_grid
from adock_amd.search import dock_problem_kernel
from helpers import (CFG, EG_FLOOR, EG_REL, GRID_FLOOR, GRID_REL, IDX, SPACING, OracleKernel, cpu_fitted_problem,
                     random_receptor, rel_err)

pytestmark = pytest.mark.gpu
NXS = len(CFG["_xs_names"])


def _types(rng, spec, n):
    """Random receptor ext types over every block of the spec."""
    xs = rng.choice([IDX[t] for t in ("C_H", "C_P", "N_D", "N_A", "N_DA", "O_A", "O_DA", "S_P", "Met_D")], n)
    return xs + NXS * rng.integers(0, spec.n_rt // NXS, n)


@pytest.mark.parametrize("case", ["vina_sc_cap", "vinardo_override"])
def test_fit_grid_matches_oracle(case):
    print(f"[gpu] fitted grid vs oracle {case}", flush=True)
    if case == "vina_sc_cap":
        spec = Spec(CFG, "vina", buried_threshold=1, rep_cap=1.5, sc_radius_scale=0.9, sc_radius_delta=0.05)
    else:
        spec = Spec(CFG, "vinardo", inter_prm={"g1_w": 0.7, "hb_bad": 0.4}, ion_rmin=2.5)
    rng = np.random.default_rng(7)
    w = rng.normal(size=len(TERM_NAMES)).astype(np.float32)
    lig_list = np.unique(rng.integers(0, spec.n_lt, 11))          # 2 passes of 8 types
    pot = FittedPotential(spec, w, lig_list, "cuda")
    xyz, _ = random_receptor(600, seed=3, extent=16.0)
    rext = _types(rng, spec, len(xyz))
    origin, n = np.array([2.5, 3.0, 1.75]), (23, 17, 29)
    g = build_grid_fit(xyz, rext, pot, origin, SPACING, n).cpu().numpy()
    o = oracle.build_grid_fit(xyz, rext, spec.rec_radius, pot.radius.cpu().numpy(), pot.W.cpu().numpy(),
                              spec.grid_prm(), origin, SPACING, n)
    assert g.shape == o.shape == (len(lig_list),) + n
    err = rel_err(g, o, GRID_FLOOR, GRID_REL)
    print(f"  max rel err {err:.2e}", flush=True)
    assert err < GRID_REL


def test_fit_grid_matches_exact_pairwise():
    """The reference's recipe: 40 atoms in [0,6]^3, buried / charged / metal receptor types, 4 ligand types, random
    weights, grid values at 4 points vs the pairwise FittedPotential energy (rel 1e-3 as there; here also 1e-5)."""
    print("[gpu] fitted grid vs exact pairwise", flush=True)
    spec = Spec(CFG, "vinardo", buried_threshold=1)
    rng = np.random.default_rng(0)
    w = rng.normal(size=len(TERM_NAMES)).astype(np.float32)
    lig_list = np.array([IDX["C_H"], IDX["N_D"] + NXS, IDX["O_A"] + 2 * NXS, IDX["O_A"]])
    pot = FittedPotential(spec, w, lig_list, "cuda")
    rxyz = rng.uniform(0, 6, (40, 3)).astype(np.float32)
    rext = rng.choice([IDX["C_H"], IDX["O_A"] + NXS * 5, IDX["N_D"] + NXS * 2, IDX["Met_D"], IDX["C_P"] + NXS], 40)
    G = build_grid_fit(rxyz, rext, pot, np.zeros(3), 0.5, 13, device="cuda")
    pts = [(0, 0, 0), (3, 4, 5), (12, 12, 12), (6, 2, 9)]
    P = torch.tensor([[0.5 * i, 0.5 * j, 0.5 * k] for i, j, k in pts], dtype=torch.float64)
    r = torch.cdist(P, torch.as_tensor(rxyz, dtype=torch.float64))
    rt = torch.as_tensor(rext)[None].expand_as(r) + pot.TL
    for q in range(len(lig_list)):
        e = pot.energy(r, torch.full_like(rt, q), rt).sum(1).numpy()
        for p, (i, j, k) in enumerate(pts):
            assert float(G[q, i, j, k]) == pytest.approx(e[p], rel=1e-5, abs=1e-5)


@pytest.mark.parametrize("base", ["vina", "vinardo"])
def test_fit_grid_base_weights_equal_vina_kernel(base):
    print(f"[gpu] fitted grid at base weights vs Vina kernel {base}", flush=True)
    spec = Spec(CFG, base)
    xyz, types = random_receptor(300, seed=5)
    lt = np.arange(NXS)
    pot = FittedPotential(spec, default_weights(spec), lt, "cuda")
    g_fit = build_grid_fit(xyz, types, pot, np.array([3.0, 3.0, 3.0]), SPACING, 21).cpu().numpy()
    g_vina = build_grid(xyz, types, sf_params(CFG, base), np.array([3.0, 3.0, 3.0]), SPACING, 21,
                        lig_types=lt).cpu().numpy()
    assert rel_err(g_fit, g_vina, GRID_FLOOR, GRID_REL) < GRID_REL


def test_fit_grid_rejects_bad_types():
    spec = Spec(CFG, "vina")
    pot = FittedPotential(spec, default_weights(spec), np.array([0]), "cuda")
    with pytest.raises(Exception):
        build_grid_fit(np.zeros((2, 3), np.float32), np.array([0, spec.n_rt]), pot, np.zeros(3), 0.5, 4)


def _exact_case(seed, L, M, A, n_rec, spec, lig_list):
    """Random complexes around two receptor sets of different sizes (padded slots -1), hydrogens (-1) in every ligand,
    poses spread over the receptor: (X, lig_types, rec_idx, rec_xyz, rec_ext)."""
    rng = np.random.default_rng(seed)
    N = max(n_rec)
    rx = np.zeros((len(n_rec), N, 3), np.float32)
    re = -np.ones((len(n_rec), N), np.int32)
    for s, n in enumerate(n_rec):
        xyz, _ = random_receptor(n, seed=seed + s, extent=16.0)
        rx[s, :n] = xyz
        re[s, :n] = _types(rng, spec, n)
    rec_idx = rng.integers(0, len(n_rec), L).astype(np.int32)
    lt = rng.integers(0, len(lig_list), (L, A)).astype(np.int32)
    lt[rng.random((L, A)) < 0.2] = -1
    X = rng.uniform(-2.0, 18.0, (L, M, A, 3)).astype(np.float32)
    return X, lt, rec_idx, rx, re


@pytest.mark.parametrize("L,M,A", [(3, 37, 70), (5, 8, 23)])
def test_inter_exact_fit_matches_oracle(L, M, A):
    """adock_inter_exact_fit against the FP64 oracle: pose counts not a multiple of the four poses per workgroup,
    more than 64 atom slots (two lane slices) and a small ligand, two receptor sets, random weights."""
    print(f"[gpu] exact fitted inter energy vs oracle L={L} M={M} A={A}", flush=True)
    from adock_amd.fit.gridfit import inter_exact_fit
    spec = Spec(CFG, "vinardo", buried_threshold=1, rep_cap=1.5, sc_radius_scale=0.9)
    rng = np.random.default_rng(11)
    w = rng.normal(size=len(TERM_NAMES)).astype(np.float32)
    lig_list = np.unique(rng.integers(0, spec.n_lt, 9))
    pot = FittedPotential(spec, w, lig_list, "cuda")
    X, lt, ri, rx, re = _exact_case(L + M, L, M, A, (600, 420), spec, lig_list)
    g = inter_exact_fit(torch.as_tensor(X).cuda(), lt, ri, rx, re, pot).cpu().numpy()
    o = oracle.inter_exact_fit(X, lt, ri, rx, re, spec.rec_radius, pot.radius.cpu().numpy(), pot.W.cpu().numpy(),
                               spec.grid_prm())
    assert g.shape == o.shape == (L, M) and np.isfinite(o).all()
    err = rel_err(g, o, EG_FLOOR, EG_REL)
    print(f"  max rel err {err:.2e}, |E| up to {np.abs(o).max():.1f}", flush=True)
    assert err < EG_REL
    again = inter_exact_fit(torch.as_tensor(X).cuda(), lt, ri, rx, re, pot).cpu().numpy()
    np.testing.assert_array_equal(g, again)


def test_inter_exact_fit_equals_grid_at_grid_points():
    """Atoms placed on grid points: the exact kernel's pose energy equals the sum of the fitted grid kernel's values at
    those points (the two kernels share the basis; the reference checks grid against exact the same way,
    tests/test_fit_features.py:103-123), and Problem.inter_exact on cuda poses takes the kernel."""
    print("[gpu] exact fitted inter energy vs fitted grid at grid points", flush=True)
    from adock_amd.fit.gridfit import inter_exact_fit
    spec = Spec(CFG, "vina", buried_threshold=1)
    rng = np.random.default_rng(2)
    w = rng.normal(size=len(TERM_NAMES)).astype(np.float32)
    lig_list = np.array([IDX["C_H"], IDX["N_D"] + NXS, IDX["O_A"] + 2 * NXS, IDX["O_A"]])
    pot = FittedPotential(spec, w, lig_list, "cuda")
    xyz, _ = random_receptor(500, seed=8, extent=14.0)
    rext = _types(rng, spec, len(xyz))
    n = 29
    G = build_grid_fit(xyz, rext, pot, np.zeros(3), SPACING, n).cpu().numpy()
    M, A = 16, 12
    ijk = rng.integers(0, n, (M, A, 3))
    X = (ijk * np.float32(SPACING)).astype(np.float32)[None]
    lt = rng.integers(0, len(lig_list), (1, A)).astype(np.int32)
    lt[0, 5] = -1
    e = inter_exact_fit(torch.as_tensor(X).cuda(), lt, np.array([0], np.int32), xyz[None],
                        rext[None].astype(np.int32), pot).cpu().numpy()
    want = np.array([sum(float(G[lt[0, a], i, j, k]) for a, (i, j, k) in enumerate(ijk[m]) if lt[0, a] >= 0)
                     for m in range(M)])
    np.testing.assert_allclose(e[0], want, rtol=1e-5, atol=1e-5)


def test_inter_exact_fit_raw_abi_edges():
    """Through the C ABI: a receptor set with no atoms gives zero energies, and a complex whose receptor-set index lies
    outside [0, n_sets) gets NaN from the kernel's own check (the Python wrapper rejects such an index before)."""
    print("[gpu] exact fitted inter energy: empty receptor, receptor index out of range", flush=True)
    from adock_amd import _lib
    spec = Spec(CFG, "vina")
    pot = FittedPotential(spec, default_weights(spec), np.array([0, IDX["N_D"]]), "cuda")
    L, M, A = 2, 5, 3
    X = torch.zeros(L, M, A, 3, device="cuda")
    lt = torch.zeros(L, A, dtype=torch.int32, device="cuda")
    rx = torch.zeros(1, 4, 3, device="cuda")
    re = torch.zeros(1, 4, dtype=torch.int32, device="cuda")
    prm = torch.as_tensor(np.asarray(spec.grid_prm(), np.float32), device="cuda")
    rr = torch.as_tensor(spec.rec_radius, dtype=torch.float32, device="cuda")
    W, rl = pot.W.float().contiguous(), pot.radius.float().contiguous()
    lib = _lib.lib()

    def call(ri, n_rec):
        out = torch.full((L, M), 7.0, device="cuda")
        ri_t = torch.as_tensor(ri, dtype=torch.int32, device="cuda")
        _lib.check(lib.adock_inter_exact_fit(_lib.ptr(X), _lib.ptr(lt), L, M, A, _lib.ptr(ri_t), _lib.ptr(rx),
                                             _lib.ptr(re), 1, n_rec, _lib.ptr(rr), _lib.ptr(rl), pot.TL, _lib.ptr(W),
                                             spec.n_rt, _lib.ptr(prm), _lib.ptr(out), None))
        torch.cuda.synchronize()
        return out.cpu().numpy()

    assert np.array_equal(call([0, 0], 0), np.zeros((L, M), np.float32))
    bad = call([0, 3], 4)
    assert np.isfinite(bad[0]).all() and np.isnan(bad[1]).all()


def test_problem_inter_exact_uses_the_kernel():
    """The fitted Problem's inter_exact on cuda poses (the kernel) equals the torch pairwise sum on the CPU."""
    print("[gpu] Problem.inter_exact kernel vs torch", flush=True)
    spec, w = load_fitted(DATA / "adock_fitted.json", CFG)
    cxs, _ = _complexes(spec)
    prob = build_problem(cxs, spec, w, CFG, "cuda")
    rng = np.random.default_rng(3)
    X = torch.as_tensor(rng.normal(0, 4, (2, 6, prob.lb.A, 3)).astype(np.float32)) + torch.as_tensor(
        np.asarray(cxs[0]["center"], np.float32))
    got = prob.inter_exact(X.cuda()).cpu().numpy()
    cpu = cpu_fitted_problem(cxs, spec, w)
    want = cpu.inter_exact(X.double()).numpy()
    assert rel_err(got, want, EG_FLOOR, EG_REL) < EG_REL


def _complexes(spec):
    w1 = synth.make_single("small", 0)
    rec = workload.receptor_dict(w1["receptor"], w1["center"], w1["size"])
    ligs = [w1["lig"], synth.make_ligand(31, 16, 3)]
    s = float(np.max(w1["size"]))
    return [synth.fitted_complex(rec, lg, np.asarray(w1["center"], float), s, spec, seed=i)
            for i, lg in enumerate(ligs)], s


def test_gridset_fit_layout():
    print("[gpu] fitted grid set layout", flush=True)
    spec, w = load_fitted(DATA / "adock_fitted.json", CFG)
    cxs, s = _complexes(spec)
    cxs[1] = dict(cxs[1], size=np.array([s - 4.0, s - 2.0, s - 3.0]))
    pot = FittedPotential(spec, w, np.array([0, IDX["N_D"], IDX["O_A"] + 2 * NXS]), "cuda")
    gs = make_gridset_fit(cxs, pot, SPACING, 10.0)
    n = int(np.ceil(s / SPACING)) + 1
    assert tuple(gs.G.shape) == (2, 3, n, n, n)
    for i, c in enumerate(cxs):
        o = np.asarray(c["center"]) - np.asarray(c["size"]) / 2
        np.testing.assert_allclose(gs.origin[i].cpu().numpy(), o, atol=1e-6)
        g1 = build_grid_fit(c["rec_xyz"], c["rec_ext"], pot, o, SPACING, n).cpu().numpy()
        np.testing.assert_array_equal(gs.G[i].cpu().numpy(), g1)


def test_build_problem_shares_identical_receptors():
    """Complexes with the same receptor and box share one fitted cube through rec_idx; a complex with another box gets
    its own; every complex's cube equals build_grid_fit of that complex alone."""
    print("[gpu] fitted problem: one cube per distinct receptor", flush=True)
    spec, w = load_fitted(DATA / "adock_fitted.json", CFG)
    w1 = synth.make_single("small", 0)
    rec = workload.receptor_dict(w1["receptor"], w1["center"], w1["size"])
    s = float(np.max(w1["size"]))
    c0 = np.asarray(w1["center"], float)
    ligs = [w1["lig"], synth.make_ligand(31, 16, 3), synth.make_ligand(32, 12, 2)]
    centers = [c0, c0, c0 + np.array([1.5, 0.0, -0.75])]
    cxs = [synth.fitted_complex(rec, lg, cc, s, spec, seed=i, rec_seed=5)
           for i, (lg, cc) in enumerate(zip(ligs, centers))]
    prob = build_problem(cxs, spec, w, CFG, "cuda")
    assert prob.rec_idx.cpu().tolist() == [0, 0, 1]
    assert prob.gs.R == 2 and prob.rec[0].shape[0] == 2
    n = prob.gs.G.shape[-1]
    for i, c in enumerate(cxs):
        r = int(prob.rec_idx[i])
        o = np.asarray(c["center"], float) - np.broadcast_to(np.asarray(c["size"], float), (3,)) / 2
        np.testing.assert_allclose(prob.gs.origin[r].cpu().numpy(), o, atol=1e-6)
        g1 = build_grid_fit(c["rec_xyz"], c["rec_ext"], prob.pot, o, SPACING, n).cpu().numpy()
        np.testing.assert_array_equal(prob.gs.G[r].cpu().numpy(), g1)


def _small(**kw):
    cfg = dict(CFG["search"]["v2_default"])
    cfg.update(n_chains=32, stageA_conformers=4, stageA_conformer_trials=32, stageA_rotations=64,
               stageA_translation_step=1.5, mc_steps=32, mc_segment=32, mc_size_ref=0, final_bfgs_iters=30,
               polish_top=6, polish_iters=20, min_total_chains=0, n_out=4)
    cfg.update(kw)
    return cfg


def _check_rescored(res, cxs, spec, w):
    """Each output's inter / intra / total equal the feature sums . weights (+ the out-of-box penalty)."""
    for r, c in zip(res, cxs):
        hv = c["lig"].types >= 0
        X = torch.as_tensor(r["coords"], dtype=torch.float64)
        F = inter_features(X[:, hv], torch.as_tensor(c["lig_ext"][hv]),
                           torch.as_tensor(c["rec_xyz"], dtype=torch.float64), torch.as_tensor(c["rec_ext"]), spec)
        inter = (F @ torch.as_tensor(w[:N_INTER], dtype=torch.float64)).numpy()
        intra = (intra_features(X, c["lig"].types, c["lig"].pairs, spec)
                 @ torch.as_tensor(w[N_INTER:], dtype=torch.float64)).numpy()
        lo = np.asarray(c["center"]) - np.asarray(c["size"]) / 2
        hi = np.asarray(c["center"]) + np.asarray(c["size"]) / 2
        Xh = X[:, hv].numpy()
        pen = 10.0 * ((np.clip(lo - Xh, 0, None) ** 2).sum((1, 2)) + (np.clip(Xh - hi, 0, None) ** 2).sum((1, 2)))
        np.testing.assert_allclose(r["inter"], inter + pen, rtol=1e-4, atol=3e-4)
        np.testing.assert_allclose(r["intra"], intra, rtol=1e-4, atol=3e-4)
        np.testing.assert_allclose(r["total"], inter + pen + intra, rtol=1e-4, atol=5e-4)


def test_dock_fitted_gpu():
    print("[gpu] fitted docking driver", flush=True)
    spec, w = load_fitted(DATA / "adock_fitted.json", CFG)
    spec = Spec(CFG, "vina", buried_threshold=150)
    cxs, s = _complexes(spec)
    cfg = {**CFG, "search": {**CFG["search"], "small": _small(), "small8": _small(mc_segment=8)}}
    stats, timings = {}, {}
    res = dock_fitted(cxs, spec, w, mode="small", seed=2, cfg=cfg, stats=stats, timings=timings)
    again = dock_fitted(cxs, spec, w, mode="small", seed=2, cfg=cfg)
    split = dock_fitted(cxs, spec, w, mode="small8", seed=2, cfg=cfg)
    for a, b, c in zip(res, again, split):
        np.testing.assert_array_equal(a["total"], b["total"])
        np.testing.assert_array_equal(a["coords"], c["coords"])
    assert stats["chain_steps"] == 2 * 32 * 32 and timings["n"] == 2
    for r, c in zip(res, cxs):
        assert r["coords"].shape[1:] == (c["lig"].n, 3) and np.all(np.diff(r["total"]) >= 0)
        assert r["total"][0] < 0 and np.isfinite(r["affinity"]).all()
    _check_rescored(res, cxs, spec, w)
    # the same driver on the oracle twin (host grids from the FP64 oracle, its own Stage A draws)
    cres = dock_problem_kernel(cpu_fitted_problem(cxs, spec, w), [c["center"] for c in cxs],
                               [c["size"] for c in cxs], kernel_cfg(_small()), seed=2, kernel_problem=OracleKernel)
    for r, cr in zip(res, cres):
        assert abs(r["total"][0] - cr["total"][0]) <= 1.5, (r["total"][0], cr["total"][0])


def test_dock_preset_v2_fitted():
    """The packaged v2_fitted preset (fitted JSON on the v2_default search), shortened."""
    print("[gpu] v2_fitted preset", flush=True)
    spec, w = load_fitted(DATA / "adock_fitted.json", CFG)
    cxs, _ = _complexes(spec)
    cfg = {**CFG, "search": {**CFG["search"], "v2_default": _small(n_chains=16, mc_steps=16, mc_segment=16)}}
    res = dock_preset(cxs[:1], mode="v2_fitted", seed=0, cfg=cfg)
    assert len(res) == 1 and res[0]["total"][0] < 0
    _check_rescored(res, cxs[:1], spec, w)
