"""Edge cases on the GPU: grids of an empty receptor (both builders), the larger torsion buckets (KM = 24, 32) against
the FP64 oracle, MC invariants on a 31-torsion ligand, the speculative search bit-identical where the LDS budget
lowers its width (64 atoms, 1,761 pairs), an empty chain list, and exact mode with its receptor neighbour lists staged
in LDS and walked through per-atom masks, staged and walked whole, and read from memory (all bit-identical)."""
import numpy as np
import pytest
import torch

from adock_amd import _lib, synth, workload
from adock_amd.config import sf_params
from adock_amd.fit.gridfit import build_grid_fit
from adock_amd.fit.terms import FittedPotential, Spec, default_weights
from adock_amd.grid import build_grid
from adock_amd.kernel import KernelProblem
from helpers import CFG, EG_FLOOR, EG_REL, SPACING, rel_err, states, vec_rel
from test_gpu_parity import MARGIN, _oracle_of

pytestmark = pytest.mark.gpu


def test_empty_receptor_grids_are_zero():
    print("[gpu] empty receptor grids", flush=True)
    sfp = sf_params(CFG, "vina")
    g = build_grid(np.zeros((0, 3), np.float32), np.zeros(0, np.int32), sfp, np.zeros(3), SPACING, (9, 7, 5),
                   lig_types=np.arange(10))
    assert g.shape == (10, 9, 7, 5) and float(g.abs().max()) == 0.0
    spec = Spec(CFG, "vina")
    pot = FittedPotential(spec, default_weights(spec), np.arange(11), "cuda")
    gf = build_grid_fit(np.zeros((0, 3), np.float32), np.zeros(0, np.int32), pot, np.zeros(3), SPACING, (9, 7, 5))
    assert gf.shape == (11, 9, 7, 5) and float(gf.abs().max()) == 0.0


@pytest.mark.parametrize("n_heavy,n_tors,KM", [(50, 20, 24), (64, 31, 32)])
def test_eg_large_torsion_buckets(n_heavy, n_tors, KM):
    print(f"[gpu] eg {n_heavy} atoms / {n_tors} torsions (KM {KM})", flush=True)
    rec = workload.receptor_dict(synth.make_receptor(3, 1500, 20.0, 6.0), np.zeros(3), 24.0)
    lig = synth.make_ligand(101, n_heavy, n_tors)
    assert lig.n_tors == n_tors
    prob = workload.gpu_problem([rec], [lig], [0])
    kp, op = KernelProblem(prob), _oracle_of(prob)
    assert kp.KM == KM
    N = 64
    st = states(77, N, np.zeros(3), 20.0, lig, kp.KM, sigma=2.0)
    cl = np.zeros(N, np.int32)
    e, g = kp.energy_grad(torch.as_tensor(cl).cuda(), torch.as_tensor(st).cuda())
    e, g = e.cpu().numpy(), g.cpu().numpy()
    eo, go, _, m = op.energy_grad(cl, st)
    keep = m >= MARGIN
    assert keep.sum() > N // 2
    D = 6 + n_tors
    assert rel_err(e[keep], eo[keep], EG_FLOOR, EG_REL) < EG_REL
    assert vec_rel(g[keep, :D], go[keep, :D], EG_FLOOR) < EG_REL


def _mc(kp, st, cl, steps, width):
    L = _lib.lib()
    _lib.check(L.adock_set_line_search_width(width))
    try:
        N = len(cl)
        cur, best = st.clone(), st.clone()
        ce, be = torch.zeros(N).cuda(), torch.full((N,), float("inf")).cuda()
        kp.eval_count.zero_()
        kp.run_mc(cl, torch.arange(N, dtype=torch.int32).cuda(), cur, ce, best, be, True, steps, 0, 1.2, 2.0, 2.0, 5,
                  0, count_evals=True)
        torch.cuda.synchronize()
        return [x.cpu().numpy() for x in (cur, ce, best, be)] + [int(kp.eval_count.item())]
    finally:
        _lib.check(L.adock_set_line_search_width(0))


def test_mc_31_torsions_widths_and_invariants():
    """64 heavy atoms, 31 torsions (KM 32), 1,761 pairs: the dense layout's LDS lowers the speculative width; every
    width gives the same chains bit for bit, best <= every start energy, and best_e re-evaluates to itself."""
    print("[gpu] mc 64 atoms / 31 torsions", flush=True)
    rec = workload.receptor_dict(synth.make_receptor(3, 1500, 20.0, 6.0), np.zeros(3), 24.0)
    lig = synth.make_ligand(101, 64, 31)
    prob = workload.gpu_problem([rec], [lig], [0])
    kp = KernelProblem(prob)
    N = 16
    st = torch.as_tensor(states(5, N, np.zeros(3), 20.0, lig, kp.KM, sigma=2.0)).cuda()
    cl = torch.zeros(N, dtype=torch.int32).cuda()
    e0 = kp.energy_grad(cl, st)[0].cpu().numpy()
    ref = _mc(kp, st, cl, 4, 1)
    for width in (2, 4):
        got = _mc(kp, st, cl, 4, width)
        assert got[4] == ref[4]
        for a, b in zip(ref[:4], got[:4]):
            np.testing.assert_array_equal(a, b)
    best, best_e = ref[2], ref[3]
    assert np.all(best_e <= e0 + 1e-4)
    eb = kp.energy_grad(cl, torch.as_tensor(best).cuda())[0].cpu().numpy()
    np.testing.assert_allclose(eb, best_e, rtol=1e-5, atol=1e-4)


def test_empty_chain_list_is_a_no_op():
    w, prob = workload.single("small", seed=0)
    kp = KernelProblem(prob)
    z = torch.zeros(0, kp.SD, device="cuda")
    ze = torch.zeros(0, device="cuda")
    zi = torch.zeros(0, dtype=torch.int32, device="cuda")
    kp.run_mc(zi, zi, z, ze, z.clone(), ze.clone(), True, 5, 0, 1.2, 2.0, 2.0, 0, 0)
    e, g = kp.energy_grad(zi, z)
    assert e.shape == (0,) and g.shape == (0, kp.DM)


@pytest.mark.parametrize("width", [1, 4])
def test_exact_staging_bit_identical(width):
    """Exact mode over receptor neighbour lists: with the lists staged in LDS and walked through per-atom masks
    (default), staged and walked whole, and read from memory, the chains end bit for bit the same, in the one-wave
    kernel and in the speculative one; enough MC steps that atoms move past the masks' margin and they are rebuilt."""
    print(f"[gpu] exact mode, masked / staged / memory neighbour lists, width {width}", flush=True)
    w, prob = workload.single("small", seed=0)
    kp = KernelProblem(prob)
    N = 24
    st = torch.as_tensor(states(9, N, w["center"], 8.0, w["lig"], kp.KM, sigma=1.0)).cuda()
    cl = torch.zeros(N, dtype=torch.int32).cuda()
    nbr, nbr_n = kp.receptor_neighbours(prob, cl, st, prob.lb.K, 2.0)
    assert 0 < nbr.shape[1] <= 1024
    L = _lib.lib()

    def run(staging, steps, nb):
        _lib.check(L.adock_set_exact_staging(staging))
        _lib.check(L.adock_set_line_search_width(width))
        try:
            cur, best = st.clone(), st.clone()
            ce, be = torch.zeros(N).cuda(), torch.full((N,), float("inf")).cuda()
            kw = dict(nbr=nbr, nbr_n=nbr_n) if nb else {}
            kp.run_mc(cl, torch.arange(N, dtype=torch.int32).cuda(), cur, ce, best, be, True, steps, 15, 1.2, 2.0,
                      2.0, 3, 0, exact=True, **kw)
            torch.cuda.synchronize()
            return [x.cpu().numpy() for x in (cur, ce, best, be)]
        finally:
            _lib.check(L.adock_set_exact_staging(2))
            _lib.check(L.adock_set_line_search_width(0))

    for steps, nb in ((2, True), (6, True), (2, False)):
        a, b, c = run(2, steps, nb), run(1, steps, nb), run(0, steps, nb)
        for x, y, z in zip(a, b, c):
            np.testing.assert_array_equal(x, y)
            np.testing.assert_array_equal(x, z)
        assert np.isfinite(a[3]).all()


@pytest.mark.parametrize("shuffle", [False, True])
def test_energy_kernel_many_states_per_wave(shuffle):
    """From 16 x CUs x 8 states, energy_grad evaluates 8 consecutive states per wave and sets a ligand up again only
    when it changes: the results equal those of 4,096-state calls (one state per wave) bit for bit, with the states of
    20 ligands in ligand order and shuffled (a new ligand at almost every state)."""
    print(f"[gpu] energy kernel, several states per wave (shuffled {shuffle})", flush=True)
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    w, prob = workload.screen(seed=5, n_ligands=20, chains=8, steps=1)
    kp = KernelProblem(prob)
    L = len(w["ligs"])
    per = max(40_000, 16 * cus * 8 + 100) // L + 1
    st, cl = workload.chain_starts(5, w["ligs"], [w["center"]] * L, [w["size"]] * L, per, kp.KM)
    if shuffle:
        perm = np.random.default_rng(0).permutation(len(cl))
        st, cl = st[perm], cl[perm]
    st, cl = torch.as_tensor(st).cuda(), torch.as_tensor(cl).cuda()
    N = len(cl)
    assert N >= 16 * cus * 8
    e, g = kp.energy_grad(cl, st)
    parts = [kp.energy_grad(cl[i:i + 4096], st[i:i + 4096]) for i in range(0, N, 4096)]
    e1, g1 = torch.cat([p[0] for p in parts]), torch.cat([p[1] for p in parts])
    torch.cuda.synchronize()
    np.testing.assert_array_equal(e.cpu().numpy(), e1.cpu().numpy())
    np.testing.assert_array_equal(g.cpu().numpy(), g1.cpu().numpy())
    assert np.isfinite(e.cpu().numpy()).all()


def test_exact_polish_neighbour_list_matches_full():
    """The reference's edge case (tests/test_mckernel.py:150-175): 16 poses near the box centre, pre-relaxed on the
    grid (100 BFGS iterations), then polished in exact mode (50 iterations) once over the whole receptor and once over
    per-chain neighbour lists (cutoff + 2 A); the polish moves poses little, so both reach the same exact energies
    within 0.05 kcal/mol, and the lists are shorter than the receptor."""
    print("[gpu] exact polish: neighbour lists vs the whole receptor", flush=True)
    from adock_amd.pose import random_quat
    w, prob = workload.single("small", seed=0)
    kp = KernelProblem(prob)
    lb = prob.lb
    g = torch.Generator(device="cuda").manual_seed(4)
    M = 16
    c = torch.as_tensor(w["center"], dtype=torch.float32, device="cuda")
    st = kp.pack(c + torch.randn((M, 3), generator=g, device="cuda") * 0.5, random_quat((M,), g, "cuda"),
                 torch.zeros(M, lb.K, device="cuda"))
    cl, ci = torch.zeros(M, dtype=torch.int32, device="cuda"), torch.arange(M, dtype=torch.int32, device="cuda")
    b0, e0 = st.clone(), torch.full((M,), float("inf"), device="cuda")
    kp.run_mc(cl, ci, st, torch.zeros(M, device="cuda"), b0, e0, True, 0, 100, 1.2, 2.0, 2.0, 0, 0)
    out = []
    for use_nb in (False, True):
        cur, best = b0.clone(), b0.clone()
        be = torch.full((M,), float("inf"), device="cuda")
        nb = kp.receptor_neighbours(prob, cl, cur, lb.K, 2.0) if use_nb else (None, None)
        kp.run_mc(cl, ci, cur, torch.zeros(M, device="cuda"), best, be, True, 0, 50, 1.2, 2.0, 2.0, 0, 0,
                  exact=True, nbr=nb[0], nbr_n=nb[1])
        out.append(kp.energy_grad(cl, best, exact=True)[0])
        if use_nb:
            assert int(nb[1].max()) < int(prob.rec[2].sum())
    torch.cuda.synchronize()
    print(f"  max |diff| {float((out[1] - out[0]).abs().max()):.2e}", flush=True)
    torch.testing.assert_close(out[1], out[0], rtol=0, atol=0.05)
