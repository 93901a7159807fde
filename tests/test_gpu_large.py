"""Ligands beyond 64 heavy atoms (two atoms per lane: AM 96 and 128) and beyond 32 torsions (64-bit torsion masks:
KM 40, 48, 56) against the FP64 oracle: per-pose energies and gradients (grid and exact mode), one BFGS iteration, MC
invariants, Stage A argmins, and the persistent queued launch bit-identical to the per-chain launch at AM 128. The
reference packs any ligand size (AM = 64 up to 64 heavy atoms, else the count rounded up to 32, KM = the torsion count
rounded up to 8; src/adock/mckernel.py:553-554)."""
import math

import numpy as np
import pytest
import torch

from adock_amd import synth, workload
from adock_amd.batch import BatchRunner
from adock_amd.kernel import KernelProblem
from adock_amd.pose import internal_coords, random_quat
from helpers import EG_FLOOR, EG_REL, OracleKernel, rel_err, states, vec_rel
from test_gpu_parity import MARGIN, _oracle_of, _run_both

pytestmark = pytest.mark.gpu

# (heavy atoms, torsions asked of the generator, ligand seed) -> buckets (AM, KM)
CASES = {
    "80 atoms / 12 torsions (AM 96)": (80, 12, 780, 96, 16),
    "120 atoms / 20 torsions (AM 128)": (120, 20, 820, 128, 24),
    "64 atoms / 40 torsions (KM 40)": (64, 40, 940, 64, 40),
    "110 atoms / 45 torsions (AM 128, KM 48)": (110, 45, 810, 128, 48),
    "128 atoms / 50 torsions (AM 128, KM 56)": (128, 50, 828, 128, 56),
}


def _problem(case, box=30.0, with_rec=True):
    n, t, seed, AM, KM = CASES[case]
    rec = workload.receptor_dict(synth.make_receptor(3, 1500, 20.0, 6.0), np.zeros(3), box)
    lig = synth.make_ligand(seed, n, t)
    prob = workload.gpu_problem([rec], [lig], [0], with_rec=with_rec)
    kp = KernelProblem(prob)
    assert (kp.AM, kp.KM) == (AM, KM), (kp.AM, kp.KM)
    return rec, lig, prob, kp


@pytest.mark.parametrize("case", sorted(CASES))
def test_energy_gradient_large_buckets(case):
    print(f"[gpu] eg {case}", flush=True)
    rec, lig, prob, kp = _problem(case)
    op = _oracle_of(prob)
    N = 96
    st = states(11, N, np.zeros(3), 22.0, lig, kp.KM, sigma=2.5)
    cl = np.zeros(N, np.int32)
    for exact in (False, True):
        e, g = kp.energy_grad(torch.as_tensor(cl).cuda(), torch.as_tensor(st).cuda(), exact=exact)
        e, g = e.cpu().numpy(), g.cpu().numpy()
        eo, go, _, m = op.energy_grad(cl, st, exact=exact)
        keep = m >= MARGIN
        assert keep.sum() > N // 2
        D = 6 + lig.n_tors
        assert rel_err(e[keep], eo[keep], EG_FLOOR, EG_REL) < EG_REL, exact
        assert vec_rel(g[keep, :D], go[keep, :D], EG_FLOOR) < EG_REL, exact


@pytest.mark.parametrize("case", sorted(CASES))
def test_bfgs_one_iteration_large_buckets(case):
    """One BFGS iteration from poses near the pocket against the FP64 oracle: >= 99 % of 128 chains within 1e-3."""
    print(f"[gpu] bfgs 1 iteration {case}", flush=True)
    rec, lig, prob, kp = _problem(case)
    op = _oracle_of(prob)
    st = states(12, 128, np.zeros(3), 22.0, lig, kp.KM, sigma=1.5)
    out = _run_both(kp, op, st, True, 0, 1)
    eg, ec = out["gpu"][3], out["cpu"][3]
    ok = np.abs(eg - ec) <= 1e-3 * np.maximum(np.abs(ec), 1)
    print(f"  within 1e-3: {ok.mean():.4f}", flush=True)
    assert ok.mean() >= 0.99, ok.mean()


@pytest.mark.parametrize("case", ["120 atoms / 20 torsions (AM 128)", "110 atoms / 45 torsions (AM 128, KM 48)"])
def test_mc_invariants_large_buckets(case):
    """Initial BFGS + 6 MC steps on 64 chains: best <= start energy after the initial optimisation, best_e
    re-evaluates to itself, the run is deterministic, and a run split in two segments equals the unsplit one."""
    print(f"[gpu] mc invariants {case}", flush=True)
    rec, lig, prob, kp = _problem(case)
    N = 64
    st = torch.as_tensor(states(13, N, np.zeros(3), 22.0, lig, kp.KM, sigma=2.0)).cuda()
    cl = torch.zeros(N, dtype=torch.int32).cuda()
    ci = torch.arange(N, dtype=torch.int32).cuda()

    def run(splits):
        cur, best = st.clone(), st.clone()
        ce, be = torch.zeros(N).cuda(), torch.full((N,), float("inf")).cuda()
        done = 0
        for k, n in enumerate(splits):
            kp.run_mc(cl, ci, cur, ce, best, be, k == 0, n, 0, 1.2, 2.0, 2.0, 77, done)
            done += n
        torch.cuda.synchronize()
        return cur, ce, best, be

    cur, ce, best, be = run([6])
    again = run([6])
    split = run([2, 4])
    for a, b, c in zip((cur, ce, best, be), again, split):
        assert torch.equal(a, b) and torch.equal(a, c)
    e_best, _ = kp.energy_grad(cl, best)
    assert torch.allclose(e_best, be, rtol=1e-4, atol=1e-4)
    assert bool(torch.isfinite(be).all()) and bool((be <= ce + 1e-6).all())


def test_stage_a_large_ligands():
    """Stage A at 80 and 120 heavy atoms (conformers from the ligand's torsions, random rotations, a lattice) against
    the FP64 oracle: argmin agreement >= 99 %, energies within 1e-4 relative where they agree."""
    print("[gpu] stage A 80 / 120 atoms", flush=True)
    rec = workload.receptor_dict(synth.make_receptor(3, 1500, 20.0, 6.0), np.zeros(3), 30.0)
    ligs = [synth.make_ligand(780, 80, 12), synth.make_ligand(820, 120, 20)]
    prob = workload.gpu_problem([rec], ligs, [0, 0], with_rec=False)
    kp, ok = KernelProblem(prob), OracleKernel(prob)
    assert kp.AM == 128
    lb = prob.lb
    gen = torch.Generator(device="cuda").manual_seed(5)
    th = (torch.rand((lb.L, 4, lb.K), generator=gen, device="cuda") * 2 - 1) * math.pi * lb.tvalid[:, None, :]
    X = internal_coords(lb, th)
    hidx = torch.zeros(lb.L, kp.AM, dtype=torch.long, device="cuda")
    for l, p in enumerate(lb.preps):
        hv = torch.as_tensor(np.flatnonzero(p.heavy), device="cuda")
        hidx[l, :len(hv)] = hv
    conf = torch.gather(X, 2, hidx[:, None, :, None].expand(lb.L, 4, kp.AM, 3)).contiguous()
    quats = random_quat((24,), gen, "cuda")
    ax = torch.arange(-6.0, 6.01, 2.0, device="cuda")
    T = torch.stack(torch.meshgrid(ax, ax, ax, indexing="ij"), -1).reshape(-1, 3)
    trans = T[None].expand(lb.L, -1, -1).contiguous()
    nT = torch.full((lb.L,), T.shape[0], dtype=torch.int32, device="cuda")
    be, bi = kp.stage_a(conf, quats, trans, nT, 0.3)
    bo, io = ok.stage_a(conf, quats, trans, nT, 0.3)
    be, bi, bo, io = be.cpu().numpy(), bi.cpu().numpy(), bo.numpy(), io.numpy()
    same = bi == io
    print(f"  argmin agrees on {same.mean():.4f} of {same.size} rows", flush=True)
    assert same.mean() >= 0.99
    assert np.abs(be - bo)[same].max() <= 1e-4 * (1 + np.abs(bo[same]).max())


def test_persistent_launch_large_bucket_bit_identical():
    """A batch whose largest ligand has 120 heavy atoms (every ligand then in the AM 128 bucket, two atoms per lane)
    together with 20- and 45-atom ligands: the persistent queued launch equals the per-chain launch bit for bit."""
    print("[gpu] persistent launch at AM 128", flush=True)
    rec = workload.receptor_dict(synth.make_receptor(3, 1500, 20.0, 6.0), np.zeros(3), 30.0)
    ligs = [synth.make_ligand(820, 120, 20), synth.make_ligand(31, 20, 4), synth.make_ligand(32, 45, 9)]
    prob = workload.gpu_problem([rec], ligs, [0, 0, 0], with_rec=False)
    kp = KernelProblem(prob)
    assert kp.AM == 128
    L, M = len(ligs), 16
    st, cl = workload.chain_starts(4, ligs, [np.zeros(3)] * L, [22.0] * L, M, kp.KM)
    st, cl = torch.as_tensor(st).cuda(), torch.as_tensor(cl).cuda()
    N = len(cl)
    ci = torch.arange(N, dtype=torch.int32).cuda()
    outs = []
    for queued in (False, True):
        cur, best = st.clone(), st.clone()
        ce, be = torch.zeros(N).cuda(), torch.full((N,), float("inf")).cuda()
        if queued:
            BatchRunner(kp, cl, n_complexes=L).launch(ci, cur, ce, best, be, True, 4, 0, 1.2, 2.0, 2.0, 9, 0)
        else:
            kp.run_mc(cl, ci, cur, ce, best, be, True, 4, 0, 1.2, 2.0, 2.0, 9, 0)
        torch.cuda.synchronize()
        outs.append([x.cpu().numpy() for x in (cur, ce, best, be)])
    for a, b in zip(*outs):
        np.testing.assert_array_equal(a, b)
