"""The benchmarked workload at its full size, through properties that hold at any size (the FP64 oracle would need
minutes for it): 256 chains x 1,000 MC steps with Vina-rule BFGS on the 4,000-atom single-complex receptor (35-atom
ligand, 8 torsions, 71^3 x 8 grids), the start states and seed of bench.py.

- best_e of every chain re-evaluates to itself (energy kernel on the best states), and so does cur_e;
- best_e <= the energy of the optimised start state, best_e <= cur_e;
- two runs are bit-identical; 2 x 500-step segments equal one 1,000-step launch bit for bit;
- the speculative line search (4 waves per chain, the benchmarked kernel) equals the one-wave search bit for bit,
  evaluation counts included.
"""
import numpy as np
import pytest
import torch

from adock_amd import _lib, synth, workload
from adock_amd.kernel import KernelProblem

pytestmark = pytest.mark.gpu
SEED64 = 0 * 1_000_003 + 12345


@pytest.fixture(scope="module")
def full():
    w, prob = workload.single("single", seed=0, device="cuda")
    kp = KernelProblem(prob)
    N = w["chains"]
    assert (N, w["steps"]) == (256, 1000)
    t, q, th = synth.start_states(1, N, w["center"], w["size"], w["lig"].n_tors, kp.KM)
    st0 = kp.pack(*(torch.as_tensor(a).cuda() for a in (t, q, th)))
    return w, kp, st0


def _run(kp, st0, steps_list, bfgs_iters, width=0):
    L = _lib.lib()
    _lib.check(L.adock_set_line_search_width(width))
    try:
        N = st0.shape[0]
        cl = torch.zeros(N, dtype=torch.int32).cuda()
        ci = torch.arange(N, dtype=torch.int32).cuda()
        cur, best = st0.clone(), st0.clone()
        ce, be = torch.zeros(N).cuda(), torch.full((N,), float("inf")).cuda()
        kp.eval_count.zero_()
        off = 0
        for i, s in enumerate(steps_list):
            kp.run_mc(cl, ci, cur, ce, best, be, i == 0, s, bfgs_iters, 1.2, 2.0, 2.0, SEED64, off,
                      count_evals=True)
            off += s
        torch.cuda.synchronize()
        return [x.cpu().numpy() for x in (cur, ce, best, be)] + [int(kp.eval_count.item())]
    finally:
        _lib.check(L.adock_set_line_search_width(0))


def test_full_size_invariants(full):
    w, kp, st0 = full
    print("[gpu] full-size single complex: 256 chains x 1000 steps, invariants", flush=True)
    cur, ce, best, be, nev = _run(kp, st0, [w["steps"]], w["bfgs_iters"])
    N = st0.shape[0]
    cl = torch.zeros(N, dtype=torch.int32).cuda()
    eb = kp.energy_grad(cl, torch.as_tensor(best).cuda())[0].cpu().numpy()
    ec = kp.energy_grad(cl, torch.as_tensor(cur).cuda())[0].cpu().numpy()
    np.testing.assert_allclose(eb, be, rtol=1e-5, atol=1e-4)
    np.testing.assert_allclose(ec, ce, rtol=1e-5, atol=1e-4)
    assert np.all(be <= ce + 1e-6) and np.isfinite(be).all()
    # the start states after the initial optimisation only (no MC step) bound best_e from above
    _, e_init, _, _, _ = _run(kp, st0, [0], w["bfgs_iters"])
    assert np.all(be <= e_init + 1e-6)
    print(f"  best_e mean {be.mean():.3f} min {be.min():.3f}; evaluations {nev}", flush=True)
    again = _run(kp, st0, [w["steps"]], w["bfgs_iters"])
    for a, b in zip((cur, ce, best, be, nev), again):
        np.testing.assert_array_equal(a, b)


def test_full_size_split_and_speculative(full):
    w, kp, st0 = full
    print("[gpu] full-size single complex: 2 x 500 segments and one-wave search vs the benchmarked launch", flush=True)
    ref = _run(kp, st0, [w["steps"]], w["bfgs_iters"])
    split = _run(kp, st0, [w["steps"] // 2, w["steps"] // 2], w["bfgs_iters"])
    assert split[4] == ref[4]
    for a, b in zip(ref[:4], split[:4]):
        np.testing.assert_array_equal(a, b)
    seq = _run(kp, st0, [w["steps"]], w["bfgs_iters"], width=1)
    assert seq[4] == ref[4]
    for a, b in zip(ref[:4], seq[:4]):
        np.testing.assert_array_equal(a, b)
