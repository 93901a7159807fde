"""Speculative line search (several waves per chain evaluate consecutive line-search trials in parallel): accepted
points, energies and evaluation counts are bit-identical to the one-wave sequential search, for a single complex and
for uneven ligands in one launch."""
import numpy as np
import pytest
import torch

from adock_amd import _lib, synth, workload
from adock_amd.kernel import KernelProblem

pytestmark = pytest.mark.gpu


def _run(kp, st, cl, steps, width, seed=77):
    L = _lib.lib()
    _lib.check(L.adock_set_line_search_width(width))
    try:
        N = len(cl)
        cur, best = st.clone(), st.clone()
        ce, be = torch.zeros(N).cuda(), torch.full((N,), float("inf")).cuda()
        kp.eval_count.zero_()
        kp.run_mc(cl, torch.arange(N, dtype=torch.int32).cuda(), cur, ce, best, be, True, steps, 0, 1.2, 2.0, 2.0,
                  seed, 0, count_evals=True)
        torch.cuda.synchronize()
        return [x.cpu().numpy() for x in (cur, ce, best, be)] + [int(kp.eval_count.item())]
    finally:
        _lib.check(L.adock_set_line_search_width(0))


@pytest.mark.parametrize("width", [2, 4])
def test_single_complex_bit_identical(width):
    print(f"[gpu] speculative line search width {width}, single complex", flush=True)
    w, prob = workload.single("single", seed=0, device="cuda")
    kp = KernelProblem(prob)
    N = 48
    t, q, th = synth.start_states(3, N, w["center"], w["size"], w["lig"].n_tors, kp.KM)
    st = kp.pack(*(torch.as_tensor(a).cuda() for a in (t, q, th)))
    cl = torch.zeros(N, dtype=torch.int32).cuda()
    ref = _run(kp, st, cl, 12, 1)
    got = _run(kp, st, cl, 12, width)
    assert got[4] == ref[4]
    for a, b in zip(ref[:4], got[:4]):
        np.testing.assert_array_equal(a, b)


def test_uneven_ligands_bit_identical():
    print("[gpu] speculative line search, uneven ligands", flush=True)
    w, prob = workload.screen(seed=2, n_ligands=12, chains=6, steps=8)
    kp = KernelProblem(prob)
    L = len(w["ligs"])
    st, cl = workload.chain_starts(2, w["ligs"], [w["center"]] * L, [w["size"]] * L, w["chains"], kp.KM)
    st, cl = torch.as_tensor(st).cuda(), torch.as_tensor(cl).cuda()
    ref = _run(kp, st, cl, w["steps"], 1)
    got = _run(kp, st, cl, w["steps"], 4)
    assert got[4] == ref[4]
    for a, b in zip(ref[:4], got[:4]):
        np.testing.assert_array_equal(a, b)
