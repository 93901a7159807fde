"""adock_mc_batch_noalloc (the allocation-free batched launch, queue workspace in the library's static device array)
against adock_mc_batch with caller-allocated workspace: two calls in a row, the second with fewer chains, queues and
complexes than the first, bit-identical states, energies and per-complex finished-chain stamps; and a call with too
many complexes is refused with ADOCK_E_ARG."""
import ctypes

import numpy as np
import pytest
import torch

from adock_amd import _lib, workload
from adock_amd.batch import BatchRunner, queue_partition
from adock_amd.kernel import KernelProblem

pytestmark = pytest.mark.gpu


def _noalloc(kp, cl, queues, st, steps, seed, times=True):
    """One adock_mc_batch_noalloc call on chains cl (device int32) with `queues` work queues."""
    dev = kp.device
    N = int(cl.numel())
    L = kp.L
    cl_h = cl.cpu().numpy()
    order, qstart = queue_partition(np.asarray(kp.cost)[cl_h], kp.grid["rec_idx"].cpu().numpy()[cl_h], queues)
    order, qstart = torch.as_tensor(order, device=dev), torch.as_tensor(qstart, device=dev)
    lig_chains = torch.as_tensor(np.bincount(cl_h, minlength=L).astype(np.int32), device=dev)
    tm = torch.zeros(1 + L, dtype=torch.int64, device=dev) if times else None
    cur, best = st.clone(), st.clone()
    ce, be = torch.zeros(N, device=dev), torch.full((N,), float("inf"), device=dev)
    ci = torch.arange(N, dtype=torch.int32, device=dev)
    args = (kp._args(False) + [_lib.ptr(cl), _lib.ptr(ci), N, None, None, 0, _lib.ptr(cur), _lib.ptr(ce),
                               _lib.ptr(best), _lib.ptr(be), 1, steps, 0, 1.2, 2.0, 2.0, seed, 0, _lib.ptr(order),
                               _lib.ptr(qstart), len(qstart) - 1, L, _lib.ptr(lig_chains), _lib.ptr(tm), 0, 0,
                               None, None, _lib.stream_of(dev)])
    _lib.check(_lib.lib().adock_mc_batch_noalloc(*args))
    torch.cuda.synchronize()
    return [x.cpu().numpy() for x in (cur, ce, best, be)], (None if tm is None else tm.cpu().numpy())


def _with_workspace(kp, cl, queues, st, steps, seed):
    N = int(cl.numel())
    r = BatchRunner(kp, cl, n_complexes=kp.L, queues=queues)
    cur, best = st.clone(), st.clone()
    ce, be = torch.zeros(N, device=kp.device), torch.full((N,), float("inf"), device=kp.device)
    r.launch(torch.arange(N, dtype=torch.int32, device=kp.device), cur, ce, best, be, True, steps, 0, 1.2, 2.0, 2.0,
             seed, 0)
    torch.cuda.synchronize()
    return [x.cpu().numpy() for x in (cur, ce, best, be)], r


def test_noalloc_matches_workspace_launch_twice():
    print("[gpu] adock_mc_batch_noalloc vs adock_mc_batch, two calls (second smaller)", flush=True)
    w, prob = workload.screen(seed=3, n_ligands=20, chains=8, steps=5)
    kp = KernelProblem(prob)
    L = len(w["ligs"])
    st, cl = workload.chain_starts(3, w["ligs"], [w["center"]] * L, [w["size"]] * L, w["chains"], kp.KM)
    st, cl = torch.as_tensor(st, device="cuda"), torch.as_tensor(cl, device="cuda")
    # call 1: all 160 chains over 8 queues; call 2: the chains of the first 7 complexes over 3 queues
    for n_lig, queues, seed in ((L, 8, 77), (7, 3, 78)):
        n = n_lig * w["chains"]
        want, r = _with_workspace(kp, cl[:n], queues, st[:n], w["steps"], seed)
        got, tm = _noalloc(kp, cl[:n], queues, st[:n], w["steps"], seed)
        for a, b in zip(want, got):
            np.testing.assert_array_equal(a, b)
        # completion stamps: every complex of this call stamped after the first wave start, the others zero
        assert tm[0] > 0 and (tm[1:1 + n_lig] >= tm[0]).all()
        assert (tm[1 + n_lig:] == 0).all()
    # the same call again without completion stamps
    got, _ = _noalloc(kp, cl[:48], 1, st[:48], w["steps"], 79, times=False)
    want, _ = _with_workspace(kp, cl[:48], 1, st[:48], w["steps"], 79)
    for a, b in zip(want, got):
        np.testing.assert_array_equal(a, b)


def test_noalloc_refuses_sizes_beyond_its_array():
    print("[gpu] adock_mc_batch_noalloc limits", flush=True)
    w, prob = workload.screen(seed=3, n_ligands=2, chains=2, steps=1)
    kp = KernelProblem(prob)
    N = 4
    st = torch.zeros(N, kp.SD, device="cuda")
    st[:, 3] = 1.0
    cl = torch.tensor([0, 0, 1, 1], dtype=torch.int32, device="cuda")
    f = torch.zeros(N, device="cuda")
    lc = torch.full((70000,), 2, dtype=torch.int32, device="cuda")
    tm = torch.zeros(70001, dtype=torch.int64, device="cuda")
    base = kp._args(False) + [_lib.ptr(cl), _lib.ptr(cl), N, None, None, 0, _lib.ptr(st), _lib.ptr(f), _lib.ptr(st),
                              _lib.ptr(f), 1, 1, 0, 1.2, 2.0, 2.0, 1, 0, None, None]
    L = _lib.lib()
    assert L.adock_mc_batch_noalloc(*base, 1, 65537, _lib.ptr(lc), _lib.ptr(tm), 0, 0, None, None,
                                    _lib.stream_of(kp.device)) == -1
    assert b"65536" in L.adock_last_error()
    assert L.adock_mc_batch_noalloc(*base, 1025, 2, _lib.ptr(lc), _lib.ptr(tm), 0, 0, None, None,
                                    _lib.stream_of(kp.device)) == -1
    torch.cuda.synchronize()
