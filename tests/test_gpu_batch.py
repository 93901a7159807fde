"""Batched persistent launch (adock_mc_batch): bit-identical to the per-chain launch, HIP-graph replays reproduce the
eager result, per-complex completion stamps are consistent, the autotuner returns a valid shape."""
import numpy as np
import pytest
import torch

from adock_amd import workload
from adock_amd.autotune import Autotuner
from adock_amd.batch import BatchRunner, GraphedBatch
from adock_amd.kernel import KernelProblem

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small_screen():
    w, prob = workload.screen(seed=1, n_ligands=24, chains=8, steps=6)
    kp = KernelProblem(prob)
    L = len(w["ligs"])
    st, cl = workload.chain_starts(1, w["ligs"], [w["center"]] * L, [w["size"]] * L, w["chains"], kp.KM)
    return w, kp, torch.as_tensor(st).cuda(), torch.as_tensor(cl).cuda()


def _run_plain(kp, st, cl, steps):
    N = len(cl)
    cur, best = st.clone(), st.clone()
    ce, be = torch.zeros(N).cuda(), torch.full((N,), float("inf")).cuda()
    kp.run_mc(cl, torch.arange(N, dtype=torch.int32).cuda(), cur, ce, best, be, True, steps, 0, 1.2, 2.0, 2.0, 55, 0)
    torch.cuda.synchronize()
    return [x.cpu().numpy() for x in (cur, ce, best, be)]


@pytest.mark.parametrize("queues", [8, 3, 1])
def test_batch_equals_per_chain_launch(small_screen, queues):
    """Bit-identical to the per-chain launch whatever the queues (one per XCD, an odd count, one global queue); with
    3 queues the complexes are put in 5 groups, as several receptor sets would be."""
    print(f"[gpu] batch vs plain, {queues} queues", flush=True)
    w, kp, st, cl = small_screen
    ref = _run_plain(kp, st, cl, w["steps"])
    groups = np.arange(len(w["ligs"])) % 5 if queues == 3 else None
    r = BatchRunner(kp, cl, n_complexes=len(w["ligs"]), groups=groups, queues=queues)
    N = len(cl)
    cur, best = st.clone(), st.clone()
    ce, be = torch.zeros(N).cuda(), torch.full((N,), float("inf")).cuda()
    r.launch(torch.arange(N, dtype=torch.int32).cuda(), cur, ce, best, be, True, w["steps"], 0, 1.2, 2.0, 2.0, 55, 0)
    torch.cuda.synchronize()
    for a, b in zip(ref, (cur, ce, best, be)):
        np.testing.assert_array_equal(a, b.cpu().numpy())
    assert (r.done.cpu().numpy() == w["chains"]).all()
    lat = r.latencies()
    assert np.isfinite(lat).all() and (lat > 0).all() and lat.max() < 600


def test_graph_replays_match_eager(small_screen):
    print("[gpu] graph replay", flush=True)
    w, kp, st, cl = small_screen
    ref = _run_plain(kp, st, cl, w["steps"])
    r = BatchRunner(kp, cl, n_complexes=len(w["ligs"]))
    N = len(cl)
    cur, best = st.clone(), st.clone()
    ce, be = torch.zeros(N).cuda(), torch.full((N,), float("inf")).cuda()
    ci = torch.arange(N, dtype=torch.int32).cuda()
    g = GraphedBatch(r, chain_id=ci, cur=cur, cur_e=ce, best=best, best_e=be, init_opt=True, mc_steps=w["steps"],
                     bfgs_iters=0, temp=1.2, amp_pos=2.0, amp_rot=2.0, seed=55, step_offset=0)
    for rep in range(3):
        cur.copy_(st); best.copy_(st); ce.zero_(); be.fill_(float("inf"))
        _ = (st * 2).sum()                  # unrelated eager work between replays
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for a, b in zip(ref, (cur, ce, best, be)):
            np.testing.assert_array_equal(a, b.cpu().numpy())
        assert (r.done.cpu().numpy() == w["chains"]).all()


def test_autotuner_picks_a_shape(small_screen):
    print("[gpu] autotune", flush=True)
    w, kp, st, cl = small_screen
    r = BatchRunner(kp, cl, n_complexes=len(w["ligs"]))
    from adock_amd import _lib
    W, blocks, layout = Autotuner().tune(r, st, probe_steps=1, reps=1)
    assert W in (1, 2, 4) and r.waves == W
    assert layout in (0, 1) and r.lds_layout == layout
    assert _lib.lib().adock_get_chains_per_block() == 0     # the shape is the runner's, not the process default
    ref = _run_plain(kp, st, cl, 2)
    N = len(cl)
    cur, best = st.clone(), st.clone()
    ce, be = torch.zeros(N).cuda(), torch.full((N,), float("inf")).cuda()
    r.launch(torch.arange(N, dtype=torch.int32).cuda(), cur, ce, best, be, True, 2, 0, 1.2, 2.0, 2.0, 55, 0)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(ref[3], be.cpu().numpy())


@pytest.mark.parametrize("waves", [4, 1])
def test_dense_lds_tile_equals_compact(small_screen, waves):
    """The persistent launch with the dense pair-gradient LDS tile gives the compact tile's results bit for bit."""
    print(f"[gpu] persistent launch: dense vs compact LDS tile, {waves} chains per workgroup", flush=True)
    w, kp, st, cl = small_screen
    N = len(cl)
    out = []
    for layout in (0, 1):
        r = BatchRunner(kp, cl, n_complexes=len(w["ligs"]), queues=3)
        r.waves, r.lds_layout = waves, layout
        cur, best = st.clone(), st.clone()
        ce, be = torch.zeros(N).cuda(), torch.full((N,), float("inf")).cuda()
        r.launch(torch.arange(N, dtype=torch.int32).cuda(), cur, ce, best, be, True, 4, 0, 1.2, 2.0, 2.0, 91, 0,
                 count_evals=True)
        torch.cuda.synchronize()
        out.append([x.cpu().numpy() for x in (cur, ce, best, be, r.lig_evals)])
        assert (r.done.cpu().numpy() == w["chains"]).all()
    for a, b in zip(*out):
        np.testing.assert_array_equal(a, b)


def test_run_mc_routes_many_chains_through_queues():
    """KernelProblem.run_mc with QUEUE_MIN_CHAINS chains or more (grid mode) takes the persistent queued launch; the
    result and the evaluation count equal adock_mc's bit for bit."""
    print("[gpu] run_mc: queued launch vs adock_mc at 4096 chains", flush=True)
    from adock_amd import kernel as kmod
    w, prob = workload.screen(seed=4, n_ligands=64, chains=64, steps=3)
    kp = KernelProblem(prob)
    L = len(w["ligs"])
    st, cl = workload.chain_starts(4, w["ligs"], [w["center"]] * L, [w["size"]] * L, w["chains"], kp.KM)
    st, cl = torch.as_tensor(st).cuda(), torch.as_tensor(cl).cuda()
    N = len(cl)
    assert N >= kmod.QUEUE_MIN_CHAINS

    def run(queues):
        kp.use_queues = queues
        try:
            cur, best = st.clone(), st.clone()
            ce, be = torch.zeros(N).cuda(), torch.full((N,), float("inf")).cuda()
            kp.eval_count.zero_()
            kp.run_mc(cl, torch.arange(N, dtype=torch.int32).cuda(), cur, ce, best, be, True, w["steps"], 0, 1.2, 2.0,
                      2.0, 9, 0, count_evals=True)
            torch.cuda.synchronize()
            assert isinstance(kp._keep[-1], BatchRunner) == queues      # the launch that ran
            return [x.cpu().numpy() for x in (cur, ce, best, be)] + [int(kp.eval_count.item())]
        finally:
            kp.use_queues = True

    ok = kp.queues_keep_occupancy()
    print(f"  resident waves per CU (persistent, per-chain): {kp._queue_occupancy}", flush=True)
    assert kp.KM == 16 and ok
    a, b = run(True), run(False)
    assert a[4] == b[4] > 0
    for x, y in zip(a[:4], b[:4]):
        np.testing.assert_array_equal(x, y)


def test_run_mc_queued_route_is_asynchronous_and_capturable():
    """run_mc's queued route builds its one longest-first queue on the device (chain_lig is never read back): the
    order equals the host partition's single queue, the route runs inside a HIP-graph capture, and three replays
    (with eager resets and synchronisations between them) equal the eager run bit for bit."""
    print("[gpu] run_mc queued route: device-built queue, graph capture", flush=True)
    from adock_amd.batch import queue_partition
    w, prob = workload.screen(seed=4, n_ligands=64, chains=64, steps=3)
    kp = KernelProblem(prob)
    L = len(w["ligs"])
    st_h, cl_h = workload.chain_starts(4, w["ligs"], [w["center"]] * L, [w["size"]] * L, w["chains"], kp.KM)
    st, cl = torch.as_tensor(st_h).cuda(), torch.as_tensor(cl_h).cuda()
    N = len(cl_h)
    r = BatchRunner.on_device(kp, cl, L)
    want, _ = queue_partition(np.asarray(kp.cost)[cl_h], np.zeros(N, np.int64), 1)
    np.testing.assert_array_equal(r.order.cpu().numpy(), want)
    np.testing.assert_array_equal(r.lig_chains.cpu().numpy(), np.bincount(cl_h, minlength=L))
    ci = torch.arange(N, dtype=torch.int32).cuda()

    def bufs():
        return st.clone(), torch.zeros(N).cuda(), st.clone(), torch.full((N,), float("inf")).cuda()

    def call(b):
        kp.run_mc(cl, ci, *b, True, w["steps"], 0, 1.2, 2.0, 2.0, 9, 0)

    ref = bufs()
    call(ref)
    torch.cuda.synchronize()
    assert isinstance(kp._keep[-1], BatchRunner)
    b = bufs()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        call(b)
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        call(b)
    keep = kp._keep
    for rep in range(3):
        b[0].copy_(st); b[1].zero_(); b[2].copy_(st); b[3].fill_(float("inf"))
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for x, y in zip(ref, b):
            assert torch.equal(x, y), rep
    del keep


def test_skewed_queues_are_stolen_and_complete(small_screen):
    """All chains in the last of 8 queues (seven empty): every workgroup finds its own queue drained and serves the
    others, every chain runs exactly once (per-complex done counts), results bit-identical to the per-chain launch."""
    print("[gpu] batch with one loaded queue out of 8", flush=True)
    w, kp, st, cl = small_screen
    ref = _run_plain(kp, st, cl, w["steps"])
    r = BatchRunner(kp, cl, n_complexes=len(w["ligs"]), queues=8)
    N = len(cl)
    r.qstart = torch.as_tensor(np.array([0] * 8 + [N], np.int32)).cuda()
    cur, best = st.clone(), st.clone()
    ce, be = torch.zeros(N).cuda(), torch.full((N,), float("inf")).cuda()
    r.launch(torch.arange(N, dtype=torch.int32).cuda(), cur, ce, best, be, True, w["steps"], 0, 1.2, 2.0, 2.0, 55, 0)
    torch.cuda.synchronize()
    for a, b in zip(ref, (cur, ce, best, be)):
        np.testing.assert_array_equal(a, b.cpu().numpy())
    assert (r.done.cpu().numpy() == w["chains"]).all()
