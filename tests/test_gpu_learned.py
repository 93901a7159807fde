"""Step-by-step search with the learned proposal: the step pipeline without the network reproduces the fused plain
kernel bit for bit (features and move logging do not touch the chain), logged moves are well formed, network-driven
search keeps the exact-score invariants, and a HIP-graph capture replays the eager result."""
import numpy as np
import pytest
import torch

from adock_amd import synth, workload
from adock_amd.kernel import KernelProblem
from adock_amd.learned import StepSearch, move_weights
from adock_amd.mlp import ProposalNet

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small_cx():
    w, prob = workload.single("small", seed=0)
    kp = KernelProblem(prob)
    N = 32
    t, q, th = synth.start_states(4, N, w["center"], w["size"], w["lig"].n_tors, kp.KM)
    st = kp.pack(torch.as_tensor(t).cuda(), torch.as_tensor(q).cuda(), torch.as_tensor(th).cuda())
    return w, kp, st


def _bufs(st):
    N = st.shape[0]
    return st.clone(), torch.zeros(N).cuda(), st.clone(), torch.full((N,), float("inf")).cuda()


def test_step_pipeline_equals_fused_plain(small_cx):
    print("[gpu] step pipeline vs fused", flush=True)
    w, kp, st = small_cx
    N, S = st.shape[0], 12
    cl = torch.zeros(N, dtype=torch.int32).cuda()
    ref = _bufs(st)
    kp.run_mc(cl, torch.arange(N, dtype=torch.int32).cuda(), *ref, True, S, 0, 1.2, 2.0, 2.0, 777, 0)
    ss = StepSearch(kp, cl, log_cap=20000)
    got = _bufs(st)
    ss.run(*got, S, seed=777, log=True)
    torch.cuda.synchronize()
    for a, b in zip(ref, got):
        np.testing.assert_array_equal(a.cpu().numpy(), b.cpu().numpy())
    X, T, M, gain = [x.cpu().numpy() for x in ss.logged()]
    assert len(X) > 0 and np.isfinite(X).all() and np.isfinite(T).all()
    nt = w["lig"].n_tors
    assert (M[:, :7 + nt] == 1).all() and (M[:, 7 + nt:] == 0).all()
    np.testing.assert_allclose(np.linalg.norm(T[:, 3:7], axis=1), 1.0, atol=1e-4)
    assert (T[:, 3] >= 0).all()
    assert (np.abs(T[:, 7:7 + nt]) <= np.pi + 1e-5).all()
    assert (np.abs(X) <= 64).all()


def test_network_search_invariants_and_graph_replay(small_cx):
    print("[gpu] network search + graph", flush=True)
    w, kp, st = small_cx
    N, S = st.shape[0], 8
    cl = torch.zeros(N, dtype=torch.int32).cuda()
    net = ProposalNet("cuda", seed=1)
    ss = StepSearch(kp, cl, net=net)
    e0, _ = kp.energy_grad(cl, st)
    eager = _bufs(st)
    ss.run(*eager, S, seed=5, use_net=True, p_plain=0.25)
    torch.cuda.synchronize()
    be = eager[3]
    assert (be <= e0 + 1e-4).all()
    e1, _ = kp.energy_grad(cl, eager[2])
    torch.testing.assert_close(e1, be, rtol=1e-4, atol=1e-4)
    bufs = _bufs(st)
    g = ss.capture(*bufs, S, seed=5, use_net=True, p_plain=0.25)
    for rep in range(2):
        bufs[0].copy_(st); bufs[1].zero_(); bufs[2].copy_(st); bufs[3].fill_(float("inf"))
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for a, b in zip(eager, bufs):
            np.testing.assert_array_equal(a.cpu().numpy(), b.cpu().numpy())


def test_training_on_logged_moves_reduces_loss(small_cx):
    print("[gpu] train on logged moves", flush=True)
    w, kp, st = small_cx
    N = st.shape[0]
    cl = torch.zeros(N, dtype=torch.int32).cuda()
    ss = StepSearch(kp, cl, log_cap=50000)
    ss.run(*_bufs(st), 40, seed=11, log=True)
    X, T, M, gain = ss.logged()
    assert X.shape[0] > 200
    W = move_weights(M, gain)
    net = ProposalNet("cuda", seed=0)
    before = net.eval_loss(X.contiguous(), T.contiguous(), W)
    net.fit(X.contiguous(), T.contiguous(), W, epochs=6, batch=256, lr=2e-3)
    after = net.eval_loss(X.contiguous(), T.contiguous(), W)
    assert after < 0.9 * before, (before, after)
