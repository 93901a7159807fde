"""Proposal-network kernels vs the FP64 numpy oracle: GEMM in every transpose combination, training forward and
backward (1e-4 relative), Adam, BF16 MFMA inference against a BF16-emulating oracle (lane layouts of the MFMA
fragments) and against FP64 (BF16 error budget), and a short fit that must reduce the loss."""
import numpy as np
import pytest
import torch

from adock_amd.mlp import SHAPES, ProposalNet, gemm
from oracle import mlp as om

pytestmark = pytest.mark.gpu


def _rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-12))


@pytest.mark.parametrize("K", [130, 3000])
@pytest.mark.parametrize("ta,tb", [(False, False), (False, True), (True, False), (True, True)])
def test_gemm_all_transposes(ta, tb, K):
    """Both GEMM paths against FP64: K = 130 (one K range) and K = 3,000 over 2 output tiles (K split over
    workgroups, partials added in order; K not a multiple of the split)."""
    print(f"[gpu] gemm ta={ta} tb={tb} K={K}", flush=True)
    g = np.random.default_rng(0)
    M, N = 77, 45
    A = g.standard_normal((K, M) if ta else (M, K)).astype(np.float32)
    B = g.standard_normal((N, K) if tb else (K, N)).astype(np.float32)
    C0 = g.standard_normal((M, N)).astype(np.float32)
    C = torch.as_tensor(C0).cuda()
    gemm(torch.as_tensor(A).cuda(), torch.as_tensor(B).cuda(), C, trans_a=ta, trans_b=tb, alpha=0.5, beta=2.0)
    want = 0.5 * ((A.T if ta else A).astype(np.float64) @ (B.T if tb else B).astype(np.float64)) + 2.0 * C0
    assert _rel(C.cpu().numpy(), want) < 1e-5


@pytest.mark.parametrize("B", [300, 2048])
def test_forward_backward_adam_match_oracle(B):
    """Forward, loss, gradients and one Adam step against the numpy FP64 oracle; with 2,048 rows the weight gradients
    take the split-K GEMM."""
    print(f"[gpu] mlp fwd/bwd B={B}", flush=True)
    net = ProposalNet("cuda", seed=3)
    p = {k: v.astype(np.float64) for k, v in net.numpy_params().items()}
    g = np.random.default_rng(4)
    X = g.standard_normal((B, 128)).astype(np.float32)
    T = g.standard_normal((B, 23)).astype(np.float32)
    M = (g.random((B, 23)) < 0.7).astype(np.float32)
    Y = net.forward(torch.as_tensor(X).cuda())
    Yo, _ = om.forward(p, X)
    assert _rel(Y.cpu().numpy(), Yo) < 1e-4
    loss = net.backward(Y, torch.as_tensor(T).cuda(), torch.as_tensor(M).cuda())
    lo, go = om.loss_and_grads(p, X, T, M)
    assert abs(float(loss.item()) - lo) <= 1e-4 * lo
    for k in SHAPES:
        gk = net.grad[k].cpu().numpy()
        assert np.linalg.norm(gk - go[k]) <= 1e-4 * max(np.linalg.norm(go[k]), 1e-8), k
    w0 = {k: v.cpu().numpy().astype(np.float64) for k, v in net.p.items()}
    net.step(lr=1e-3)
    for k in SHAPES:
        w1, _, _ = om.adam(w0[k], net.grad[k].cpu().numpy().astype(np.float64), 0.0, 0.0, 1, 1e-3)
        np.testing.assert_allclose(net.p[k].cpu().numpy(), w1, rtol=1e-5, atol=1e-7)


def test_mfma_inference_matches_bf16_oracle():
    print("[gpu] mlp mfma inference", flush=True)
    net = ProposalNet("cuda", seed=5)
    p = {k: v for k, v in net.numpy_params().items()}
    g = np.random.default_rng(6)
    for B in (1, 64, 1000, 4096):
        X = g.standard_normal((B, 128)).astype(np.float32)
        Y = net.infer(torch.as_tensor(X).cuda()).cpu().numpy()
        Yb = om.forward_bf16(p, X)
        Y64, _ = om.forward({k: v.astype(np.float64) for k, v in p.items()}, X)
        assert _rel(Y, Yb) < 2e-3, B            # same roundings; only accumulation order and hidden re-rounding differ
        assert _rel(Y, Y64) < 5e-2, B           # BF16 error budget vs full precision


def test_fit_reduces_loss():
    print("[gpu] mlp fit", flush=True)
    teacher = ProposalNet("cuda", seed=11)
    g = np.random.default_rng(12)
    X = torch.as_tensor(g.standard_normal((16384, 128)).astype(np.float32)).cuda()
    T = teacher.forward(X).detach().clone()
    net = ProposalNet("cuda", seed=13)
    losses = net.fit(X, T, epochs=6, batch=1024, lr=2e-3)
    assert losses[-1] < 0.3 * losses[0], losses
