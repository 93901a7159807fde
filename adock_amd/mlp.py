"""Learned MC proposal network: 128 features -> 256 -> 256 -> 23 move components (ReLU hidden layers).

Training runs on the GPU with this repository's kernels only (FP32 tiled GEMM, bias/ReLU, masked MSE, Adam);
inference runs the fused BF16 MFMA kernel on BF16 copies of the weights, refreshed after every optimiser step.
Weights are stored [out, in] (W1 [256,128], W2 [256,256], W3 [23,256]).
"""
import math

import numpy as np
import torch

from . import _lib

IN, HID, OUT = 128, 256, 23
SHAPES = {"W1": (HID, IN), "b1": (HID,), "W2": (HID, HID), "b2": (HID,), "W3": (OUT, HID), "b3": (OUT,)}


def _s(dev):
    return _lib.stream_of(dev)


SPLIT_TILES = 128       # below this many 64x64 output tiles and with K >= SPLIT_MIN_K, K is split over workgroups
SPLIT_MIN_K = 1024
_WORK = {}


def _work(dev, n):
    """Split-K workspace of at least n floats on dev, kept for reuse (only the most recent launch reads it)."""
    w = _WORK.get(dev)
    if w is None or w.numel() < n:
        w = _WORK[dev] = torch.empty(n, dtype=torch.float32, device=dev)
    return w


def gemm(A, B, C, trans_a=False, trans_b=False, alpha=1.0, beta=0.0):
    """C = alpha op(A) op(B) + beta C on the GPU (row-major FP32, contiguous). Products with few output tiles and a
    long K (the weight gradients, K = the batch) split K over workgroups and add the partials in a fixed order."""
    M = A.shape[1] if trans_a else A.shape[0]
    K = A.shape[0] if trans_a else A.shape[1]
    N = B.shape[0] if trans_b else B.shape[1]
    assert C.shape == (M, N) and (B.shape[1] if trans_b else B.shape[0]) == K
    tiles = ((M + 63) // 64) * ((N + 63) // 64)
    if tiles < SPLIT_TILES and K >= SPLIT_MIN_K:
        splits = min(max(1, 256 // tiles), K // 256)
        w = _work(C.device, splits * M * N)
        _lib.check(_lib.lib().adock_gemm_f32_splitk(
            int(trans_a), int(trans_b), M, N, K, float(alpha), _lib.ptr(A), A.shape[1], _lib.ptr(B), B.shape[1],
            float(beta), _lib.ptr(C), C.shape[1], splits, _lib.ptr(w), _s(C.device)))
        return C
    _lib.check(_lib.lib().adock_gemm_f32(int(trans_a), int(trans_b), M, N, K, float(alpha), _lib.ptr(A), A.shape[1],
                                         _lib.ptr(B), B.shape[1], float(beta), _lib.ptr(C), C.shape[1], _s(C.device)))
    return C


class ProposalNet:
    def __init__(self, device="cuda", seed=0, params=None):
        self.device = torch.device(device)
        g = np.random.default_rng(seed)
        if params is None:
            params = {}
            for k, shp in SHAPES.items():
                if k.startswith("W"):
                    params[k] = (g.standard_normal(shp) * math.sqrt(2.0 / shp[1])).astype(np.float32)
                else:
                    params[k] = np.zeros(shp, np.float32)
            params["W3"] *= 0.1
        self.p = {k: torch.as_tensor(np.ascontiguousarray(v, np.float32), device=self.device) for k, v in params.items()}
        self.grad = {k: torch.zeros_like(v) for k, v in self.p.items()}
        self.m = {k: torch.zeros_like(v) for k, v in self.p.items()}
        self.v = {k: torch.zeros_like(v) for k, v in self.p.items()}
        self.bf = {k: torch.zeros(v.shape, dtype=torch.int16, device=self.device) for k, v in self.p.items()
                   if k.startswith("W")}
        self.t = 0
        self.loss = torch.zeros(1, device=self.device)
        self.refresh_bf16()

    def numpy_params(self):
        return {k: v.cpu().numpy() for k, v in self.p.items()}

    def refresh_bf16(self):
        for k, b in self.bf.items():
            _lib.check(_lib.lib().adock_to_bf16(self.p[k].numel(), _lib.ptr(self.p[k]), _lib.ptr(b), _s(self.device)))

    # ---- inference (BF16 MFMA) ----
    def infer(self, X, Y=None):
        B = X.shape[0]
        Y = torch.empty(B, OUT, device=self.device) if Y is None else Y
        _lib.check(_lib.lib().adock_mlp_infer(B, _lib.ptr(X), _lib.ptr(self.bf["W1"]), _lib.ptr(self.p["b1"]),
                                              _lib.ptr(self.bf["W2"]), _lib.ptr(self.p["b2"]), _lib.ptr(self.bf["W3"]),
                                              _lib.ptr(self.p["b3"]), _lib.ptr(Y), _s(self.device)))
        return Y

    # ---- training (FP32) ----
    def forward(self, X):
        B = X.shape[0]
        L = _lib.lib()
        H1 = gemm(X, self.p["W1"], torch.empty(B, HID, device=self.device), trans_b=True)
        _lib.check(L.adock_bias_act(B, HID, _lib.ptr(H1), _lib.ptr(self.p["b1"]), 1, _s(self.device)))
        H2 = gemm(H1, self.p["W2"], torch.empty(B, HID, device=self.device), trans_b=True)
        _lib.check(L.adock_bias_act(B, HID, _lib.ptr(H2), _lib.ptr(self.p["b2"]), 1, _s(self.device)))
        Y = gemm(H2, self.p["W3"], torch.empty(B, OUT, device=self.device), trans_b=True)
        _lib.check(L.adock_bias_act(B, OUT, _lib.ptr(Y), _lib.ptr(self.p["b3"]), 0, _s(self.device)))
        self._cache = (X, H1, H2)
        return Y

    def backward(self, Y, T, mask=None):
        """Masked MSE between Y and targets T (mean over valid entries): fills self.grad, returns the loss tensor."""
        X, H1, H2 = self._cache
        B = Y.shape[0]
        L = _lib.lib()
        s = _s(self.device)
        n_valid = float(mask.sum().item()) if mask is not None else float(Y.numel())
        dY = torch.empty_like(Y)
        self.loss.zero_()
        _lib.check(L.adock_mse(Y.numel(), _lib.ptr(Y), _lib.ptr(T), _lib.ptr(mask), 1.0 / max(n_valid, 1.0),
                               _lib.ptr(dY), _lib.ptr(self.loss), s))
        g = self.grad
        gemm(dY, H2, g["W3"], trans_a=True)
        _lib.check(L.adock_colsum(B, OUT, _lib.ptr(dY), _lib.ptr(g["b3"]), s))
        dH2 = gemm(dY, self.p["W3"], torch.empty(B, HID, device=self.device))
        _lib.check(L.adock_relu_bwd(dH2.numel(), _lib.ptr(dH2), _lib.ptr(H2), s))
        gemm(dH2, H1, g["W2"], trans_a=True)
        _lib.check(L.adock_colsum(B, HID, _lib.ptr(dH2), _lib.ptr(g["b2"]), s))
        dH1 = gemm(dH2, self.p["W2"], torch.empty(B, HID, device=self.device))
        _lib.check(L.adock_relu_bwd(dH1.numel(), _lib.ptr(dH1), _lib.ptr(H1), s))
        gemm(dH1, X, g["W1"], trans_a=True)
        _lib.check(L.adock_colsum(B, HID, _lib.ptr(dH1), _lib.ptr(g["b1"]), s))
        return self.loss

    def eval_loss(self, X, T, mask=None, batch=4096):
        """Masked MSE of the FP32 forward pass over all rows (no parameter update)."""
        L = _lib.lib()
        s = _s(self.device)
        n_valid = float(mask.sum().item()) if mask is not None else float(T.numel())
        acc = torch.zeros(1, device=self.device)
        for b0 in range(0, X.shape[0], batch):
            xb, tb = X[b0:b0 + batch].contiguous(), T[b0:b0 + batch].contiguous()
            mb = mask[b0:b0 + batch].contiguous() if mask is not None else None
            y = self.forward(xb)
            dY = torch.empty_like(y)
            _lib.check(L.adock_mse(y.numel(), _lib.ptr(y), _lib.ptr(tb), _lib.ptr(mb), 1.0 / max(n_valid, 1.0),
                                   _lib.ptr(dY), _lib.ptr(acc), s))
        return float(acc.item())

    def step(self, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
        self.t += 1
        L = _lib.lib()
        for k in self.p:
            _lib.check(L.adock_adam(self.p[k].numel(), _lib.ptr(self.p[k]), _lib.ptr(self.grad[k]),
                                    _lib.ptr(self.m[k]), _lib.ptr(self.v[k]), float(lr), float(beta1), float(beta2),
                                    float(eps), self.t, _s(self.device)))
        self.refresh_bf16()

    def fit(self, X, T, mask=None, epochs=5, batch=4096, lr=1e-3, seed=0, log=None):
        """Minibatch Adam over (X [n,128], T [n,23], mask [n,23]) device tensors. Returns per-epoch mean losses."""
        n = X.shape[0]
        gen = torch.Generator(device=self.device).manual_seed(seed)
        out = []
        for ep in range(epochs):
            perm = torch.randperm(n, generator=gen, device=self.device)
            tot, nb = 0.0, 0
            for b0 in range(0, n - batch + 1 if n >= batch else 1, batch):
                idx = perm[b0:b0 + batch]
                xb, tb = X[idx].contiguous(), T[idx].contiguous()
                mb = mask[idx].contiguous() if mask is not None else None
                y = self.forward(xb)
                loss = self.backward(y, tb, mb)
                self.step(lr)
                tot += float(loss.item())
                nb += 1
            out.append(tot / max(nb, 1))
            if log:
                log(f"epoch {ep}: loss {out[-1]:.5f}")
        return out

    def save(self, path):
        np.savez(path, **self.numpy_params())

    @classmethod
    def load(cls, path, device="cuda"):
        d = np.load(path)
        return cls(device, params={k: d[k] for k in SHAPES})
