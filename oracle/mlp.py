"""TEST INFRASTRUCTURE ONLY: numpy FP64 reference of the proposal network (forward, masked-MSE backward, Adam) and a
BF16-emulating forward that rounds exactly where the MFMA inference kernel rounds (inputs, weights and hidden
activations to bfloat16 with round-to-nearest-even; accumulation, biases and the output in wide precision)."""
import numpy as np


def relu(x):
    return np.maximum(x, 0.0)


def forward(p, X):
    X = np.asarray(X, np.float64)
    H1 = relu(X @ p["W1"].T.astype(np.float64) + p["b1"])
    H2 = relu(H1 @ p["W2"].T.astype(np.float64) + p["b2"])
    Y = H2 @ p["W3"].T.astype(np.float64) + p["b3"]
    return Y, (X, H1, H2)


def loss_and_grads(p, X, T, M=None):
    Y, (X, H1, H2) = forward(p, X)
    M = np.ones_like(Y) if M is None else np.asarray(M, np.float64)
    n = max(M.sum(), 1.0)
    D = (Y - T) * M
    loss = (D * D).sum() / n
    dY = 2.0 * D / n
    g = {"W3": dY.T @ H2, "b3": dY.sum(0)}
    dH2 = (dY @ p["W3"].astype(np.float64)) * (H2 > 0)
    g["W2"], g["b2"] = dH2.T @ H1, dH2.sum(0)
    dH1 = (dH2 @ p["W2"].astype(np.float64)) * (H1 > 0)
    g["W1"], g["b1"] = dH1.T @ X, dH1.sum(0)
    return loss, g


def adam(w, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    w = w - lr * (m / (1 - b1 ** t)) / (np.sqrt(v / (1 - b2 ** t)) + eps)
    return w, m, v


def bf16(x):
    """float32 -> bfloat16 -> float32, round to nearest even."""
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def forward_bf16(p, X):
    X = bf16(X).astype(np.float64)
    W = {k: bf16(p[k]).astype(np.float64) for k in ("W1", "W2", "W3")}
    H1 = bf16(relu(X @ W["W1"].T + p["b1"])).astype(np.float64)
    H2 = bf16(relu(H1 @ W["W2"].T + p["b2"])).astype(np.float64)
    return H2 @ W["W3"].T + p["b3"]
