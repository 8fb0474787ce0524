"""NumPy restatement of the learned first pass's network (TEST INFRASTRUCTURE ONLY: the
checker for the HIP MLP kernels; imported by tests/ and bench.py's reporting, never by
the product package).

The network has no counterpart in the reference; this file fixes its arithmetic:
  xb = bf16(x);  z1 = xb @ bf16(W1) + b1;  h1 = bf16(relu(z1))
  z2 = h1 @ bf16(W2) + b2;  h2 = relu(z2);  z3 = h2 @ bf16(w3) + b3;  p = sigmoid(z3)
with bf16 = round-to-nearest-even to 8 mantissa bits, and the loss
  L = mean_i w_i * (y_i * softplus(-z3_i) + (1 - y_i) * softplus(z3_i)),
  w_i = sample_w_i * (pos_weight if y_i else 1).
Gradients go straight through the bf16 roundings.  Parameters: W1[16][64], b1[64],
W2[64][64], b2[64], w3[64], b3 (5313 floats).
"""
import numpy as np

N_IN, N_H = 16, 64
OFF_W1 = 0
OFF_B1 = OFF_W1 + N_IN * N_H
OFF_W2 = OFF_B1 + N_H
OFF_B2 = OFF_W2 + N_H * N_H
OFF_W3 = OFF_B2 + N_H
OFF_B3 = OFF_W3 + N_H
N_PARAM = OFF_B3 + 1


def bf16(x):
    """float32 -> nearest bf16 value (ties to even), returned as float32."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def unpack(w):
    w = np.asarray(w)
    return (w[OFF_W1:OFF_B1].reshape(N_IN, N_H), w[OFF_B1:OFF_W2], w[OFF_W2:OFF_B2].reshape(N_H, N_H),
            w[OFF_B2:OFF_W3], w[OFF_W3:OFF_B3], w[OFF_B3])


def forward(w, X, rounding=True):
    """Returns p and the intermediates (float64 math on the rounded operands)."""
    r = bf16 if rounding else (lambda a: np.asarray(a, dtype=np.float64))
    W1, b1, W2, b2, w3, b3 = unpack(w)
    xb = r(X).astype(np.float64)
    z1 = xb @ r(W1).astype(np.float64) + b1.astype(np.float64)
    h1 = r(np.maximum(z1, 0.0).astype(np.float32) if rounding else np.maximum(z1, 0.0)).astype(np.float64)
    z2 = h1 @ r(W2).astype(np.float64) + b2.astype(np.float64)
    h2 = np.maximum(z2, 0.0)
    z3 = h2 @ r(w3).astype(np.float64) + float(b3)
    p = 1.0 / (1.0 + np.exp(-z3))
    return p, (xb, z1, h1, z2, h2, z3)


def softplus(t):
    return np.maximum(t, 0.0) + np.log1p(np.exp(-np.abs(t)))


def loss_and_grad(w, X, y, sample_w=None, pos_weight=1.0, rounding=True):
    """(loss, gradient) of the weighted mean BCE."""
    n = X.shape[0]
    W1, b1, W2, b2, w3, b3 = unpack(w)
    r = bf16 if rounding else (lambda a: np.asarray(a, dtype=np.float64))
    p, (xb, z1, h1, z2, h2, z3) = forward(w, X, rounding)
    yy = np.asarray(y, dtype=np.float64)
    wt = (np.ones(n) if sample_w is None else np.asarray(sample_w, dtype=np.float64)) * np.where(yy > 0, pos_weight, 1.0)
    loss = float(np.sum(wt * (yy * softplus(-z3) + (1 - yy) * softplus(z3))) / max(n, 1))
    dz3 = wt * (p - yy) / max(n, 1)
    g = np.zeros(N_PARAM)
    g[OFF_W3:OFF_B3] = h2.T @ dz3
    g[OFF_B3] = dz3.sum()
    dz2 = np.outer(dz3, r(w3).astype(np.float64)) * (z2 > 0)
    g[OFF_W2:OFF_B2] = (h1.T @ dz2).reshape(-1)
    g[OFF_B2:OFF_W3] = dz2.sum(axis=0)
    dz1 = (dz2 @ r(W2).astype(np.float64).T) * (z1 > 0)
    g[OFF_W1:OFF_B1] = (xb.T @ dz1).reshape(-1)
    g[OFF_B1:OFF_W2] = dz1.sum(axis=0)
    return loss, g


def init_params(seed):
    """A float32 parameter vector with He-style scales (for tests; the device keeps its own
    seeded initialisation)."""
    rng = np.random.default_rng(seed)
    w = np.zeros(N_PARAM, dtype=np.float32)
    w[OFF_W1:OFF_B1] = rng.normal(0, np.sqrt(2 / N_IN), N_IN * N_H)
    w[OFF_W2:OFF_B2] = rng.normal(0, np.sqrt(2 / N_H), N_H * N_H)
    w[OFF_W3:OFF_B3] = rng.normal(0, np.sqrt(1 / N_H), N_H)
    w[OFF_B1:OFF_W2] = rng.normal(0, 0.1, N_H)
    w[OFF_B2:OFF_W3] = rng.normal(0, 0.1, N_H)
    w[OFF_B3] = 0.05
    return w
