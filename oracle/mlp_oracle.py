"""numpy restatement of the step-proposer MLP (12 -> 64 -> 64 -> 1, ReLU,
FP32 weights) and of its weighted mean-squared-error gradient.

TEST INFRASTRUCTURE ONLY: the checker of the HIP forward/backward/Adam
kernels (abie_mlp_* in libabie_hip.so).  The network is the project's own
addition (no counterpart in the reference); this oracle is pinned by finite
differences in tests/test_mlp_oracle.py.

Layout of the flat weight vector: W1[12][64], b1[64], W2[64][64], b2[64],
w3[64], b3[1], W[i][j] connecting input i to unit j.
"""
import numpy as np

IN, H = 12, 64
NPARAM = IN * H + H + H * H + H + H + 1


def unpack(w):
    w = np.asarray(w, dtype=np.float64)
    o = 0
    W1 = w[o:o + IN * H].reshape(IN, H); o += IN * H
    b1 = w[o:o + H]; o += H
    W2 = w[o:o + H * H].reshape(H, H); o += H * H
    b2 = w[o:o + H]; o += H
    w3 = w[o:o + H]; o += H
    b3 = w[o]
    return W1, b1, W2, b2, w3, b3


def init(seed=0):
    rng = np.random.default_rng(seed)
    parts = [rng.normal(0, np.sqrt(2.0 / IN), IN * H), np.zeros(H),
             rng.normal(0, np.sqrt(2.0 / H), H * H), np.zeros(H),
             rng.normal(0, np.sqrt(1.0 / H), H), np.zeros(1)]
    return np.concatenate(parts).astype(np.float32)


def forward(w, x):
    W1, b1, W2, b2, w3, b3 = unpack(w)
    x = np.asarray(x, dtype=np.float64)
    z1 = x @ W1 + b1
    h1 = np.maximum(z1, 0)
    z2 = h1 @ W2 + b2
    h2 = np.maximum(z2, 0)
    return h2 @ w3 + b3, (x, z1, h1, z2, h2)


def loss_grad(w, x, t, sw=None):
    """L = sum_s sw_s (y_s - t_s)^2 / sum(sw) and dL/dw (float64)."""
    W1, b1, W2, b2, w3, b3 = unpack(w)
    y, (x, z1, h1, z2, h2) = forward(w, x)
    t = np.asarray(t, dtype=np.float64)
    sw = np.ones_like(t) if sw is None else np.asarray(sw, dtype=np.float64)
    sw = sw / sw.sum()
    r = y - t
    loss = float(np.sum(sw * r * r))
    dy = 2 * sw * r
    gw3 = h2.T @ dy
    gb3 = dy.sum()
    d2 = np.outer(dy, w3) * (z2 > 0)
    gW2 = h1.T @ d2
    gb2 = d2.sum(0)
    d1 = (d2 @ W2.T) * (z1 > 0)
    gW1 = x.T @ d1
    gb1 = d1.sum(0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gw3, [gb3]])
    return loss, g


def adam(w, g, m, v, step, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return w - lr * mh / (np.sqrt(vh) + eps), m, v
