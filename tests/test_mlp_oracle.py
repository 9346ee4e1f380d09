"""Pin of the numpy MLP oracle by central finite differences (float64)."""
import numpy as np


def test_gradient_matches_finite_differences():
    from oracle import mlp_oracle as M
    rng = np.random.default_rng(1)
    w = M.init(3).astype(np.float64) + rng.normal(0, 0.05, M.NPARAM)
    x = rng.normal(size=(40, M.IN))
    t = rng.normal(size=40)
    sw = rng.uniform(0.5, 2.0, 40)
    loss, g = M.loss_grad(w, x, t, sw)
    idx = rng.choice(M.NPARAM, 60, replace=False)
    for i in idx:
        e = np.zeros(M.NPARAM); e[i] = 1e-6
        lp, _ = M.loss_grad(w + e, x, t, sw)
        lm, _ = M.loss_grad(w - e, x, t, sw)
        fd = (lp - lm) / 2e-6
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i]))


def test_forward_shapes_and_relu():
    from oracle import mlp_oracle as M
    w = np.zeros(M.NPARAM); w[-1] = 0.25
    y, _ = M.forward(w, np.ones((5, M.IN)))
    assert np.allclose(y, 0.25)


def test_adam_constant_gradient_closed_form():
    """With a constant gradient g, Adam's bias-corrected moments are exactly
    m_hat = g and v_hat = g^2 at every step t, so w_t = w_0 - t lr g / (|g| + eps)
    (Kingma & Ba 2015, Algorithm 1)."""
    from oracle import mlp_oracle as M
    rng = np.random.default_rng(7)
    w0 = rng.normal(size=50)
    g = rng.normal(size=50) * 10.0 ** rng.uniform(-3, 2, 50)
    lr, b1, b2, eps = 1e-2, 0.8, 0.95, 1e-6
    w, m, v = w0.copy(), np.zeros(50), np.zeros(50)
    for t in range(1, 21):
        w, m, v = M.adam(w, g, m, v, t, lr=lr, b1=b1, b2=b2, eps=eps)
        expect = w0 - t * lr * g / (np.abs(g) + eps)
        assert np.allclose(w, expect, rtol=0, atol=1e-12 * t)


def test_adam_two_steps_by_hand():
    """Two steps with a changing gradient, worked by hand with beta1 = 1/2,
    beta2 = 3/4, lr = 1/10, eps = 0, g = 2 then 4:
      step 1: m = 1, v = 1, m_hat = 2, v_hat = 4, w -= 0.1 * 2 / 2 = 0.1;
      step 2: m = 1/2 + 2 = 5/2, v = 3/4 + 4 = 19/4,
              m_hat = (5/2) / (3/4) = 10/3, v_hat = (19/4) / (7/16) = 76/7,
              w -= 0.1 * (10/3) / sqrt(76/7).
    Swapping the two betas gives the same first step (bias correction makes
    it lr sign(g)) but a different second one."""
    from oracle import mlp_oracle as M
    w, m, v = np.array([1.0]), np.zeros(1), np.zeros(1)
    w, m, v = M.adam(w, np.array([2.0]), m, v, 1, lr=0.1, b1=0.5, b2=0.75, eps=0.0)
    assert w[0] == 0.9 and m[0] == 1.0 and v[0] == 1.0
    w, m, v = M.adam(w, np.array([4.0]), m, v, 2, lr=0.1, b1=0.5, b2=0.75, eps=0.0)
    assert m[0] == 2.5 and v[0] == 4.75
    assert abs(w[0] - (0.9 - 0.1 * (10.0 / 3.0) / np.sqrt(76.0 / 7.0))) < 1e-15
    ws, ms, vs = M.adam(np.array([1.0]), np.array([2.0]), np.zeros(1), np.zeros(1), 1, lr=0.1, b1=0.75, b2=0.5,
                        eps=0.0)
    ws, _, _ = M.adam(ws, np.array([4.0]), ms, vs, 2, lr=0.1, b1=0.75, b2=0.5, eps=0.0)
    assert abs(ws[0] - w[0]) > 1e-3


def test_forward_one_hot_network_by_hand():
    """A network with one path: input 2 -> hidden-1 unit 5 (weight 2, bias -1)
    -> hidden-2 unit 7 (weight 3, bias 1/2) -> output (weight -2, bias 1/4).
    y = 1/4 - 2 relu(3 relu(2 x_2 - 1) + 1/2); other units stay at relu(0).
    A transposed weight matrix (W[j][i] read as W[i][j]) breaks the path."""
    from oracle import mlp_oracle as M
    W1 = np.zeros((M.IN, M.H)); W1[2, 5] = 2.0
    b1 = np.zeros(M.H); b1[5] = -1.0
    W2 = np.zeros((M.H, M.H)); W2[5, 7] = 3.0
    b2 = np.zeros(M.H); b2[7] = 0.5
    w3 = np.zeros(M.H); w3[7] = -2.0
    w = np.concatenate([W1.ravel(), b1, W2.ravel(), b2, w3, [0.25]])
    x = np.zeros((3, M.IN))
    x[0, 2] = 1.5    # h1 = 2, h2 = 6.5, y = 0.25 - 13
    x[1, 2] = 0.25   # h1 = relu(-0.5) = 0, h2 = 0.5, y = 0.25 - 1
    x[2, 5] = 9.0    # off the path: y as for x = 0
    y, _ = M.forward(w, x)
    assert y.tolist() == [-12.75, -0.75, -0.75]
