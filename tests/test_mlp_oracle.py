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
