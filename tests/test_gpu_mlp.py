"""HIP MLP kernels (forward, loss gradient, Adam) against the numpy oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_forward_grad_adam_match_oracle():
    import torch
    from abie_amd import engine
    from oracle import mlp_oracle as M
    rng = np.random.default_rng(4)
    w = M.init(2) + rng.normal(0, 0.02, M.NPARAM).astype(np.float32)
    x = rng.normal(size=(3000, M.IN)).astype(np.float32)
    t = rng.normal(size=3000).astype(np.float32)
    sw = rng.uniform(0.5, 2.0, 3000).astype(np.float32)
    dev = lambda a: torch.tensor(a, device="cuda")
    wd, xd, td, swd = dev(w), dev(x), dev(t), dev(sw)
    y = engine.mlp_forward(wd, xd).cpu().numpy()
    yr, _ = M.forward(w, x)
    assert np.max(np.abs(y - yr)) <= 1e-5 * max(1.0, np.abs(yr).max())
    g, loss = engine.mlp_loss_grad(wd, xd, td, swd)
    lr, gr = M.loss_grad(w, x, t, sw)
    assert float(loss.cpu()) == pytest.approx(lr, rel=1e-4)
    g = g.cpu().numpy()
    assert np.linalg.norm(g - gr) <= 1e-4 * np.linalg.norm(gr)
    # deterministic: same inputs, same bits
    g2, _ = engine.mlp_loss_grad(wd, xd, td, swd)
    assert np.array_equal(g, g2.cpu().numpy())
    m = torch.zeros_like(wd); v = torch.zeros_like(wd)
    gd = torch.tensor(g, device="cuda")
    engine.mlp_adam(wd, gd, m, v, 1, lr=1e-3)
    w1, _, _ = M.adam(w.astype(np.float64), g.astype(np.float64), np.zeros(M.NPARAM), np.zeros(M.NPARAM), 1)
    assert np.max(np.abs(wd.cpu().numpy() - w1)) < 1e-6
