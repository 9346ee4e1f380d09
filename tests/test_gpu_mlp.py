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


def test_adam_many_steps_match_oracle():
    """12 HIP Adam steps with a different gradient each step against the
    pinned oracle (distinct beta1, beta2: a mix-up shows from step 2 on)."""
    import torch
    from abie_amd import engine
    from oracle import mlp_oracle as M
    rng = np.random.default_rng(11)
    w = rng.normal(size=M.NPARAM).astype(np.float32)
    wd = torch.tensor(w, device="cuda")
    md, vd = torch.zeros_like(wd), torch.zeros_like(wd)
    wr, mr, vr = w.astype(np.float64), np.zeros(M.NPARAM), np.zeros(M.NPARAM)
    for step in range(1, 13):
        g = (rng.normal(size=M.NPARAM) * 10.0 ** rng.uniform(-2, 1)).astype(np.float32)
        engine.mlp_adam(wd, torch.tensor(g, device="cuda"), md, vd, step, lr=3e-3, beta1=0.85, beta2=0.97,
                        eps=1e-7)
        wr, mr, vr = M.adam(wr, g.astype(np.float64), mr, vr, step, lr=3e-3, b1=0.85, b2=0.97, eps=1e-7)
    assert np.max(np.abs(wd.cpu().numpy() - wr)) < 2e-6
    assert np.max(np.abs(md.cpu().numpy() - mr)) <= 1e-5 * np.abs(mr).max()
    assert np.max(np.abs(vd.cpu().numpy() - vr)) <= 1e-5 * np.abs(vr).max()


def test_forward_one_hot_network():
    """The HIP forward kernel on the hand-computed one-path network of the
    oracle's pin (tests/test_mlp_oracle.py)."""
    import torch
    from abie_amd import engine
    from oracle import mlp_oracle as M
    W1 = np.zeros((M.IN, M.H)); W1[2, 5] = 2.0
    b1 = np.zeros(M.H); b1[5] = -1.0
    W2 = np.zeros((M.H, M.H)); W2[5, 7] = 3.0
    b2 = np.zeros(M.H); b2[7] = 0.5
    w3 = np.zeros(M.H); w3[7] = -2.0
    w = np.concatenate([W1.ravel(), b1, W2.ravel(), b2, w3, [0.25]]).astype(np.float32)
    x = np.zeros((3, M.IN), np.float32)
    x[0, 2], x[1, 2], x[2, 5] = 1.5, 0.25, 9.0
    y = engine.mlp_forward(torch.tensor(w, device="cuda"), torch.tensor(x, device="cuda")).cpu().numpy()
    assert y.tolist() == [-12.75, -0.75, -0.75]


def test_in_kernel_proposer_matches_oracle():
    """The network the learned controller evaluates inside the GR15 kernel
    (one wave, hidden layer through LDS) on the committed trained weights:
    its logged inputs match the features recomputed from the same step-log
    fields, and its logged output matches the oracle's forward pass on those
    inputs to FP32 accuracy."""
    import os
    from abie_amd import engine, ics, proposer
    from oracle import mlp_oracle as M
    w = np.load(os.path.join(os.path.dirname(__file__), "..", "abie_amd", "data", "proposer.npz"))["weights"]
    systems = ics.ensemble(48, seed=20261203 + 777000)
    for fm in ("fp64", "mixed"):
        e = engine.Ensemble(systems, force_mode=fm, controller="learned")
        e.set_proposer(w)
        e.set_steplog(200000)
        e.run(0.5)
        rec = e.steplog()
        cnt = e.counters()
        e.close()
        prop = np.isfinite(rec[:, proposer.C_PROP_Y])
        # one logged proposal per accepted attempt, none on rejected ones
        assert prop.sum() == cnt[:, 4].sum() > 200
        assert np.all(rec[prop, proposer.C_ACC] == 1) and np.all(~prop[rec[:, proposer.C_ACC] == 0])
        xk = rec[prop, proposer.C_FEAT:proposer.C_FEAT + M.IN]
        xh = proposer.features(rec[prop]).astype(np.float64)
        assert np.all(np.isfinite(xk))
        assert np.max(np.abs(xk - xh)) <= 2e-6 * max(1.0, np.abs(xh).max())
        yk = rec[prop, proposer.C_PROP_Y]
        yr, _ = M.forward(w, xk.astype(np.float32))
        assert np.max(np.abs(yk - yr)) <= 1e-5 * max(1.0, np.abs(yr).max())
