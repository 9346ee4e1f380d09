"""The learned GR15 step proposer: data, training and evaluation.

After an accepted step the plain controller tries h_next = dtreq = h / err
next (integrator_gauss_radau15.c:349-355).  That attempt is rejected when
the next step's own dtreq is smaller, which is what makes ~2/3 of the
reference's attempts fail.  The proposer predicts

    y = log10(dtreq_{i+1} / dtreq_i)

from the state of the accepted step i, and the engine tries
h_{i+1} = dtreq_i * 10^(y_hat - margin) instead; GR15's own error test still
accepts or rejects it, and a rejection falls back to the plain rule.

Features (ABIE_MLP_IN = 12, computed identically on the device, see
gr15_impl.hpp): log10(h_i/dtreq_{i-1}), log10 err_i, log10 err_{i-1},
log10(h_i/h_{i-1}), pc_i/12, log10(h_i/tau) with tau = min over massive
pairs of r_ij/|v_ij| (the shortest encounter timescale of the new state),
log10 n_massive, log10(n_test+1) (at most log10 2049, the training range),
log10(1 + rejections before step i),
log10(dtreq_i/dtreq_{i-1}), pc_{i-1}/12, log10 of the mass ratio (sum of
the other massive masses / largest).  Every feature is invariant under a
change of time unit: an absolute step size is out of distribution near
close encounters, where a network that saw it could keep shrinking h.
The device clamps the output to [-0.3, 0] dex: a proposal is never longer
than the plain controller's dtreq.

Training runs on the GPU with the project's HIP kernels (abie_mlp_loss_grad,
abie_mlp_adam); standardisation is folded into the first layer afterwards so
the device evaluates raw features.
"""
import numpy as np

from . import _lib

EPS = 1e-300
# step-log columns (gr15_impl.hpp)
(C_SYS, C_T, C_H, C_EST, C_ERR, C_DTREQ, C_ACC, C_PC, C_REJ, C_PH, C_PERR, C_PDT, C_PPC, C_NM, C_NT, C_HNEXT,
 C_RATE2, C_MRATIO) = range(18)
# the proposer's inputs as the kernel evaluated them ([C_FEAT, C_FEAT + 12))
# and its raw output; NaN on attempts without a proposal
C_FEAT, C_PROP_Y = 18, 30


def features(rec):
    """Feature rows (n, 12) of step-log records (one per attempt)."""
    r = np.asarray(rec, dtype=np.float64)
    h, err, est, dtreq = r[:, C_H], r[:, C_ERR], r[:, C_EST], r[:, C_DTREQ]
    ph, perr, pdt = r[:, C_PH], r[:, C_PERR], r[:, C_PDT]
    x = np.empty((len(r), _lib.MLP_IN), dtype=np.float64)
    x[:, 0] = np.log10(np.where(pdt > 0, h / np.where(pdt > 0, pdt, 1.0), 1.0))
    x[:, 1] = np.log10(err + EPS)
    x[:, 2] = np.log10(perr + EPS)
    x[:, 3] = np.log10(h / np.where(ph > 0, ph, h))
    x[:, 4] = r[:, C_PC] / 12.0
    x[:, 5] = np.log10(h * np.sqrt(r[:, C_RATE2]) + 1e-30)
    x[:, 6] = np.log10(r[:, C_NM])
    x[:, 7] = np.minimum(np.log10(r[:, C_NT] + 1.0), np.log10(2049.0))  # the training range, as the engine
    x[:, 8] = np.log10(1.0 + r[:, C_REJ])
    x[:, 9] = np.log10(np.where(pdt > 0, dtreq / np.where(pdt > 0, pdt, 1.0), 1.0))
    x[:, 10] = r[:, C_PPC] / 12.0
    x[:, 11] = np.log10(r[:, C_MRATIO] + 1e-12)
    return x.astype(np.float32)


def dataset(rec):
    """(X, y) pairs: every accepted attempt that has a following attempt of
    the same system; y = log10(dtreq_next / dtreq)."""
    r = np.asarray(rec)
    X, Y = [], []
    for s in np.unique(r[:, C_SYS]):
        rs = r[r[:, C_SYS] == s]
        acc = np.nonzero(rs[:-1, C_ACC] > 0)[0]
        if len(acc) == 0:
            continue
        X.append(features(rs[acc]))
        Y.append(np.log10(rs[acc + 1, C_DTREQ] / rs[acc, C_DTREQ]))
    X = np.concatenate(X)
    Y = np.concatenate(Y).astype(np.float32)
    ok = np.all(np.isfinite(X), axis=1) & np.isfinite(Y)
    return X[ok], Y[ok]


def init_weights(seed=0):
    rng = np.random.default_rng(seed)
    IN, H = _lib.MLP_IN, _lib.MLP_HIDDEN
    return np.concatenate([rng.normal(0, np.sqrt(2.0 / IN), IN * H), np.zeros(H),
                           rng.normal(0, np.sqrt(2.0 / H), H * H), np.zeros(H),
                           rng.normal(0, np.sqrt(1.0 / H), H), np.zeros(1)]).astype(np.float32)


def fold_standardisation(w, mu, sd):
    """Weights acting on raw x from weights trained on (x - mu) / sd."""
    IN, H = _lib.MLP_IN, _lib.MLP_HIDDEN
    w = np.array(w, dtype=np.float64)
    W1 = w[: IN * H].reshape(IN, H)
    b1 = w[IN * H: IN * H + H]
    W1n = W1 / sd[:, None]
    b1n = b1 - (mu / sd) @ W1
    out = w.copy()
    out[: IN * H] = W1n.ravel()
    out[IN * H: IN * H + H] = b1n
    return out.astype(np.float32)


def train(X, y, epochs=300, lr=3e-3, batch=8192, seed=0, log_every=0):
    """Adam on the GPU with the HIP MLP kernels; returns (weights for raw
    features, standardisation, loss history)."""
    import torch
    from . import engine
    mu = X.mean(0).astype(np.float64)
    sd = X.std(0).astype(np.float64)
    const = sd < 1e-6  # constant inputs pass through unchanged
    mu[const] = 0.0
    sd[const] = 1.0
    Z = ((X - mu) / sd).astype(np.float32)
    dev = "cuda"
    Zd = torch.tensor(Z, device=dev)
    yd = torch.tensor(y, device=dev)
    w = torch.tensor(init_weights(seed), device=dev)
    m = torch.zeros_like(w)
    v = torch.zeros_like(w)
    rng = np.random.default_rng(seed)
    n = len(Z)
    step = 0
    hist = []
    for ep in range(epochs):
        order = torch.tensor(rng.permutation(n), device=dev)
        for b0 in range(0, n, batch):
            idx = order[b0:b0 + batch]
            g, loss = engine.mlp_loss_grad(w, Zd[idx].contiguous(), yd[idx].contiguous())
            step += 1
            engine.mlp_adam(w, g, m, v, step, lr=lr)
        if log_every and (ep % log_every == 0 or ep == epochs - 1):
            _, loss = engine.mlp_loss_grad(w, Zd, yd)
            hist.append((ep, float(loss.cpu())))
    wr = fold_standardisation(w.cpu().numpy(), mu, sd)
    return wr, (mu, sd), hist


def predict(w, X):
    import torch
    from . import engine
    return engine.mlp_forward(torch.tensor(w, device="cuda"), torch.tensor(X, device="cuda")).cpu().numpy()


def with_margin(w, margin):
    """Shift the output bias by -margin (dex)."""
    w = np.array(w, dtype=np.float32)
    w[-1] -= margin
    return w
