"""What the trained step proposer gains over the same engine without it.

The committed network (abie_amd/data/proposer.npz, trained on step logs of
1,024 seed systems) against the massive-only controller it extends, on
systems drawn from the ensemble distribution with seeds disjoint from the
training set: fewer force evaluations per simulated year, fewer rejected
attempts, energy error of the same order."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

G = 4 * np.pi ** 2
WEIGHTS = os.path.join(os.path.dirname(__file__), "..", "abie_amd", "data", "proposer.npz")


def run(systems, ctrl, w, years):
    from abie_amd import engine
    ens = engine.Ensemble(systems, G=G, controller=ctrl)
    if w is not None:
        ens.set_proposer(w)
    e0 = ens.energy()
    assert ens.run(years, max_seconds=60)
    _, _, t, _ = ens.state()
    c = ens.counters()
    e1 = ens.energy()
    ens.close()
    return c, t, np.abs((e1 - e0) / e0)


def test_trained_proposer_saves_evaluations():
    from abie_amd import ics
    from abie_amd.train_proposer import EVAL_SEED
    w = np.load(WEIGHTS)["weights"]
    systems = ics.ensemble(128, seed=EVAL_SEED + 77)
    cm, tm, dem = run(systems, "massive", None, 1.0)
    cl, tl, del_ = run(systems, "learned", w, 1.0)
    ev_m = cm[:, 0].sum() / tm.sum()
    ev_l = cl[:, 0].sum() / tl.sum()
    rej_m = 1 - cm[:, 2].sum() / cm[:, 1].sum()
    rej_l = 1 - cl[:, 2].sum() / cl[:, 1].sum()
    print("evals/yr massive %.1f learned %.1f (ratio %.3f); rejections %.3f -> %.3f; "
          "max |dE/E| %.2e -> %.2e" % (ev_m, ev_l, ev_m / ev_l, rej_m, rej_l, dem.max(), del_.max()))
    assert np.all(cl[:, 4] > 0)          # every system consulted the network
    assert ev_l < 0.9 * ev_m             # at least 10 % fewer force evaluations
    assert rej_l < rej_m
    assert np.median(del_) < 1e-13 and np.quantile(del_, 0.95) < 1e-11
