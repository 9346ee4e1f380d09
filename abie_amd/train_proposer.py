"""Train the learned GR15 step proposer and measure what it gains.

    python -m abie_amd.train_proposer [--systems 1024] [--years 3] [--dagger 2] [--out abie_amd/data/proposer.npz]

(the defaults are the settings of the committed abie_amd/data/proposer.npz)

1. Step logs: ``--systems`` seed systems (drawn like the ensemble workload:
   massive U{3..64}, test particles U{0..2048}, seeds disjoint from the
   benchmark's) are integrated by the device engine with the massive-only
   controller and the step log on.
2. Dataset: every accepted attempt that has a successor gives one sample
   (features of the accepted step, target log10(dtreq_next / dtreq)).
3. Training: the HIP kernels abie_mlp_loss_grad / abie_mlp_adam, on 90 % of
   the systems; the other 10 % are held out.
4. Margin: the output shift that maximises predicted step length per
   attempt on the held-out samples (a rejected attempt costs its step).
5. Aggregation (``--dagger`` rounds): fresh systems are integrated with the
   current network, their steps join the data set, and 3-4 are repeated.
6. Evaluation: a fresh set of systems integrated with and without the
   proposer; evaluations per simulated year, rejection rate, |dE/E|.
"""
import argparse
import json
import os
import time

import numpy as np

from . import engine, ics, proposer

G = 4 * np.pi ** 2
TRAIN_SEED = 20261203 + 500000
EVAL_SEED = 20261203 + 900000


def collect_logs(n_sys, years, seed, force_mode="fp64", cap=None, weights=None):
    """Step logs of n_sys seed systems under the massive-only controller, or
    under the learned one driven by ``weights``."""
    systems = ics.ensemble(n_sys, seed=seed)
    ens = engine.Ensemble(systems, G=G, force_mode=force_mode,
                          controller="massive" if weights is None else "learned")
    if weights is not None:
        ens.set_proposer(weights)
    cap = cap or n_sys * 4000
    ens.set_steplog(cap)
    ens.run(years)
    rec = ens.steplog()
    ens.close()
    return rec


CLAMP_LO, CLAMP_HI = -0.3, 0.0  # the device's clamp of the proposer output (gr15_impl.hpp)


def split(rec):
    """(train X, y, held-out X, y): systems with index % 10 == 9 held out."""
    hold = rec[:, proposer.C_SYS] % 10 == 9
    Xtr, ytr = proposer.dataset(rec[~hold])
    Xte, yte = proposer.dataset(rec[hold])
    return Xtr, ytr, Xte, yte


def choose_margin(pred, y, grid=np.linspace(-0.2, 0.3, 51)):
    """Margin that maximises expected step length per attempt: a proposal
    dtreq*10^(p-m) (clamped as on the device) is accepted when p - m <= y
    (the next dtreq ratio)."""
    best, best_m = -1.0, 0.0
    for m in grid:
        prop = np.clip(pred - m, CLAMP_LO, CLAMP_HI)
        ok = prop <= y
        gain = np.mean(np.where(ok, 10.0 ** prop, 0.0))
        if gain > best:
            best, best_m = gain, m
    return float(best_m), float(best)


def evaluate(weights, n_sys, years, force_mode="fp64", max_seconds=120.0):
    """Massive-only controller against the learned one on the same systems.
    Each run is bounded by ``max_seconds``; systems that did not reach
    ``years`` are reported (``unfinished``) and rates use the time reached."""
    out = {}
    systems = ics.ensemble(n_sys, seed=EVAL_SEED)
    for name, ctrl, w in (("massive", "massive", None), ("learned", "learned", weights)):
        ens = engine.Ensemble(systems, G=G, force_mode=force_mode, controller=ctrl)
        if w is not None:
            ens.set_proposer(w)
        e0 = ens.energy()
        t0 = time.perf_counter()
        done = ens.run(years, max_seconds=max_seconds)
        wall = time.perf_counter() - t0
        _, _, t, _ = ens.state()
        cnt = ens.counters()
        e1 = ens.energy()
        ens.close()
        late = np.nonzero(t < years)[0]
        out[name] = {"done": bool(done), "wall_s": wall, "system_yr_per_s": float(t.sum() / wall),
                     "evals_per_yr": float(cnt[:, 0].sum() / t.sum()),
                     "attempts": int(cnt[:, 1].sum()), "accepted": int(cnt[:, 2].sum()),
                     "reject_rate": float(1 - cnt[:, 2].sum() / max(cnt[:, 1].sum(), 1)),
                     "proposals": int(cnt[:, 4].sum()), "proposals_rejected": int(cnt[:, 5].sum()),
                     "max_dE_E": float(np.nanmax(np.abs((e1 - e0) / e0))),
                     "unfinished": [(int(k), float(t[k])) for k in late[:20]], "n_unfinished": int(len(late))}
        print(name, json.dumps(out[name]), flush=True)
    out["evals_ratio"] = out["massive"]["evals_per_yr"] / out["learned"]["evals_per_yr"]
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--systems", type=int, default=1024)
    ap.add_argument("--years", type=float, default=3.0)
    ap.add_argument("--epochs", type=int, default=60)
    ap.add_argument("--eval-systems", type=int, default=256)
    ap.add_argument("--eval-years", type=float, default=2.0)
    ap.add_argument("--eval-seconds", type=float, default=60.0)
    ap.add_argument("--dagger", type=int, default=2, help="rounds of data aggregation under the learned controller")
    ap.add_argument("--out", default=os.path.join(os.path.dirname(__file__), "data", "proposer.npz"))
    args = ap.parse_args()
    t0 = time.perf_counter()
    rec = collect_logs(args.systems, args.years, TRAIN_SEED)
    print("step log: %d attempts from %d systems (%.1f s)" % (len(rec), args.systems, time.perf_counter() - t0),
          flush=True)
    Xtr, ytr, Xte, yte = split(rec)
    q = np.quantile(ytr, [0.001, 0.01, 0.5, 0.99, 0.999])
    print("samples: train %d, held out %d; target quantiles (0.1%%, 1%%, 50%%, 99%%, 99.9%%) %s"
          % (len(Xtr), len(Xte), np.array2string(q, precision=3)), flush=True)
    # Rounds of dataset aggregation: the learned controller visits states the
    # massive-only controller's logs do not contain (h below dtreq, fewer
    # rejections), so fresh systems are integrated with the current network
    # and their steps added to the training set.
    for rnd in range(args.dagger + 1):
        w, (mu, sd), hist = proposer.train(Xtr, ytr, epochs=args.epochs, log_every=max(1, args.epochs // 6))
        pred = proposer.predict(w, Xte)
        mse = float(np.mean((pred - yte) ** 2))
        base = float(np.mean(yte ** 2))
        margin, gain = choose_margin(pred, yte)
        base_gain = float(np.mean(np.where(0.0 <= yte, 1.0, 0.0)))
        print("round %d: loss %.4e, held-out mse %.4e (predict-0 %.4e); margin %.3f dex, step/attempt %.3f vs plain %.3f"
              % (rnd, hist[-1][1], mse, base, margin, gain, base_gain), flush=True)
        wm = proposer.with_margin(w, margin)
        if rnd == args.dagger:
            break
        rec2 = collect_logs(args.systems // 2, args.years, TRAIN_SEED + 100_000 * (rnd + 1), weights=wm)
        X2, y2, X2e, y2e = split(rec2)
        Xtr, ytr = np.concatenate([Xtr, X2]), np.concatenate([ytr, y2])
        Xte, yte = np.concatenate([Xte, X2e]), np.concatenate([yte, y2e])
        print("round %d: +%d samples under the learned controller" % (rnd + 1, len(X2)), flush=True)
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    report = {"train_systems": args.systems, "train_years": args.years, "samples": int(len(Xtr)),
              "dagger_rounds": args.dagger,
              "heldout_mse": mse, "heldout_mse_zero": base, "margin_dex": margin,
              "heldout_step_per_attempt": gain, "heldout_step_per_attempt_plain": base_gain,
              "target_quantiles": q.tolist(), "loss_history": hist}
    np.savez(args.out, weights=wm, mu=mu, sd=sd, report=json.dumps(report))
    print("saved", args.out, "; evaluating", flush=True)
    ev = evaluate(wm, args.eval_systems, args.eval_years, max_seconds=args.eval_seconds)
    print(json.dumps(ev, indent=1), flush=True)
    report["evaluation"] = ev
    np.savez(args.out, weights=wm, mu=mu, sd=sd, report=json.dumps(report))
    print("saved", args.out)


if __name__ == "__main__":
    main()
