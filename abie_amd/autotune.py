"""Autotuner of the GR15 engine's group shapes.

    python -m abie_amd.autotune [--out abie_amd/data/gr15_tuning.json]

For each size bucket (upper bound on bodies per system) it builds an ensemble
of identical copies of one planetary system (32 massive bodies, the rest test
particles, sized at the bucket's upper bound; identical copies do identical
work, so the wall time measures the engine rather than the slowest system),
runs it with every engine configuration (threads per group x bodies per
thread) that holds it, and keeps the configuration with the most force
evaluations per second.  The
result is a rules table {max_bodies -> configuration} that the library uses
through abie_gr15_set_tuning (applied automatically when the table exists).
"""
import argparse
import ctypes
import json
import os
import time

import numpy as np

from . import _lib, engine, ics

BUCKETS = [64, 128, 256, 512, 768, 1024, 1536, 2048, 2560]
DEFAULT_OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "gr15_tuning.json")


def configs():
    lib = _lib.load()
    out = []
    for c in range(lib.abie_gr15_num_configs()):
        nt, bpt = ctypes.c_int(), ctypes.c_int()
        _lib.check(lib.abie_gr15_config(c, ctypes.byref(nt), ctypes.byref(bpt)))
        out.append((nt.value, bpt.value))
    return out


def set_rules(rules):
    lib = _lib.load()
    mx = (ctypes.c_int * len(rules))(*[r[0] for r in rules])
    cf = (ctypes.c_int * len(rules))(*[r[1] for r in rules])
    _lib.check(lib.abie_gr15_set_tuning(len(rules), mx, cf), "abie_gr15_set_tuning")


def time_config(systems, cfg, n_bodies, years, controller="plain", force_mode="fp64", min_wall=0.4):
    """Force evaluations per second of one configuration; the timed run is
    extended until it lasts at least ``min_wall`` seconds."""
    set_rules([(n_bodies, cfg)])
    ens = engine.Ensemble(systems, force_mode=force_mode, controller=controller)
    t = years * 0.1
    ens.run(t)  # warm: first launch, code object load
    ev0 = int(ens.counters()[:, 0].sum())
    span = years
    t0 = time.perf_counter()
    while True:
        t += span
        ens.run(t)
        wall = time.perf_counter() - t0
        if wall >= min_wall:
            break
        span *= 2.0
    evals = int(ens.counters()[:, 0].sum()) - ev0
    ens.close()
    return evals / wall, wall


def tune(n_copies=1024, years=0.01, nm=32, verbose=True):
    cfgs = configs()
    rules, table = [], {}
    for b in BUCKETS:
        m = min(nm, b // 2)
        systems = [ics.planetary_system(3000 + b, m, b - m)] * n_copies
        best = None
        table[b] = {}
        for c, (nt, bpt) in enumerate(cfgs):
            if b > nt * bpt or nt * bpt >= 4 * b or m > nt:
                continue  # does not hold it, or far too large a group
            try:
                rate, wall = time_config(systems, c, b, years)
            except _lib.AbieHipError as e:
                if verbose:
                    print("bucket %d cfg %d failed: %s" % (b, c, e))
                continue
            table[b]["%dx%d" % (nt, bpt)] = rate
            if verbose:
                print("bodies <= %5d  cfg %2d (%4d x %d): %.3e evals/s (%.2f s)" % (b, c, nt, bpt, rate, wall),
                      flush=True)
            if best is None or rate > best[1]:
                best = (c, rate)
        if best is not None:
            rules.append((b, best[0]))
    return rules, table


def load_rules(path=DEFAULT_OUT):
    """Apply a saved rules table, if present; returns the rules or None."""
    try:
        with open(path) as f:
            rules = [tuple(r) for r in json.load(f)["rules"]]
    except (OSError, ValueError, KeyError):
        return None
    set_rules(rules)
    return rules


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=DEFAULT_OUT)
    ap.add_argument("--copies", type=int, default=1024)
    ap.add_argument("--years", type=float, default=0.01)
    args = ap.parse_args()
    rules, table = tune(args.copies, args.years)
    os.makedirs(os.path.dirname(args.out), exist_ok=True)
    cfgs = configs()
    with open(args.out, "w") as f:
        json.dump({"rules": rules, "configs": cfgs, "evals_per_s": table,
                   "method": "copies=%d, years=%g, 32 massive bodies, plain controller, fp64" %
                             (args.copies, args.years)}, f, indent=1)
    print("rules:", rules)


if __name__ == "__main__":
    main()
