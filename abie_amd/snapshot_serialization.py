"""Merge the ``Step#k`` groups of a snapshot file into flat datasets
(abie/snapshot_serialization.py).

    python -m abie_amd.snapshot_serialization -f data.hdf5 [--dataset x] [-o out.hdf5]

Every dataset of the step groups is concatenated along time, in step order;
per-body datasets keep the columns of the first step's bodies, matched by
particle hash, so bodies removed by a merge read NaN afterwards
(abie/snapshot_serialization.py:47-130).  Datasets outside the step groups
are copied.  ``hash`` is written as int64, everything else as float64.  The
output goes to ``-o`` or, as in the reference, to the input's base name with
its last extension removed (``p_sys_0.hdf5.unfinished`` → ``p_sys_0.hdf5``)
in the current directory.
"""
import argparse
import glob
import os

import numpy as np

from . import h5file


def _step_ids(f):
    out = []
    for name in f.keys("/"):
        if name.startswith("Step#"):
            out.append(int(name.split("#")[1]))
    return sorted(out)


def serialize(in_path, out_path, dataset=None):
    """Write the merged form of ``in_path`` to ``out_path``; returns the dict
    of merged arrays."""
    with h5file.File(in_path, "r") as f:
        steps = _step_ids(f)
        merged = {}
        loose = {}
        for name in f.keys("/"):
            if not name.startswith("Step#") and not f.is_group(name):
                loose[name] = f.read(name)
        if steps:
            first = None
            for s in steps:
                h = f.read("Step#%d/hash" % s)
                if h.size:
                    first = h[0]
                    break
            lens = [f.read("Step#%d/x" % s).shape[0] for s in steps]
            total = int(np.sum(lens))
            cursor = 0
            for s, n in zip(steps, lens):
                g = "Step#%d" % s
                h = f.read(g + "/hash")
                if h.size == 0:
                    continue
                names = [dataset] if dataset is not None else f.keys(g)
                cur = h[0]
                for key in names:
                    d = f.read("%s/%s" % (g, key)).astype(np.float64)
                    if key not in merged:
                        merged[key] = np.full((total,) + ((len(first),) if d.ndim == 2 else ()), np.nan)
                    if d.ndim == 2:
                        where = np.nonzero(np.isin(first, cur))[0]
                        src = np.array([int(np.nonzero(cur == first[w])[0][0]) for w in where], dtype=np.int64)
                        merged[key][cursor:cursor + n, where] = d[:, src]
                    else:
                        merged[key][cursor:cursor + n] = d
                cursor += n
    with h5file.File(out_path, "w") as out:
        for key, val in loose.items():
            out.create_dataset(key, val)
        for key, val in merged.items():
            if key == "hash":
                out.create_dataset(key, np.nan_to_num(val, nan=-1).astype(np.int64))
            else:
                out.create_dataset(key, val)
    return merged


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("-f", "--file", required=True, help="snapshot file(s), glob allowed")
    ap.add_argument("--dataset", default=None, help="merge only this dataset (e.g. mass)")
    ap.add_argument("-o", "--output", default=None, help="output file (one input only)")
    args = ap.parse_args(argv)
    paths = sorted(glob.glob(args.file))
    if args.output is not None and len(paths) > 1:
        ap.error("-o takes one input file")
    for path in paths:
        out = args.output or os.path.splitext(os.path.basename(path))[0]
        print("Processing %s -> %s" % (path, out))
        serialize(path, out, args.dataset)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
