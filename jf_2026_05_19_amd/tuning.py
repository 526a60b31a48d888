"""Persisted table of tuned solve-product launch shapes.

The device autotuner (``sol105_solve_autotune``) picks, per tree height and
direction, the product launch shape (kmax, ks, pch, kg) that runs fastest on a
model's own factor.  The best shape depends only on the height's shape
class: how many fronts it has and how large they are.  This module keys the
tuned shapes by that class and keeps them in a JSON file, so a new model
whose heights fall in known classes gets tuned shapes without a tuning run:

    key = "<fwd|bwd>:<fronts bucket>:<max m bucket>:<max ns bucket>"

with power-of-two buckets for the front count and 64-wide buckets for the
sizes.  ``apply`` sets the stored shapes for the classes it knows (defaults
elsewhere); ``tune`` runs the autotuner and merges its choices into the
table.
"""
from __future__ import annotations

import json
import os

import numpy as np

__all__ = ["shape_keys", "load", "save", "apply", "tune"]

DEFAULT = (384, 128, 0, 1)


def _bucket_pow2(n):
    n = max(int(n), 1)
    return 1 << (n - 1).bit_length()


def shape_keys(heights):
    """Table keys of each height (rows: fronts, max m, max ns, nonzeros) for
    the forward and backward launches."""
    keys = []
    for fr, mm, ns, _ in np.asarray(heights, dtype=np.int64):
        cls = "%d:%d:%d" % (_bucket_pow2(fr), (int(mm) + 63) // 64, (int(ns) + 63) // 64)
        keys.append(("fwd:" + cls, "bwd:" + cls))
    return keys


def load(path):
    if not path or not os.path.exists(path):
        return {}
    with open(path) as fh:
        data = json.load(fh)
    # tables written before the k-group count existed hold three values
    return {k: tuple(int(x) for x in v) + (1,) * (4 - len(v)) for k, v in data.get("shapes", {}).items()}


def save(path, table):
    tmp = path + ".tmp%d" % os.getpid()
    with open(tmp, "w") as fh:
        json.dump({"version": 2, "shapes": {k: list(v) for k, v in sorted(table.items())}}, fh, indent=1)
    os.replace(tmp, path)


def apply(dm, table):
    """Set the tabled shapes on a device model (or batch's model) that has
    been solved once; returns the number of heights the table covered."""
    keys = shape_keys(dm.solve_heights())
    cfg = np.zeros((len(keys), 2, 4), np.int32)
    hit = 0
    for h, (kf, kb) in enumerate(keys):
        for d, k in enumerate((kf, kb)):
            v = table.get(k)
            cfg[h, d] = v if v is not None else DEFAULT
            hit += v is not None
    dm.set_solve_cfg(cfg)
    return hit


def tune(dm, table=None, reps=7):
    """Autotune dm and merge its shapes into table (a dict, updated and
    returned with the autotuner's report)."""
    table = {} if table is None else table
    report = dm.autotune(reps)
    keys = shape_keys(dm.solve_heights())
    cfg = dm.solve_cfg()
    for h, (kf, kb) in enumerate(keys):
        table[kf] = tuple(int(x) for x in cfg[h, 0])
        table[kb] = tuple(int(x) for x in cfg[h, 1])
    return table, report
