"""Golden Local_Alignment vectors (tests/golden/la/) and comparison helpers."""

from __future__ import annotations

import gzip
import json
import os

import numpy as np

from marvel_amd import latasks

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "la")
SETS = sorted(f[:-5] for f in os.listdir(GOLD) if f.endswith(".json"))


def load(name):
    meta = json.load(open(os.path.join(GOLD, name + ".json")))
    kw = {k: meta[k] for k in meta if k not in ("task_sha256", "producer")}
    if "freq" in kw:
        kw["freq"] = tuple(kw["freq"])
    ts = latasks.make_self_tasks(**kw) if kw.pop("self", False) else latasks.make_tasks(**kw)
    assert ts.digest() == meta["task_sha256"], "task generator no longer reproduces the golden inputs"
    res = latasks.parse_results(gzip.open(os.path.join(GOLD, name + ".res.gz")).read(), ts.n)
    return ts, res


def diff(got, want, status=None, limit=5):
    """List of human-readable mismatches (empty when bit-identical)."""
    bad = []
    for i, (g, w) in enumerate(zip(got, want)):
        if status is not None and status[i] != 0:
            bad.append(f"task {i}: status {status[i]}")
        elif tuple(g[0]) != tuple(w[0]) or not np.array_equal(g[1], w[1]) or not np.array_equal(g[2], w[2]):
            bad.append(f"task {i}: got {tuple(g[0])} want {tuple(w[0])}")
        if len(bad) >= limit:
            break
    return bad
