"""Writes the golden vectors of tests/golden/ from the compiled reference.

Run in the container that has /root/reference (after `make -C oracle ref`):
    python tests/golden/make_goldens.py
Inputs are regenerated from seeds by marvel_amd.latasks / marvel_amd.sim, so
only their SHA-256 and the reference's outputs are committed.

* la/<set>.res.gz  : oracle/_ref/la_vectors (reference dalign/align.c
                     Local_Alignment) results for the task set <set>
* la/<set>.json    : generator arguments + task-file digest
* las/*.json       : reference daligner .las SHA-256 per workload + command
"""

from __future__ import annotations

import gzip
import json
import os
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from marvel_amd import latasks  # noqa: E402

REF = os.path.join(ROOT, "oracle", "_ref")

LA_SETS = {
    # config-3 style (seed offset up to 200 bp off the true diagonal), short reads
    "off200": dict(n=400, seed=101, lmin=1000, lmax=8000),
    # seeds on the true diagonal, 1-20 kb
    "exact": dict(n=400, seed=102, lmin=1000, lmax=20000, offmax=0),
    # long reads 20-60 kb
    "long": dict(n=40, seed=103, lmin=20000, lmax=60000, offmax=0),
    # high divergence
    "hierr": dict(n=200, seed=104, lmin=1000, lmax=10000, emin=0.20, emax=0.35, offmax=5),
    # 16-bit trace regime (spacing > 125) and a different correlation
    "ts250": dict(n=200, seed=105, lmin=1000, lmax=10000, offmax=0, tspace=250, ave_corr=0.75),
    # biased base composition (bias factor below 1)
    "biased": dict(n=200, seed=106, lmin=1000, lmax=10000, offmax=3, freq=(0.35, 0.15, 0.15, 0.35)),
    # very short reads: seeds close to both ends, clipping paths
    "short": dict(n=400, seed=107, lmin=40, lmax=400, emin=0.0, emax=0.25, offmax=30),
    # a read against itself (the same bases): the selfie bounds keep the wave off the main diagonal
    "selfie": dict(self=True, n=400, seed=108),
}


def la_set(name):
    kw = dict(LA_SETS[name])
    if kw.pop("self", False):
        return latasks.make_self_tasks(**kw)
    return latasks.make_tasks(**kw)


def write_la(name):
    ts = la_set(name)
    with tempfile.TemporaryDirectory() as d:
        tf = os.path.join(d, "t.bin")
        rf = os.path.join(d, "r.bin")
        with open(tf, "wb") as f:
            f.write(ts.task_file_bytes())
        subprocess.check_call([os.path.join(REF, "la_vectors"), tf, rf])
        res = open(rf, "rb").read()
    os.makedirs(os.path.join(HERE, "la"), exist_ok=True)
    with open(os.path.join(HERE, "la", name + ".res.gz"), "wb") as raw:
        with gzip.GzipFile(fileobj=raw, mode="wb", compresslevel=9, mtime=0) as f:
            f.write(res)
    meta = dict(LA_SETS[name])
    meta["task_sha256"] = ts.digest()
    meta["n"] = ts.n
    meta["producer"] = "oracle/_ref/la_vectors (reference dalign/align.c Local_Alignment)"
    with open(os.path.join(HERE, "la", name + ".json"), "w") as f:
        json.dump(meta, f, indent=1, sort_keys=True)
    return ts.n, len(res)


if __name__ == "__main__":
    which = sys.argv[1:] or list(LA_SETS)
    for nm in which:
        if nm in LA_SETS:
            print(nm, write_la(nm))
