"""GPU Compute_Trace_PTS (mi_trace_pts_host): on the A paths the GPU aligner
computes for every Local_Alignment golden set, the traces of all three
modes are word for word the reference's own Compute_Trace_PTS output
(SHA-256 digests in tests/golden/trace/)."""

import hashlib
import json
import os

import numpy as np
import pytest

import la_golden
from marvel_amd.align import la_batch, trace_pts

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "trace")


@pytest.mark.gpu
@pytest.mark.parametrize("name", la_golden.SETS)
def test_gpu_trace_matches_reference(name):
    print(f"[gpu trace] set {name}", flush=True)
    meta = json.load(open(os.path.join(GOLD, name + ".json")))
    ts, _ = la_golden.load(name)
    paths, status = la_batch(ts)
    kept = [paths[i] if status[i] == 0 else None for i in range(ts.n)]
    modes = [trace_pts(ts, kept, m) for m in (-1, 0, 1)]
    words = []
    for i in range(ts.n):
        if status[i] != 0:
            words += [3, 0]
            continue
        words += [0, int(paths[i][0][5])]
        for m in range(3):
            d, tr = modes[m][i]
            words += [d, len(tr)] + [int(x) for x in tr]
    got = np.array(words, dtype=np.int32)
    assert len(got) == meta["words"]
    assert hashlib.sha256(got.tobytes()).hexdigest() == meta["sha256"]
