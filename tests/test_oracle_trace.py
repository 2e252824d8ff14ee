"""The oracle's Compute_Trace_PTS restatement (oracle/trace_oracle.c):
identical output words to the reference's own Compute_Trace_PTS on every
Local_Alignment golden set (digests in tests/golden/trace/), and segment
distances equal to a unit-cost edit-distance DP."""

import hashlib
import json
import os

import numpy as np
import pytest

import la_golden
import oracle_lib

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "trace")


@pytest.mark.parametrize("name", la_golden.SETS)
def test_oracle_trace_matches_reference(name):
    meta = json.load(open(os.path.join(GOLD, name + ".json")))
    ts, _ = la_golden.load(name)
    got = oracle_lib.run_trace_tasks(ts)
    assert len(got) == meta["words"]
    assert hashlib.sha256(got.tobytes()).hexdigest() == meta["sha256"]


def levenshtein(a, b):
    prev = np.arange(len(b) + 1)
    for i in range(1, len(a) + 1):
        cur = np.empty_like(prev)
        cur[0] = i
        for j in range(1, len(b) + 1):
            cur[j] = min(prev[j] + 1, cur[j - 1] + 1, prev[j - 1] + (a[i - 1] != b[j - 1]))
        prev = cur
    return int(prev[-1])


def apply_trace(a, b, trace):
    """Walk the alignment a trace describes and count its differences.

    Trace values (dalign/align.c:4897-4912): the path runs along its diagonal
    (match or substitution) up to the position a value names, then takes one
    indel there: p > 0 -- at B position p-1 an A base is aligned to a gap;
    p < 0 -- at A position -p-1 a B base is aligned to a gap."""
    i = j = 0
    diffs = 0
    for t in list(trace) + [None]:
        if t is None:
            ti, tj = len(a), len(b)
        elif t > 0:
            tj = t - 1
            ti = i + (tj - j)
        else:
            ti = -t - 1
            tj = j + (ti - i)
        while i < ti and j < tj:
            diffs += int(a[i] != b[j])
            i += 1
            j += 1
        if t is None:
            break
        if t > 0:
            i += 1
        else:
            j += 1
        diffs += 1
    assert i == len(a) and j == len(b)
    return diffs


@pytest.mark.parametrize("seed", range(12))
@pytest.mark.parametrize("mode", [-1, 0, 1])
def test_segment_distance_is_edit_distance(seed, mode):
    rng = np.random.default_rng(seed)
    a = rng.integers(0, 4, size=int(rng.integers(5, 60))).astype(np.uint8)
    b = a.copy()
    for _ in range(int(rng.integers(0, 12))):   # random edits
        op = rng.integers(0, 3)
        p = int(rng.integers(0, len(b) + 1))
        if op == 0 and p < len(b):
            b[p] = (b[p] + rng.integers(1, 4)) & 3
        elif op == 1:
            b = np.insert(b, p, rng.integers(0, 4)).astype(np.uint8)
        elif len(b) > 1 and p < len(b):
            b = np.delete(b, p)
    d, trace = oracle_lib.trace_segment(a, b, 100, mode)
    assert d == levenshtein(a, b)
    assert apply_trace(a, b, trace) == d
