"""Batched Local_Alignment at the config-3 task model (read-pair lengths
log-uniform in [1 kb, 60 kb], pair error U[5 %, 20 %] with ins:del:sub =
9:4:2, seed offset up to 200 along B): 10,000 tasks on the GPU, each result
checked against size-independent properties of a daligner alignment, and a
sample of the same model bit-exact against the oracle restatement."""

import time

import numpy as np
import pytest

import la_golden
import oracle_lib
from marvel_amd import latasks
from marvel_amd.align import la_batch

pytestmark = pytest.mark.gpu


def _segments(beg, end, ts):
    """trace segments of [beg, end) between trace points at multiples of ts"""
    if end <= beg:
        return 0
    return (end - 1) // ts - beg // ts + 1


def test_c3_batch_properties():
    ts = latasks.make_tasks(10000, 3)
    print(f"[la c3] n={ts.n}", flush=True)
    t0 = time.perf_counter()
    got, status, (cells, snake) = la_batch(ts, return_counters=True)
    dt = time.perf_counter() - t0
    ok = np.flatnonzero(status == 0)
    print(f"[la c3] {len(ok)} aligned in {dt:.2f} s (host call), W_align {cells + snake}", flush=True)
    assert len(ok) >= 0.99 * ts.n
    assert set(np.unique(status)) <= {0, 3}
    T = ts.tspace
    nonempty = 0
    for i in ok:
        check_task(ts, i, got[i], T)
        nonempty += got[i][0][2] > got[i][0][0]
    assert nonempty >= 0.4 * len(ok)   # seeds off the true diagonal by up to 200 often find no path


def check_task(ts, i, res, T):
    """Properties every daligner alignment has (dalign/align.c:970-1107,
    :1709-1901, :2004-2023): traces with one pair per trace-point segment,
    the A view's pairs summing to the path's differences and B extent, the B
    view's advances summing to the A extent, and the B view's coordinates
    the A view's seen from B (complemented for COMP)."""
    alen, blen, comp = int(ts.par[i][0]), int(ts.par[i][1]), int(ts.par[i][2])
    (ab, bb, ae, be, diffs, atl, bab, bbb, bae, bbe, _, btl), at, bt = res
    at = np.asarray(at, dtype=np.int64)
    bt = np.asarray(bt, dtype=np.int64)
    assert 0 <= ab <= ae <= alen and 0 <= bb <= be <= blen, i
    assert len(at) == atl and len(bt) == btl and atl % 2 == 0 and btl % 2 == 0, i
    if comp:
        assert (bab, bae, bbb, bbe) == (blen - be, blen - bb, alen - ae, alen - ab), i
    else:
        assert (bab, bae, bbb, bbe) == (bb, be, ab, ae), i
    assert at[1::2].sum() == be - bb and at[0::2].sum() == diffs, i
    assert bt[1::2].sum() == ae - ab, i
    if ae > ab and be > bb:
        assert atl // 2 == _segments(ab, ae, T), i
        assert btl // 2 == _segments(bab, bae, T), i   # B's own trace points (B-view coordinates)
    # a segment's differences never exceed its two lengths together
    assert (at[0::2] <= T + at[1::2]).all(), i


def test_c3_sample_matches_oracle():
    ts = latasks.make_tasks(60, 33)
    print(f"[la c3 oracle] n={ts.n}", flush=True)
    want, st = oracle_lib.run_tasks(ts)
    got, status = la_batch(ts)
    keep = [i for i in range(ts.n) if st[i] == 0]
    assert len(keep) > 0.9 * ts.n
    bad = la_golden.diff([got[i] for i in keep], [want[i] for i in keep], status[keep])
    assert not bad, "\n".join(bad)
