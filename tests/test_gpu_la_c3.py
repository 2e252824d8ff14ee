"""Batched Local_Alignment at the batched-aligner task model (read-pair
lengths log-uniform in [1 kb, 60 kb], pair error U[5 %, 20 %] with
ins:del:sub = 9:4:2, seed offset up to 200 along B).

* 10,000 host-built tasks through mi_la_batch_host, each result checked
  against size-independent properties of a daligner alignment, and a sample
  bit-exact against the oracle restatement;
* the device generator against its CPU restatement, byte for byte;
* the full workload: 2,000,000 tasks generated in HBM and aligned from one
  device queue (mi_la_batch_device), every result checked against the same
  properties on the GPU, and a sample of 1,500 regenerated on the CPU and
  aligned by the oracle, bit-exact;
* a launch stopped by the host's stop flag reports its unrun tasks."""

import ctypes
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import la_golden
import oracle_lib
from marvel_amd import latasks
from marvel_amd.align import gen_tasks_device, la_batch, la_batch_device
from marvel_amd.lib import lib

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


def _oracle_pairs(seed, idx):
    """CPU-regenerated tasks `idx` as a TaskSet (oracle/la_gen_oracle.c)."""
    reads, par = [], []
    for i in idx:
        a, b, p, _ = oracle_lib.gen_la_pair(seed, int(i))
        reads += [a, b]
        par.append(p)
    return latasks.TaskSet(0.70, 100, (0.25, 0.25, 0.25, 0.25), reads, np.array(par, dtype=np.int32))


def test_c3_generator_matches_oracle():
    n = 3000
    print(f"[la c3 gen] n={n}", flush=True)
    dt = gen_tasks_device(n, 3)
    f = dt.task_fields()
    words = dt.words()
    sh = (2 * np.arange(32, dtype=np.uint64))[None, :]

    def unpack(w0, ln):
        nw = (ln + 31) // 32
        return ((words[w0:w0 + nw, None] >> sh) & np.uint64(3)).astype(np.uint8).ravel()[:ln]

    for i in range(n):
        a, b, par, _ = oracle_lib.gen_la_pair(3, i)
        aw = int(f[i][0]) | (int(f[i][1]) << 32)
        bw = int(f[i][2]) | (int(f[i][3]) << 32)
        assert tuple(f[i][4:12]) == tuple(par), (i, f[i], par)
        assert np.array_equal(unpack(aw, len(a)), a), i
        assert np.array_equal(unpack(bw, len(b)), b), i
        # guard words around every read are zero
        assert words[aw - 1] == 0 and words[bw - 1] == 0
        assert words[aw + (len(a) + 31) // 32] == 0 and words[bw + (len(b) + 31) // 32 + 1] == 0


def _check_all(tf, res, trace, T):
    """The alignment properties of check_task, vectorised over every task."""
    st = res["status"]
    ok = st == 0
    ab, bb, ae, be, diffs = (res[k].astype(np.int64) for k in ("abpos", "bbpos", "aepos", "bepos", "diffs"))
    bab, bbb, bae, bbe = (res[k].astype(np.int64) for k in ("b_abpos", "b_bbpos", "b_aepos", "b_bepos"))
    atl, btl = res["atlen"].astype(np.int64), res["btlen"].astype(np.int64)
    at, bt = res["atrace"], res["btrace"]
    alen, blen, comp = tf[:, 4].astype(np.int64), tf[:, 5].astype(np.int64), tf[:, 6]
    bad = lambda m: int((ok & ~m).sum())
    assert bad((ab >= 0) & (ab <= ae) & (ae <= alen) & (bb >= 0) & (bb <= be) & (be <= blen)) == 0
    assert bad((atl % 2 == 0) & (btl % 2 == 0)) == 0
    c = comp == 1
    vb = np.where(c, (bab == blen - be) & (bae == blen - bb) & (bbb == alen - ae) & (bbe == alen - ab),
                  (bab == bb) & (bae == be) & (bbb == ab) & (bbe == ae))
    assert bad(vb) == 0
    used = len(trace)
    assert bad((at >= 0) & (at % 2 == 0) & (bt == at + atl) & (bt + btl <= used)) == 0
    tr = trace.astype(np.int64)
    even = tr.copy()
    even[1::2] = 0
    odd = tr - even
    ce = np.concatenate([[0], np.cumsum(even)])
    co = np.concatenate([[0], np.cumsum(odd)])
    cl = lambda v: np.clip(v, 0, used)
    assert bad((co[cl(at + atl)] - co[cl(at)]) == be - bb) == 0
    assert bad((ce[cl(at + atl)] - ce[cl(at)]) == diffs) == 0
    assert bad((co[cl(bt + btl)] - co[cl(bt)]) == ae - ab) == 0
    seg = lambda b, e: np.where(e > b, (e - 1) // T - b // T + 1, 0)
    ne = (ae > ab) & (be > bb)
    assert bad(~ne | ((atl // 2) == seg(ab, ae))) == 0
    assert bad(~ne | ((btl // 2) == seg(bab, bae))) == 0
    return int(ok.sum()), int((ok & ne & ((ae - ab) * 2 > alen)).sum())


def test_c3_full_size_batch():
    """2,000,000 tasks from one device queue."""
    n = 2_000_000
    print(f"[la c3 full] n={n} generating", flush=True)
    t0 = time.perf_counter()
    dt = gen_tasks_device(n, 3)
    t1 = time.perf_counter()
    print(f"[la c3 full] generated {dt.npacked * 32 / 1e9:.1f} Gbases of words in {t1 - t0:.2f} s; aligning",
          flush=True)
    r = la_batch_device(dt)
    t2 = time.perf_counter()
    res = r.fields()
    st = np.bincount(res["status"], minlength=8)
    print(f"[la c3 full] aligned in {t2 - t1:.2f} s, statuses {st.tolist()}, work {r.work}, "
          f"{r.used} trace values", flush=True)
    assert st[0] + st[3] == n, st   # every task finished (3: the reference's undefined wave state)
    assert st[0] >= 0.99 * n
    trace = r.traces()
    tf = dt.task_fields()
    nok, nlong = _check_all(tf, res, trace, 100)
    print(f"[la c3 full] properties hold for {nok} tasks ({nlong} aligned over half of A)", flush=True)
    assert nlong >= 0.15 * n

    # a sample regenerated on the CPU and aligned by the oracle
    rng = np.random.default_rng(33)
    idx = np.sort(rng.choice(n, 1500, replace=False))

    def run(chunk):
        return oracle_lib.run_tasks(_oracle_pairs(3, chunk))

    with ThreadPoolExecutor(12) as ex:
        outs = list(ex.map(run, np.array_split(idx, 12)))
    want = [w for o in outs for w in o[0]]
    wst = np.concatenate([o[1] for o in outs])
    got, gst = [], []
    for i in idx:
        x = res[i]
        f = (int(x["abpos"]), int(x["bbpos"]), int(x["aepos"]), int(x["bepos"]), int(x["diffs"]), int(x["atlen"]),
             int(x["b_abpos"]), int(x["b_bbpos"]), int(x["b_aepos"]), int(x["b_bepos"]), int(x["diffs"]),
             int(x["btlen"]))
        got.append((f, trace[x["atrace"]:x["atrace"] + x["atlen"]], trace[x["btrace"]:x["btrace"] + x["btlen"]]))
        gst.append(int(x["status"]))
    gst = np.array(gst)
    keep = [j for j in range(len(idx)) if wst[j] == 0]
    assert len(keep) >= 0.99 * len(idx)
    assert (gst[wst == 3] == 3).all() and (gst[wst == 0] == 0).all()
    bad = la_golden.diff([got[j] for j in keep], [want[j] for j in keep], gst[keep])
    assert not bad, "\n".join(bad[:20])
    print(f"[la c3 full] {len(keep)} sampled tasks bit-identical to the oracle", flush=True)


def test_stop_flag_reports_unrun_tasks():
    """With the stop flag raised from the host the batch kernel exits at once
    (normally) and the call reports the tasks it left, instead of returning
    stale records; lowered again, the same call succeeds."""
    ts = latasks.make_tasks(50, 8, lmin=1000, lmax=3000, offmax=0)
    print("[la stop flag] raising", flush=True)
    assert lib.mi_abort_device_work(1) == 0
    try:
        with pytest.raises(RuntimeError, match="-4"):
            la_batch(ts)
    finally:
        assert lib.mi_abort_device_work(0) == 0
    got, status = la_batch(ts)
    want, st = oracle_lib.run_tasks(ts)
    keep = [i for i in range(ts.n) if st[i] == 0]
    assert not la_golden.diff([got[i] for i in keep], [want[i] for i in keep], status[keep])
