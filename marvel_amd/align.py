"""Batched Local_Alignment on the GPU (host wrapper of mi_la_batch_host)."""

from __future__ import annotations

import ctypes as C

import numpy as np

from .lib import MiLaResult, MiLaTask, MiTraceResult, MiTraceTask, lib, params


def la_batch(ts, return_counters: bool = False):
    """Run every task of a latasks.TaskSet on the GPU.

    Returns a list of (fields, atrace, btrace) with fields = (abpos, bbpos,
    aepos, bepos, diffs, atlen, b_abpos, b_bbpos, b_aepos, b_bepos, diffs,
    btlen), the layout of oracle/la_vectors, plus the status array."""
    words, aw, bw = ts.packed()
    n = ts.n
    tasks = (MiLaTask * n)()
    for i in range(n):
        p = ts.par[i]
        t = tasks[i]
        t.aword, t.bword = int(aw[i]), int(bw[i])
        t.alen, t.blen, t.comp, t.low, t.hgh, t.anti, t.lbord, t.hbord = (int(x) for x in p)
        t.selfie = int(ts.selfie[i]) if ts.selfie is not None else 0
    res = (MiLaResult * n)()
    cap = int(sum(4 * (max(int(p[0]), int(p[1])) // ts.tspace + 2) for p in ts.par)) + 16
    tr = np.zeros(cap, dtype=np.uint16)
    prm = params(ts.ave_corr, ts.tspace, ts.freq)
    rc = lib.mi_la_batch_host(C.byref(prm), words.ctypes.data, words.shape[0], tasks, n, res, tr.ctypes.data, cap)
    if rc != 0:
        raise RuntimeError(f"mi_la_batch_host failed ({rc}): {lib.last_error()}")
    out = []
    status = np.zeros(n, dtype=np.int32)
    cells = snake = 0
    for i in range(n):
        r = res[i]
        status[i] = r.status
        f = (r.abpos, r.bbpos, r.aepos, r.bepos, r.diffs, r.atlen, r.b_abpos, r.b_bbpos, r.b_aepos, r.b_bepos,
             r.diffs, r.btlen)
        at = tr[r.atrace:r.atrace + r.atlen].copy()
        bt = tr[r.btrace:r.btrace + r.btlen].copy()
        out.append((f, at, bt))
        cells += r.cells
        snake += r.snake
    if return_counters:
        return out, status, (cells, snake)
    return out, status


def trace_pts(ts, paths, mode: int):
    """Compute_Trace_PTS on the GPU for the A-view paths of a TaskSet's
    alignments.  `paths` is the list la_batch returns (tasks whose status is
    not 0 are skipped: pass their entry as None).  Returns, per task, None or
    (diffs, int trace)."""
    words, aw, bw = ts.packed()
    idx = [i for i, p in enumerate(paths) if p is not None]
    n = len(idx)
    tasks = (MiTraceTask * max(n, 1))()
    pts = []
    off = 0
    cap = 16
    for q, i in enumerate(idx):
        f, at, _ = paths[i]
        t = tasks[q]
        t.aword, t.bword = int(aw[i]), int(bw[i])
        t.alen, t.blen = int(ts.par[i][0]), int(ts.par[i][1])
        t.abpos, t.bbpos, t.aepos, t.bepos = f[0], f[1], f[2], f[3]
        t.tlen = int(f[5])
        t.points = off
        t.selfie = 0
        pts.append(np.asarray(at, dtype=np.uint16))
        off += int(f[5])
        cap += (f[2] - f[0]) + (f[3] - f[1]) + 16
    points = np.concatenate(pts) if pts else np.zeros(1, dtype=np.uint16)
    res = (MiTraceResult * max(n, 1))()
    tr = np.zeros(cap, dtype=np.int32)
    rc = lib.mi_trace_pts_host(ts.tspace, mode, words.ctypes.data, words.shape[0], points.ctypes.data, points.shape[0],
                               tasks, n, res, tr.ctypes.data, cap)
    if rc != 0:
        raise RuntimeError(f"mi_trace_pts_host failed ({rc}): {lib.last_error()}")
    out = [None] * len(paths)
    for q, i in enumerate(idx):
        r = res[q]
        if r.status != 0:
            raise RuntimeError(f"task {i}: trace status {r.status}")
        out[i] = (r.diffs, tr[r.trace:r.trace + r.ntrace].copy())
    return out
