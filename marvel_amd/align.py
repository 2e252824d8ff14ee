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


# ---- device-resident batches (the batched-aligner workload) ---------------------

TASK_INTS = C.sizeof(MiLaTask) // 4
RES_INTS = C.sizeof(MiLaResult) // 4


class DeviceTasks:
    """Tasks and their packed reads in device memory, as mi_gen_la_tasks writes them."""

    def __init__(self, packed, npacked, tasks, n, params=None):
        self.packed = packed      # DeviceBuffer, npacked uint64 words
        self.npacked = npacked
        self.tasks = tasks        # DeviceBuffer, n MiLaTask records
        self.n = n
        self.params = params

    def task_fields(self, count=None):
        """Host copy of the first `count` task records as int32 [count, 14]
        (aword and bword as two ints each, then alen, blen, comp, low, hgh,
        anti, lbord, hbord, selfie, pad)."""
        import numpy as np
        n = self.n if count is None else count
        return self.tasks.download(np.int32, n * TASK_INTS).reshape(n, TASK_INTS)

    def words(self):
        import numpy as np
        return self.packed.download(np.uint64, self.npacked)


def gen_tasks_device(n: int, seed: int, lmin: int = 1000, lmax: int = 60000, emin: float = 0.05,
                     emax: float = 0.20, offmax: int = 200) -> DeviceTasks:
    """n Local_Alignment tasks generated on the GPU (mi_gen_la_tasks)."""
    from .lib import DeviceBuffer, MiGenTaskParams

    g = MiGenTaskParams()
    g.seed, g.lmin, g.lmax = seed, lmin, lmax
    g.emin_ppm, g.emax_ppm, g.offmax = int(round(emin * 1e6)), int(round(emax * 1e6)), offmax
    npk = C.c_longlong(0)
    rc = lib.mi_gen_la_tasks(C.byref(g), n, None, 0, C.byref(npk), None)
    if rc != 0:
        raise RuntimeError(f"mi_gen_la_tasks failed ({rc}): {lib.last_error()}")
    packed = DeviceBuffer(8 * npk.value)
    tasks = DeviceBuffer(n * C.sizeof(MiLaTask))
    rc = lib.mi_gen_la_tasks(C.byref(g), n, packed.ptr, npk.value, C.byref(npk), tasks.ptr)
    if rc != 0:
        raise RuntimeError(f"mi_gen_la_tasks failed ({rc}): {lib.last_error()}")
    return DeviceTasks(packed, npk.value, tasks, n, g)


def trace_bound(fields, tspace: int) -> int:
    """Trace values a task set can produce: 4 (max(alen, blen) / tspace + 2) per task."""
    import numpy as np
    mx = np.maximum(fields[:, 4], fields[:, 5]).astype(np.int64)
    return int((4 * (mx // tspace + 2)).sum()) + 64


class DeviceBatchResult:
    def __init__(self, results, n, arena, used, work):
        self.results = results    # DeviceBuffer, n MiLaResult records
        self.n = n
        self.arena = arena        # DeviceBuffer of uint16 trace values
        self.used = used          # trace values written
        self.work = work          # executed cells, snake, kernel us per level (4), re-runs

    def fields(self):
        """Host copy of the results as a structured numpy array (MiLaResult fields)."""
        import numpy as np
        dt = np.dtype([(name, np.int32 if ct is C.c_int else np.int64) for name, ct in MiLaResult._fields_])
        return self.results.download(np.uint8, self.n * C.sizeof(MiLaResult)).view(dt)

    def traces(self):
        import numpy as np
        return self.arena.download(np.uint16, self.used)


def la_batch_device(dt: DeviceTasks, ave_corr: float = 0.70, tspace: int = 100, freq=(0.25, 0.25, 0.25, 0.25),
                    order=None, count=None) -> DeviceBatchResult:
    """Runs the first `count` tasks (all by default) of a DeviceTasks on the
    GPU (mi_la_batch_device); the trace arena is sized by trace_bound."""
    import numpy as np
    from .lib import DeviceBuffer

    n = dt.n if count is None else count
    bound = trace_bound(dt.task_fields(n), tspace)
    res = DeviceBuffer(n * C.sizeof(MiLaResult))
    arena = DeviceBuffer(2 * bound)
    used = DeviceBuffer(8)
    used.zero()
    work = (C.c_longlong * 8)()
    prm = params(ave_corr, tspace, freq)
    rc = lib.mi_la_batch_device(C.byref(prm), dt.packed.ptr, dt.npacked, dt.tasks.ptr, n,
                                None if order is None else order.ptr, res.ptr, arena.ptr, bound, used.ptr, work)
    if rc != 0:
        raise RuntimeError(f"mi_la_batch_device failed ({rc}): {lib.last_error()}")
    return DeviceBatchResult(res, n, arena, int(used.download(np.int64, 1)[0]), list(work))
