"""The allocation-free batched Local_Alignment entry point
(mi_la_batch_device_noalloc) against the allocating one (mi_la_batch_device)
on the same device-resident tasks: every task's status, path fields and
traces are identical, on two calls in a row, the second with fewer tasks
(the shared scratch keeps the first call's contents)."""

import ctypes as C

import numpy as np
import pytest

from marvel_amd.align import gen_tasks_device, la_batch_device, trace_bound
from marvel_amd.lib import DeviceBuffer, MiLaResult, lib, params

pytestmark = pytest.mark.gpu


def _noalloc(dt, n, tspace=100):
    h = lib.handle()
    bound = trace_bound(dt.task_fields(n), tspace)
    res = DeviceBuffer(n * C.sizeof(MiLaResult))
    arena = DeviceBuffer(2 * bound)
    used = DeviceBuffer(8)
    used.zero()
    prm = params(0.70, tspace, (0.25, 0.25, 0.25, 0.25))
    rc = h.mi_la_batch_device_noalloc(C.byref(prm), dt.packed.ptr, dt.npacked, dt.tasks.ptr, n, res.ptr, arena.ptr,
                                      bound, used.ptr)
    assert rc == 0, lib.last_error()
    dtp = np.dtype([(name, np.int32 if ct is C.c_int else np.int64) for name, ct in MiLaResult._fields_])
    u = int(used.download(np.int64, 1)[0])
    return res.download(np.uint8, n * C.sizeof(MiLaResult)).view(dtp), arena.download(np.uint16, u)


def _same(a_res, a_tr, b_res, b_tr, n):
    keys = ["status", "abpos", "bbpos", "aepos", "bepos", "diffs", "b_abpos", "b_bbpos", "b_aepos", "b_bepos",
            "atlen", "btlen", "cells", "snake"]
    for k in keys:
        assert np.array_equal(a_res[k][:n], b_res[k][:n]), k
    for i in range(n):
        x, y = a_res[i], b_res[i]
        for o, l in (("atrace", "atlen"), ("btrace", "btlen")):
            assert np.array_equal(a_tr[x[o]:x[o] + x[l]], b_tr[y[o]:y[o] + y[l]]), (i, o)


def test_noalloc_matches_device_batch_twice():
    print("[no-alloc] n=4000 then 700", flush=True)
    dt = gen_tasks_device(4000, 5)
    ref = la_batch_device(dt)
    ref_res, ref_tr = ref.fields(), ref.traces()
    assert (ref_res["status"] == 0).sum() >= 0.99 * dt.n
    got_res, got_tr = _noalloc(dt, dt.n)
    _same(ref_res, ref_tr, got_res, got_tr, dt.n)
    # second call, fewer tasks: the first 700 of the same set
    got2_res, got2_tr = _noalloc(dt, 700)
    _same(ref_res, ref_tr, got2_res, got2_tr, 700)


def test_noalloc_rejects_too_many_tasks():
    h = lib.handle()
    x = DeviceBuffer(256)
    prm = params(0.70, 100, (0.25, 0.25, 0.25, 0.25))
    rc = h.mi_la_batch_device_noalloc(C.byref(prm), x.ptr, 16, x.ptr, (1 << 20) + 1, x.ptr, x.ptr, 0, x.ptr)
    assert rc == -1 and "1048576" in lib.last_error()
