"""ctypes binding of the oracle restatement (oracle/_bin/liboracle.so).

Test infrastructure: the checker for the HIP path, never the product."""

from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_SO = os.path.join(ORACLE_DIR, "_bin", "liboracle.so")
ORACLE_CLI = os.path.join(ORACLE_DIR, "_bin", "daligner_oracle")
REF_DIR = os.path.join(ORACLE_DIR, "_ref")

_h = None


def ensure_built():
    if not (os.path.exists(ORACLE_SO) and os.path.exists(ORACLE_CLI)):
        subprocess.check_call(["make", "-s", "-C", ORACLE_DIR, "oracle"])


def handle():
    global _h
    if _h is None:
        ensure_built()
        _h = C.CDLL(ORACLE_SO)
        _h.or_la_tasks.restype = C.c_int
        _h.or_la_tasks.argtypes = [C.c_double, C.c_int, C.POINTER(C.c_float), C.c_void_p, C.c_int,
                                   C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                   C.c_longlong]
    return _h


def run_tasks(ts):
    """Oracle Local_Alignment over a TaskSet -> (list of (fields12, atrace, btrace), status[n])."""
    h = handle()
    seq, ao, bo = ts.blob()
    n = ts.n
    par = np.ascontiguousarray(ts.par, dtype=np.int32)
    res = np.zeros((n, 13), dtype=np.int32)
    cap = int(sum(4 * (max(int(p[0]), int(p[1])) // ts.tspace + 2) for p in ts.par)) + 16
    tr = np.zeros(cap, dtype=np.uint16)
    toff = np.zeros(n + 1, dtype=np.int64)
    freq = (C.c_float * 4)(*ts.freq)
    rc = h.or_la_tasks(ts.ave_corr, ts.tspace, freq, seq.ctypes.data, n, ao.ctypes.data, bo.ctypes.data,
                       par.ctypes.data, res.ctypes.data, tr.ctypes.data, toff.ctypes.data, cap)
    if rc < 0:
        raise RuntimeError("oracle trace capacity")
    out = []
    for i in range(n):
        f = tuple(int(x) for x in res[i, :12])
        o = int(toff[i])
        at = tr[o:o + f[5]].copy()
        bt = tr[o + f[5]:o + f[5] + f[11]].copy()
        out.append((f, at, bt))
    return out, res[:, 12].copy()
