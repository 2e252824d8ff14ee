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


class OrPath(C.Structure):
    _fields_ = [("tlen", C.c_int), ("diffs", C.c_int), ("abpos", C.c_int), ("bbpos", C.c_int),
                ("aepos", C.c_int), ("bepos", C.c_int)]


def _trace_handle():
    h = handle()
    if not getattr(h, "_trace_ready", False):
        h.or_trace_tasks.restype = C.c_int64
        h.or_trace_tasks.argtypes = [C.c_double, C.c_int, C.POINTER(C.c_float), C.c_void_p, C.c_int, C.c_void_p,
                                     C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64]
        h.or_trace_pts.restype = C.c_int
        h.or_trace_pts.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(OrPath), C.c_void_p, C.c_int, C.c_int,
                                   C.c_void_p, C.c_int64, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        h._trace_ready = True
    return h


def run_trace_tasks(ts) -> np.ndarray:
    """Oracle Local_Alignment + Compute_Trace_PTS (3 modes) over a TaskSet, in
    the result layout of oracle/trace_vectors.c (int32 words)."""
    h = _trace_handle()
    seq, ao, bo = ts.blob()
    par = np.ascontiguousarray(ts.par, dtype=np.int32)
    cap = int(sum(2 + 3 * (2 + 2 * (int(p[0]) + int(p[1]))) for p in ts.par)) + 64
    res = np.zeros(cap, dtype=np.int32)
    freq = (C.c_float * 4)(*ts.freq)
    n = h.or_trace_tasks(ts.ave_corr, ts.tspace, freq, seq.ctypes.data, ts.n, ao.ctypes.data, bo.ctypes.data,
                         par.ctypes.data, res.ctypes.data, cap)
    if n < 0:
        raise RuntimeError("oracle trace capacity")
    return res[:n].copy()


def trace_segment(a: np.ndarray, b: np.ndarray, tspace: int, mode: int):
    """Compute_Trace_PTS of a one-segment path a[0..M) x b[0..N) (M <= tspace)
    with the given recorded difference count -> (diffs, trace)."""
    h = _trace_handle()
    buf = np.concatenate([[4], a, [4], b, [4]]).astype(np.uint8)
    M, N = len(a), len(b)
    path = OrPath(2, 0, 0, 0, M, N)
    pts = np.array([M + N, N], dtype=np.uint16)
    out = np.zeros(M + N + 16, dtype=np.int32)
    nt, df = C.c_int(), C.c_int()
    rc = h.or_trace_pts(buf.ctypes.data + 1, buf.ctypes.data + 2 + M, C.byref(path), pts.ctypes.data, tspace, mode,
                        out.ctypes.data, len(out), C.byref(nt), C.byref(df))
    if rc:
        raise RuntimeError(f"or_trace_pts {rc}")
    return df.value, out[:nt.value].copy()


def gen_la_pair(seed, task, lmin=1000, lmax=60000, emin_ppm=50000, emax_ppm=200000, offmax=200):
    """The batched-aligner workload generator restated on the CPU: (A bases,
    B bases, par[8], stats[10]) of one task (oracle/la_gen_oracle.c)."""
    h = handle()
    if not getattr(h, "_gen_ready", False):
        h.or_gen_la_pair.restype = C.c_int
        h.or_gen_la_pair.argtypes = [C.c_uint64, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p,
                                     C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
        h._gen_ready = True
    cap = 2 * lmax + 16
    a = np.zeros(cap, dtype=np.uint8)
    b = np.zeros(cap, dtype=np.uint8)
    par = np.zeros(8, dtype=np.int32)
    st = np.zeros(10, dtype=np.int32)
    rc = h.or_gen_la_pair(seed, task, lmin, lmax, emin_ppm, emax_ppm, offmax, a.ctypes.data, cap, b.ctypes.data, cap,
                          par.ctypes.data, st.ctypes.data)
    if rc != 0:
        raise RuntimeError("or_gen_la_pair capacity")
    return a[:par[0]].copy(), b[:par[1]].copy(), par, st
