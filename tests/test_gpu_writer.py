"""Write_Overlap_Buffer (dalign/align.c:5815-6043): every buffered record,
thread buffers in order, stably sorted by (aread, bread, COMP, abpos) and
written as one .las.  The buffers are filled through AddOverlapToBuffer with
sorted runs full of equal keys (the merge path) and with many short runs (the
full stable sort); the file must equal Python's stable sort of the records
in buffer order.  (New_Align_Spec places its tables on the GPU, hence the
marker.)"""

import ctypes as C
import os
import struct

import numpy as np
import pytest


class Path(C.Structure):
    _fields_ = [("trace", C.c_void_p), ("tlen", C.c_int), ("diffs", C.c_int), ("abpos", C.c_int),
                ("bbpos", C.c_int), ("aepos", C.c_int), ("bepos", C.c_int)]


class Overlap(C.Structure):
    _fields_ = [("path", Path), ("flags", C.c_uint32), ("aread", C.c_int), ("bread", C.c_int)]


class IOBuf(C.Structure):
    _fields_ = [("tbytes", C.c_int), ("tmax", C.c_uint64), ("ttop", C.c_uint64), ("trace", C.c_void_p),
                ("no_trace", C.c_int), ("omax", C.c_int), ("otop", C.c_int), ("ovls", C.c_void_p)]


def _runs(rng, nbuf, per_run, nruns):
    """nruns sorted runs spread over nbuf buffers; keys drawn from a small
    range so that equal (aread, bread, comp, abpos) keys are frequent."""
    bufs = [[] for _ in range(nbuf)]
    serial = 0
    for r in range(nruns):
        keys = sorted((int(rng.integers(0, 40)), int(rng.integers(0, 30)), int(rng.integers(0, 2)),
                       int(rng.integers(0, 8))) for _ in range(per_run))
        for ar, br, cp, ab in keys:
            tl = 2 * int(rng.integers(0, 6))
            tr = bytes(int(x) for x in rng.integers(0, 256, tl))
            # diffs carries a serial number: the order of equal keys is visible in the file
            bufs[r % nbuf].append((ar, br, cp, ab, serial, tr))
            serial += 1
    return bufs


def _check(tmp_path, nbuf, per_run, nruns, seed):
    from marvel_amd.daligner import Daligner
    dal = Daligner()
    h = dal.h
    h.OVL_IO_Buffer.restype = C.POINTER(IOBuf)
    h.OVL_IO_Buffer.argtypes = [C.c_void_p]
    h.AddOverlapToBuffer.restype = C.c_int
    h.AddOverlapToBuffer.argtypes = [C.POINTER(IOBuf), C.POINTER(Overlap), C.c_int]
    spec = dal.new_align_spec(0.70, 100, [0.25] * 4, nthreads=nbuf)
    io = h.OVL_IO_Buffer(spec)
    rng = np.random.default_rng(seed)
    bufs = _runs(rng, nbuf, per_run, nruns)
    keep = []
    for t, recs in enumerate(bufs):
        for ar, br, cp, ab, serial, tr in recs:
            tb = C.create_string_buffer(tr, max(len(tr), 1))
            keep.append(tb)
            o = Overlap()
            o.path.trace = C.cast(tb, C.c_void_p).value if tr else None
            o.path.tlen = len(tr)
            o.path.diffs = serial
            o.path.abpos, o.path.aepos = ab, ab + 100
            o.path.bbpos, o.path.bepos = 0, 100
            o.flags, o.aread, o.bread = cp, ar, br
            assert h.AddOverlapToBuffer(C.pointer(io[t]), C.byref(o), 1) == 0
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        os.makedirs("001_00001", exist_ok=True)
        dal.write_overlap_buffer(spec, 1, b"X.1", b"X.1", 10**6)
        raw = open(os.path.join("001_00001", "X.1.X.1.las"), "rb").read()
    finally:
        os.chdir(cwd)
        dal.reset_overlap_buffer(spec)
    want = sorted((r for recs in bufs for r in recs), key=lambda r: (r[0], r[1], r[2], r[3]))
    n, ts = struct.unpack_from("<qi", raw, 0)
    assert n == len(want) and ts == 100
    o = 12
    for ar, br, cp, ab, serial, tr in want:
        tlen, diffs, abpos, _, aepos, _, flags, aread, bread, pad = struct.unpack_from("<iiiiiiIiii", raw, o)
        assert (aread, bread, flags, abpos, diffs, tlen, pad) == (ar, br, cp, ab, serial, len(tr), 0)
        o += 40
        assert raw[o:o + tlen] == tr
        o += tlen
    assert o == len(raw)


@pytest.mark.gpu
def test_writer_merges_sorted_runs_stably(tmp_path):
    print("[writer] 3 sorted runs, 3 x 60000 records", flush=True)
    _check(tmp_path, nbuf=4, per_run=60000, nruns=3, seed=1)


@pytest.mark.gpu
def test_writer_sorts_many_runs_stably(tmp_path):
    print("[writer] 40 sorted runs", flush=True)
    _check(tmp_path, nbuf=8, per_run=700, nruns=40, seed=2)
