"""The background .las writer (host code, no GPU call): records added with
AddOverlapToBuffer, written with Write_Overlap_Buffer while the next records
are appended, give the files of a synchronous writer -- stable order by
(aread, bread, COMP, abpos), thread buffers in order -- including a second
write without a reset in between (the old records are written again, as in
the reference) and a process that exits without flushing."""

import ctypes as C
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import las_tools_oracle  # noqa: E402

from marvel_amd.lib import LIB_PATH  # noqa: E402


class Path(C.Structure):
    _fields_ = [("trace", C.c_void_p), ("tlen", C.c_int), ("diffs", C.c_int), ("abpos", C.c_int), ("bbpos", C.c_int),
                ("aepos", C.c_int), ("bepos", C.c_int)]


class Overlap(C.Structure):
    _fields_ = [("path", Path), ("flags", C.c_uint32), ("aread", C.c_int), ("bread", C.c_int)]


class IOBuf(C.Structure):   # Overlap_IO_Buffer
    _fields_ = [("tbytes", C.c_int), ("tmax", C.c_uint64), ("ttop", C.c_uint64), ("trace", C.c_void_p),
                ("no_trace", C.c_int), ("omax", C.c_int), ("otop", C.c_int), ("ovls", C.c_void_p)]


def _lib():
    h = C.CDLL(LIB_PATH)
    h.New_Align_Spec.argtypes = [C.c_double, C.c_int, C.POINTER(C.c_float), C.c_int, C.c_int, C.c_int, C.c_int]
    h.New_Align_Spec.restype = C.c_void_p
    h.OVL_IO_Buffer.argtypes = [C.c_void_p]
    h.OVL_IO_Buffer.restype = C.c_void_p
    h.AddOverlapToBuffer.argtypes = [C.c_void_p, C.POINTER(Overlap), C.c_int]
    h.Write_Overlap_Buffer.argtypes = [C.c_void_p, C.c_int, C.c_char_p, C.c_char_p, C.c_int]
    h.Reset_Overlap_Buffer.argtypes = [C.c_void_p]
    h.Free_Align_Spec.argtypes = [C.c_void_p]
    return h


def _records(seed, n):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        tl = 2 * int(rng.integers(0, 8))
        out.append(dict(aread=int(rng.integers(0, 30)), bread=int(rng.integers(0, 30)), flags=int(rng.integers(0, 2)),
                        abpos=int(rng.integers(0, 5)) * 100, bbpos=int(rng.integers(0, 900)),
                        aepos=int(rng.integers(1000, 3000)), bepos=int(rng.integers(1000, 3000)),
                        diffs=int(rng.integers(0, 99)), trace=bytes(rng.integers(0, 200, size=tl, dtype=np.uint8))))
    return out


def _add(h, spec, recs, nthreads):
    buf = h.OVL_IO_Buffer(spec)
    size = C.sizeof(IOBuf)
    keep = []
    for i, r in enumerate(recs):
        t = C.create_string_buffer(r["trace"], max(1, len(r["trace"])))
        keep.append(t)
        o = Overlap(Path(C.cast(t, C.c_void_p), len(r["trace"]), r["diffs"], r["abpos"], r["bbpos"], r["aepos"],
                         r["bepos"]), r["flags"], r["aread"], r["bread"])
        assert h.AddOverlapToBuffer(buf + size * (i % nthreads), C.byref(o), 1) == 0


def _image(recs, nthreads):
    """The .las a synchronous writer makes: records in thread-buffer order,
    then the stable (aread, bread, COMP, abpos) order."""
    order = [r for t in range(nthreads) for i, r in enumerate(recs) if i % nthreads == t]
    body = b""
    for r in order:
        body += np.array([len(r["trace"]), r["diffs"], r["abpos"], r["bbpos"], r["aepos"], r["bepos"], r["flags"],
                          r["aread"], r["bread"], 0], dtype="<i4").tobytes() + r["trace"]
    raw = np.array([len(order)], "<i8").tobytes() + np.array([100], "<i4").tobytes() + body
    return las_tools_oracle.las_merge([raw, np.array([0], "<i8").tobytes() + np.array([100], "<i4").tobytes()])


def test_overlapped_writes_match_synchronous_files(tmp_path):
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        h = _lib()
        f = (C.c_float * 4)(.25, .25, .25, .25)
        nt = 4
        spec = h.New_Align_Spec(0.7, 100, f, nt, 1, 0, 0)
        r1, r2, r3 = _records(1, 3000), _records(2, 2000), _records(3, 500)
        _add(h, spec, r1, nt)
        h.Write_Overlap_Buffer(spec, 1, b"X", b"X", 0)        # runs in the background
        h.Reset_Overlap_Buffer(spec)                           # trades buffers with the spare set
        _add(h, spec, r2, nt)                                  # appends while the first write may run
        h.Write_Overlap_Buffer(spec, 1, b"Y", b"Y", 0)
        _add(h, spec, r3, nt)                                  # no reset: appends wait for the write
        h.Write_Overlap_Buffer(spec, 1, b"Z", b"Z", 0)        # holds r2 and r3, as in the reference
        h.Reset_Overlap_Buffer(spec)
        assert h.mi_flush_overlap_writes() == 0
        h.Free_Align_Spec(spec)
        assert (tmp_path / "X.X.las").read_bytes() == _image(r1, nt)
        assert (tmp_path / "Y.Y.las").read_bytes() == _image(r2, nt)
        assert (tmp_path / "Z.Z.las").read_bytes() == _image(r2 + r3, nt)
    finally:
        os.chdir(cwd)


def test_exit_waits_for_the_last_write(tmp_path):
    """A host that exits right after Write_Overlap_Buffer (as the reference
    daligner.c does) still leaves the whole file."""
    code = f"""
import sys, os
sys.path.insert(0, {os.path.join(ROOT, 'tests')!r}); sys.path.insert(0, {ROOT!r})
import ctypes as C
import test_writer_async as t
os.chdir({str(tmp_path)!r})
h = t._lib()
spec = h.New_Align_Spec(0.7, 100, (C.c_float * 4)(.25, .25, .25, .25), 2, 1, 0, 0)
t._add(h, spec, t._records(7, 20000), 2)
h.Write_Overlap_Buffer(spec, 1, b"E", b"E", 0)
sys.exit(0)
"""
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    assert (tmp_path / "E.E.las").read_bytes() == _image(_records(7, 20000), 2)
