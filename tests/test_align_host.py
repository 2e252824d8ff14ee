"""Host entry points of align.h in libmarvl_align (Complement_Seq,
Flip_Alignment, Print_Overlap) against the reference's own functions: the
commands of tests/golden/flip/cases.txt and what the reference printed for
them (tests/golden/make_flip_goldens.py).  No GPU call is made."""

import ctypes as C
import os

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "marvel_amd", "libmarvl_align.so")
GOLD = os.path.join(ROOT, "tests", "golden", "flip")


class Path(C.Structure):
    _fields_ = [("trace", C.c_void_p), ("tlen", C.c_int), ("diffs", C.c_int), ("abpos", C.c_int),
                ("bbpos", C.c_int), ("aepos", C.c_int), ("bepos", C.c_int)]


class Alignment(C.Structure):
    _fields_ = [("path", C.POINTER(Path)), ("flags", C.c_uint32), ("aseq", C.c_void_p), ("bseq", C.c_void_p),
                ("alen", C.c_int), ("blen", C.c_int)]


class Overlap(C.Structure):
    _fields_ = [("path", Path), ("flags", C.c_uint32), ("aread", C.c_int), ("bread", C.c_int)]


def _run(lib, libc, line):
    f = line.split()
    op, v = f[0], [int(x) for x in f[1:]]
    if op == "C":
        n = v[0]
        a = (C.c_char * max(n, 1))(*[bytes([x]) for x in v[1:1 + n]])
        lib.Complement_Seq(a, n)
        return " ".join(str(a[i][0]) for i in range(n)) + "\n"
    if op == "F":
        comp, full, alen, blen, ab, ae, bb, be, tlen = v[:9]
        t = (C.c_int * max(tlen, 1))(*v[9:9 + tlen])
        p = Path(C.cast(t, C.c_void_p), tlen, 0, ab, bb, ae, be)
        sa, sb = C.create_string_buffer(1), C.create_string_buffer(1)
        al = Alignment(C.pointer(p), comp, C.cast(sa, C.c_void_p), C.cast(sb, C.c_void_p), alen, blen)
        lib.Flip_Alignment(C.byref(al), full)
        swapped = int(al.aseq == C.cast(sb, C.c_void_p).value and al.bseq == C.cast(sa, C.c_void_p).value)
        vals = [al.alen, al.blen, p.abpos, p.aepos, p.bbpos, p.bepos, swapped] + [t[i] for i in range(tlen)]
        return " ".join(map(str, vals)) + "\n"
    tbytes, indent, aread, bread, flags, ab, ae, bb, be, diffs, tlen = v[:11]
    ty = C.c_uint8 if tbytes == 1 else C.c_uint16
    t = (ty * max(tlen, 1))(*v[11:11 + tlen])
    o = Overlap(Path(C.cast(t, C.c_void_p) if tlen else None, tlen, diffs, ab, bb, ae, be), flags, aread, bread)
    buf, size = C.c_char_p(), C.c_size_t()
    fp = libc.open_memstream(C.byref(buf), C.byref(size))
    lib.Print_Overlap(fp, C.byref(o), tbytes, indent)
    libc.fclose(fp)
    out = C.string_at(buf, size.value).decode()
    libc.free(buf)
    return out


def test_host_entry_points_match_reference():
    lib = C.CDLL(LIB)
    libc = C.CDLL(None)
    libc.open_memstream.restype = C.c_void_p
    libc.open_memstream.argtypes = [C.POINTER(C.c_char_p), C.POINTER(C.c_size_t)]
    libc.fclose.argtypes = [C.c_void_p]
    libc.free.argtypes = [C.c_void_p]
    lib.Print_Overlap.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int]
    cases = open(os.path.join(GOLD, "cases.txt")).read().splitlines()
    want = open(os.path.join(GOLD, "expected.txt")).read().split("%\n")[:-1]
    assert len(cases) == len(want) == 107
    kinds = {"C": 0, "F": 0, "P": 0}
    for line, w in zip(cases, want):
        got = _run(lib, libc, line)
        if line[0] == "C" and line.split()[1] == "0":
            got = "\n"
        assert got == w, (line[:60], got[:200], w[:200])
        kinds[line[0]] += 1
    assert kinds == {"C": 7, "F": 60, "P": 40}


def test_complement_seq_is_an_involution():
    lib = C.CDLL(LIB)
    import numpy as np
    rng = np.random.default_rng(3)
    for n in (1, 2, 5, 64, 999):
        x = rng.integers(0, 4, n).astype(np.uint8)
        a = C.create_string_buffer(x.tobytes(), n)
        lib.Complement_Seq(a, n)
        got = np.frombuffer(a.raw[:n], dtype=np.uint8)
        assert np.array_equal(got, (3 - x[::-1]).astype(np.uint8))
        lib.Complement_Seq(a, n)
        assert np.array_equal(np.frombuffer(a.raw[:n], dtype=np.uint8), x)
