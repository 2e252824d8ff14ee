"""Python host side of the GPU daligner: the reference's filter interface
(dalign/filter.h Set_Filter_Params / Sort_Kmers / Match_Filter and
dalign/align.h New_Align_Spec / Write_Overlap_Buffer / Reset_Overlap_Buffer)
over libmarvl_align's C-ABI, plus its device-resident variants.

    dal = Daligner()                                   # loads the library
    dal.configure(nthreads=8, mem_limit_gb=32)         # driver globals, -j / -M
    dal.set_filter_params(14, 6, 0, 35)                # -k -w -t -h
    A = dal.open_block("DIR/C2.1")
    spec = dal.new_align_spec(.70, 100, A.freq)        # -e -s
    dal.compare_self(A, spec, run_id=1, name="C2.1")   # both strands + .las

Errors raise DalignerError with the library's message; nothing here falls
back to the CPU.
"""

from __future__ import annotations

import ctypes as C
import os
import time

import numpy as np

from .lib import LIB_PATH, LibraryMissing


class DalignerError(RuntimeError):
    pass


class HITS_READ(C.Structure):   # db/DB.h:164-171
    _fields_ = [("rlen", C.c_int), ("boff", C.c_int64), ("coff", C.c_int64), ("flags", C.c_int)]


class HITS_DB(C.Structure):     # db/DB.h:213-234
    _fields_ = [("ureads", C.c_int), ("freq", C.c_float * 4), ("maxlen", C.c_int), ("totlen", C.c_int64),
                ("nreads", C.c_int), ("part", C.c_int), ("ufirst", C.c_int), ("path", C.c_char_p),
                ("loaded", C.c_int), ("bases", C.c_void_p), ("reads", C.POINTER(HITS_READ)),
                ("tracks", C.c_void_p)]


class HITS_TRACK(C.Structure):  # db/DB.h:179-185
    pass


HITS_TRACK._fields_ = [("next", C.POINTER(HITS_TRACK)), ("name", C.c_char_p), ("size", C.c_int),
                       ("anno", C.c_void_p), ("data", C.c_void_p)]


class MiFilterStats(C.Structure):
    _fields_ = [("nkmers_a", C.c_longlong), ("nkmers_b", C.c_longlong), ("nhits", C.c_longlong),
                ("npairs", C.c_longlong), ("njobs", C.c_longlong), ("naligned", C.c_longlong),
                ("nrecords", C.c_longlong), ("cells", C.c_longlong), ("snake", C.c_longlong),
                ("edge_events", C.c_longlong), ("edge_fixups", C.c_longlong), ("retries", C.c_longlong),
                ("limit", C.c_int), ("t_filter", C.c_double), ("t_report", C.c_double), ("t_copy", C.c_double),
                ("t_total", C.c_double), ("t_report_kernel", C.c_double), ("report_launches", C.c_longlong),
                ("t_seed_sort", C.c_double), ("seed_sort_passes", C.c_longlong), ("njobs_kept", C.c_longlong),
                ("wave_steps", C.c_longlong), ("cand_scored", C.c_longlong), ("cand_skipped", C.c_longlong),
                ("redund", C.c_longlong * 9), ("graphs_captured", C.c_longlong), ("graphs_replayed", C.c_longlong),
                ("launches_replayed", C.c_longlong)]

    def as_dict(self):
        v = {k: getattr(self, k) for k, _ in self._fields_}
        v["redund"] = list(v["redund"])
        return v


assert C.sizeof(HITS_READ) == 32 and C.sizeof(HITS_DB) == 88


class Daligner:
    def __init__(self):
        if not os.path.exists(LIB_PATH):
            raise LibraryMissing(f"{LIB_PATH} not built: run `python -m marvel_amd.build`")
        h = C.CDLL(LIB_PATH)
        vp, i, db = C.c_void_p, C.c_int, C.POINTER(HITS_DB)
        sig = {
            "Set_Filter_Params": ([i, i, i, i], i),
            "New_Align_Spec": ([C.c_double, i, C.POINTER(C.c_float), i, i, i, i], vp),
            "Free_Align_Spec": ([vp], None),
            "Sort_Kmers": ([db, C.POINTER(i)], vp),
            "mi_match_filter": ([C.c_char_p, db, C.c_char_p, db, vp, vp, i, vp], i),
            "mi_filter_stats": ([C.POINTER(MiFilterStats)], None),
            "mi_sort_kmers_complement": ([vp, db, C.POINTER(i)], vp),
            "mi_block_device": ([db], vp),
            "mi_sort_kmers_device": ([vp, C.POINTER(i)], vp),
            "mi_free_kmer_index": ([vp], None),
            "mi_device_synchronize": ([], i),
            "mi_open_block": ([C.c_char_p, db], i),
            "mi_close_block": ([db], None),
            "mi_complement_header": ([db, db], None),
            "mi_load_mask_tracks": ([db, C.POINTER(C.c_char_p), i], i),
            "mi_release_tracks": ([db], None),
            "Write_Overlap_Buffer": ([vp, i, C.c_char_p, C.c_char_p, i], None),
            "Reset_Overlap_Buffer": ([vp], None),
            "mi_flush_overlap_writes": ([], i),
            "mi_last_error": ([], C.c_char_p),
        }
        for name, (args, res) in sig.items():
            f = getattr(h, name)
            f.argtypes = args
            f.restype = res
        self.h = h

    # ---- driver globals (dalign/daligner.c:60-71) ----
    def configure(self, nthreads=4, mem_limit_gb=None, minover=2000, symmetric=1, hgap_min=0, verbose=0, biased=0,
                  identity=0):
        """The driver's globals (dalign/daligner.c:60-71): -j, -M, -l, -A, -H, -v, -b, -I.  -O is
        new_align_spec(..., only_identity=1) with identity=1."""
        if nthreads < 1 or nthreads & (nthreads - 1):
            raise DalignerError("nthreads must be a power of 2")
        C.c_int.in_dll(self.h, "NTHREADS").value = nthreads
        C.c_int.in_dll(self.h, "NSHIFT").value = nthreads.bit_length() - 1
        if mem_limit_gb is not None:
            C.c_uint64.in_dll(self.h, "MEM_LIMIT").value = int(mem_limit_gb) * 0x40000000
        C.c_int.in_dll(self.h, "MINOVER").value = minover
        C.c_int.in_dll(self.h, "SYMMETRIC").value = symmetric
        C.c_int.in_dll(self.h, "HGAP_MIN").value = hgap_min
        C.c_int.in_dll(self.h, "VERBOSE").value = verbose
        C.c_int.in_dll(self.h, "BIASED").value = biased
        C.c_int.in_dll(self.h, "IDENTITY").value = identity

    def _err(self, what):
        raise DalignerError(f"{what}: {self.h.mi_last_error().decode(errors='replace')}")

    def set_filter_params(self, kmer=14, binshift=6, suppress=0, hitmin=35):
        if self.h.Set_Filter_Params(kmer, binshift, suppress, hitmin):
            self._err("Set_Filter_Params")

    def open_block(self, path: str) -> HITS_DB:
        b = HITS_DB()
        if self.h.mi_open_block(path.encode(), C.byref(b)):
            self._err(f"open {path}")
        return b

    def close_block(self, b: HITS_DB):
        self.h.mi_close_block(C.byref(b))

    def load_mask_tracks(self, b: HITS_DB, names):
        """read_DB's -m tracks (dalign/daligner.c:373-431): b.tracks becomes
        their union (one track)."""
        arr = (C.c_char_p * len(names))(*[n.encode() for n in names])
        if self.h.mi_load_mask_tracks(C.byref(b), arr, len(names)):
            self._err("load mask tracks")

    def release_tracks(self, b: HITS_DB):
        self.h.mi_release_tracks(C.byref(b))

    @staticmethod
    def mask_points(b: HITS_DB):
        """Per read the point list of b's mask track (None without one)."""
        if not b.tracks:
            return None
        t = C.cast(b.tracks, C.POINTER(HITS_TRACK)).contents
        anno = np.ctypeslib.as_array(C.cast(t.anno, C.POINTER(C.c_int64)), shape=(b.nreads + 1,)).copy()
        n = int(anno[-1] - anno[0])
        pts = np.ctypeslib.as_array(C.cast(t.data, C.POINTER(C.c_int32)), shape=(max(n, 1),))[:n].copy()
        return [pts[anno[i] - anno[0]:anno[i + 1] - anno[0]] for i in range(b.nreads)]

    def new_align_spec(self, ave_corr, tspace, freq, nthreads=4, symmetric=1, only_identity=0, no_trace=0):
        f = (C.c_float * 4)(*freq)
        return self.h.New_Align_Spec(ave_corr, tspace, f, nthreads, symmetric, only_identity, no_trace)

    def sort_kmers(self, b: HITS_DB):
        n = C.c_int()
        t = self.h.Sort_Kmers(C.byref(b), C.byref(n))
        if not t:
            self._err("Sort_Kmers")
        return t

    def block_device(self, b: HITS_DB):
        t = self.h.mi_block_device(C.byref(b))
        if not t:
            self._err("mi_block_device")
        return t

    def sort_kmers_device(self, dev):
        t = self.h.mi_sort_kmers_device(dev, None)
        if not t:
            self._err("mi_sort_kmers_device")
        return t

    def sort_kmers_complement(self, table, cblock: HITS_DB):
        t = self.h.mi_sort_kmers_complement(table, C.byref(cblock), None)
        if not t:
            self._err("mi_sort_kmers_complement")
        return t

    def complement_header(self, b: HITS_DB) -> HITS_DB:
        c = HITS_DB()
        self.h.mi_complement_header(C.byref(b), C.byref(c))
        return c

    def match_filter(self, aname, A, bname, B, atable, btable, comp, spec) -> dict:
        """Match_Filter; `aname is bname` makes it a self comparison (the
        reference compares the name pointers).  Returns the pass statistics."""
        if self.h.mi_match_filter(aname, C.byref(A), bname, C.byref(B), atable, btable, comp, spec):
            self._err("Match_Filter")
        st = MiFilterStats()
        self.h.mi_filter_stats(C.byref(st))
        return st.as_dict()

    def free_index(self, t):
        self.h.mi_free_kmer_index(t)

    def write_overlap_buffer(self, spec, run_id, aname: bytes, bname: bytes, last_read: int, wait: bool = True):
        """Write_Overlap_Buffer; the write runs in the background, so with
        wait=False the .las files are complete only after flush_writes()."""
        self.h.Write_Overlap_Buffer(spec, run_id, aname, bname, last_read)
        if wait:
            self.flush_writes()

    def flush_writes(self):
        if self.h.mi_flush_overlap_writes():
            raise DalignerError("a .las write failed")

    def reset_overlap_buffer(self, spec):
        self.h.Reset_Overlap_Buffer(spec)

    def synchronize(self):
        if self.h.mi_device_synchronize():
            self._err("device synchronize")

    def compare_self(self, A: HITS_DB, spec, run_id: int, name: str, dev=None, cA=None, wait: bool = True):
        """One block against itself, both strands, .las written in the
        current directory's run folder (dalign/daligner.c:994-1024).  With
        `dev` (a mi_block_device handle) the index is built from reads
        already in HBM.  wait=False leaves the .las write running in the
        background (overlapping the next call; flush_writes() waits for it).
        Returns (forward stats, complement stats)."""
        nm = C.create_string_buffer(name.encode())
        os.makedirs("%03d_%05d" % (run_id, A.part), exist_ok=True)   # createSubdir (dalign/daligner.c:583-611)
        t = [time.perf_counter()]
        a = self.sort_kmers_device(dev) if dev is not None else self.sort_kmers(A)
        t.append(time.perf_counter())
        try:
            s1 = self.match_filter(nm, A, nm, A, a, a, 0, spec)
            t.append(time.perf_counter())
            own = cA is None
            cA = self.complement_header(A) if own else cA
            try:
                c = self.sort_kmers_complement(a, cA)
                t.append(time.perf_counter())
                s2 = self.match_filter(nm, A, nm, cA, a, c, 1, spec)   # consumes c
                t.append(time.perf_counter())
            finally:
                if own:
                    self.release_tracks(cA)   # the complemented mask, if A has one
        finally:
            self.free_index(a)
        self.write_overlap_buffer(spec, run_id, nm, nm, A.ufirst + A.nreads - 1, wait=wait)
        self.reset_overlap_buffer(spec)
        t.append(time.perf_counter())
        self.last_phases = {k: round(t[i + 1] - t[i], 4) for i, k in
                            enumerate(("index", "forward", "complement_index", "complement", "write"))}
        return s1, s2
