"""CPU restatement of the .las utilities after the overlap path (test
infrastructure only: the checker of the GPU LAsort and the host LAcat, never
the product path).

  las_sort  dalign/LAsort.c:60-86, :183-235 -- records ordered by (aread, bread,
            abpos); qsort is glibc's merge sort, so equal keys keep their input
            order (a stable sort)
  las_merge utils/LAmerge.c:306-506 (heap merge), :17-133 (-s), SORT_OVL at
            utils/LAmergeUtils.c:50-80 -- the records of all inputs in
            (aread, bread, COMP, abpos) order; for sorted inputs whose equal
            keys lie in one file, and for -s, equal keys keep input order
  las_cat   dalign/LAcat.c:66-190 -- ROOT.1.las, ROOT.2.las, ... up to the first
            missing file; header (sum of counts, first file's tspace); the
            records of each file in order
"""

from __future__ import annotations

import os
import struct

import numpy as np

REC = np.dtype([("tlen", "<i4"), ("diffs", "<i4"), ("abpos", "<i4"), ("bbpos", "<i4"), ("aepos", "<i4"),
                ("bepos", "<i4"), ("flags", "<u4"), ("aread", "<i4"), ("bread", "<i4"), ("pad", "<i4")])


def split_records(raw: bytes):
    """(novl, tspace, list of record byte strings) of a .las image."""
    novl, tspace = struct.unpack_from("<qi", raw, 0)
    tb = 1 if tspace <= 125 else 2
    out, o = [], 12
    for _ in range(novl):
        tlen = struct.unpack_from("<i", raw, o)[0]
        n = 40 + tlen * tb
        out.append(raw[o:o + n])
        o += n
    return novl, tspace, out


def las_sort(raw: bytes) -> bytes:
    novl, tspace, recs = split_records(raw)
    if not recs:
        return raw[:12]
    hdr = np.frombuffer(b"".join(r[:40] for r in recs), dtype=REC)
    order = np.lexsort((hdr["abpos"], hdr["bread"], hdr["aread"]))   # stable, last key primary
    return raw[:12] + b"".join(recs[i] for i in order)


def las_cat(root: str) -> bytes:
    """The bytes LAcat writes to stdout for source ROOT (with or without .las)."""
    if root.endswith(".las"):
        root = root[:-4]
    parts, i = [], 1
    while os.path.exists(f"{root}.{i}.las"):
        parts.append(open(f"{root}.{i}.las", "rb").read())
        i += 1
    novl, tspace, body = 0, 0, []
    for k, raw in enumerate(parts):
        n, ts, recs = split_records(raw)
        if k == 0:
            tspace = ts
        elif ts != tspace:
            raise ValueError(f"PT-point spacing conflict ({tspace} vs {ts})")
        novl += n
        body += recs
    return struct.pack("<qi", novl, tspace) + b"".join(body)


def las_merge(raws) -> bytes:
    """LAmerge's output for the .las images `raws` (in input order)."""
    if len(raws) == 1:
        return raws[0]
    novl, tspace, recs = 0, None, []
    for raw in raws:
        n, ts, r = split_records(raw)
        if tspace is None:
            tspace = ts
        elif ts != tspace:
            raise ValueError(f"PT-point spacing conflict ({tspace} vs {ts})")
        novl += n
        recs += r
    hdr = struct.pack("<qi", novl, tspace)
    if not recs:
        return hdr
    h = np.frombuffer(b"".join(r[:40] for r in recs), dtype=REC)
    order = np.lexsort((h["abpos"], h["flags"] & 1, h["bread"], h["aread"]))
    return hdr + b"".join(recs[i] for i in order)
