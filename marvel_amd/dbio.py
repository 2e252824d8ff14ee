"""MARVEL DB files for synthetic reads, and .las reading.

write_db() produces the three files daligner reads for a DB ROOT
(db/DB.h:164-249), as FA2db + DBsplit would for a FASTA of the same reads
in the same order with one read per PacBio well:

  ROOT.db     text stub: "files = %9d", one "  %9d %s %s" line per FASTA file,
              then the DBsplit part "blocks = %9d", "size = %9lld" and
              blocks+1 lines " %9d" with the first read of every block
              (db/DBsplit.c:200-243: a block closes once its bases reach
              size * 10^6)
  .ROOT.idx   88-byte DB record (ureads, freq[4], maxlen, totlen; the rest
              zero here where FA2db writes in-memory pointers) followed by one
              32-byte record per read (rlen, boff into .bps, coff = -1,
              flags = DB_BEST), db/FA2db.c:573-586
  .ROOT.bps   per read ceil(rlen/4) bytes, base i at bits 6-2(i%4)
              (Compress_Read, db/DB.c:261-281)
freq[c] = (float)(count[c] / totlen) (db/FA2db.c:1101-1105).
"""

from __future__ import annotations

import os
import struct

import numpy as np

DB_BEST = 0x0800


def pack_bps(read: np.ndarray) -> bytes:
    n = read.shape[0]
    m = (n + 3) // 4
    pad = np.zeros(m * 4, dtype=np.uint8)
    pad[:n] = read
    q = pad.reshape(m, 4)
    return ((q[:, 0] << 6) | (q[:, 1] << 4) | (q[:, 2] << 2) | q[:, 3]).astype(np.uint8).tobytes()


def block_starts(lengths, size_mbp: int):
    """First read of every block, DBsplit's rule (db/DBsplit.c:204-236)."""
    size = size_mbp * 1000000
    starts = [0]
    tot = 0
    ireads = 0
    for i, L in enumerate(lengths):
        ireads += 1
        tot += int(L)
        if tot >= size:
            starts.append(i + 1)
            tot = 0
            ireads = 0
    if ireads > 0:
        starts.append(len(lengths))
    return starts


def write_db(root: str, reads, size_mbp: int = 400, prolog: str = "sim", fasta: str = "sim") -> dict:
    """Writes ROOT.db, .ROOT.idx, .ROOT.bps; returns the DB statistics."""
    d, name = os.path.split(root)
    d = d or "."
    lens = np.array([len(r) for r in reads], dtype=np.int64)
    counts = np.zeros(4, dtype=np.int64)
    for r in reads:
        counts += np.bincount(r, minlength=4)[:4]
    totlen = int(lens.sum())
    maxlen = int(lens.max()) if len(reads) else 0
    freq = [np.float32((1.0 * int(c)) / totlen) if totlen else np.float32(0) for c in counts]
    with open(os.path.join(d, "." + name + ".bps"), "wb") as fb, open(os.path.join(d, "." + name + ".idx"), "wb") as fi:
        hdr = struct.pack("<i4fiq", len(reads), *[float(f) for f in freq], maxlen, totlen)
        fi.write(hdr + b"\0" * (88 - len(hdr)))
        off = 0
        recs = []
        for r in reads:
            b = pack_bps(r)
            fb.write(b)
            recs.append(struct.pack("<iiqqii", len(r), 0, off, -1, DB_BEST, 0))
            off += len(b)
        fi.write(b"".join(recs))
    starts = block_starts(lens, size_mbp)
    with open(os.path.join(d, name + ".db"), "w") as fs:
        fs.write("files = %9d\n" % 1)
        fs.write("  %9d %s %s\n" % (len(reads), prolog, fasta))
        fs.write("blocks = %9d\n" % (len(starts) - 1))
        fs.write("size = %9d\n" % size_mbp)
        for s in starts:
            fs.write(" %9d\n" % s)
    return {"nreads": len(reads), "totlen": totlen, "maxlen": maxlen, "freq": [float(f) for f in freq],
            "blocks": len(starts) - 1}


def read_idx(root: str):
    """(header fields, per-read (rlen, boff) array) of a DB's .idx."""
    d, name = os.path.split(root)
    with open(os.path.join(d or ".", "." + name + ".idx"), "rb") as f:
        raw = f.read()
    ureads, f0, f1, f2, f3, maxlen, totlen = struct.unpack_from("<i4fiq", raw, 0)
    rec = np.frombuffer(raw, dtype=np.dtype([("rlen", "<i4"), ("pad", "<i4"), ("boff", "<i8"), ("coff", "<i8"),
                                             ("flags", "<i4"), ("pad2", "<i4")]), count=ureads, offset=88)
    return {"ureads": ureads, "freq": [f0, f1, f2, f3], "maxlen": maxlen, "totlen": totlen}, rec


# ---- interval tracks (the -m masks of daligner) -------------------------------
#
# Two on-disk layouts, both named after the DB's hidden prefix (".ROOT."):
#   Dazzler  .ROOT.NAME.anno: int nreads, int size (= 8), int64 byte offsets
#            [nreads+1] into .ROOT.NAME.data, which holds int32 points
#            (written by db/DBdust.c:195-216, read by Load_Track,
#            db/DB.c:967-1120).  A per-block track is .ROOT.<b>.NAME.anno.
#   MARVEL   .ROOT.NAME.a2: a 64-byte header (u16 version = 2, u16 size = 8,
#            u32 pad, u64 len = nreads, u64 clen, u64 cdlen, 4 x u64 reserved)
#            then the anno offsets in zlib chunks; .ROOT.NAME.d2: the points
#            in zlib chunks.  A chunk is a u64 compressed length followed by
#            the zlib stream of at most 8 MiB of input (lib/tracks.h:40-56,
#            lib/tracks.c:60-170, lib/compression.c:11-100).
# A read's points come in pairs (begin, end) of masked intervals.

TRACK_CHUNK = 8 * 1024 * 1024


def _zchunks(raw: bytes) -> bytes:
    import zlib
    out = []
    for o in range(0, len(raw), TRACK_CHUNK):
        c = zlib.compress(raw[o:o + TRACK_CHUNK])
        out.append(struct.pack("<Q", len(c)) + c)
    return b"".join(out)


def _unzchunks(raw: bytes) -> bytes:
    import zlib
    out, o = [], 0
    while o < len(raw):
        (n,) = struct.unpack_from("<Q", raw, o)
        out.append(zlib.decompress(raw[o + 8:o + 8 + n]))
        o += 8 + n
    return b"".join(out)


def write_track(root: str, name: str, points, fmt: str = "dazzler", block: int | None = None) -> None:
    """Writes interval track NAME of DB ROOT; points[i] = the flat point list
    (b0, e0, b1, e1, ...) of read i.  fmt "dazzler" or "marvel"; block = a
    per-block Dazzler track (.ROOT.<block>.NAME.*)."""
    d, base = os.path.split(root)
    d = d or "."
    stem = "." + base + (f".{block}" if block else "") + "." + name
    data = np.concatenate([np.asarray(p, dtype=np.int32) for p in points]) if len(points) else np.zeros(0, np.int32)
    anno = np.zeros(len(points) + 1, dtype=np.int64)
    anno[1:] = np.cumsum([4 * len(p) for p in points])
    if fmt == "dazzler":
        with open(os.path.join(d, stem + ".anno"), "wb") as f:
            f.write(struct.pack("<ii", len(points), 8) + anno.tobytes())
        with open(os.path.join(d, stem + ".data"), "wb") as f:
            f.write(data.tobytes())
    elif fmt == "marvel":
        if block:
            raise ValueError("MARVEL tracks are whole-DB tracks")
        ca, cd = _zchunks(anno.tobytes()), _zchunks(data.tobytes())
        hdr = struct.pack("<HHIQQQ4Q", 2, 8, 0, len(points), len(ca), len(cd), 0, 0, 0, 0)
        with open(os.path.join(d, stem + ".a2"), "wb") as f:
            f.write(hdr + ca)
        with open(os.path.join(d, stem + ".d2"), "wb") as f:
            f.write(cd)
    else:
        raise ValueError(f"unknown track format {fmt!r}")


def read_track(root: str, name: str):
    """Point lists of every read of a whole-DB track (either layout)."""
    d, base = os.path.split(root)
    stem = os.path.join(d or ".", "." + base + "." + name)
    if os.path.exists(stem + ".a2"):
        raw = open(stem + ".a2", "rb").read()
        _, _, _, n, clen, cdlen = struct.unpack_from("<HHIQQQ", raw, 0)
        anno = np.frombuffer(_unzchunks(raw[64:64 + clen]), dtype=np.int64)
        data = np.frombuffer(_unzchunks(open(stem + ".d2", "rb").read()[:cdlen]), dtype=np.int32)
    else:
        raw = open(stem + ".anno", "rb").read()
        n, size = struct.unpack_from("<ii", raw, 0)
        if size != 8:
            raise ValueError(f"{stem}.anno: annotation size {size}, expected 8")
        anno = np.frombuffer(raw, dtype=np.int64, count=n + 1, offset=8)
        data = np.fromfile(stem + ".data", dtype=np.int32)
    return [data[anno[i] // 4:anno[i + 1] // 4].copy() for i in range(n)]


# ---- .las (dalign/align.c:3325-3350) -----------------------------------------

LAS_REC = np.dtype([("tlen", "<i4"), ("diffs", "<i4"), ("abpos", "<i4"), ("bbpos", "<i4"), ("aepos", "<i4"),
                    ("bepos", "<i4"), ("flags", "<u4"), ("aread", "<i4"), ("bread", "<i4"), ("pad", "<i4")])


def read_las(path: str):
    """(tspace, records structured array, list of trace byte strings)."""
    with open(path, "rb") as f:
        raw = f.read()
    n, tspace = struct.unpack_from("<qi", raw, 0)
    tbytes = 1 if tspace <= 125 else 2
    o = 12
    recs = np.empty(n, dtype=LAS_REC)
    traces = []
    for i in range(n):
        r = np.frombuffer(raw, dtype=LAS_REC, count=1, offset=o)[0]
        recs[i] = r
        o += 40
        tb = int(r["tlen"]) * tbytes
        traces.append(raw[o:o + tb])
        o += tb
    if o != len(raw):
        raise ValueError(f"{path}: {len(raw) - o} trailing bytes")
    return tspace, recs, traces


def write_las(path: str, tspace: int, recs, traces) -> None:
    """Writes a .las file: header (novl, tspace), then each record's 40 bytes
    and its trace bytes (1 byte per value for tspace <= 125, else 2)."""
    recs = np.asarray(recs, dtype=LAS_REC)
    with open(path, "wb") as f:
        f.write(struct.pack("<qi", len(recs), tspace))
        for r, t in zip(recs, traces):
            f.write(r.tobytes())
            f.write(bytes(t))


def first_difference(path_a: str, path_b: str):
    """None when the files are byte-identical, else a description of the first differing record."""
    with open(path_a, "rb") as f:
        a = f.read()
    with open(path_b, "rb") as f:
        b = f.read()
    if a == b:
        return None
    ta, ra, xa = read_las(path_a)
    tb, rb, xb = read_las(path_b)
    if ta != tb:
        return f"trace spacing {ta} != {tb}"
    for i in range(min(len(ra), len(rb))):
        if ra[i].tobytes() != rb[i].tobytes() or xa[i] != xb[i]:
            return f"record {i}: {ra[i]} vs {rb[i]}" + ("" if xa[i] == xb[i] else " (traces differ)")
    return f"record counts {len(ra)} != {len(rb)}"
