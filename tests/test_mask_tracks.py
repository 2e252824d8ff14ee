"""The -m mask tracks as libmarvl_align loads them (host code only, no GPU
call): points of either file layout, per-block files, the union of several
tracks, the complement, and the errors for unusable tracks.

The union is pinned by what it must mean for the index: a k-mer covering
bases [s, e] of a read is indexed iff no interval [b, f) of any track has
s < f and e >= b (an empty interval b == f cuts the read between b-1 and b).
That rule is evaluated here by brute force over the track files and compared
with the k-mers the loaded track admits under tuple_thread's extraction
(dalign/filter.c:416-452)."""

import os

import numpy as np
import pytest

import las_cases
from marvel_amd import dbio
from marvel_amd.daligner import Daligner, DalignerError

K = 14


@pytest.fixture(scope="module")
def dal():
    return Daligner()


def admitted(points, rlen, k=K):
    """Last-base positions of the k-mers tuple_thread emits for one read."""
    out = []
    b, f = 0, len(points)
    for a in range(b, f + 1, 2):
        p = 0 if a == b else int(points[a - 1])
        q = rlen if a == f else int(points[a])
        if p + k <= q:
            out.extend(range(p + k - 1, q))
    return out


def brute(tracks_pts, rlen, k=K):
    """Last-base positions of the k-mers no interval of any track touches."""
    ok = np.ones(max(rlen - k + 1, 0), dtype=bool)   # index s = first base
    for pts in tracks_pts:
        for b, f in zip(pts[0::2], pts[1::2]):
            s = np.arange(len(ok))
            ok &= ~((s < f) & (s + k - 1 >= b))
    return list(np.nonzero(ok)[0] + k - 1)


def _db(tmp_path, case):
    las_cases.make_db(case, str(tmp_path))
    return os.path.join(str(tmp_path), "C1")


def test_single_track_points_both_layouts(dal, tmp_path):
    root = _db(tmp_path, "rep_mdust")
    want = dbio.read_track(root, "dust")
    dbio.write_track(root, "dm", want, fmt="marvel")
    for name in ("dust", "dm"):
        b = dal.open_block(root + ".1")
        dal.load_mask_tracks(b, [name])
        got = dal.mask_points(b)
        assert len(got) == len(want) and all(np.array_equal(g, w) for g, w in zip(got, want)), name
        dal.close_block(b)


def test_block_slices_and_block_files(dal, tmp_path):
    """Blocks 2 and 3 take their slice of whole-DB tracks; block 2 reads its
    own .C1.2.pt.* file instead of the whole-DB one."""
    root = _db(tmp_path, "rep_mask_blocks")
    lens = [len(r) for r in las_cases.reads_of("small_rep")]
    starts = dbio.block_starts(lens, 1)
    full = {n: dbio.read_track(root, n) for n in ("dust", "rnd", "pt")}
    own = las_cases.random_points(lens[starts[1]:starts[2]], 9)
    for blk in (1, 2, 3):
        lo, hi = starts[blk - 1], starts[blk]
        b = dal.open_block(f"{root}.{blk}")
        assert b.nreads == hi - lo
        for name in ("dust", "pt"):
            dal.load_mask_tracks(b, [name])
            got = dal.mask_points(b)
            want = own if (name == "pt" and blk == 2) else full[name][lo:hi]
            assert all(np.array_equal(g, np.asarray(w, dtype=np.int32)) for g, w in zip(got, want)), (blk, name)
        dal.close_block(b)


@pytest.mark.parametrize("case,names", [("rep_mdust_mrnd", ["dust", "rnd"]),
                                        ("rep_mask_blocks", ["dust", "rnd", "pt"])])
def test_union_admits_exactly_the_untouched_kmers(dal, tmp_path, case, names):
    root = _db(tmp_path, case)
    lens = [len(r) for r in las_cases.reads_of("small_rep")]
    tracks = [dbio.read_track(root, n) for n in names]
    b = dal.open_block(root + ".1")
    dal.load_mask_tracks(b, names)
    got = dal.mask_points(b)
    for i in range(b.nreads):
        assert admitted(got[i], lens[i]) == brute([t[i] for t in tracks], lens[i]), i
    dal.close_block(b)


def test_union_with_touching_and_empty_intervals(dal, tmp_path):
    root = os.path.join(str(tmp_path), "T")
    rng = np.random.default_rng(3)
    reads = [rng.integers(0, 4, size=n, dtype=np.uint8) for n in (300, 300, 300, 300)]
    dbio.write_db(root, reads)
    t1 = [[10, 50], [0, 0, 100, 100], [0, 300], [20, 40, 40, 60]]
    t2 = [[50, 90], [100, 100], [], [40, 40, 60, 80]]
    t3 = [[50, 50], [], [150, 150], [0, 20]]
    for n, t in (("a", t1), ("b", t2), ("c", t3)):
        dbio.write_track(root, n, t)
    b = dal.open_block(root)
    dal.load_mask_tracks(b, ["a", "b", "c"])
    got = dal.mask_points(b)
    for i in range(4):
        assert admitted(got[i], 300) == brute([t1[i], t2[i], t3[i]], 300), i
    dal.close_block(b)


def test_complement_of_the_mask(dal, tmp_path):
    root = _db(tmp_path, "rep_mdust_mrnd")
    b = dal.open_block(root + ".1")
    dal.load_mask_tracks(b, ["rnd"])
    fwd = dal.mask_points(b)
    c = dal.complement_header(b)
    rev = dal.mask_points(c)
    for i in range(b.nreads):
        assert np.array_equal(rev[i], b.reads[i].rlen - fwd[i][::-1]), i
    dal.release_tracks(c)
    assert not c.tracks
    dal.close_block(b)


def test_unusable_tracks_are_errors(dal, tmp_path):
    root = os.path.join(str(tmp_path), "T")
    rng = np.random.default_rng(4)
    dbio.write_db(root, [rng.integers(0, 4, size=200, dtype=np.uint8) for _ in range(3)])
    dbio.write_track(root, "odd", [[5], [], []])
    dbio.write_track(root, "down", [[50, 40], [], []])
    dbio.write_track(root, "out", [[], [190, 201], []])
    dbio.write_track(root, "ok", [[], [], []])
    dbio.write_track(root + "X", "short", [[], []])   # another DB's size
    os.rename(os.path.join(str(tmp_path), ".TX.short.anno"), os.path.join(str(tmp_path), ".T.short.anno"))
    os.rename(os.path.join(str(tmp_path), ".TX.short.data"), os.path.join(str(tmp_path), ".T.short.data"))
    for names, msg in ((["odd"], "odd number"), (["down"], "not increasing"), (["out"], "not increasing"),
                       (["none"], "does not exist"), (["ok", "ok"], "given twice"), (["short"], "size")):
        b = dal.open_block(root)
        with pytest.raises(DalignerError, match=msg):
            dal.load_mask_tracks(b, names)
        assert not b.tracks
        dal.close_block(b)
