"""marvel_amd.dbio writes the DB files the reference FA2db + DBsplit write
for the same reads (checked against the compiled reference tools when they
are built, oracle/_ref), and reads .las files back."""

import os
import subprocess
import tempfile

import numpy as np
import pytest

from marvel_amd import dbio, sim

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FA2DB = os.path.join(ROOT, "oracle", "_ref", "FA2db")
DBSPLIT = os.path.join(ROOT, "oracle", "_ref", "DBsplit")


def test_bps_packing_is_msb_first():
    # bases 0,1,2,3 -> 0b00011011; a 5th base 3 -> 0b11000000
    assert dbio.pack_bps(np.array([0, 1, 2, 3, 3], dtype=np.uint8)) == bytes([0x1B, 0xC0])


def test_block_split_rule():
    # a block closes as soon as its bases reach size * 10^6
    assert dbio.block_starts([600000, 500000, 100, 1000000, 5], 1) == [0, 2, 4, 5]
    assert dbio.block_starts([10, 20], 1) == [0, 2]


@pytest.mark.skipif(not (os.path.exists(FA2DB) and os.path.exists(DBSPLIT)), reason="reference tools not built")
@pytest.mark.parametrize("split", [400, 2])
def test_db_matches_reference_fa2db_dbsplit(split):
    reads = sim.config_small(1).reads[:300]
    with tempfile.TemporaryDirectory() as d:
        sim.write_fasta(os.path.join(d, "sim.fasta"), reads)
        subprocess.run([FA2DB, "-x0", "R", "sim.fasta"], cwd=d, check=True, capture_output=True)
        subprocess.run([DBSPLIT, f"-s{split}", "R"], cwd=d, check=True, capture_output=True)
        os.mkdir(os.path.join(d, "o"))
        dbio.write_db(os.path.join(d, "o", "R"), reads, size_mbp=split)
        assert open(os.path.join(d, "R.db")).read() == open(os.path.join(d, "o", "R.db")).read()
        assert open(os.path.join(d, ".R.bps"), "rb").read() == open(os.path.join(d, "o", ".R.bps"), "rb").read()
        hr, rr = dbio.read_idx(os.path.join(d, "R"))
        ho, ro = dbio.read_idx(os.path.join(d, "o", "R"))
        assert hr == ho
        for f in ("rlen", "boff", "coff", "flags"):
            assert np.array_equal(rr[f], ro[f]), f


def test_las_roundtrip():
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "x.las")
        import struct
        recs = [(4, 3, 0, 5, 200, 210, 1, 7, 9, 0), (2, 0, 150, 0, 180, 31, 0, 7, 8, 0)]
        tr = [bytes([1, 100, 2, 110]), bytes([0, 31])]
        with open(p, "wb") as f:
            f.write(struct.pack("<qi", 2, 100))
            for r, t in zip(recs, tr):
                f.write(struct.pack("<iiiiiiIiii", *r) + t)
        ts, got, traces = dbio.read_las(p)
        assert ts == 100 and len(got) == 2
        assert got[0]["aepos"] == 200 and got[1]["bread"] == 8 and traces == tr
        assert dbio.first_difference(p, p) is None
