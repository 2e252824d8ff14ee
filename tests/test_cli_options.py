"""Command-line checks of the GPU daligner that end before any GPU call:
options it does not support and invalid k-mer lengths stop it with a
message; a .dam block opens like a .db block (host code only)."""

import os
import subprocess

import numpy as np

from marvel_amd import dbio
from marvel_amd.daligner import Daligner

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "marvel_amd", "bin", "daligner")


def _db(d, name="T", stub=".db"):
    rng = np.random.default_rng(3)
    dbio.write_db(os.path.join(d, name), [rng.integers(0, 4, size=500, dtype=np.uint8) for _ in range(5)])
    if stub == ".dam":
        os.replace(os.path.join(d, name + ".db"), os.path.join(d, name + ".dam"))


def test_dynamic_mask_server_is_refused(tmp_path):
    _db(str(tmp_path))
    r = subprocess.run([EXE, "-D127.0.0.1:1", "T", "T"], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "not supported" in r.stderr


def test_kmer_length_bounds(tmp_path):
    _db(str(tmp_path))
    for k in ("-k32", "-k1"):
        r = subprocess.run([EXE, k, "T", "T"], cwd=tmp_path, capture_output=True, text=True, timeout=60)
        assert r.returncode == 1 and "k-mer length" in r.stderr, (k, r.stderr)


def test_dam_block_opens_like_db(tmp_path):
    _db(str(tmp_path), "D", ".db")
    _db(str(tmp_path), "M", ".dam")
    dal = Daligner()
    a = dal.open_block(str(tmp_path / "D"))
    b = dal.open_block(str(tmp_path / "M.dam"))
    c = dal.open_block(str(tmp_path / "M"))
    assert a.nreads == b.nreads == c.nreads == 5
    assert [a.reads[i].rlen for i in range(5)] == [b.reads[i].rlen for i in range(5)]
    dbio.write_track(str(tmp_path / "M"), "x", [[1, 20], [], [0, 500], [100, 100], []])
    dal.load_mask_tracks(b, ["x"])
    assert [list(p) for p in dal.mask_points(b)] == [[1, 20], [], [0, 500], [100, 100], []]
    for blk in (a, b, c):
        dal.close_block(blk)
