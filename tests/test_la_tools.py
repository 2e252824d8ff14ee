"""The .las utilities after the path against the compiled reference's
outputs (tests/golden/la_tools): the oracle restatement (CPU), the host
LAcat (CPU), and the GPU LAsort (gpu marker)."""

import hashlib
import json
import os
import subprocess
import sys

import pytest

import la_tool_cases

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import las_tools_oracle  # noqa: E402

GOLD = json.load(open(os.path.join(ROOT, "tests", "golden", "la_tools", "goldens.json")))
BIN = os.path.join(ROOT, "marvel_amd", "bin")


def _sha(b: bytes) -> str:
    return hashlib.sha256(b).hexdigest()


@pytest.mark.parametrize("case", sorted(la_tool_cases.CASES))
def test_oracle_matches_reference(case, tmp_path):
    tool, files, args = la_tool_cases.CASES[case]
    la_tool_cases.make_inputs(case, str(tmp_path))
    if tool == "LAcat":
        got = {"stdout": _sha(las_tools_oracle.las_cat(str(tmp_path / args[0])))}
    else:
        got = {}
        for a in args:
            root = a[:-4] if a.endswith(".las") else a
            got[root + ".S.las"] = _sha(las_tools_oracle.las_sort(open(tmp_path / (root + ".las"), "rb").read()))
    assert got == GOLD[case]


@pytest.mark.parametrize("case", sorted(c for c, v in la_tool_cases.CASES.items() if v[0] == "LAcat"))
def test_lacat_matches_reference(case, tmp_path):
    exe = os.path.join(BIN, "LAcat")
    assert os.path.exists(exe), "LAcat not built (python -m marvel_amd.build)"
    la_tool_cases.make_inputs(case, str(tmp_path))
    assert la_tool_cases.run(exe, case, str(tmp_path)) == GOLD[case]


def test_lacat_tspace_conflict(tmp_path):
    from marvel_amd import dbio
    for i, ts in ((1, 100), (2, 200)):
        recs, tr = la_tool_cases.random_las(i, 10, ts, 5, 5)
        dbio.write_las(str(tmp_path / f"c.{i}.las"), ts, recs, tr)
    r = subprocess.run([os.path.join(BIN, "LAcat"), "c"], cwd=tmp_path, capture_output=True)
    assert r.returncode == 1 and b"spacing conflict" in r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("case", sorted(c for c, v in la_tool_cases.CASES.items() if v[0] == "LAsort"))
def test_gpu_lasort_matches_reference(case, tmp_path):
    exe = os.path.join(BIN, "LAsort")
    assert os.path.exists(exe), "LAsort not built (python -m marvel_amd.build)"
    print(f"[gpu LAsort] case {case}", flush=True)
    la_tool_cases.make_inputs(case, str(tmp_path))
    assert la_tool_cases.run(exe, case, str(tmp_path)) == GOLD[case]


@pytest.mark.gpu
def test_gpu_lasort_of_daligner_output(tmp_path):
    """LAsort of a .las the GPU daligner wrote (byte-identical to the
    reference's, c1_j8) = the oracle's LAsort of the same bytes."""
    import las_cases
    las_cases.make_db("c1_j8", str(tmp_path))
    _, _, flags, blocks = las_cases.CASES["c1_j8"]
    r = subprocess.run([os.path.join(BIN, "daligner")] + flags + blocks, cwd=tmp_path, capture_output=True, timeout=240)
    assert r.returncode == 0, r.stderr[-2000:]
    src = tmp_path / "001_00001" / "C1.1.C1.1.las"
    want = las_tools_oracle.las_sort(src.read_bytes())
    r = subprocess.run([os.path.join(BIN, "LAsort"), str(src)], capture_output=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert (tmp_path / "001_00001" / "C1.1.C1.1.S.las").read_bytes() == want


@pytest.mark.gpu
def test_gpu_lasort_large_is_a_stable_sort(tmp_path):
    """3 M records (about 150 MB): the output holds the same records, in
    (aread, bread, abpos) order, equal keys in input order (each record
    carries its input index in `diffs`)."""
    import numpy as np
    from marvel_amd import dbio
    n = 3_000_000
    rng = np.random.default_rng(12)
    recs = np.zeros(n, dtype=dbio.LAS_REC)
    recs["aread"] = rng.integers(0, 40000, size=n)
    recs["bread"] = rng.integers(0, 40000, size=n)
    recs["abpos"] = rng.integers(0, 8, size=n) * 1000
    recs["diffs"] = np.arange(n)
    recs["tlen"] = 2 * rng.integers(0, 12, size=n)
    tr = [bytes([i & 255]) * int(t) for i, t in enumerate(recs["tlen"])]
    dbio.write_las(str(tmp_path / "big.las"), 100, recs, tr)
    r = subprocess.run([os.path.join(BIN, "LAsort"), "big"], cwd=tmp_path, capture_output=True, timeout=300)
    assert r.returncode == 0, r.stderr
    _, got, gtr = dbio.read_las(str(tmp_path / "big.S.las"))
    idx = got["diffs"].astype(np.int64)
    assert np.array_equal(np.sort(idx), np.arange(n))
    assert np.array_equal(got, recs[idx])
    assert all(gtr[j] == tr[i] for j, i in enumerate(idx[::997]))   # sampled trace bytes follow their record
    k = (got["aread"].astype(np.int64) << 40) | (got["bread"].astype(np.int64) << 16) | got["abpos"]
    assert np.all(np.diff(k) >= 0)
    tie = np.diff(k) == 0
    assert np.all(np.diff(idx)[tie] > 0)
