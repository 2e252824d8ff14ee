"""The .las utilities after the path against the compiled reference's
outputs (tests/golden/la_tools): the oracle restatement (CPU), the host
LAcat (CPU), and the GPU LAsort / LAmerge (gpu marker)."""

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
    elif tool == "LAmerge":
        names = [f[0] for f in files]   # inputs in the order the cases name them (bread ranges differ per file)
        got = {la_tool_cases.MERGE_OUT[case]: _sha(las_tools_oracle.las_merge(
            [open(tmp_path / n, "rb").read() for n in names]))}
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
@pytest.mark.parametrize("case", sorted(c for c, v in la_tool_cases.CASES.items() if v[0] == "LAmerge"))
def test_gpu_lamerge_matches_reference(case, tmp_path):
    exe = os.path.join(BIN, "LAmerge")
    assert os.path.exists(exe), "LAmerge not built (python -m marvel_amd.build)"
    print(f"[gpu LAmerge] case {case}", flush=True)
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


def _las_image(recs, tspace, chunk=200_000):
    """.las bytes of records whose trace values are all (diffs & 255),
    built with array operations (1-byte values)."""
    import numpy as np
    n = len(recs)
    sz = 40 + recs["tlen"].astype(np.int64)
    off = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(sz, out=off[1:])
    buf = np.empty(12 + int(off[-1]), dtype=np.uint8)
    buf[:12] = np.frombuffer(np.array([n], "<i8").tobytes() + np.array([tspace], "<i4").tobytes(), np.uint8)
    raw = recs.view(np.uint8).reshape(n, 40)
    for a in range(0, n, chunk):
        b = min(n, a + chunk)
        buf[(12 + off[a:b, None] + np.arange(40)).ravel()] = raw[a:b].ravel()
        tl = recs["tlen"][a:b].astype(np.int64)
        starts = np.repeat(12 + off[a:b] + 40, tl)
        within = np.arange(int(tl.sum())) - np.repeat(np.cumsum(tl) - tl, tl)
        buf[starts + within] = np.repeat((recs["diffs"][a:b] & 255).astype(np.uint8), tl)
    return buf.tobytes()


@pytest.mark.gpu
def test_gpu_lasort_large_is_a_stable_sort(tmp_path):
    """3 M records (about 150 MB, many equal keys): the output is byte for
    byte the records in a stable (aread, bread, abpos) order -- numpy's
    stable lexsort, the order the oracle's las_sort uses."""
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
    (tmp_path / "big.las").write_bytes(_las_image(recs, 100))
    order = np.lexsort((recs["abpos"], recs["bread"], recs["aread"]))
    want = _las_image(recs[order], 100)
    print("[gpu LAsort] 3M records", flush=True)
    r = subprocess.run([os.path.join(BIN, "LAsort"), "big"], cwd=tmp_path, capture_output=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert (tmp_path / "big.S.las").read_bytes() == want


def test_las_image_matches_the_writer(tmp_path):
    """The array builder above writes what dbio.write_las writes."""
    import numpy as np
    from marvel_amd import dbio
    recs, _ = la_tool_cases.random_las(13, 500, 100, 20, 20)
    recs["tlen"] = 2 * (np.arange(500) % 7)
    tr = [bytes([int(d) & 255]) * int(t) for d, t in zip(recs["diffs"], recs["tlen"])]
    dbio.write_las(str(tmp_path / "w.las"), 100, recs, tr)
    assert (tmp_path / "w.las").read_bytes() == _las_image(recs, 100, chunk=77)
