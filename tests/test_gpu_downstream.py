"""A downstream caller of the align.h drop-ins: the reference's LAcheck
(utils/LAcheck.c, with its DB and pass-through code) linked against
libmarvl_align checks a golden .las with -a (Complement_Seq and
Compute_Trace_PTS on the GPU for every overlap) and prints what the same
LAcheck built on the reference's align.c prints.  The per-call time of
the drop-ins is reported."""

import os
import subprocess
import time

import pytest

import las_cases

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_DIR = os.path.join(ROOT, "oracle", "_ref")
EXE = os.path.join(ROOT, "marvel_amd", "bin", "daligner")
OURS = os.path.join(REF_DIR, "lacheck_on_libmarvl")
THEIRS = os.path.join(REF_DIR, "LAcheck")

pytestmark = pytest.mark.gpu


@pytest.mark.skipif(not (os.path.exists(OURS) and os.path.exists(THEIRS)),
                    reason="LAcheck builds absent (make -C oracle ref integ)")
@pytest.mark.parametrize("case", ["c1_j8", "rep_j8"])
def test_lacheck_on_library(case, tmp_path):
    print(f"[lacheck] {case}", flush=True)
    las_cases.make_db(case, str(tmp_path))
    _, _, flags, blocks = las_cases.CASES[case]
    r = subprocess.run([EXE] + flags + blocks, cwd=tmp_path, capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr[-2000:]
    las = os.path.join("001_00001", f"{blocks[0]}.{blocks[1]}.las")
    db = blocks[0].split(".")[0]
    args = ["-a", "-p", "-s", "-d", db, las]
    t0 = time.perf_counter()
    want = subprocess.run([THEIRS] + args, cwd=tmp_path, capture_output=True, text=True, timeout=600)
    t1 = time.perf_counter()
    got = subprocess.run([OURS] + args, cwd=tmp_path, capture_output=True, text=True, timeout=600)
    t2 = time.perf_counter()
    import struct
    n = struct.unpack("<q", open(os.path.join(tmp_path, las), "rb").read(8))[0]
    print(f"[lacheck] {case}: {n} overlaps; reference align.c {t1 - t0:.2f} s, library {t2 - t1:.2f} s "
          f"({(t2 - t1) / max(n, 1) * 1e6:.0f} us per overlap)", flush=True)
    assert want.returncode == 0, want.stderr[-1000:]
    assert got.returncode == want.returncode, got.stderr[-2000:]
    assert got.stdout == want.stdout
    assert got.stderr == want.stderr
