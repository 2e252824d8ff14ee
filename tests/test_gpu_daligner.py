"""GPU daligner (marvel_amd/bin/daligner over libmarvl_align) writes .las
files byte-identical to the compiled reference daligner's for every golden
case; on a mismatch the oracle restatement's output names the first
differing record."""

import hashlib
import json
import os
import subprocess
import tempfile

import pytest

import las_cases
import oracle_lib
from marvel_amd import dbio

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "marvel_amd", "bin", "daligner")
GOLD = json.load(open(os.path.join(ROOT, "tests", "golden", "las", "goldens.json")))


def _shas(d):
    out = {}
    for dp, _, fs in os.walk(d):
        for f in fs:
            if f.endswith(".las"):
                p = os.path.join(dp, f)
                out[os.path.relpath(p, d)] = hashlib.sha256(open(p, "rb").read()).hexdigest()
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("case", sorted(las_cases.CASES))
def test_gpu_daligner_matches_reference(case):
    assert os.path.exists(EXE), "GPU daligner not built (python -m marvel_amd.build)"
    print(f"[gpu daligner] case {case}", flush=True)
    with tempfile.TemporaryDirectory() as d:
        las_cases.make_db(case, d)
        _, _, flags, blocks = las_cases.CASES[case]
        r = subprocess.run([EXE] + flags + blocks, cwd=d, capture_output=True, text=True, timeout=240)
        assert r.returncode == 0, r.stderr[-2000:]
        got = _shas(d)
        if got != GOLD[case]["las"]:
            oracle_lib.ensure_built()
            with tempfile.TemporaryDirectory() as d2:
                for f in os.listdir(d):
                    if f.endswith((".db", ".idx", ".bps")):
                        os.symlink(os.path.join(d, f), os.path.join(d2, f))
                subprocess.run([oracle_lib.ORACLE_CLI] + flags + blocks, cwd=d2, capture_output=True, timeout=600)
                diffs = {rel: dbio.first_difference(os.path.join(d, rel), os.path.join(d2, rel))
                         for rel in GOLD[case]["las"] if os.path.exists(os.path.join(d, rel))}
            pytest.fail(f"{case}: {got} != {GOLD[case]['las']}; first differences vs oracle: {diffs}")
