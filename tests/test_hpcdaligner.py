"""marvel_amd/bin/HPCdaligner writes the same block-pair plans as the
reference HPCdaligner (goldens: tests/golden/hpc/plans.json, from
tests/golden/make_hpc_goldens.py) for every option set, including its quirks
and its rejections of bad arguments."""

import json
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "marvel_amd", "bin", "HPCdaligner")
GOLD = json.load(open(os.path.join(ROOT, "tests", "golden", "hpc", "plans.json")))


@pytest.mark.parametrize("case", sorted(GOLD["cases"]))
def test_plan_matches_reference(case, tmp_path):
    if not os.path.exists(EXE):
        pytest.skip("HPCdaligner not built (python -m marvel_amd.build)")
    g = GOLD["cases"][case]
    (tmp_path / "sub").mkdir()
    for p in ("H.db", os.path.join("sub", "H.db")):
        (tmp_path / p).write_text(GOLD["stub"])
    r = subprocess.run([EXE] + g["args"], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert r.returncode == g["rc"], r.stderr
    if g["rc"] == 0:
        assert r.stdout == g["stdout"]
        for name, text in g["files"].items():
            assert (tmp_path / name).read_text() == text
