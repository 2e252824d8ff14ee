"""Writes tests/golden/hpc/plan_run.json: SHA-256 of every .las the reference
daligner (oracle/_ref/daligner) writes when it runs the daligner lines of the
reference HPCdaligner plan (oracle/_ref/HPCdaligner --dal 2 -j4 -M16) for the
small read set split into 3 blocks of ~1.5 Mbp, and of the block files the
reference LAmerge (oracle/_ref/LAmerge) then writes for the plan's merge
lines.

    python tests/golden/make_plan_goldens.py
"""

from __future__ import annotations

import hashlib
import json
import os
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from marvel_amd import dbio, sim  # noqa: E402

REF_HPC = os.path.join(ROOT, "oracle", "_ref", "HPCdaligner")
REF_DAL = os.path.join(ROOT, "oracle", "_ref", "daligner")
REF_MRG = os.path.join(ROOT, "oracle", "_ref", "LAmerge")
OUT = os.path.join(HERE, "hpc", "plan_run.json")
PLAN_ARGS = ["--dal", "2", "-j4", "-M16", "P"]
SPLIT_MBP = 1.5


def make_db(d: str) -> None:
    dbio.write_db(os.path.join(d, "P"), sim.config_small(1).reads, size_mbp=SPLIT_MBP)


def daligner_lines(hpc: str, d: str) -> list[list[str]]:
    out = subprocess.run([hpc] + PLAN_ARGS, cwd=d, capture_output=True, text=True, check=True).stdout
    return [l.split()[1:] for l in out.splitlines() if l.startswith("daligner ")]


def merge_lines(hpc: str, d: str) -> list[list[str]]:
    out = subprocess.run([hpc] + PLAN_ARGS, cwd=d, capture_output=True, text=True, check=True).stdout
    return [l.split()[1:] for l in out.splitlines() if l.startswith("LAmerge ")]


def run_merges(mrg: str, lines: list[list[str]], d: str) -> dict:
    """Runs the merge lines; SHA-256 of each output (the argument after the DB)."""
    shas = {}
    for args in lines:
        subprocess.run([mrg] + args, cwd=d, capture_output=True, check=True, timeout=600)
        out = [a for a in args if not a.startswith("-") and not a.isdigit()][1]
        shas[out] = hashlib.sha256(open(os.path.join(d, out), "rb").read()).hexdigest()
    return shas


def run_plan(dal: str, lines: list[list[str]], d: str) -> dict:
    for args in lines:
        subprocess.run([dal] + args, cwd=d, capture_output=True, check=True, timeout=600)
    shas = {}
    for dp, _, fs in os.walk(d):
        for f in fs:
            if f.endswith(".las"):
                p = os.path.join(dp, f)
                shas[os.path.relpath(p, d)] = hashlib.sha256(open(p, "rb").read()).hexdigest()
    return shas


def main():
    with tempfile.TemporaryDirectory() as d:
        make_db(d)
        lines = daligner_lines(REF_HPC, d)
        shas = run_plan(REF_DAL, lines, d)
        mlines = merge_lines(REF_HPC, d)
        merged = run_merges(REF_MRG, mlines, d)
    gold = {"plan_args": PLAN_ARGS, "split_mbp": SPLIT_MBP, "daligner_lines": lines, "las": shas,
            "merge_lines": mlines, "merged": merged}
    with open(OUT, "w") as f:
        json.dump(gold, f, indent=1, sort_keys=True)
    print(len(lines), "daligner jobs,", len(shas), "las files")


if __name__ == "__main__":
    main()
