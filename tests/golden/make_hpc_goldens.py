"""Writes tests/golden/hpc/plans.json: the plans the reference HPCdaligner
(oracle/_ref/HPCdaligner, compiled from /root/reference by
`make -C oracle ref`) prints for a set of option lists, on a 9-block DB stub
(only the stub's block count matters to the plan).

    python tests/golden/make_hpc_goldens.py
"""

from __future__ import annotations

import json
import os
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(ROOT, "oracle", "_ref", "HPCdaligner")
OUT = os.path.join(HERE, "hpc", "plans.json")

# .db stub of a DB split into 9 blocks (db/DB.h:245-249 layout)
STUB = ("files =         1\n       1000 H H\nblocks =         9\nsize =        500 cutoff =         0 all = 0\n"
        + "".join(f" {i * 110:9d} {i * 440000:9d}\n" for i in range(10)))

CASES = {
    "defaults": ["H"],
    "all_daligner_options": ["-v", "-A", "-k12", "-w5", "-h30", "-e.80", "-l500", "-s200", "-M32", "-T", "-j8",
                             "-r3", "--dal", "2", "--mrg", "4", "H"],
    "consecutive": ["-X", "H"],
    "consecutive_asym": ["-X", "-A", "--dal", "3", "H"],
    "consecutive_opts": ["-X", "-M16", "-T", "-vv", "H"],
    "diag_jobs": ["--dal", "3", "--dalDiag", "2", "H"],
    "dust_server": ["-D", "localhost:1234", "H"],
    "dust_server_noport": ["-D", "localhost", "--dal", "5", "H"],
    "range": ["H", "3-7"],
    "single_block": ["H", "5"],
    "tracks_and_flags": ["-m", "dust", "-m", "tan", "-t30", "-H5000", "-b", "-I", "-K", "-S", "H"],
    "dir_path": ["sub/H"],
    "dot_path": ["./H.db"],
    "out_files": ["-o", "plan", "--dal", "2", "H"],
    "bad_threads": ["-j3", "H"],
    "bad_cor": ["-e.4", "H"],
    "bad_block": ["H", "12"],
    "bad_range": ["H", "6-3"],
}


def run(exe: str, args: list[str]) -> dict:
    with tempfile.TemporaryDirectory() as d:
        os.makedirs(os.path.join(d, "sub"))
        for p in ("H.db", os.path.join("sub", "H.db")):
            with open(os.path.join(d, p), "w") as f:
                f.write(STUB)
        r = subprocess.run([exe] + args, cwd=d, capture_output=True, text=True, timeout=60)
        files = {}
        for name in ("plan.dalign.plan", "plan.merge.plan"):
            p = os.path.join(d, name)
            if os.path.exists(p):
                files[name] = open(p).read()
        return {"rc": r.returncode, "stdout": r.stdout if r.returncode == 0 else "", "files": files}


def main():
    gold = {"stub": STUB, "cases": {}}
    for name, args in CASES.items():
        gold["cases"][name] = {"args": args, **run(REF, args)}
        print(name, gold["cases"][name]["rc"], len(gold["cases"][name]["stdout"]), flush=True)
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    with open(OUT, "w") as f:
        json.dump(gold, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
