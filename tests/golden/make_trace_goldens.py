"""Writes tests/golden/trace/<set>.json: SHA-256 of the reference's own
Compute_Trace_PTS output (oracle/_ref/trace_vectors, compiled from
/root/reference: Local_Alignment on every task of a Local_Alignment golden
set, then Compute_Trace_PTS in the LOWERMOST, GREEDIEST and UPPERMOST modes).

Run in the container that has /root/reference (after `make -C oracle ref`):
    python tests/golden/make_trace_goldens.py
Inputs are the task sets of tests/golden/la/, regenerated from their seeds.
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
sys.path.insert(0, os.path.dirname(HERE))

import la_golden  # noqa: E402

EXE = os.path.join(ROOT, "oracle", "_ref", "trace_vectors")
OUT = os.path.join(HERE, "trace")


def main():
    os.makedirs(OUT, exist_ok=True)
    for name in la_golden.SETS:
        ts, _ = la_golden.load(name)
        with tempfile.TemporaryDirectory() as d:
            tf, rf = os.path.join(d, "t.bin"), os.path.join(d, "r.bin")
            open(tf, "wb").write(ts.task_file_bytes())
            subprocess.run([EXE, tf, rf], check=True)
            raw = open(rf, "rb").read()
        meta = {"la_set": name, "words": len(raw) // 4, "sha256": hashlib.sha256(raw).hexdigest(),
                "producer": "oracle/_ref/trace_vectors (reference dalign/align.c Compute_Trace_PTS)"}
        json.dump(meta, open(os.path.join(OUT, name + ".json"), "w"), indent=1)
        print(name, meta["words"], meta["sha256"][:16])


if __name__ == "__main__":
    main()
