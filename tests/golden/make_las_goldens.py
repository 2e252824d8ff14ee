"""Writes tests/golden/las/goldens.json: SHA-256 of every .las the reference
daligner (oracle/_ref/daligner, compiled from /root/reference by
`make -C oracle ref`) writes for the cases of tests/las_cases.py.

Run in the container that has /root/reference:
    python tests/golden/make_las_goldens.py [case ...]
The DBs are regenerated from seeds with marvel_amd.dbio.write_db, whose
files are checked against the reference FA2db + DBsplit by
tests/test_dbio.py.
"""

from __future__ import annotations

import json
import os
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

import las_cases  # noqa: E402

REF = os.path.join(ROOT, "oracle", "_ref", "daligner")
OUT = os.path.join(HERE, "las", "goldens.json")


def main(names):
    gold = json.load(open(OUT)) if os.path.exists(OUT) else {}
    for name in names or sorted(las_cases.CASES):
        with tempfile.TemporaryDirectory() as d:
            las_cases.make_db(name, d)
            shas = las_cases.run(REF, name, d)
        data, split, flags, blocks = las_cases.CASES[name]
        gold[name] = {"data": data, "split_mbp": split, "flags": flags, "blocks": blocks, "las": shas}
        print(name, shas, flush=True)
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    with open(OUT, "w") as f:
        json.dump(gold, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main(sys.argv[1:])
