"""Writes tests/golden/dust/goldens.json: SHA-256 of the track files the
reference DBdust (oracle/_ref/DBdust, compiled from /root/reference by
`make -C oracle ref`) writes for the cases of tests/dust_cases.py.

    python tests/golden/make_dust_goldens.py
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

import dust_cases  # noqa: E402

OUT = os.path.join(HERE, "dust", "goldens.json")


def main():
    gold = {}
    for name in sorted(dust_cases.CASES):
        with tempfile.TemporaryDirectory() as d:
            gold[name] = dust_cases.run(os.path.join(ROOT, "oracle", "_ref", "DBdust"), name, d)
        print(name, gold[name], flush=True)
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    with open(OUT, "w") as f:
        json.dump(gold, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
