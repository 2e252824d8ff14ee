"""Writes tests/golden/la_tools/goldens.json: SHA-256 of what the reference
LAsort / LAcat / LAmerge (oracle/_ref, compiled from /root/reference by
`make -C oracle ref`) write for the inputs of tests/la_tool_cases.py.

    python tests/golden/make_la_tool_goldens.py
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

import la_tool_cases  # noqa: E402

OUT = os.path.join(HERE, "la_tools", "goldens.json")


def main():
    gold = {}
    for name, (tool, _, args) in sorted(la_tool_cases.CASES.items()):
        with tempfile.TemporaryDirectory() as d:
            la_tool_cases.make_inputs(name, d)
            gold[name] = la_tool_cases.run(os.path.join(ROOT, "oracle", "_ref", tool), name, d)
        print(name, gold[name], flush=True)
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    with open(OUT, "w") as f:
        json.dump(gold, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
