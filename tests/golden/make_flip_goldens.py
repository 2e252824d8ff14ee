"""Writes tests/golden/flip/{cases,expected}.txt: seeded commands for
Complement_Seq, Flip_Alignment and Print_Overlap and what the reference's
own functions (dalign/align.c, linked into oracle/_ref/flip_vectors by
`make -C oracle ref`) print for them.

    python tests/golden/make_flip_goldens.py
"""

import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
EXE = os.path.join(ROOT, "oracle", "_ref", "flip_vectors")


def cases(seed=17):
    rng = np.random.default_rng(seed)
    out = []
    for n in (0, 1, 2, 3, 7, 64, 1001):
        out.append("C %d %s" % (n, " ".join(str(int(b)) for b in rng.integers(0, 4, n))))
    for _ in range(60):
        comp, full = int(rng.integers(0, 2)), int(rng.integers(0, 2))
        alen, blen = int(rng.integers(100, 20000)), int(rng.integers(100, 20000))
        ab = int(rng.integers(0, alen // 2))
        ae = int(rng.integers(ab, alen + 1))
        bb = int(rng.integers(0, blen // 2))
        be = int(rng.integers(bb, blen + 1))
        tlen = int(rng.integers(0, 40))
        # int trace of Compute_Trace_*: gaps in A as -(B position), gaps in B as A position
        t = [int(-rng.integers(1, blen + 1)) if rng.random() < 0.5 else int(rng.integers(1, alen + 1))
             for _ in range(tlen)]
        out.append("F %d %d %d %d %d %d %d %d %d %s" % (comp, full, alen, blen, ab, ae, bb, be, tlen,
                                                        " ".join(map(str, t))))
    for k in range(40):
        tb = 1 if k % 2 == 0 else 2
        tlen = 2 * int(rng.integers(0, 25))
        hi = 256 if tb == 1 else 65536
        t = [int(x) for x in rng.integers(0, hi, tlen)]
        out.append("P %d %d %d %d %d %d %d %d %d %d %d %s" % (
            tb, int(rng.integers(0, 9)), int(rng.integers(0, 10**6)), int(rng.integers(0, 10**6)),
            int(rng.integers(0, 4)), int(rng.integers(0, 5000)), int(rng.integers(5000, 90000)),
            int(rng.integers(0, 5000)), int(rng.integers(5000, 90000)), int(rng.integers(0, 9000)), tlen,
            " ".join(map(str, t))))
    return out


def main():
    cs = cases()
    text = "\n".join(cs) + "\n"
    r = subprocess.run([EXE], input=text, capture_output=True, text=True, check=True)
    open(os.path.join(HERE, "flip", "cases.txt"), "w").write(text)
    open(os.path.join(HERE, "flip", "expected.txt"), "w").write(r.stdout)


if __name__ == "__main__":
    main()
