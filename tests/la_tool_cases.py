"""Inputs of the .las utilities (LAsort, LAcat, LAmerge): synthetic .las files from
seeds, with many records sharing (aread, bread) and (aread, bread, abpos) so
that the order of equal keys is exercised.  The goldens
(tests/golden/la_tools/goldens.json) hold the SHA-256 of what the compiled
reference tools write for them."""

from __future__ import annotations

import hashlib
import os
import subprocess

import numpy as np

from marvel_amd import dbio

# name -> (tool, [(file name, seed, records, tspace, reads, positions)], arguments)
CASES = {
    "sort_s100": ("LAsort", [("x.las", 1, 20000, 100, 300, 60)], ["x"]),
    "sort_s200": ("LAsort", [("y.las", 2, 6000, 200, 50, 40)], ["y.las"]),
    "sort_ties": ("LAsort", [("t.las", 3, 5000, 100, 4, 3)], ["t"]),
    "sort_empty": ("LAsort", [("e.las", 4, 0, 100, 1, 1)], ["e"]),
    "sort_two": ("LAsort", [("a.las", 5, 3000, 100, 100, 30), ("b.las", 6, 2000, 100, 100, 30)], ["a", "b.las"]),
    "cat_3": ("LAcat", [("z.1.las", 7, 4000, 100, 200, 1000), ("z.2.las", 8, 0, 100, 1, 1),
                        ("z.3.las", 9, 3000, 100, 200, 1000)], ["z"]),
    "cat_s200": ("LAcat", [("w.1.las", 10, 2000, 200, 80, 500), ("w.2.las", 11, 1500, 200, 80, 500)], ["w.las"]),
    "cat_none": ("LAcat", [], ["q"]),
    # LAmerge: (file, seed, records, tspace, reads, positions, bread offset, sorted); DB stub D.db of 4 blocks
    "merge_files": ("LAmerge", [("D.1.D.1.las", 20, 3000, 100, 100, 50, 0, True),
                                ("D.1.D.2.las", 21, 2000, 100, 100, 50, 100, True),
                                ("D.1.D.3.las", 22, 2500, 100, 100, 50, 200, True)],
                    ["-n", "2", "D", "D.1.las", "D.1.D.1.las", "D.1.D.2.las", "D.1.D.3.las"]),
    "merge_dir": ("LAmerge", [("001_00001/D.1.D.%d.las" % j, 30 + j, 1500 + 100 * j, 100, 80, 40, 80 * j, True)
                              for j in range(1, 6)],
                  ["-v", "-n", "4", "D.db", "D.2.las", "001_00001"]),
    "merge_sort": ("LAmerge", [("u1.las", 40, 2000, 100, 20, 10, 0, False), ("u2.las", 41, 2000, 100, 20, 10, 0, False),
                               ("u3.las", 42, 1000, 100, 20, 10, 0, False)],
                   ["-s", "D", "M.las", "u1.las", "u2.las", "u3.las"]),
    "merge_suffix": ("LAmerge", [("D.%d.las" % j, 50 + j, 800, 100, 60, 30, 60 * j, True) for j in range(1, 5)],
                     ["D", "all.las"]),
    "merge_one": ("LAmerge", [("one_in.las", 60, 500, 100, 30, 30, 0, False)], ["D", "one.las", "one_in.las"]),
    "merge_s200": ("LAmerge", [("w1.las", 70, 1200, 200, 50, 40, 0, True), ("w2.las", 71, 900, 200, 50, 40, 50, True)],
                   ["-v", "D", "W", "w1.las", "w2.las"]),
}
# output file of each LAmerge case
MERGE_OUT = {"merge_files": "D.1.las", "merge_dir": "D.2.las", "merge_sort": "M.las", "merge_suffix": "all.las",
             "merge_one": "one.las", "merge_s200": "W.las"}


def db_stub(path: str, nblocks: int = 4) -> None:
    """A .db stub with one FASTA line and `nblocks` blocks (all LAmerge reads of it)."""
    with open(path, "w") as f:
        f.write("files = %9d\n" % 1)
        f.write("  %9d %s %s\n" % (400, "sim", "sim"))
        f.write("blocks = %9d\n" % nblocks)
        f.write("size = %9d\n" % 1)
        for i in range(nblocks + 1):
            f.write(" %9d\n" % (100 * i))


def random_las(seed: int, n: int, tspace: int, nreads: int, npos: int):
    """n records: aread, bread in [0, nreads), abpos in [0, npos) * 37, COMP
    flag at random, 0..40 trace values (consistent tlen), random contents."""
    rng = np.random.default_rng(seed)
    recs = np.zeros(n, dtype=dbio.LAS_REC)
    recs["aread"] = rng.integers(0, nreads, size=n)
    recs["bread"] = rng.integers(0, nreads, size=n)
    recs["abpos"] = rng.integers(0, npos, size=n) * 37
    recs["aepos"] = recs["abpos"] + rng.integers(1, 5000, size=n)
    recs["bbpos"] = rng.integers(0, 9000, size=n)
    recs["bepos"] = recs["bbpos"] + rng.integers(1, 5000, size=n)
    recs["diffs"] = rng.integers(0, 900, size=n)
    recs["flags"] = rng.integers(0, 2, size=n)
    recs["tlen"] = 2 * rng.integers(0, 21, size=n)
    tb = 1 if tspace <= 125 else 2
    traces = [rng.integers(0, 256, size=int(t) * tb, dtype=np.uint8).tobytes() for t in recs["tlen"]]
    return recs, traces


def make_inputs(case: str, d: str) -> None:
    os.makedirs(d, exist_ok=True)
    tool = CASES[case][0]
    if tool == "LAmerge":
        db_stub(os.path.join(d, "D.db"))
    for spec in CASES[case][1]:
        name, seed, n, tspace, nreads, npos = spec[:6]
        recs, traces = random_las(seed, n, tspace, nreads, npos)
        if tool == "LAmerge":
            recs["bread"] += spec[6]
            if spec[7]:   # in LAmerge's order (aread, bread, COMP, abpos)
                order = np.lexsort((recs["abpos"], recs["flags"] & 1, recs["bread"], recs["aread"]))
                recs, traces = recs[order], [traces[i] for i in order]
        os.makedirs(os.path.dirname(os.path.join(d, name)), exist_ok=True)
        dbio.write_las(os.path.join(d, name), tspace, recs, traces)


def run(exe: str, case: str, d: str) -> dict:
    """Runs a LAsort- or LAcat-compatible binary in d; {output name: sha256}."""
    tool, _, args = CASES[case]
    r = subprocess.run([exe] + args, cwd=d, capture_output=True, timeout=300)
    if r.returncode != 0:
        raise RuntimeError(f"{exe} {case}: rc {r.returncode}\n{r.stderr[-2000:]!r}")
    if tool == "LAcat":
        return {"stdout": hashlib.sha256(r.stdout).hexdigest()}
    if tool == "LAmerge":
        f = MERGE_OUT[case]
        return {f: hashlib.sha256(open(os.path.join(d, f), "rb").read()).hexdigest()}
    out = {}
    for f in sorted(os.listdir(d)):
        if f.endswith(".S.las"):
            out[f] = hashlib.sha256(open(os.path.join(d, f), "rb").read()).hexdigest()
    return out
