"""DBdust workloads with goldens from the compiled reference DBdust
(tests/golden/dust/goldens.json: SHA-256 of the .dust.anno / .dust.data it
writes).  A case names a read set (regenerated from its seed), the DBsplit
size, the DBdust options and the DB argument; "grow" cases first dust the DB
holding only the first `grow` reads, then rewrite the DB with all reads and
dust again (the reference appends the new reads' intervals)."""

from __future__ import annotations

import hashlib
import os
import subprocess

import las_cases
from marvel_amd import dbio

# name -> (data, DBsplit Mbp, options, DB argument, grow)
CASES = {
    "small_default": ("small", 400, [], "C1", 0),
    "rep_default": ("small_rep", 400, [], "C1", 0),
    "rep_w32_t15_m5": ("small_rep", 400, ["-w32", "-t1.5", "-m5"], "C1", 0),
    "rep_w100_t3": ("small_rep", 400, ["-w100", "-t3", "-m20"], "C1.db", 0),
    "rep_w256_t12_m4": ("small_rep", 400, ["-w256", "-t1.2", "-m4"], "C1", 0),
    "rep_biased": ("small_rep", 400, ["-b"], "C1", 0),
    "rep_biased_w40": ("small_rep", 400, ["-b", "-w40", "-t1.2"], "C1", 0),
    "rep_block2": ("small_rep", 1, [], "C1.2", 0),
    "rep_grow": ("small_rep", 400, [], "C1", 400),
    "edges_default": ("edges", 400, ["-m3"], "E", 0),
    # the 5 Mbp / 30x read set (about 15,000 reads, 150 Mbp)
    "c2_default": ("5mbp", 400, [], "C2", 0),
    "c2_biased": ("5mbp", 400, ["-b", "-w48"], "C2", 0),
}


def stem(case: str) -> str:
    arg = CASES[case][3]
    return "." + (arg[:-3] if arg.endswith(".db") else arg) + ".dust"


def run(exe: str, case: str, d: str, env: dict | None = None, log: list | None = None) -> dict:
    """Runs `exe` on the case in directory d (extra environment `env`; the
    runs' stderr appended to `log`) and returns the SHA-256 of the track files."""
    data, split, opts, arg, grow = CASES[case]
    full_env = dict(os.environ, **(env or {}))
    root = os.path.join(d, {"edges": "E", "5mbp": "C2"}.get(data, "C1"))
    reads = las_cases.reads_of(data)
    if grow:
        dbio.write_db(root, reads[:grow], size_mbp=split)
        r = subprocess.run([exe] + opts + [arg], cwd=d, capture_output=True, timeout=300, env=full_env)
        if log is not None:
            log.append(r.stderr.decode(errors="replace"))
        if r.returncode != 0:
            raise RuntimeError(f"{exe} {case} (first {grow} reads): rc {r.returncode} {r.stderr[-2000:]!r}")
    dbio.write_db(root, reads, size_mbp=split)
    r = subprocess.run([exe] + opts + [arg], cwd=d, capture_output=True, timeout=300, env=full_env)
    if log is not None:
        log.append(r.stderr.decode(errors="replace"))
    if r.returncode != 0:
        raise RuntimeError(f"{exe} {case}: rc {r.returncode} {r.stderr[-2000:]!r}")
    return {ext: hashlib.sha256(open(os.path.join(d, stem(case) + "." + ext), "rb").read()).hexdigest()
            for ext in ("anno", "data")}
