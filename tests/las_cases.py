"""daligner workloads with a .las golden from the compiled reference.

Each case names a read set (marvel_amd.sim, regenerated from its seed), the
DBsplit block size and the daligner arguments.  The goldens
(tests/golden/las/goldens.json) hold the SHA-256 of every .las the reference
daligner wrote for the case; tests run the oracle restatement (CPU) and the
GPU daligner on the same DB and compare.
"""

from __future__ import annotations

import hashlib
import os
import subprocess

from marvel_amd import dbio, sim

DATA = {
    "small": lambda: sim.config_small(1),
    "small_rep": lambda: sim.config_small_repeats(11),
    "5mbp": lambda: sim.config_5mbp(2),
    "edges": lambda: sim.config_band_edges(21),
    "at_rich": lambda: sim.config_at_rich(31),
    "self_rep": lambda: sim.config_self_repeats(41),
    # the first 200 Mbp block of the 50 Mbp repeat genome (200 families x 40 copies, 1 % tandems) at 30x, 12 %
    "repeats_50mbp": lambda: sim.config_repeats_50mbp(4),
    # the first two 400 Mbp blocks of the 100 Mbp genome at 30x, nanopore-like reads (mean 25 kb, 8 %)
    "nanopore_100mbp": lambda: sim.config_nanopore_100mbp(5),
}

# name -> (data, DBsplit size in Mbp, daligner flags, block arguments)
CASES = {
    "c1_j8": ("small", 400, ["-j8", "-M16"], ["C1.1", "C1.1"]),
    "c1_j1": ("small", 400, ["-j1", "-M16"], ["C1.1", "C1.1"]),
    "c1_j64": ("small", 400, ["-j64", "-M16"], ["C1.1", "C1.1"]),
    "c1_asym": ("small", 400, ["-A", "-j4", "-M16"], ["C1.1", "C1.1"]),
    "c1_s200": ("small", 400, ["-s200", "-j8", "-M16"], ["C1.1", "C1.1"]),
    "c1_notrace": ("small", 400, ["-T", "-j8", "-M16"], ["C1.1", "C1.1"]),
    "c1_k12": ("small", 400, ["-k12", "-w5", "-h30", "-l500", "-e.80", "-j8", "-M16"], ["C1.1", "C1.1"]),
    "c1_hgap": ("small", 400, ["-H6000", "-j8", "-M16"], ["C1.1", "C1.1"]),
    # longer k-mers (the block's 64-bit keys hold 2k + 10 read bits + 15 position bits up to k = 19)
    "c1_k16": ("small", 400, ["-k16", "-j8", "-M16"], ["C1.1", "C1.1"]),
    "c1_k19": ("small", 400, ["-k19", "-h50", "-j4", "-M16"], ["C1.1", "C1.1"]),
    "c1_blocks_12": ("small", 2, ["-j8", "-M16"], ["C1.1", "C1.2"]),
    "c1_blocks_21": ("small", 2, ["-j4", "-M16"], ["C1.2", "C1.1"]),
    "c1_blocks_asym": ("small", 2, ["-A", "-j8", "-M16"], ["C1.2", "C1.1", "C1.2"]),
    "rep_j8": ("small_rep", 400, ["-j8", "-M16"], ["C1.1", "C1.1"]),
    "rep_t30": ("small_rep", 400, ["-t30", "-j8", "-M16"], ["C1.1", "C1.1"]),
    "rep_j16_l500": ("small_rep", 400, ["-j16", "-l500", "-M16"], ["C1.1", "C1.1"]),
    # pairs in the outermost diagonal bands (the band cell shared by lastp[maxdiag] and lasta[mindiag-1])
    "edge_j1": ("edges", 400, ["-j1", "-M16"], ["E.1", "E.1"]),
    "edge_j4": ("edges", 400, ["-j4", "-M16"], ["E.1", "E.1"]),
    # soft masks (-m): the reference DBdust track of the tandem-repeat reads,
    # random interval tracks merged with it, a MARVEL-format track, and
    # per-block loading (a block's own track file beside the whole-DB one)
    "rep_mdust": ("small_rep", 400, ["-mdust", "-j8", "-M16"], ["C1.1", "C1.1"]),
    "rep_mdust_mrnd": ("small_rep", 400, ["-mdust", "-mrnd", "-j4", "-M16"], ["C1.1", "C1.1"]),
    "rep_mm2": ("small_rep", 400, ["-mm2", "-j8", "-M16"], ["C1.1", "C1.1"]),
    "rep_mask_blocks": ("small_rep", 1, ["-mdust", "-mrnd", "-mpt", "-j8", "-M16"], ["C1.2", "C1.1", "C1.3"]),
    # biased k-mers (-b): variable-length k-mers of high specificity under skewed base frequencies (on
    # uniform ones they are the plain 14-mers); unbiased for comparison, with a mask (k-mers can end one past
    # a gap), and over blocks (the weights come from the first block indexed, A)
    "c1_b": ("small", 400, ["-b", "-j8", "-M16"], ["C1.1", "C1.1"]),
    "at_j8": ("at_rich", 400, ["-j8", "-M16"], ["C1.1", "C1.1"]),
    "at_b": ("at_rich", 400, ["-b", "-j8", "-M16"], ["C1.1", "C1.1"]),
    "at_b_k16": ("at_rich", 400, ["-b", "-k16", "-j8", "-M16"], ["C1.1", "C1.1"]),
    "at_b_k19": ("at_rich", 400, ["-b", "-k19", "-h50", "-j4", "-M16"], ["C1.1", "C1.1"]),
    "at_b_mrnd": ("at_rich", 400, ["-b", "-mrnd", "-j4", "-M16"], ["C1.1", "C1.1"]),
    "at_b_blocks": ("at_rich", 1, ["-b", "-j8", "-M16"], ["C1.3", "C1.1", "C1.2"]),
    # identity mode on reads holding copies of their own segments (direct and inverted): each read against
    # itself too (-I), only those overlaps (-O), asymmetric, over blocks.  (The reference's -I crashes on the
    # tandem-repeat reads of small_rep.)
    "srep_I": ("self_rep", 400, ["-I", "-j8", "-M16"], ["C1.1", "C1.1"]),
    "srep_O": ("self_rep", 400, ["-O", "-j4", "-M16"], ["C1.1", "C1.1"]),
    "srep_I_A": ("self_rep", 400, ["-I", "-A", "-j8", "-M16"], ["C1.1", "C1.1"]),
    "srep_I_blocks": ("self_rep", 2, ["-I", "-j8", "-M16"], ["C1.2", "C1.1", "C1.2"]),
    # a .dam database (the stub named ROOT.dam, as FA2dam writes it), two blocks
    "c1_dam": ("small", 2, ["-j8", "-M16"], ["C1.2", "C1.1"]),
    # the 5 Mbp / 30x workload (one ~158 Mbp block against itself)
    "c2_j8": ("5mbp", 400, ["-j8", "-M32"], ["C2.1", "C2.1"]),
    # the same with the reference DBdust track of its reads as a soft mask
    "c2_mdust": ("5mbp", 400, ["-mdust", "-j8", "-M32"], ["C2.1", "C2.1"]),
    # the same reads in ~40 Mbp blocks: block 2 against blocks 1-4 (symmetric: the .las of both sides)
    "c2_blocks": ("5mbp", 40, ["-j8", "-M32"], ["C2.2", "C2.1", "C2.2", "C2.3", "C2.4"]),
    # a memory limit that caps k-mer groups below MAXGRAM (dalign/filter.c:2107-2174)
    "c2_M4": ("5mbp", 400, ["-j8", "-M4"], ["C2.1", "C2.1"]),
    # the repeat-rich workload of the learned scorer: one 200 Mbp block against itself
    "c4_j8": ("repeats_50mbp", 200, ["-j8", "-M32"], ["R.1", "R.1"]),
    # the nanopore-like block pair: 400 Mbp against 400 Mbp at -e.75
    "c5_e75": ("nanopore_100mbp", 400, ["-e.75", "-j8", "-M32"], ["N.1", "N.2"]),
}
# minutes on one CPU core: the oracle runs them only when MI_SLOW=1
LARGE = {"c2_j8", "c2_mdust", "c2_blocks", "c2_M4", "c4_j8", "c5_e75"}

# -m tracks a case's DB carries: name -> (layout, source).  "dust" is the
# reference DBdust output for the reads (a fixture); "rnd<seed>" random
# intervals; "part<seed>" writes random intervals for block 2 in its own
# file (.ROOT.2.NAME.*) and different ones in the whole-DB file.
MASKS = {
    "rep_mdust": {"dust": ("dazzler", "dust")},
    "rep_mdust_mrnd": {"dust": ("dazzler", "dust"), "rnd": ("dazzler", "rnd5")},
    "rep_mm2": {"m2": ("marvel", "rnd6")},
    "rep_mask_blocks": {"dust": ("dazzler", "dust"), "rnd": ("dazzler", "rnd7"), "pt": ("dazzler", "part8")},
    "c2_mdust": {"dust": ("dazzler", "dust")},
    "at_b_mrnd": {"rnd": ("dazzler", "rnd9")},
}
DAM = {"c1_dam"}   # cases whose DB stub is ROOT.dam
MASK_FIXTURES = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "mask")

_reads = {}


def random_points(lens, seed: int):
    """Per read a point list of masked intervals: none (10 %), the whole read
    (3 %), or up to 6 intervals of mean length 150 starting anywhere; they
    may touch, be empty, or start at 0 / end at the read's end."""
    import numpy as np
    rng = np.random.default_rng(seed)
    out = []
    for L in lens:
        u = rng.random()
        if u < 0.10:
            out.append([])
            continue
        if u < 0.13:
            out.append([0, int(L)])
            continue
        n = int(rng.integers(1, 7))
        starts = np.sort(rng.integers(0, L + 1, size=n))
        if rng.random() < 0.2:
            starts[0] = 0
        pts, prev = [], 0
        for b in starts:
            b = max(int(b), prev)
            e = min(int(L), b + int(rng.geometric(1 / 150.0)) - 1)
            if rng.random() < 0.1:
                e = b   # empty interval
            pts += [b, e]
            prev = e
        if rng.random() < 0.2:
            pts[-1] = int(L)
        out.append(pts)
    return out


def reads_of(data: str):
    if data not in _reads:
        _reads[data] = DATA[data]().reads
    return _reads[data]


def make_db(case: str, d: str) -> None:
    data, split, _, blocks = CASES[case]
    os.makedirs(d, exist_ok=True)
    root = os.path.join(d, blocks[0].split(".")[0])
    reads = reads_of(data)
    dbio.write_db(root, reads, size_mbp=split)
    if case in DAM:
        os.replace(root + ".db", root + ".dam")
    lens = [len(r) for r in reads]
    for name, (fmt, src) in MASKS.get(case, {}).items():
        if src == "dust":
            for ext in ("anno", "data"):
                with open(os.path.join(MASK_FIXTURES, f"{data}.dust.{ext}"), "rb") as f, \
                        open(f"{os.path.join(d, '.' + os.path.basename(root))}.{name}.{ext}", "wb") as g:
                    g.write(f.read())
        elif src.startswith("rnd"):
            dbio.write_track(root, name, random_points(lens, int(src[3:])), fmt=fmt)
        else:   # part<seed>: block 2's own file plus a different whole-DB track
            seed = int(src[4:])
            dbio.write_track(root, name, random_points(lens, seed), fmt=fmt)
            first = dbio.block_starts(lens, split)
            dbio.write_track(root, name, random_points(lens[first[1]:first[2]], seed + 1), fmt=fmt, block=2)


def run(exe: str, case: str, d: str, timeout: float = 600) -> dict:
    """Runs `exe` (a daligner-compatible binary) in DB directory d; returns
    {relative .las path: sha256}."""
    _, _, flags, blocks = CASES[case]
    r = subprocess.run([exe] + flags + blocks, cwd=d, capture_output=True, text=True, timeout=timeout)
    if r.returncode != 0:
        raise RuntimeError(f"{exe} {case}: rc {r.returncode}\n{r.stdout[-2000:]}\n{r.stderr[-2000:]}")
    out = {}
    for dp, _, fs in os.walk(d):
        for f in fs:
            if f.endswith(".las"):
                p = os.path.join(dp, f)
                out[os.path.relpath(p, d)] = hashlib.sha256(open(p, "rb").read()).hexdigest()
    return out
