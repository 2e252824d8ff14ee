"""Seeded synthetic genomes and long reads for the daligner overlap workloads.

Everything here is deterministic in its seed (numpy PCG64), so the same call
produces the same reads on any machine with the same numpy.  Bases are coded
0..3 = A,C,G,T as in the DB format (db/DB.c:261-281).

Error model (per template base, independent):
  * insertion: with probability p_ins a uniformly random base is emitted
    before the template base;
  * deletion:  with probability p_del the template base is dropped;
  * substitution: with probability p_sub the template base is replaced by one
    of the three other bases (uniform);
  * otherwise the template base is copied.
Deletion and substitution are exclusive outcomes of one uniform draw; the
insertion draw is separate.  Reads are taken from a uniform start on the
genome and reverse-complemented with probability 1/2.
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np


@dataclass
class ReadSet:
    reads: list            # list of np.uint8 arrays (0..3)
    starts: np.ndarray     # template start on the genome
    lens_t: np.ndarray     # template length
    strand: np.ndarray     # 1 = reverse complement


def random_genome(length: int, seed: int) -> np.ndarray:
    """iid uniform ACGT genome of `length` bases."""
    rng = np.random.Generator(np.random.PCG64(seed))
    return rng.integers(0, 4, size=length, dtype=np.uint8)


def repeat_genome(length: int, seed: int, families=(), tandems=()) -> np.ndarray:
    """iid genome with interspersed repeat families and tandem repeats written over it.

    families: (consensus length, copies, divergence) -- every copy is the
              consensus through mutate() with the divergence split 9:4:2 into
              insertion / deletion / substitution rates, at a uniform position;
    tandems:  (period, length) -- a random unit of `period` bases repeated to
              `length` bases at a uniform position.
    """
    rng = np.random.Generator(np.random.PCG64(seed))
    g = rng.integers(0, 4, size=length, dtype=np.uint8)
    for cl, copies, div in families:
        cons = rng.integers(0, 4, size=cl, dtype=np.uint8)
        for _ in range(copies):
            c = mutate(cons, rng, div * 9 / 15, div * 4 / 15, div * 2 / 15)
            at = int(rng.integers(0, length - c.shape[0]))
            g[at:at + c.shape[0]] = c
    for period, tl in tandems:
        unit = rng.integers(0, 4, size=period, dtype=np.uint8)
        t = np.tile(unit, tl // period + 1)[:tl]
        at = int(rng.integers(0, length - tl))
        g[at:at + tl] = t
    return g


def lognormal_params(mean: float, sd: float):
    s2 = math.log(1.0 + (sd * sd) / (mean * mean))
    return math.log(mean) - s2 / 2.0, math.sqrt(s2)


def mutate(template: np.ndarray, rng, p_ins: float, p_del: float, p_sub: float) -> np.ndarray:
    n = template.shape[0]
    u = rng.random(n)
    ins = rng.random(n) < p_ins
    ins_base = rng.integers(0, 4, size=n, dtype=np.uint8)
    sub_off = rng.integers(1, 4, size=n, dtype=np.uint8)
    keep = u >= p_del
    sub = (u >= p_del) & (u < p_del + p_sub)
    base = template.copy()
    base[sub] = (base[sub] + sub_off[sub]) & 3
    # interleave: [ins_base if ins] then [base if keep], per template position
    cnt = ins.astype(np.int64) + keep.astype(np.int64)
    out = np.empty(int(cnt.sum()), dtype=np.uint8)
    pos = np.cumsum(cnt) - cnt
    out[pos[ins]] = ins_base[ins]
    out[(pos + ins.astype(np.int64))[keep]] = base[keep]
    return out


def homopolymer_mask(t: np.ndarray, min_run: int) -> np.ndarray:
    """True at every position of a run of one base at least `min_run` long."""
    n = t.shape[0]
    if n == 0:
        return np.zeros(0, dtype=bool)
    edges = np.flatnonzero(np.diff(t)) + 1
    starts = np.concatenate([[0], edges])
    runs = np.diff(np.concatenate([starts, [n]]))
    return np.repeat(runs >= min_run, runs)


def mutate_hp(template: np.ndarray, rng, p_ins: float, p_del: float, p_sub: float, hp_bias: float,
              hp_min: int = 4) -> np.ndarray:
    """mutate() with insertions and deletions `hp_bias` times as likely inside
    homopolymer runs of at least `hp_min` bases as elsewhere, scaled so the
    read's expected indel rates stay p_ins and p_del."""
    n = template.shape[0]
    w = np.where(homopolymer_mask(template, hp_min), hp_bias, 1.0)
    w /= w.mean() if n else 1.0
    u = rng.random(n)
    ins = rng.random(n) < np.minimum(p_ins * w, 1.0)
    ins_base = rng.integers(0, 4, size=n, dtype=np.uint8)
    sub_off = rng.integers(1, 4, size=n, dtype=np.uint8)
    pd = np.minimum(p_del * w, 1.0)
    keep = u >= pd
    sub = keep & (u < pd + p_sub)
    base = template.copy()
    base[sub] = (base[sub] + sub_off[sub]) & 3
    cnt = ins.astype(np.int64) + keep.astype(np.int64)
    out = np.empty(int(cnt.sum()), dtype=np.uint8)
    pos = np.cumsum(cnt) - cnt
    out[pos[ins]] = ins_base[ins]
    out[(pos + ins.astype(np.int64))[keep]] = base[keep]
    return out


def simulate_reads(genome: np.ndarray, coverage: float, mean_len: float, sd_len: float,
                   min_len: int, seed: int, p_ins=0.09, p_del=0.04, p_sub=0.02,
                   max_len: int | None = None, hp_bias: float | None = None,
                   blocks: tuple | None = None) -> ReadSet:
    """Reads at `coverage`x with lognormal template lengths (mean, sd), clipped
    below at `min_len` by resampling and above at `max_len` (and the genome).
    hp_bias: indels that many times as likely in homopolymer runs >= 4
    (mutate_hp).  blocks = (size_mbp, n): stop once the reads fill n DBsplit
    blocks of size_mbp (db/DBsplit.c:204-236) -- the first n blocks of the
    full read set, read for read."""
    rng = np.random.Generator(np.random.PCG64(seed))
    mu, sigma = lognormal_params(mean_len, sd_len)
    G = genome.shape[0]
    target = coverage * G
    total = 0
    btot, bdone = 0, 0
    reads, starts, lens, strands = [], [], [], []
    while total < target:
        L = int(rng.lognormal(mu, sigma))
        if L < min_len or L > G or (max_len is not None and L > max_len):
            continue
        s = int(rng.integers(0, G - L + 1))
        t = genome[s:s + L]
        if hp_bias is None:
            r = mutate(t, rng, p_ins, p_del, p_sub)
        else:
            r = mutate_hp(t, rng, p_ins, p_del, p_sub, hp_bias)
        st = int(rng.integers(0, 2))
        if st:
            r = (3 - r[::-1]).astype(np.uint8)
        reads.append(r)
        starts.append(s)
        lens.append(L)
        strands.append(st)
        total += L
        if blocks is not None:
            btot += len(r)
            if btot >= blocks[0] * 1000000:
                btot = 0
                bdone += 1
                if bdone == blocks[1]:
                    break
    return ReadSet(reads, np.array(starts), np.array(lens), np.array(strands))


def repeat_model_genome(length: int, seed: int, families: int = 200, family_len: int = 6000, copies: int = 40,
                        divergence: float = 0.10, tandem_frac: float = 0.01, periods=(2, 50),
                        tandem_lens=(500, 5000)) -> np.ndarray:
    """iid genome with `families` interspersed repeat families (a random
    consensus of family_len bases, `copies` copies each at `divergence`
    split 9:4:2 into insertions / deletions / substitutions, at uniform
    positions) and tandem repeats (period uniform in `periods`, array length
    uniform in `tandem_lens`) until they cover tandem_frac of the genome."""
    rng = np.random.Generator(np.random.PCG64(seed + 7919))
    tandems = []
    covered = 0
    while covered < tandem_frac * length:
        per = int(rng.integers(periods[0], periods[1] + 1))
        tl = int(rng.integers(tandem_lens[0], tandem_lens[1] + 1))
        tandems.append((per, tl))
        covered += tl
    return repeat_genome(length, seed, families=[(family_len, copies, divergence)] * families, tandems=tandems)


def write_fasta(path: str, reads, prefix: str = "sim", width: int = 80) -> None:
    """PacBio-style headers `>prefix/<well>/0_<len>` so FA2db keeps one read per well."""
    letters = np.frombuffer(b"ACGT", dtype=np.uint8)
    with open(path, "wb") as f:
        for i, r in enumerate(reads):
            f.write(b">%s/%d/0_%d\n" % (prefix.encode(), i + 1, len(r)))
            s = letters[r].tobytes()
            for j in range(0, len(s), width):
                f.write(s[j:j + width] + b"\n")


# ---- the named workloads -------------------------------------------------------

def config_small(seed: int = 1):
    """200 kbp iid genome, 20x, lognormal(4 kb, 2 kb, min 500), 9/4/2 % errors."""
    g = random_genome(200_000, seed)
    return simulate_reads(g, 20.0, 4000.0, 2000.0, 500, seed + 1000)


def config_small_repeats(seed: int = 11):
    """config_small's read model over a 200 kbp genome with repeat families
    (3 kb x 6 at 5 %, 1.5 kb x 10 at 2 %) and tandem repeats (periods 2, 7, 30)."""
    g = repeat_genome(200_000, seed, families=((3000, 6, 0.05), (1500, 10, 0.02)),
                      tandems=((2, 3000), (7, 2000), (30, 4000)))
    return simulate_reads(g, 20.0, 4000.0, 2000.0, 500, seed + 1000)


def config_at_rich(seed: int = 31):
    """config_small's read model over a 200 kbp genome of skewed composition
    (A, C, G, T at 35 / 15 / 15 / 35 %): the base frequencies daligner -b
    weighs k-mers by."""
    rng = np.random.default_rng(seed)
    g = rng.choice(4, size=200_000, p=[0.35, 0.15, 0.15, 0.35]).astype(np.uint8)
    return simulate_reads(g, 20.0, 4000.0, 2000.0, 500, seed + 1000)


def config_self_repeats(seed: int = 41):
    """config_small's read model with repeats inside reads: 30 % of the reads
    (>= 8 kb) hold a copy of one of their own 1-2.5 kb segments (4 % point
    changes) at least 3 kb further on, 15 % an inverted copy (reverse
    complement); the seeds daligner -I pairs a read with itself on."""
    rs = config_small(seed)
    rng = np.random.default_rng(seed + 7)
    out = []
    for r in rs.reads:
        r = r.copy()
        u = rng.random()
        if len(r) >= 8000 and u < 0.45:
            ln = int(rng.integers(1000, 2500))
            s0 = int(rng.integers(0, len(r) - ln - 3000 - ln))
            d0 = int(rng.integers(s0 + ln + 3000, len(r) - ln + 1))
            seg = r[s0:s0 + ln].copy()
            if u >= 0.30:
                seg = (3 - seg[::-1]).astype(np.uint8)
            flip = rng.random(ln) < 0.04
            seg[flip] = (seg[flip] + rng.integers(1, 4, size=int(flip.sum()))) & 3
            r[d0:d0 + ln] = seg
        out.append(r)
    return ReadSet(out, rs.starts, rs.lens_t, rs.strand)


def config_band_edges(seed: int = 21):
    """Reads that drive report_thread's outermost diagonal bands: two reads of
    the block's maximum length X1, X2 (4,990 bp); Y begins with X1's last 60 bp
    (the pair A=Y, B=X1 lies in the lowest band, whose widened span reaches
    the band-array cell shared with the highest band's lastp); W begins with
    X2's last 60 bp (the later pair A=X2, B=W scores seeds in the highest
    band); error-free, with random filler reads around them.  Read order
    R0 < X1 < Y < W < X2 < R5.."""
    rng = np.random.default_rng(seed)
    rnd = lambda n: rng.integers(0, 4, size=n, dtype=np.uint8)
    x1, x2 = rnd(4990), rnd(4990)
    y = np.concatenate([x1[-60:], rnd(2940)])
    w = np.concatenate([x2[-60:], rnd(2940)])
    reads = [rnd(3000), x1, y, w, x2] + [rnd(int(rng.integers(1500, 4000))) for _ in range(6)]
    n = len(reads)
    return ReadSet(reads, np.zeros(n, dtype=np.int64), np.array([len(r) for r in reads]), np.zeros(n, dtype=np.int64))


def config_5mbp(seed: int = 2):
    """5 Mbp iid genome, 30x, lognormal(10 kb, 6 kb, min 1 kb), 9/4/2 % errors."""
    g = random_genome(5_000_000, seed)
    return simulate_reads(g, 30.0, 10000.0, 6000.0, 1000, seed + 1000)


# ---- the repeat-rich and long-read workloads --------------------------------------

def config_repeats_50mbp(seed: int = 4, blocks: int = 1):
    """50 Mbp iid genome with 200 repeat families (6 kb consensus, 40 copies at
    10 % divergence) and 1 % tandem repeats (period 2-50); reads at 30x,
    lognormal(10 kb, 6 kb, min 1 kb), 12 % error split 9:4:2.  Only the
    first `blocks` DBsplit blocks of 200 Mbp are simulated (the full set's
    first reads, read for read)."""
    g = repeat_model_genome(50_000_000, seed)
    e = 0.12
    return simulate_reads(g, 30.0, 10000.0, 6000.0, 1000, seed + 1000, p_ins=e * 9 / 15, p_del=e * 4 / 15,
                          p_sub=e * 2 / 15, blocks=(200, blocks))


def config_nanopore_100mbp(seed: int = 5, blocks: int = 2, block_mbp: int = 400):
    """100 Mbp iid genome with the same repeat model; reads at 30x,
    lognormal(25 kb, 20 kb) up to 200 kb (min 1 kb), 8 % error split 9:4:2
    with indels 3x as likely in homopolymer runs >= 4.  Only the first
    `blocks` DBsplit blocks of block_mbp are simulated."""
    g = repeat_model_genome(100_000_000, seed)
    e = 0.08
    return simulate_reads(g, 30.0, 25000.0, 20000.0, 1000, seed + 1000, p_ins=e * 9 / 15, p_del=e * 4 / 15,
                          p_sub=e * 2 / 15, max_len=200_000, hp_bias=3.0, blocks=(block_mbp, blocks))


def config_repeats_train(seed: int = 14):
    """Training data for the learned scorer with the repeat model of
    config_repeats_50mbp on another, smaller genome: 20 Mbp iid with 80
    repeat families (6 kb consensus, 40 copies at 10 %) and 1 % tandems,
    reads at 30x, lognormal(10 kb, 6 kb, min 1 kb), 12 % error; the first
    100 Mbp block."""
    g = repeat_model_genome(20_000_000, seed, families=80)
    e = 0.12
    return simulate_reads(g, 30.0, 10000.0, 6000.0, 1000, seed + 1000, p_ins=e * 9 / 15, p_del=e * 4 / 15,
                          p_sub=e * 2 / 15, blocks=(100, 1))
