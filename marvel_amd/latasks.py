"""Seeded Local_Alignment task sets (the batched-aligner workload).

A task is one daligner Local_Alignment call: two reads, a COMP flag and a
seed point (diagonal `low == hgh`, anti-diagonal `anti`).  Reads of a pair are
two noisy copies of one random template; the seed sits on the true local
diagonal, optionally shifted by an offset along B (the `offset` knob models
seeds that are off the true diagonal).

Layouts produced (both consumed by the C ABI / the test harnesses):
  * `seq`   : uint8 bytes, every read preceded and followed by the value 4
              (the in-memory DB layout of db/DB.c:1349-1410);
  * `packed`: uint64 words, 2 bits per base, base i of a read at bits
              2*(i&31) of word (word0 + (i>>5)), one zero guard word before
              each read and two after it.
"""

from __future__ import annotations

import hashlib
import math
import struct
from dataclasses import dataclass

import numpy as np


def mutate_map(t: np.ndarray, rng, p_ins: float, p_del: float, p_sub: float):
    """Noisy copy of template t and, per template position i, the number of
    output bases emitted before position i."""
    n = t.shape[0]
    u = rng.random(n)
    ins = rng.random(n) < p_ins
    ins_base = rng.integers(0, 4, size=n, dtype=np.uint8)
    sub_off = rng.integers(1, 4, size=n, dtype=np.uint8)
    keep = u >= p_del
    sub = (u >= p_del) & (u < p_del + p_sub)
    base = t.copy()
    base[sub] = (base[sub] + sub_off[sub]) & 3
    cnt = ins.astype(np.int64) + keep.astype(np.int64)
    start = np.cumsum(cnt) - cnt
    out = np.empty(int(cnt.sum()), dtype=np.uint8)
    out[start[ins]] = ins_base[ins]
    out[(start + ins.astype(np.int64))[keep]] = base[keep]
    return out, start


def pack_reads(reads):
    """2-bit pack a list of uint8 reads; returns (words, word0 per read)."""
    offs = []
    total = 1
    for r in reads:
        offs.append(total)
        total += (len(r) + 31) // 32 + 2
    words = np.zeros(total + 1, dtype=np.uint64)
    for r, o in zip(reads, offs):
        n = len(r)
        nw = (n + 31) // 32
        pad = np.zeros(nw * 32, dtype=np.uint64)
        pad[:n] = r
        sh = (2 * np.arange(32, dtype=np.uint64))[None, :]
        words[o:o + nw] = np.bitwise_or.reduce(pad.reshape(nw, 32) << sh, axis=1)
    return words, np.array(offs, dtype=np.int64)


def seq_blob(reads):
    """Byte layout with a 4 before and after every read; returns (bytes, offset of each read)."""
    offs = []
    parts = [np.array([4], dtype=np.uint8)]
    pos = 1
    for r in reads:
        offs.append(pos)
        parts.append(r.astype(np.uint8))
        parts.append(np.array([4], dtype=np.uint8))
        pos += len(r) + 1
    return np.concatenate(parts), np.array(offs, dtype=np.int64)


@dataclass
class TaskSet:
    ave_corr: float
    tspace: int
    freq: tuple
    reads: list               # [a0, b0, a1, b1, ...]
    par: np.ndarray           # int32 [n, 8]: alen, blen, comp, low, hgh, anti, lbord, hbord
    selfie: np.ndarray = None  # bool [n]: B is A itself (same bases), Local_Alignment's selfie case

    @property
    def n(self) -> int:
        return self.par.shape[0]

    def _self_offsets(self, ao, bo):
        if self.selfie is not None:
            bo = np.where(self.selfie, ao, bo)
        return ao, bo

    def blob(self):
        seq, offs = seq_blob(self.reads)
        ao, bo = self._self_offsets(offs[0::2].copy(), offs[1::2].copy())
        return seq, ao, bo

    def packed(self):
        w, offs = pack_reads(self.reads)
        ao, bo = self._self_offsets(offs[0::2].copy(), offs[1::2].copy())
        return w, ao, bo

    def task_file_bytes(self) -> bytes:
        """The oracle/la_vectors task-file encoding."""
        seq, ao, bo = self.blob()
        out = [struct.pack("<idi4f", self.n, self.ave_corr, self.tspace, *self.freq),
               struct.pack("<q", seq.shape[0]), seq.tobytes()]
        for i in range(self.n):
            out.append(struct.pack("<qq8i", int(ao[i]), int(bo[i]), *[int(x) for x in self.par[i]]))
        return b"".join(out)

    def digest(self) -> str:
        return hashlib.sha256(self.task_file_bytes()).hexdigest()


def make_tasks(n: int, seed: int, lmin: int = 1000, lmax: int = 60000, emin: float = 0.05,
               emax: float = 0.20, offmax: int = 200, ave_corr: float = 0.70, tspace: int = 100,
               comp_frac: float = 0.5, freq=(0.25, 0.25, 0.25, 0.25)) -> TaskSet:
    """n tasks: template length log-uniform in [lmin, lmax], pair error e ~ U[emin, emax]
    split ins:del:sub = 9:4:2 and half into each read, seed offset along B uniform in
    [-offmax, offmax]."""
    rng = np.random.Generator(np.random.PCG64(seed))
    reads = []
    par = np.zeros((n, 8), dtype=np.int32)
    for i in range(n):
        L = int(math.exp(rng.uniform(math.log(lmin), math.log(lmax))))
        e = rng.uniform(emin, emax) / 2.0
        t = rng.integers(0, 4, size=L, dtype=np.uint8)
        a, sa = mutate_map(t, rng, e * 9 / 15, e * 4 / 15, e * 2 / 15)
        b, sb = mutate_map(t, rng, e * 9 / 15, e * 4 / 15, e * 2 / 15)
        p = int(rng.integers(L // 4, max(L // 4 + 1, 3 * L // 4)))
        apos = int(sa[p])
        bpos = int(sb[p])
        off = int(rng.integers(-offmax, offmax + 1)) if offmax > 0 else 0
        bpos = min(max(bpos + off, 0), len(b))
        comp = 1 if rng.random() < comp_frac else 0
        reads += [a, b]
        par[i] = (len(a), len(b), comp, apos - bpos, apos - bpos, apos + bpos, -1, -1)
    return TaskSet(ave_corr, tspace, tuple(freq), reads, par)


def make_self_tasks(n: int, seed: int, lmin: int = 3000, lmax: int = 20000, emin: float = 0.02,
                    emax: float = 0.15, tandem_frac: float = 0.5) -> TaskSet:
    """n tasks aligning a read against itself (B = A, the same bases: the
    reference's selfie case, dalign/align.c:1966-1985).  Each read holds a
    mutated copy (error e ~ U[emin, emax]) of one of its own segments
    further on: far away, or (tandem_frac of the tasks) right after a 20-300
    bp segment, whose wave comes near the main diagonal the selfie bounds
    keep it off.  The seed sits on the copy's diagonal, above the main
    diagonal for half of the tasks and below it (A and B roles swapped) for
    the others."""
    rng = np.random.Generator(np.random.PCG64(seed))
    reads = []
    par = np.zeros((n, 8), dtype=np.int32)
    for i in range(n):
        L = int(math.exp(rng.uniform(math.log(lmin), math.log(lmax))))
        t = rng.integers(0, 4, size=L, dtype=np.uint8)
        if rng.random() < tandem_frac:
            ln = int(rng.integers(20, 300))
            s0 = int(rng.integers(0, L - 3 * ln))
            d0 = s0 + ln
        else:
            ln = int(rng.integers(L // 8, L // 3))
            s0 = int(rng.integers(0, L - 2 * ln - 100))
            d0 = int(rng.integers(s0 + ln + 50, L - ln))
        e = rng.uniform(emin, emax)
        copy, cmap = mutate_map(t[s0:s0 + ln], rng, e * 9 / 15, e * 4 / 15, e * 2 / 15)
        a = np.concatenate([t[:d0], copy, t[d0 + ln:]]).astype(np.uint8)
        x = int(rng.integers(ln // 4, 3 * ln // 4))
        apos, bpos = d0 + int(cmap[x]), s0 + x
        if rng.random() < 0.5:
            apos, bpos = bpos, apos
        reads += [a, a]
        par[i] = (len(a), len(a), 0, apos - bpos, apos - bpos, apos + bpos, -1, -1)
    return TaskSet(0.70, 100, (0.25, 0.25, 0.25, 0.25), reads, par, np.ones(n, dtype=bool))


def parse_results(buf: bytes, n: int):
    """Decode an oracle/la_vectors result file: list of (fields[12], atrace, btrace)."""
    out = []
    o = 0
    for _ in range(n):
        f = struct.unpack_from("<12i", buf, o)
        o += 48
        at = np.frombuffer(buf, dtype=np.uint16, count=f[5], offset=o)
        o += 2 * f[5]
        bt = np.frombuffer(buf, dtype=np.uint16, count=f[11], offset=o)
        o += 2 * f[11]
        out.append((f, at.copy(), bt.copy()))
    if o != len(buf):
        raise ValueError("trailing bytes in result file")
    return out


def encode_results(res) -> bytes:
    """Inverse of parse_results."""
    parts = []
    for f, at, bt in res:
        parts.append(struct.pack("<12i", *f))
        parts.append(np.asarray(at, dtype=np.uint16).tobytes())
        parts.append(np.asarray(bt, dtype=np.uint16).tobytes())
    return b"".join(parts)
