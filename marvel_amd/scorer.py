"""Learned candidate scorer for the daligner report path (no reference counterpart).

A candidate is a read pair that passes the exact pre-check of
report_thread's band gate (dalign/filter.c:1740-1778): a pair the plain path
would hand to the report kernel.  k_job_features (filter_kernels.hpp)
computes 16 features per candidate from its seeds (FEATURES): the seed
count and the strongest diagonal bins -- seeds, apos span, mean and sd of the
seed spacing, the overlap the bin's diagonal implies -- the read lengths and
the share of short-period (tandem-repeat) k-mers.  In gating mode an MLP
16 -> 64 -> 64 -> 1 (ReLU, bf16 weights and activations, fp32 accumulation,
gfx950 MFMA) scores every candidate before alignment and the pairs below the
threshold are dropped; every pair it keeps runs the exact report kernel.

Labels come from the plain path itself (capture mode): the records each
pair's exact run writes and the aligner work (band cells + snake steps) it
takes.  Training runs on the GPU (mi_scorer_train: bf16 MFMA forward and
backward kernels, fp32 Adam): positives weighted by their records, negatives
by their aligner work, and the threshold is the highest that keeps the
target share of the training records.

Workflow:
    python -m marvel_amd.scorer capture --config 5mbp --out X.npz          (GPU: exact run + features)
    python -m marvel_amd.scorer train X1.npz X2.npz --out W.json           (GPU training)
    python -m marvel_amd.scorer evaluate --config repeats_50mbp            (GPU: record recall and time of gating)

The exact path (mode 0) is untouched by all of this; gating trades recall for
fewer alignments and is reported separately.
"""

from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import shutil
import struct
import sys
import tempfile
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_WEIGHTS = os.path.join(HERE, "data", "scorer_weights.json")
NFEAT, NHID = 16, 64
NPARAM = NFEAT * NHID + NHID + NHID * NHID + NHID + NHID + 1   # w1, b1, w2, b2, w3, b3
FEATURES = [
    "log2(1 + seeds of the pair) / 16",
    "seeds in the strongest diagonal bin / seeds examined (at most 256, evenly spaced)",
    "log2(1 + seeds in the strongest bin) / 16",
    "log2(1 + seeds in the runner-up bin) / 16",
    "apos span of the strongest bin / alen",
    "log2(1 + apos span of the strongest bin) / 16",
    "log2(1 + mean apos spacing of its seeds) / 16",
    "log2(1 + sd of that spacing) / 16",
    "log2(1 + alen) / 16",
    "log2(1 + blen) / 16",
    "log2(1 + overlap implied by the bin's diagonal) / 16",
    "apos span / that overlap (capped at 2)",
    "COMP",
    "short-period k-mers in the strongest bin / its seeds",
    "bins with >= minhit seeds / 8 (8 bins tracked)",
    "short-period k-mers / seeds examined",
]


@dataclass
class ScorerWeights:
    w1: np.ndarray          # (16, 64) input-major
    b1: np.ndarray          # (64,)
    w2: np.ndarray          # (64, 64)
    b2: np.ndarray          # (64,)
    w3: np.ndarray          # (64,)
    b3: float
    threshold: float
    meta: dict = field(default_factory=dict)

    def save(self, path: str):
        d = {k: np.asarray(getattr(self, k), dtype=np.float32).ravel().tolist() for k in ("w1", "b1", "w2", "b2", "w3")}
        d.update(b3=float(self.b3), threshold=float(self.threshold), meta=self.meta, features=FEATURES)
        os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
        with open(path, "w") as f:
            json.dump(d, f)

    @staticmethod
    def load(path: str = DEFAULT_WEIGHTS) -> "ScorerWeights":
        with open(path) as f:
            d = json.load(f)
        a = lambda k, shape: np.asarray(d[k], dtype=np.float32).reshape(shape)
        return ScorerWeights(a("w1", (NFEAT, NHID)), a("b1", (NHID,)), a("w2", (NHID, NHID)), a("b2", (NHID,)),
                             a("w3", (NHID,)), float(d["b3"]), float(d["threshold"]), d.get("meta", {}))


def bf16_round(x: np.ndarray) -> np.ndarray:
    """float32 -> nearest-even bf16, returned as float32 (finite inputs)."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def reference_logits(X: np.ndarray, W: ScorerWeights, kernel_rounding: bool = True) -> np.ndarray:
    """float32 forward pass.  With kernel_rounding the inputs, weights and
    hidden activations are rounded to bf16 where the MFMA kernel rounds them
    (products and sums stay float32)."""
    r = bf16_round if kernel_rounding else (lambda v: np.asarray(v, dtype=np.float32))
    x = r(X.astype(np.float32))
    h1 = r(np.maximum(x @ r(W.w1) + W.b1, 0))
    h2 = r(np.maximum(h1 @ r(W.w2) + W.b2, 0))
    return (h2 @ r(W.w3) + np.float32(W.b3)).astype(np.float32)


def sample_weights(nrec: np.ndarray, work: np.ndarray) -> np.ndarray:
    """Training weights of captured pairs: a pair with records weighs its
    record count (recall is counted in records), a pair without weighs the
    aligner work it took (what dropping it saves); each class sums to half
    the number of rows."""
    nrec = np.asarray(nrec, dtype=np.float64)
    work = np.asarray(work, dtype=np.float64)
    pos = nrec > 0
    w = np.zeros(len(nrec))
    n = len(nrec)
    if pos.any():
        w[pos] = nrec[pos] / nrec[pos].sum() * (n / 2)
    if (~pos).any():
        wk = np.maximum(work[~pos], 1.0)
        w[~pos] = wk / wk.sum() * (n / 2)
    return w.astype(np.float32)


def init_params(seed: int) -> np.ndarray:
    """He-initialised weights and zero biases as a parameter vector."""
    rng = np.random.default_rng(seed)
    W = ScorerWeights(rng.normal(0, np.sqrt(2 / NFEAT), (NFEAT, NHID)).astype(np.float32), np.zeros(NHID, np.float32),
                      rng.normal(0, np.sqrt(2 / NHID), (NHID, NHID)).astype(np.float32), np.zeros(NHID, np.float32),
                      rng.normal(0, np.sqrt(1 / NHID), NHID).astype(np.float32), 0.0, 0.0)
    return params_vector(W)


def params_vector(W: ScorerWeights) -> np.ndarray:
    """The parameter vector of the GPU trainer: w1, b1, w2, b2, w3, b3."""
    return np.concatenate([np.ravel(W.w1), W.b1, np.ravel(W.w2), W.b2, W.w3, [W.b3]]).astype(np.float32)


def weights_from_vector(p: np.ndarray, threshold: float = 0.0) -> ScorerWeights:
    o = [0, 1024, 1088, 5184, 5248, 5312]
    return ScorerWeights(p[o[0]:o[1]].reshape(NFEAT, NHID).copy(), p[o[1]:o[2]].copy(),
                         p[o[2]:o[3]].reshape(NHID, NHID).copy(), p[o[3]:o[4]].copy(), p[o[4]:o[5]].copy(),
                         float(p[o[5]]), threshold)


def choose_threshold(logits: np.ndarray, y: np.ndarray, recall: float) -> float:
    """Largest threshold that keeps at least `recall` of the positives'
    weight: y is 0 / 1 or a record count per row (logit >= threshold keeps)."""
    y = np.asarray(y, dtype=np.float64)
    pos = y > 0
    if not pos.any():
        return float("-inf")
    order = np.argsort(logits[pos], kind="stable")
    lg, wt = logits[pos][order], y[pos][order]
    total = wt.sum()
    lost = np.concatenate([[0.0], np.cumsum(wt)])   # weight below index k
    k = int(np.flatnonzero(total - lost[:-1] >= recall * total - 1e-9 * total).max())
    return float(np.nextafter(np.float32(lg[k]), np.float32(-np.inf)))


# ---- GPU binding ---------------------------------------------------------------

def _bind(h):
    fp = C.POINTER(C.c_float)
    sig = {
        "mi_scorer_load": ([fp, fp, fp, fp, fp, C.c_float, C.c_float], C.c_int),
        "mi_scorer_mode": ([C.c_int], C.c_int),
        "mi_scorer_captured": ([], C.c_longlong),
        "mi_scorer_take": ([fp, C.POINTER(C.c_int), C.POINTER(C.c_ulonglong), C.c_longlong], C.c_longlong),
        "mi_mlp_score": ([fp, C.c_longlong, fp], C.c_int),
        "mi_scorer_grad": ([fp, fp, fp, C.c_longlong, fp, fp, fp], C.c_int),
        "mi_scorer_train": ([fp, fp, fp, C.c_longlong, C.c_int, C.c_int, C.c_float, C.c_ulonglong, fp, fp], C.c_int),
    }
    for name, (args, res) in sig.items():
        f = getattr(h, name)
        f.argtypes = args
        f.restype = res


def _fp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class GpuScorer:
    """Scorer entry points of libmarvl_align (the handle of a Daligner)."""

    def __init__(self, h):
        _bind(h)
        self.h = h
        self._keep = None

    def _check(self, rc, what):
        if rc != 0:
            raise RuntimeError(f"{what}: {self.h.mi_last_error().decode(errors='replace')}")

    def load(self, W: ScorerWeights, threshold: float | None = None):
        arrs = [np.ascontiguousarray(np.asarray(a, dtype=np.float32)) for a in (W.w1, W.b1, W.w2, W.b2, W.w3)]
        self._keep = arrs
        thr = W.threshold if threshold is None else threshold
        self._check(self.h.mi_scorer_load(*[_fp(a) for a in arrs], C.c_float(W.b3), C.c_float(thr)),
                    "mi_scorer_load")

    def mode(self, m: int):
        self._check(self.h.mi_scorer_mode(m), "mi_scorer_mode")

    def take(self):
        """Captured rows: (features, records of the exact run, aligner work)."""
        n = self.h.mi_scorer_captured()
        X = np.zeros((max(n, 1), NFEAT), dtype=np.float32)
        nrec = np.zeros(max(n, 1), dtype=np.int32)
        work = np.zeros(max(n, 1), dtype=np.uint64)
        got = self.h.mi_scorer_take(_fp(X), nrec.ctypes.data_as(C.POINTER(C.c_int)),
                                    work.ctypes.data_as(C.POINTER(C.c_ulonglong)), n)
        return X[:got], nrec[:got], work[:got]

    def grad(self, X: np.ndarray, y: np.ndarray, w: np.ndarray, params: np.ndarray):
        """Loss gradient over all rows at `params` (5313 floats) on the GPU: (grad, loss)."""
        X = np.ascontiguousarray(X, dtype=np.float32)
        y = np.ascontiguousarray(y, dtype=np.float32)
        w = np.ascontiguousarray(w, dtype=np.float32)
        p = np.ascontiguousarray(params, dtype=np.float32)
        g = np.zeros(NPARAM, dtype=np.float32)
        loss = C.c_float(0)
        self._check(self.h.mi_scorer_grad(_fp(X), _fp(y), _fp(w), len(y), _fp(p), _fp(g), C.byref(loss)),
                    "mi_scorer_grad")
        return g, float(loss.value)

    def train(self, X: np.ndarray, y: np.ndarray, w: np.ndarray, params: np.ndarray, epochs: int = 30,
              batch: int = 4096, lr: float = 2e-3, seed: int = 0):
        """Adam on the GPU from `params`; returns (trained params, loss per epoch)."""
        X = np.ascontiguousarray(X, dtype=np.float32)
        y = np.ascontiguousarray(y, dtype=np.float32)
        w = np.ascontiguousarray(w, dtype=np.float32)
        p = np.array(params, dtype=np.float32)
        el = np.zeros(epochs, dtype=np.float32)
        self._check(self.h.mi_scorer_train(_fp(X), _fp(y), _fp(w), len(y), epochs, batch, C.c_float(lr), seed,
                                           _fp(p), _fp(el)), "mi_scorer_train")
        return p, el

    def logits(self, X: np.ndarray) -> np.ndarray:
        X = np.ascontiguousarray(X, dtype=np.float32)
        out = np.zeros(len(X), dtype=np.float32)
        if len(X):
            self._check(self.h.mi_mlp_score(_fp(X), len(X), _fp(out)), "mi_mlp_score")
        return out


# ---- record recall -----------------------------------------------------------------

def las_keys(path: str) -> list:
    """(aread, bread, flags, abpos, aepos, bbpos, bepos, diffs) of every record of a .las."""
    with open(path, "rb") as f:
        raw = f.read()
    n, tspace = struct.unpack_from("<qi", raw, 0)
    tb = 1 if tspace <= 125 else 2
    rec = struct.Struct("<iiiiiiIiii")
    o, out = 12, []
    for _ in range(n):
        tlen, diffs, ab, bb, ae, be, fl, ar, br, _ = rec.unpack_from(raw, o)
        out.append((ar, br, fl, ab, ae, bb, be, diffs))
        o += 40 + tlen * tb
    return out


def record_recall(exact_las: str, gated_las: str) -> tuple[float, int, int]:
    """Fraction of the exact run's records that the gated run reproduces."""
    from collections import Counter
    e, g = Counter(las_keys(exact_las)), Counter(las_keys(gated_las))
    hit = sum(min(c, g[k]) for k, c in e.items())
    tot = sum(e.values())
    return (hit / tot if tot else 1.0), hit, tot


# ---- workloads ---------------------------------------------------------------------

def _reads(config: str):
    from . import sim
    if config == "small":
        return sim.config_small(1).reads
    if config == "small_repeats":
        return sim.config_small_repeats(11).reads
    if config == "repeats_b":   # same repeat model, another genome (held out from training)
        return sim.config_small_repeats(12).reads
    if config == "5mbp":
        return sim.config_5mbp(2).reads
    if config == "repeats_50mbp":   # the first 200 Mbp block of the 50 Mbp repeat genome at 30x, 12 % error
        return sim.config_repeats_50mbp(4).reads
    if config == "repeats_train":   # the same repeat model on another 20 Mbp genome (training data)
        return sim.config_repeats_train(14).reads
    raise ValueError(f"unknown config {config}")


def train_gpu(sc: GpuScorer, X: np.ndarray, nrec: np.ndarray, work: np.ndarray, recall: float = 0.999,
              epochs: int = 40, batch: int = 4096, lr: float = 2e-3, seed: int = 0,
              groups: np.ndarray | None = None) -> ScorerWeights:
    """Trains the MLP on the GPU from captured rows (sample_weights) and sets
    the threshold that keeps `recall` of the training records of every group
    (the training configs: the lowest of the per-group thresholds, so a small
    config is not outweighed by a large one)."""
    w = sample_weights(nrec, work)
    y = (np.asarray(nrec) > 0).astype(np.float32)
    p, losses = sc.train(X, y, w, init_params(seed), epochs=epochs, batch=batch, lr=lr, seed=seed)
    W = weights_from_vector(p)
    sc.load(W)
    lg = sc.logits(X)
    nrec = np.asarray(nrec)
    groups = np.zeros(len(nrec), dtype=np.int64) if groups is None else np.asarray(groups)
    W.threshold = min(choose_threshold(lg[groups == g], nrec[groups == g], recall) for g in np.unique(groups))
    keep = lg >= W.threshold
    work = np.asarray(work, dtype=np.float64)
    W.meta = {"epochs": epochs, "seed": seed, "n": int(len(y)), "pos": int(y.sum()), "recall_target": recall,
              "loss_first": float(losses[0]), "loss_last": float(losses[-1]),
              "train_record_recall": float(np.asarray(nrec)[keep].sum() / max(1, np.asarray(nrec).sum())),
              "train_record_recall_per_group": [float(nrec[keep & (groups == g)].sum() /
                                                      max(1, nrec[groups == g].sum())) for g in np.unique(groups)],
              "train_pairs_dropped": float(1 - keep.mean()),
              "train_work_dropped": float(work[~keep].sum() / max(1.0, work.sum()))}
    return W


def run_self(config: str, mode: int, W: ScorerWeights | None = None, workdir: str | None = None,
             mem_gb: int | None = None):
    """One block of `config` against itself (daligner -k14 -w6 -h35 -e.70
    -l1000 -s100 -j8, -M16 for the small configs, -M32 otherwise) with scorer
    mode 0 / 1 / 2.  Returns (las path, (forward stats, complement stats),
    captured (X, nrec, work) or None)."""
    from . import dbio
    from .daligner import Daligner
    d = workdir or tempfile.mkdtemp(prefix="miscorer_", dir="/dev/shm" if os.access("/dev/shm", os.W_OK) else None)
    big = config in ("5mbp", "repeats_50mbp", "repeats_train")
    dbio.write_db(os.path.join(d, "S"), _reads(config), size_mbp=400)
    dal = Daligner()
    sc = GpuScorer(dal.h)
    mem = mem_gb if mem_gb is not None else (32 if big else 16)
    dal.configure(nthreads=8, mem_limit_gb=mem, minover=2000, symmetric=1)
    dal.set_filter_params(14, 6, 0, 35)
    A = dal.open_block(os.path.join(d, "S.1"))
    spec = dal.new_align_spec(0.70, 100, list(A.freq), nthreads=8)
    if W is not None:
        sc.load(W)
    cwd = os.getcwd()
    os.chdir(d)
    try:
        sc.mode(mode)
        stats = dal.compare_self(A, spec, 1, "S.1")
    finally:
        sc.mode(0)
        os.chdir(cwd)
    cap = sc.take() if mode == 1 else None
    return os.path.join(d, "001_00001", "S.1.S.1.las"), stats, cap


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m marvel_amd.scorer")
    sub = ap.add_subparsers(dest="cmd", required=True)
    c = sub.add_parser("capture", help="exact GPU run recording every candidate's features and outcome")
    c.add_argument("--config", default="5mbp")
    c.add_argument("--out", required=True)
    t = sub.add_parser("train", help="train on captured .npz files (GPU)")
    t.add_argument("inputs", nargs="+")
    t.add_argument("--out", default=DEFAULT_WEIGHTS)
    t.add_argument("--recall", type=float, default=0.999,
                   help="share of each input's training records the threshold keeps")
    t.add_argument("--epochs", type=int, default=40)
    e = sub.add_parser("evaluate", help="record recall and time of gating (GPU)")
    e.add_argument("--config", default="repeats_50mbp")
    e.add_argument("--weights", default=DEFAULT_WEIGHTS)
    a = ap.parse_args(argv)
    if a.cmd == "capture":
        las, stats, (X, nrec, work) = run_self(a.config, 1)
        np.savez_compressed(a.out, X=X, nrec=nrec, work=work, config=np.array(a.config))
        print(json.dumps({"config": a.config, "candidates": int(len(nrec)), "with_records": int((nrec > 0).sum()),
                          "records": int(nrec.sum()), "work": int(work.sum())}))
        shutil.rmtree(os.path.dirname(os.path.dirname(las)), ignore_errors=True)
    elif a.cmd == "train":
        from .daligner import Daligner
        Xs, rs, ws, gs, cfgs = [], [], [], [], []
        for i, p in enumerate(a.inputs):
            z = np.load(p)
            Xs.append(z["X"])
            rs.append(z["nrec"])
            ws.append(z["work"])
            gs.append(np.full(len(z["nrec"]), i))
            cfgs.append(str(z["config"]))
        sc = GpuScorer(Daligner().h)
        W = train_gpu(sc, np.concatenate(Xs), np.concatenate(rs), np.concatenate(ws), recall=a.recall,
                      epochs=a.epochs, groups=np.concatenate(gs))
        W.meta["configs"] = cfgs
        W.save(a.out)
        print(json.dumps(W.meta))
    else:
        W = ScorerWeights.load(a.weights)
        ex, s_ex, _ = run_self(a.config, 0)
        gd, s_gd, _ = run_self(a.config, 2, W)
        rec, hit, tot = record_recall(ex, gd)
        print(json.dumps({"config": a.config, "record_recall": rec, "records_exact": tot, "records_found": hit,
                          "pairs_scored": sum(s["cand_scored"] for s in s_gd),
                          "pairs_dropped": sum(s["cand_skipped"] for s in s_gd),
                          "report_s_exact": sum(s["t_report"] for s in s_ex),
                          "report_s_gated": sum(s["t_report"] for s in s_gd),
                          "total_s_exact": sum(s["t_total"] for s in s_ex),
                          "total_s_gated": sum(s["t_total"] for s in s_gd)}))
        for p in (ex, gd):
            shutil.rmtree(os.path.dirname(os.path.dirname(p)), ignore_errors=True)


if __name__ == "__main__":
    main()
