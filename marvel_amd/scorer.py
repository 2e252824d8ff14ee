"""Learned candidate scorer for the daligner report path (no reference counterpart).

A candidate is a seed that passes report_thread's band gate
(dalign/filter.c:1772-1778): the point where the reference calls
Local_Alignment.  The GPU report kernel computes 16 features per candidate
(FEATURES) and, in gating mode, evaluates an MLP 16 -> 64 -> 64 -> 1 (ReLU,
bf16 weights and activations, fp32 accumulation, gfx950 MFMA) and skips the
alignment when the logit is below the threshold.  Label 1 = the alignment
from that candidate was kept ((aepos-abpos)+(bepos-bbpos) >= 2*minlen).

Workflow:
    python -m marvel_amd.scorer capture --config small_repeats --out X.npz   (GPU: exact run + features)
    python -m marvel_amd.scorer train X1.npz X2.npz --out W.json          (CPU: torch training)
    python -m marvel_amd.scorer evaluate --config small --weights W.json  (GPU: record recall of gating)

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
FEATURES = [
    "gate score (score[b] + max(score[b-1], score[b+1])) / h",
    "score[b] / h",
    "max(score[b-1], score[b+1]) / h",
    "log2(1 + seeds of the pair) / 16",
    "seed index within the pair / seeds of the pair",
    "apos / alen",
    "bpos / blen",
    "log2(1 + alen) / 16",
    "log2(1 + blen) / 16",
    "comp",
    "log2(1 + overlap length implied by the diagonal) / 16",
    "alignments of the pair so far / 8 (capped at 16)",
    "kept alignments of the pair so far / 8 (capped at 16)",
    "lasta[b] > 0",
    "seeds of the panel / seeds of the pair",
    "log2(1 + distance of the seed to the nearest read end) / 16",
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


def train(X: np.ndarray, y: np.ndarray, epochs: int = 40, seed: int = 0, lr: float = 2e-3,
          batch: int = 2048) -> ScorerWeights:
    """Trains the MLP with torch on the CPU (binary cross-entropy with the
    classes balanced).  Threshold left at 0; see choose_threshold."""
    import torch

    torch.manual_seed(seed)
    rng = np.random.default_rng(seed)
    Xt = torch.from_numpy(X.astype(np.float32))
    yt = torch.from_numpy(y.astype(np.float32))
    model = torch.nn.Sequential(torch.nn.Linear(NFEAT, NHID), torch.nn.ReLU(), torch.nn.Linear(NHID, NHID),
                                torch.nn.ReLU(), torch.nn.Linear(NHID, 1))
    npos = float(y.sum())
    nneg = float(len(y) - npos)
    pos_weight = torch.tensor([nneg / max(npos, 1.0)])
    lossf = torch.nn.BCEWithLogitsLoss(pos_weight=pos_weight)
    opt = torch.optim.Adam(model.parameters(), lr=lr)
    n = len(y)
    for ep in range(epochs):
        perm = torch.from_numpy(rng.permutation(n))
        for i in range(0, n, batch):
            idx = perm[i:i + batch]
            opt.zero_grad()
            loss = lossf(model(Xt[idx]).squeeze(1), yt[idx])
            loss.backward()
            opt.step()
    W = weights_from_torch(model)
    W.meta = {"epochs": epochs, "seed": seed, "n": int(n), "pos": int(npos)}
    return W


def weights_from_torch(model) -> ScorerWeights:
    """ScorerWeights of a torch Sequential(Linear(16,64), ReLU, Linear(64,64), ReLU, Linear(64,1))."""
    l1, l2, l3 = model[0], model[2], model[4]
    g = lambda t: t.detach().numpy().astype(np.float32)
    return ScorerWeights(g(l1.weight).T.copy(), g(l1.bias), g(l2.weight).T.copy(), g(l2.bias), g(l3.weight)[0].copy(),
                         float(g(l3.bias)[0]), 0.0)


def choose_threshold(logits: np.ndarray, y: np.ndarray, recall: float) -> float:
    """Largest threshold that keeps at least `recall` of the positives."""
    pos = np.sort(logits[y == 1])
    if len(pos) == 0:
        return float("-inf")
    k = int(np.floor((1.0 - recall) * len(pos)))   # positives that may fall below
    return float(np.nextafter(pos[k], -np.inf))


# ---- GPU binding ---------------------------------------------------------------

def _bind(h):
    fp = C.POINTER(C.c_float)
    sig = {
        "mi_scorer_load": ([fp, fp, fp, fp, fp, C.c_float, C.c_float], C.c_int),
        "mi_scorer_mode": ([C.c_int], C.c_int),
        "mi_scorer_captured": ([], C.c_longlong),
        "mi_scorer_take": ([fp, C.POINTER(C.c_int), C.c_longlong], C.c_longlong),
        "mi_mlp_score": ([fp, C.c_longlong, fp], C.c_int),
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
        n = self.h.mi_scorer_captured()
        X = np.zeros((max(n, 1), NFEAT), dtype=np.float32)
        y = np.zeros(max(n, 1), dtype=np.int32)
        got = self.h.mi_scorer_take(_fp(X), y.ctypes.data_as(C.POINTER(C.c_int)), n)
        return X[:got], y[:got]

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
    raise ValueError(f"unknown config {config}")


def run_self(config: str, mode: int, W: ScorerWeights | None = None, workdir: str | None = None,
             mem_gb: int | None = None):
    """One block of `config` against itself (daligner -k14 -w6 -h35 -e.70
    -l1000 -s100 -j8, -M16 for the small configs and -M32 for 5mbp) with
    scorer mode 0 / 1 / 2.  Returns (las path, (forward stats, complement
    stats), captured X, y)."""
    from . import dbio
    from .daligner import Daligner
    d = workdir or tempfile.mkdtemp(prefix="miscorer_", dir="/dev/shm" if os.access("/dev/shm", os.W_OK) else None)
    dbio.write_db(os.path.join(d, "S"), _reads(config), size_mbp=400)
    dal = Daligner()
    sc = GpuScorer(dal.h)
    mem = mem_gb if mem_gb is not None else (32 if config == "5mbp" else 16)
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
    X, y = sc.take() if mode == 1 else (None, None)
    return os.path.join(d, "001_00001", "S.1.S.1.las"), stats, X, y


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m marvel_amd.scorer")
    sub = ap.add_subparsers(dest="cmd", required=True)
    c = sub.add_parser("capture", help="exact GPU run recording candidate features and outcomes")
    c.add_argument("--config", default="small_repeats")
    c.add_argument("--out", required=True)
    t = sub.add_parser("train", help="train on captured .npz files (CPU)")
    t.add_argument("inputs", nargs="+")
    t.add_argument("--out", default=DEFAULT_WEIGHTS)
    t.add_argument("--recall", type=float, default=0.999, help="candidate-level recall the threshold keeps")
    t.add_argument("--epochs", type=int, default=40)
    e = sub.add_parser("evaluate", help="record recall and skipped alignments of gating (GPU)")
    e.add_argument("--config", default="small")
    e.add_argument("--weights", default=DEFAULT_WEIGHTS)
    a = ap.parse_args(argv)
    if a.cmd == "capture":
        las, stats, X, y = run_self(a.config, 1)
        np.savez_compressed(a.out, X=X, y=y, config=np.array(a.config))
        print(json.dumps({"config": a.config, "candidates": int(len(y)), "kept": int(y.sum())}))
        shutil.rmtree(os.path.dirname(os.path.dirname(las)), ignore_errors=True)
    elif a.cmd == "train":
        Xs, ys, cfgs = [], [], []
        for p in a.inputs:
            z = np.load(p)
            Xs.append(z["X"])
            ys.append(z["y"])
            cfgs.append(str(z["config"]))
        X, y = np.concatenate(Xs), np.concatenate(ys)
        W = train(X, y, epochs=a.epochs)
        lg = reference_logits(X, W)
        W.threshold = choose_threshold(lg, y, a.recall)
        keep = lg >= W.threshold
        W.meta.update(recall_target=a.recall, configs=cfgs,
                      train_recall=float(keep[y == 1].mean()), train_skip=float(1 - keep.mean()),
                      neg_skipped=float(1 - keep[y == 0].mean()) if (y == 0).any() else None)
        W.save(a.out)
        print(json.dumps(W.meta))
    else:
        W = ScorerWeights.load(a.weights)
        ex, s_ex, _, _ = run_self(a.config, 0)
        gd, s_gd, _, _ = run_self(a.config, 2, W)
        rec, hit, tot = record_recall(ex, gd)
        print(json.dumps({"config": a.config, "record_recall": rec, "records_exact": tot, "records_found": hit,
                          "scored": sum(s["cand_scored"] for s in s_gd),
                          "skipped": sum(s["cand_skipped"] for s in s_gd),
                          "aligned_exact": sum(s["naligned"] for s in s_ex),
                          "aligned_gated": sum(s["naligned"] for s in s_gd)}))
        for p in (ex, gd):
            shutil.rmtree(os.path.dirname(os.path.dirname(p)), ignore_errors=True)


if __name__ == "__main__":
    main()
