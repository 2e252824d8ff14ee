"""Candidate scorer on the GPU: the MFMA MLP against a float32 reference,
capture mode leaves the .las bit-exact and labels every candidate pair with
the records its exact run wrote, gating with an open threshold is the exact
path, and the shipped weights keep >= 99.5 % of the records on data they
were not trained on."""

import hashlib
import json
import os
import shutil

import numpy as np
import pytest

import scorer_ref as R
from marvel_amd import scorer as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = json.load(open(os.path.join(ROOT, "tests", "golden", "las", "goldens.json")))
CASE = {"small": "c1_j8", "small_repeats": "rep_j8"}


def _sha(p):
    return hashlib.sha256(open(p, "rb").read()).hexdigest()


def _gold(config):
    return next(iter(GOLD[CASE[config]]["las"].values()))


def _cleanup(las):
    shutil.rmtree(os.path.dirname(os.path.dirname(las)), ignore_errors=True)


@pytest.mark.gpu
def test_mlp_kernel_matches_fp32_reference():
    from marvel_amd.daligner import Daligner
    print("[scorer] MFMA MLP vs float32 reference", flush=True)
    sc = S.GpuScorer(Daligner().h)
    p = R.init_params(7)
    rng = np.random.default_rng(8)
    p[1024:1088] = rng.normal(0, 0.2, 64)
    p[5184:5248] = rng.normal(0, 0.2, 64)
    W = S.weights_from_vector(p)
    sc.load(W)
    for n in (1, 31, 33, 1000, 4099):
        X = rng.uniform(0, 2, (n, 16)).astype(np.float32)
        got = sc.logits(X)
        ref32 = R.forward(X, p, dtype=np.float32)[3]
        refk = S.reference_logits(X, W)
        # against the reference that rounds to bf16 where the kernel does:
        # only the float32 summation order differs
        assert np.all(np.abs(got - refk) <= 1e-3 * (1 + np.abs(refk))), (n, np.abs(got - refk).max())
        # against plain float32: bf16 inputs, weights and activations (8-bit mantissas)
        assert np.all(np.abs(got - ref32) <= 0.03 * (1 + np.abs(ref32))), (n, np.abs(got - ref32).max())


@pytest.mark.gpu
@pytest.mark.parametrize("config", ["small", "small_repeats"])
def test_capture_keeps_exact_output(config):
    print(f"[scorer] capture {config}", flush=True)
    las, stats, (X, nrec, work) = S.run_self(config, 1)
    try:
        assert _sha(las) == _gold(config)
        assert X.shape == (len(nrec), 16) and np.isfinite(X).all()
        # one row per pair past the gate pre-check; their records are the .las records
        assert len(nrec) == sum(s["njobs_kept"] for s in stats)
        assert int(nrec.sum()) == len(S.las_keys(las))
        assert (work[nrec > 0] > 0).all()
        assert 0 < (nrec > 0).sum() < len(nrec)
        # features in their ranges: shares in [0, 1], COMP 0 / 1 (forward then complement pass)
        for j in (1, 13, 14, 15):
            assert (X[:, j] >= 0).all() and (X[:, j] <= 1).all(), j
        assert set(np.unique(X[:, 12])) == {0.0, 1.0}
    finally:
        _cleanup(las)


@pytest.mark.gpu
def test_gate_with_open_threshold_is_exact():
    print("[scorer] gate, threshold -inf", flush=True)
    W = S.weights_from_vector(R.init_params(9))
    W.threshold = -3.0e38
    las, stats, _ = S.run_self("small", 2, W)
    try:
        assert _sha(las) == _gold("small")
        assert sum(s["cand_skipped"] for s in stats) == 0
        assert sum(s["cand_scored"] for s in stats) == sum(s["njobs_kept"] for s in stats)
    finally:
        _cleanup(las)


@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(S.DEFAULT_WEIGHTS), reason="no trained weights shipped")
@pytest.mark.parametrize("config", ["small", "repeats_b"])
def test_gate_recall_held_out(config):
    W = S.ScorerWeights.load()
    if config in W.meta.get("configs", []):
        pytest.skip(f"{config} is training data")
    print(f"[scorer] gate recall {config}", flush=True)
    ex, _, _ = S.run_self(config, 0)
    gd, stats, _ = S.run_self(config, 2, W)
    try:
        rec, hit, tot = S.record_recall(ex, gd)
        print(f"recall {rec:.5f} ({hit}/{tot}), skipped {sum(s['cand_skipped'] for s in stats)} of "
              f"{sum(s['cand_scored'] for s in stats)} candidates", flush=True)
        assert rec >= 0.995
    finally:
        _cleanup(ex)
        _cleanup(gd)


@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(S.DEFAULT_WEIGHTS), reason="no trained weights shipped")
def test_trained_network_gains_over_single_feature_rules():
    """What the trained network gains over the method without it: on held-out
    repeat-rich reads, at the record recall the shipped threshold reaches,
    the MLP drops more aligner work than the best threshold on any single
    feature (either sign) keeping at least that recall."""
    from marvel_amd.daligner import Daligner

    W = S.ScorerWeights.load()
    assert "repeats_b" not in W.meta.get("configs", [])
    print("[scorer] gain over single-feature rules, repeats_b", flush=True)
    las, _, (X, nrec, work) = S.run_self("repeats_b", 1)
    _cleanup(las)
    sc = S.GpuScorer(Daligner().h)
    sc.load(W)
    lg = sc.logits(X)
    nrec = nrec.astype(np.float64)
    work = work.astype(np.float64)
    keep = lg >= W.threshold
    rec = nrec[keep].sum() / nrec.sum()
    mlp = work[~keep].sum() / work.sum()
    best = 0.0
    for j in range(X.shape[1]):
        for sgn in (1.0, -1.0):
            v = (sgn * X[:, j]).astype(np.float32)
            k = v >= S.choose_threshold(v, nrec, rec)
            if nrec[k].sum() / nrec.sum() >= rec - 1e-12:
                best = max(best, work[~k].sum() / work.sum())
    print(f"recall {rec:.5f}: MLP drops {mlp:.5f} of the aligner work, the best single feature {best:.5f}", flush=True)
    assert rec >= 0.995
    assert mlp > best
