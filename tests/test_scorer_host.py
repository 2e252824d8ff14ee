"""Host side of the candidate scorer: bf16 rounding, weight conventions,
threshold choice, training, weight files (CPU only)."""

import numpy as np
import pytest
import torch

from marvel_amd import scorer as S


def test_bf16_round_matches_torch():
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.normal(0, 3, 20000), rng.normal(0, 1e-3, 2000), [0.0, -0.0, 1.0, 65504.0, 1e-30]])
    x = x.astype(np.float32)
    # exact ties: the low 16 bits are 0x8000 (round half to even)
    ties = (np.arange(1, 200, dtype=np.uint32) << 16 | 0x8000).view(np.float32)
    x = np.concatenate([x, ties, -ties])
    want = torch.from_numpy(x).to(torch.bfloat16).float().numpy()
    assert np.array_equal(S.bf16_round(x), want)


def test_reference_logits_follow_the_cpu_forward_pass():
    import scorer_ref as R
    p = R.init_params(3)
    p[1024:1088] = np.random.default_rng(1).normal(0, 0.2, 64)
    W = S.weights_from_vector(p)
    assert np.array_equal(S.params_vector(W), p)
    X = np.random.default_rng(2).normal(0, 1, (500, 16)).astype(np.float32)
    want = R.forward(X, p, dtype=np.float32)[3]
    got = S.reference_logits(X, W, kernel_rounding=False)
    assert np.allclose(got, want, atol=1e-5, rtol=1e-5)
    # bf16 rounding moves the logits by a small relative amount only
    rb = S.reference_logits(X, W, kernel_rounding=True)
    assert np.abs(rb - want).max() < 0.05 + 0.02 * np.abs(want).max()
    assert np.allclose(rb, R.forward(X, p, rnd=R.bf16, dtype=np.float32)[3], atol=1e-4, rtol=1e-4)


def test_choose_threshold_keeps_recall():
    rng = np.random.default_rng(4)
    lg = rng.normal(0, 1, 5000).astype(np.float32)
    y = (rng.random(5000) < 0.7).astype(np.int32)
    for r in (0.9, 0.99, 0.999, 1.0):
        t = S.choose_threshold(lg, y, r)
        kept = (lg[y == 1] >= t).mean()
        assert kept >= r
        # a threshold just above the chosen one's next positive drops below the target
        pos = np.sort(lg[y == 1])
        above = pos[pos >= t]
        if r < 1.0 and len(above):
            assert (lg[y == 1] >= np.nextafter(above[0], np.inf)).mean() < r
    assert S.choose_threshold(lg, np.zeros_like(y), 0.99) == float("-inf")


def test_threshold_on_record_counts():
    """Recall counted in records: a pair with many records weighs more."""
    lg = np.array([0.1, 0.2, 0.3, 0.4, 0.5], dtype=np.float32)
    nrec = np.array([10, 0, 1, 0, 89])
    # keeping >= 99 % of 100 records: the pair with 1 record (logit 0.3) may be lost
    # 99 %: dropping the 10-record pair (the lowest positive) loses too much -> keep everything
    t = S.choose_threshold(lg, nrec, 0.99)
    assert t <= 0.1 and nrec[lg >= t].sum() == 100
    # 90 %: the 10-record pair may go (90 kept), the 1-record pair not as well (89)
    t2 = S.choose_threshold(lg, nrec, 0.90)
    assert 0.1 < t2 <= 0.3 and nrec[lg >= t2].sum() == 90


def test_sample_weights_balance_classes():
    nrec = np.array([0, 0, 3, 1, 0, 6])
    work = np.array([100, 300, 5, 5, 600, 9], dtype=np.uint64)
    w = S.sample_weights(nrec, work)
    pos = nrec > 0
    assert np.isclose(w[pos].sum(), 3.0) and np.isclose(w[~pos].sum(), 3.0)
    assert np.allclose(w[pos], np.array([3, 1, 6]) / 10 * 3)
    assert np.allclose(w[~pos], np.array([100, 300, 600]) / 1000 * 3)


def test_weights_roundtrip(tmp_path):
    import scorer_ref as R
    W = S.weights_from_vector(R.init_params(6))
    W.threshold = -1.25
    W.meta = {"k": 1}
    p = tmp_path / "w.json"
    W.save(str(p))
    V = S.ScorerWeights.load(str(p))
    for k in ("w1", "b1", "w2", "b2", "w3"):
        assert np.array_equal(getattr(V, k), getattr(W, k))
    assert V.b3 == pytest.approx(W.b3) and V.threshold == -1.25 and V.meta == {"k": 1}


def test_shipped_weights_load():
    import os
    if not os.path.exists(S.DEFAULT_WEIGHTS):
        pytest.skip("no trained weights shipped yet")
    W = S.ScorerWeights.load()
    assert W.w1.shape == (16, 64) and W.w2.shape == (64, 64) and W.w3.shape == (64,)
    assert np.isfinite(W.threshold)


def test_training_gradient_closed_form_matches_finite_differences():
    """The CPU reference of the training gradient (tests/scorer_ref.py), in
    float64 without rounding, against central differences of its loss: a
    dropped term, a missing ReLU mask or a transposed product fails here."""
    import scorer_ref as R
    rng = np.random.default_rng(11)
    n = 64
    X = rng.normal(0.5, 1.0, (n, 16))
    y = (rng.random(n) < 0.4).astype(np.float64)
    w = rng.uniform(0.5, 3.0, n)
    p = R.init_params(3).astype(np.float64)
    p[1024:1088] = rng.normal(0, 0.1, 64)      # non-zero biases so every term is exercised
    p[5184:5248] = rng.normal(0, 0.1, 64)
    p[5312] = 0.3
    g = R.grad(X, y, w, p)
    eps = 1e-6
    idx = np.concatenate([rng.choice(1024, 40, replace=False), 1024 + np.arange(0, 64, 7),
                          1088 + rng.choice(4096, 60, replace=False), 5184 + np.arange(0, 64, 5),
                          5248 + np.arange(0, 64, 3), [5312]])
    for i in idx:
        e = np.zeros_like(p)
        e[i] = eps
        fd = (R.loss(X, y, w, p + e) - R.loss(X, y, w, p - e)) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])
