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


def _model(seed):
    torch.manual_seed(seed)
    return torch.nn.Sequential(torch.nn.Linear(16, 64), torch.nn.ReLU(), torch.nn.Linear(64, 64), torch.nn.ReLU(),
                               torch.nn.Linear(64, 1))


def test_reference_logits_follow_torch_model():
    m = _model(3)
    W = S.weights_from_torch(m)
    X = np.random.default_rng(2).normal(0, 1, (500, 16)).astype(np.float32)
    want = m(torch.from_numpy(X)).detach().numpy()[:, 0]
    got = S.reference_logits(X, W, kernel_rounding=False)
    assert np.allclose(got, want, atol=1e-5, rtol=1e-5)
    # bf16 rounding moves the logits by a small relative amount only
    rb = S.reference_logits(X, W, kernel_rounding=True)
    assert np.abs(rb - want).max() < 0.05 + 0.02 * np.abs(want).max()
    assert not np.array_equal(rb, got)


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


def test_train_separates_a_simple_rule():
    rng = np.random.default_rng(5)
    X = rng.random((4000, 16)).astype(np.float32)
    y = (X[:, 0] + 0.5 * X[:, 3] > 0.8).astype(np.int32)
    W = S.train(X, y, epochs=30, seed=0, batch=256)
    lg = S.reference_logits(X, W)
    acc = ((lg >= 0) == (y == 1)).mean()
    assert acc > 0.95


def test_weights_roundtrip(tmp_path):
    W = S.weights_from_torch(_model(6))
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
