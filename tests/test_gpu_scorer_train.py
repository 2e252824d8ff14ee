"""Scorer training on the GPU (bf16 MFMA forward and backward, fp32 Adam):
the gradient against the CPU closed form that rounds where the kernel rounds
(tests/scorer_ref.py, itself pinned by finite differences), and training
that learns a rule."""

import numpy as np
import pytest

import scorer_ref as R
from marvel_amd import scorer as S

pytestmark = pytest.mark.gpu

BLOCKS = [("w1", 0, 1024), ("b1", 1024, 1088), ("w2", 1088, 5184), ("b2", 5184, 5248), ("w3", 5248, 5312),
          ("b3", 5312, 5313)]


def _scorer():
    from marvel_amd.daligner import Daligner
    return S.GpuScorer(Daligner().h)


def _problem(n, seed):
    rng = np.random.default_rng(seed)
    X = np.concatenate([rng.normal(0.5, 1.0, (n // 2, 16)), rng.uniform(0, 2, (n - n // 2, 16))]).astype(np.float32)
    y = (rng.random(n) < 0.35).astype(np.float32)
    w = np.where(y > 0, 2.0, 1.0).astype(np.float32) * rng.uniform(0.5, 1.5, n).astype(np.float32)
    p = R.init_params(seed)
    p[1024:1088] = rng.normal(0, 0.2, 64)
    p[5184:5248] = rng.normal(0, 0.2, 64)
    p[5312] = -0.2
    return X, y, w, p.astype(np.float32)


@pytest.mark.parametrize("n", [32, 1000, 20011])
def test_gradient_matches_cpu_reference(n):
    print(f"[scorer train] gradient n={n}", flush=True)
    X, y, w, p = _problem(n, 5 + n)
    g, loss = _scorer().grad(X, y, w, p)
    ref = R.grad(X, y, w, p, rnd=R.bf16, dtype=np.float32)
    ref64 = R.grad(X, y, w, p)
    assert abs(loss - R.loss(X, y, w, p, rnd=R.bf16, dtype=np.float32)) < 1e-4 * (1 + abs(loss))
    for name, a, b in BLOCKS:
        d = g[a:b] - ref[a:b]
        scale = np.abs(ref[a:b]).max() + 1e-12
        # same roundings, other float32 summation order: a bf16 error signal may
        # round one unit the other way (2^-8 relative), nothing more
        assert np.abs(d).max() <= 1e-2 * scale, (name, np.abs(d).max(), scale)
        assert np.linalg.norm(d) <= 3e-3 * np.linalg.norm(ref[a:b]) + 1e-9, name
        # and the float64 gradient without any rounding is near: bf16 operands move
        # pre-activations by ~2^-9 and flip the ReLU mask of units near zero, whose
        # effect on a sum over n rows shrinks like 1 / sqrt(n)
        tol = 0.05 + 1.0 / np.sqrt(n)
        assert np.linalg.norm(g[a:b] - ref64[a:b]) <= tol * np.linalg.norm(ref64[a:b]) + 1e-9, name


def test_training_learns_a_rule():
    print("[scorer train] learning", flush=True)
    rng = np.random.default_rng(21)
    n = 30000
    X = rng.random((n, 16)).astype(np.float32)
    y = ((X[:, 0] + 0.5 * X[:, 3] - 0.7 * X[:, 9]) > 0.4).astype(np.float32)
    w = np.ones(n, dtype=np.float32)
    sc = _scorer()
    p0 = R.init_params(4)
    p, losses = sc.train(X, y, w, p0, epochs=25, batch=2048, lr=3e-3, seed=1)
    print("losses", losses[[0, 5, 10, 24]], flush=True)
    assert losses[-1] < 0.5 * losses[0] and losses[-1] < 0.12
    W = S.weights_from_vector(p)
    sc.load(W)
    acc = ((sc.logits(X) >= 0) == (y > 0)).mean()
    assert acc > 0.97, acc
    # the trained vector through the inference kernel = the CPU forward pass
    lg = sc.logits(X[:500])
    ref = R.forward(X[:500], p, rnd=R.bf16, dtype=np.float32)[3]
    assert np.abs(lg - ref).max() <= 1e-3 * (1 + np.abs(ref).max())
