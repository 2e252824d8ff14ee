"""The batched-aligner workload generator, restated on the CPU
(oracle/la_gen_oracle.c), against the workload it must produce: template
lengths log-uniform in [1 kb, 60 kb], pair error uniform in [5 %, 20 %] split
half into each read as insertion : deletion : substitution = 9 : 4 : 2, seed
offset along B uniform in [-200, 200], COMP on half of the pairs.  With the
error set to zero both reads are the template and the seed sits exactly at
the drawn template position, which fixes the seed bookkeeping."""

import math

import numpy as np

import oracle_lib


def _draw(n, seed=3, **kw):
    return [oracle_lib.gen_la_pair(seed, i, **kw) for i in range(n)]


def test_lengths_log_uniform():
    st = np.array([p[3] for p in _draw(4000)])
    L = st[:, 0]
    assert L.min() >= 1000 and L.max() < 60000
    # Kolmogorov-Smirnov distance of log(L) to the uniform law on [log 1000, log 60000]
    u = np.sort((np.log(L) - math.log(1000)) / (math.log(60000) - math.log(1000)))
    n = len(u)
    ks = max(np.max(np.arange(1, n + 1) / n - u), np.max(u - np.arange(n) / n))
    assert ks < 1.36 / math.sqrt(n) + 0.002, ks   # 5 % critical value (+ quantile-table slack)


def test_error_model():
    d = _draw(2000)
    st = np.array([p[3] for p in d]).astype(np.int64)
    L, e = st[:, 0], st[:, 1]
    assert e.min() >= 50000 and e.max() <= 200000
    assert abs(e.mean() - 125000) < 4 * 43301 / math.sqrt(len(e))   # U[5 %, 20 %]: sd 4.33 %
    expect = (L * e / 2e6)   # expected events of one read, all kinds together
    for r in range(2):
        ins, dele, sub = st[:, 4 + 3 * r], st[:, 5 + 3 * r], st[:, 6 + 3 * r]
        for got, frac in ((ins, 9 / 15), (dele, 4 / 15), (sub, 2 / 15)):
            ratio = got.sum() / (expect * frac).sum()
            assert abs(ratio - 1) < 0.01, (r, frac, ratio)
    # read lengths: template + insertions - deletions
    par = np.array([p[2] for p in d])
    assert (par[:, 0] == L + st[:, 4] - st[:, 5]).all()
    assert (par[:, 1] == L + st[:, 7] - st[:, 8]).all()


def test_offsets_and_comp():
    d = _draw(4000)
    st = np.array([p[3] for p in d])
    par = np.array([p[2] for p in d])
    off = st[:, 2]
    assert off.min() >= -200 and off.max() <= 200
    assert abs(off.mean()) < 4 * 116 / math.sqrt(len(off))
    assert abs(par[:, 2].mean() - 0.5) < 0.04
    assert (par[:, 3] == par[:, 4]).all() and (par[:, 6] == -1).all() and (par[:, 7] == -1).all()
    seedp, L = st[:, 3], st[:, 0]
    assert (seedp >= L // 4).all() and (seedp < L // 4 + np.maximum(1, L // 2)).all()


def test_error_free_pairs_fix_the_seed():
    """With no error both reads equal the template and apos is the drawn
    template position; bpos is that position moved by the offset, clamped."""
    for a, b, par, st in _draw(300, seed=9, emin_ppm=0, emax_ppm=0):
        L, off, seedp = int(st[0]), int(st[2]), int(st[3])
        assert len(a) == len(b) == L and np.array_equal(a, b)
        assert (st[4:] == 0).all()
        apos = (par[5] + par[3]) // 2
        bpos = (par[5] - par[3]) // 2
        assert apos == seedp
        assert bpos == min(max(seedp + off, 0), L)


def test_pairs_share_the_template():
    """At 5 % error the reads agree on most of their k-mers: the fraction of
    A's 12-mers found in B is far above chance (4^-12 per k-mer)."""
    for a, b, _, _ in _draw(20, emin_ppm=50000, emax_ppm=50000):
        ka = {a[i:i + 12].tobytes() for i in range(0, len(a) - 12, 7)}
        kb = {b[i:i + 12].tobytes() for i in range(len(b) - 12)}
        assert len(ka & kb) > 0.5 * len(ka)


def test_deterministic():
    x = oracle_lib.gen_la_pair(3, 12345)
    y = oracle_lib.gen_la_pair(3, 12345)
    z = oracle_lib.gen_la_pair(4, 12345)
    assert np.array_equal(x[0], y[0]) and np.array_equal(x[2], y[2])
    assert not (len(x[0]) == len(z[0]) and np.array_equal(x[0], z[0]))
