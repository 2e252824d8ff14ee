"""Wave64 building blocks of the aligner on the GPU against CPU semantics."""

import numpy as np
import pytest

from marvel_amd import latasks
from marvel_amd.lib import lib

pytestmark = pytest.mark.gpu


def _run(which, inp=None, inp2=None, q=None, arg=0, nout=0):
    out = np.zeros(nout, dtype=np.int32)
    a = inp if inp is not None else np.zeros(1, dtype=np.int32)
    b = inp2 if inp2 is not None else np.zeros(1, dtype=np.int32)
    qq = q if q is not None else np.zeros((1, 4), dtype=np.int32)
    rc = lib.mi_selftest(which, a.ctypes.data, a.nbytes if inp is not None else 0, b.ctypes.data,
                         b.nbytes if inp2 is not None else 0, qq.ctypes.data, qq.shape[0] if q is not None else 0,
                         arg, out.ctypes.data, out.nbytes)
    assert rc == 0, lib.last_error()
    return out


def test_persistent_loop():
    print("[prim] persistent", flush=True)
    out = _run(0, arg=37, nout=5000)
    assert np.array_equal(out, 3 * np.arange(5000) + 1)


def test_lane_ops():
    print("[prim] lanes", flush=True)
    rng = np.random.default_rng(5)
    waves = 64
    v = rng.integers(0, 1000, size=waves * 64).astype(np.int32)
    out = _run(1, inp=v, arg=990, nout=waves * 64 * 8).reshape(waves, 64, 8)
    for w in range(waves):
        x = v[w * 64:(w + 1) * 64]
        for ln in range(64):
            o = out[w, ln]
            assert o[0] == (x[ln + 1:].max() if ln < 63 else -1)
            assert o[1] == (x[:ln].min() if ln > 0 else 2**31 - 1)
            assert o[2] == x[(ln - 1) % 64] and o[3] == x[(ln + 1) % 64]
            odd = np.nonzero(x & 1)[0]
            assert o[4] == (x[odd[0]] if len(odd) else -7)
            assert o[5] == int(((x[:ln] & 1) != 0).sum())
            assert o[6] == int((x > 990).any())
            assert o[7] == len(odd)


def test_snake_matcher():
    print("[prim] snake", flush=True)
    rng = np.random.default_rng(9)
    L = 3000
    a = rng.integers(0, 4, size=L, dtype=np.uint8)
    b = a.copy()
    flip = rng.random(L) < 0.05
    b[flip] = (b[flip] + 1) & 3
    wa, oa = latasks.pack_reads([a])
    wb, ob = latasks.pack_reads([b])
    assert oa[0] == 1 and ob[0] == 1
    qs = []
    for _ in range(4000):
        d = int(rng.integers(0, 2)) * 2 - 1
        x = int(rng.integers(0, L + 1))
        y = x
        lim = (L - x) if d > 0 else x
        lim = int(rng.integers(0, lim + 1))
        qs.append((x, y, lim, d))
    q = np.array(qs, dtype=np.int32)
    out = _run(2, inp=wa, inp2=wb, q=q, nout=2 * len(qs)).reshape(-1, 2)
    for i, (x, y, lim, d) in enumerate(qs):
        s = 0
        if d > 0:
            while s < lim and a[x + s] == b[y + s]:
                s += 1
        else:
            while s < lim and a[x - 1 - s] == b[y - 1 - s]:
                s += 1
        assert out[i, 0] == s and out[i, 1] == int(s < lim), (i, x, y, lim, d, out[i], s)


def test_dpp_wave_rotations_and_scans():
    """lane_below / lane_above (DPP wave_ror:1 / wave_rol:1) read lanes L-1 /
    L+1 mod 64; wave_suffix_max (values >= -1, the forward pass's
    antidiagonals and sentinel) and wave_prefix_min (values >= 0, the reverse
    pass's, sentinel INT32_MAX) are the inclusive suffix max / prefix min,
    both for the ds_bpermute suffix max (64-lane windows) and the DPP one
    (wider windows); pick3 reads lane (L + sel) mod 64."""
    print("[prim] dpp", flush=True)
    rng = np.random.default_rng(9)
    waves = 8
    v = rng.integers(-1, 2**30, size=waves * 64).astype(np.int32)
    v[64:128] = np.arange(64, dtype=np.int32)            # increasing
    v[128:192] = np.arange(64, 0, -1, dtype=np.int32)    # decreasing
    v[192:256] = 7                                       # constant
    v[256:320] = -1                                      # all sentinels
    v[320::5] = -1                                       # scattered sentinels
    v[330] = 2**30
    out = _run(3, inp=v, nout=waves * 64 * 8).reshape(waves, 64, 8)
    for w in range(waves):
        x = v[w * 64:(w + 1) * 64]
        xr = np.where(x < 0, np.iinfo(np.int32).max, x).astype(np.int32)
        assert np.array_equal(out[w, :, 0], np.roll(x, 1))
        assert np.array_equal(out[w, :, 1], np.roll(x, -1))
        assert np.array_equal(out[w, :, 2], np.maximum.accumulate(x[::-1])[::-1])
        assert np.array_equal(out[w, :, 3], np.minimum.accumulate(xr))
        assert np.array_equal(out[w, :, 4], np.append(x[1:], -7))   # wave_shl:1, lane 63 keeps the edge value
        assert np.array_equal(out[w, :, 5], np.insert(x[:-1], 0, -9))   # wave_shr:1, lane 0 keeps it
        assert np.array_equal(out[w, :, 6], np.maximum.accumulate(x[::-1])[::-1])
        lanes = np.arange(64)
        assert np.array_equal(out[w, :, 7], x[(lanes + lanes % 3 - 1) % 64])
