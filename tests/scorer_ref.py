"""CPU reference of the scorer's training gradient (test infrastructure).

grad(X, y, w, params, rnd) is the closed-form gradient of the weighted
binary cross-entropy of the MLP 16 -> 64 -> 64 -> 1 with ReLU hidden layers,
in the parameter layout of mi_scorer_grad (w1, b1, w2, b2, w3, b3).
`rnd` is applied where the GPU kernel rounds to bf16 (inputs, weights,
hidden activations, the back-propagated error signals); with the identity
and float64 it is the textbook gradient, which finite differences pin."""

import numpy as np

NF, NH = 16, 64
SIZES = [("w1", NF * NH), ("b1", NH), ("w2", NH * NH), ("b2", NH), ("w3", NH), ("b3", 1)]
NPARAM = sum(s for _, s in SIZES)


def unpack(p):
    out, o = {}, 0
    for name, s in SIZES:
        out[name] = p[o:o + s]
        o += s
    out["w1"] = out["w1"].reshape(NF, NH)
    out["w2"] = out["w2"].reshape(NH, NH)
    out["b3"] = out["b3"][0]
    return out


def pack(d):
    return np.concatenate([np.ravel(d[name]) if name != "b3" else np.array([d["b3"]]) for name, _ in SIZES])


def forward(X, p, rnd=lambda a: a, dtype=np.float64):
    q = unpack(np.asarray(p, dtype=dtype))
    x = rnd(np.asarray(X, dtype=dtype))
    h1 = rnd(np.maximum(x @ rnd(q["w1"]) + q["b1"], 0))
    h2 = rnd(np.maximum(h1 @ rnd(q["w2"]) + q["b2"], 0))
    z = h2 @ rnd(q["w3"]) + q["b3"]
    return x, h1, h2, z, q


def loss(X, y, w, p, rnd=lambda a: a, dtype=np.float64):
    *_, z, _ = forward(X, p, rnd, dtype)
    sp = np.maximum(z, 0) + np.log1p(np.exp(-np.abs(z)))
    return float(np.sum(w * (sp - y * z)) / np.sum(w))


def grad(X, y, w, p, rnd=lambda a: a, dtype=np.float64):
    x, h1, h2, z, q = forward(X, p, rnd, dtype)
    y = np.asarray(y, dtype=dtype)
    w = np.asarray(w, dtype=dtype)
    dz = w * (1 / (1 + np.exp(-z)) - y) / np.sum(w)
    g = {"w3": h2.T @ dz, "b3": np.sum(dz)}
    dh2 = rnd(np.where(h2 > 0, dz[:, None] * q["w3"][None, :], 0))
    g["b2"] = dh2.sum(0)
    g["w2"] = h1.T @ dh2
    dh1 = rnd(np.where(h1 > 0, dh2 @ rnd(q["w2"]).T, 0))
    g["b1"] = dh1.sum(0)
    g["w1"] = x.T @ dh1
    return pack(g)


def bf16(a):
    """round to the nearest bfloat16 (ties to even), as float32"""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def init_params(seed):
    """He-initialised weights, zero biases (the layout of mi_scorer_grad)."""
    rng = np.random.default_rng(seed)
    return pack({"w1": rng.normal(0, np.sqrt(2 / NF), (NF, NH)), "b1": np.zeros(NH),
                 "w2": rng.normal(0, np.sqrt(2 / NH), (NH, NH)), "b2": np.zeros(NH),
                 "w3": rng.normal(0, np.sqrt(1 / NH), NH), "b3": 0.0}).astype(np.float32)
