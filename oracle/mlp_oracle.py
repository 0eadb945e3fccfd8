"""CPU restatement of the learned rebuild scheduler's network -- TEST
INFRASTRUCTURE ONLY (tests/ may import it; the product never does).

Network: 16 -> 64 -> 64 -> 1, ReLU after both hidden layers, inputs
standardised with the model's featMean / featStd.  Loss per row: squared
error of the prediction (log2 steps); a censored row (the true value is only
known to exceed the target) contributes only when the prediction is below its
target.  Gradients are sums over rows.

Two forms:
  exact      float64 throughout; pinned by central finite differences
             (tests/test_mlp_oracle.py)
  bf16       rounds to bfloat16 (round to nearest even) exactly where the
             gfx950 kernels feed MFMA operands (standardised input, W1, W2,
             first hidden activations, the two pre-activation gradients) and
             keeps the rest in float32, so it matches the device up to
             float32 accumulation order; the rounding is pinned against
             torch's bfloat16 conversion.

There is no reference counterpart: the learned scheduler is specified by the
task, not by HiDEM (it replaces the check of HiDEMLib.h:544-559).
"""
import numpy as np

NF, NH = 16, 64
SHAPES = (("W1", (NF, NH)), ("b1", (NH,)), ("W2", (NH, NH)), ("b2", (NH,)), ("w3", (NH,)), ("b3", ()))
N_PARAMS = NF * NH + NH + NH * NH + NH + NH + 1
N_WEIGHTS = N_PARAMS + 2 * NF


def unpack(flat):
    """Views of a flat hdm_mlp_weights vector (struct order)."""
    flat = np.asarray(flat)
    out, o = {}, 0
    for name, shp in SHAPES:
        size = int(np.prod(shp)) if shp else 1
        out[name] = flat[o:o + size].reshape(shp) if shp else flat[o]
        o += size
    out["mean"] = flat[o:o + NF]
    out["std"] = flat[o + NF:o + 2 * NF]
    return out


def init_weights(seed=0, mean=None, std=None):
    """He-initialised weights (float32, struct order) with the given standardisation."""
    r = np.random.default_rng(seed)
    w = np.zeros(N_WEIGHTS, np.float32)
    u = unpack(w)
    u["W1"][...] = r.normal(0, np.sqrt(2.0 / NF), (NF, NH))
    u["W2"][...] = r.normal(0, np.sqrt(2.0 / NH), (NH, NH))
    u["w3"][...] = r.normal(0, np.sqrt(1.0 / NH), NH)
    u["mean"][...] = 0.0 if mean is None else mean
    u["std"][...] = 1.0 if std is None else std
    return w


def bf16(x):
    """float32 -> bfloat16 (round to nearest, ties to even) -> float32."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    nan = np.isnan(a)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    out = u.astype(np.uint32).view(np.float32).reshape(a.shape)
    return np.where(nan, np.float32(np.nan), out)


def _forward(w, X, emulate):
    u = unpack(w)
    if emulate:
        f = np.float32
        xn = bf16((X.astype(f) - u["mean"].astype(f)) / u["std"].astype(f))
        W1, W2 = bf16(u["W1"]), bf16(u["W2"])
        a1 = (xn.astype(np.float64) @ W1.astype(np.float64)).astype(f) + u["b1"].astype(f)
        H1 = bf16(np.maximum(a1, 0))
        a2 = (H1.astype(np.float64) @ W2.astype(np.float64)).astype(f) + u["b2"].astype(f)
        H2 = np.maximum(a2, 0).astype(f)
        y = (H2.astype(np.float64) @ u["w3"].astype(np.float64)).astype(f) + f(u["b3"])
        return y, (xn, W1, W2, H1, H2)
    d = np.float64
    xn = (X.astype(d) - u["mean"].astype(d)) / u["std"].astype(d)
    W1, W2 = u["W1"].astype(d), u["W2"].astype(d)
    H1 = np.maximum(xn @ W1 + u["b1"].astype(d), 0)
    H2 = np.maximum(H1 @ W2 + u["b2"].astype(d), 0)
    y = H2 @ u["w3"].astype(d) + d(u["b3"])
    return y, (xn, W1, W2, H1, H2)


def forward(w, X, emulate_bf16=False):
    return _forward(w, X, emulate_bf16)[0]


def loss_and_grad(w, X, target, censored=None, emulate_bf16=False):
    """(loss summed over rows, gradient [N_PARAMS] summed over rows)."""
    y, (xn, W1, W2, H1, H2) = _forward(w, X, emulate_bf16)
    u = unpack(w)
    e = y - target
    keep = np.ones_like(e, dtype=bool) if censored is None else ~((np.asarray(censored) != 0) & (e > 0))
    dy = np.where(keep, 2.0 * e, 0.0)
    loss = float(np.sum(np.where(keep, e * e, 0.0)))
    g = {}
    if emulate_bf16:
        f = np.float32
        dy = dy.astype(f)
        g["w3"] = (H2.T.astype(np.float64) @ dy.astype(np.float64)).astype(f)
        g["b3"] = f(np.sum(dy.astype(np.float64)))
        G2f = (dy[:, None] * u["w3"][None, :].astype(f)) * (H2 > 0)
        g["b2"] = G2f.astype(np.float64).sum(0).astype(f)
        G2 = bf16(G2f)
        g["W2"] = (H1.T.astype(np.float64) @ G2.astype(np.float64)).astype(f)
        G1f = (G2.astype(np.float64) @ W2.T.astype(np.float64)).astype(f) * (H1 > 0)
        g["b1"] = G1f.astype(np.float64).sum(0).astype(f)
        G1 = bf16(G1f)
        g["W1"] = (xn.T.astype(np.float64) @ G1.astype(np.float64)).astype(f)
    else:
        g["w3"] = H2.T @ dy
        g["b3"] = np.sum(dy)
        G2 = (dy[:, None] * u["w3"][None, :].astype(np.float64)) * (H2 > 0)
        g["b2"] = G2.sum(0)
        g["W2"] = H1.T @ G2
        G1 = (G2 @ W2.T) * (H1 > 0)
        g["b1"] = G1.sum(0)
        g["W1"] = xn.T @ G1
    flat = np.concatenate([np.ravel(g[name]) for name, _ in SHAPES])
    return loss, flat
