"""Float64 CPU restatement of the learned densification proposal (the MLP and its
inputs), for the tests only.

TEST INFRASTRUCTURE: only the tests load this module, as the checker. The MLP has
no counterpart in the reference; it is pinned by central finite differences of
its own loss (tests/test_oracle_mlp.py) and by the BF16 rounding rule.

Layout of the 5,252 parameters: W1 [12][64], b1 [64], W2 [64][64], b2 [64],
W3 [64][4], b3 [4], row-major, input index first. Classes: 0 keep, 1 clone,
2 split, 3 prune.
"""
import numpy as np

INPUTS, HIDDEN, CLASSES = 12, 64, 4
SHAPES = [(INPUTS, HIDDEN), (HIDDEN,), (HIDDEN, HIDDEN), (HIDDEN,), (HIDDEN, CLASSES), (CLASSES,)]
PARAMS = sum(int(np.prod(s)) for s in SHAPES)
SH_C0 = 0.28209479177387814


def unpack(w):
    w = np.asarray(w, dtype=np.float64).reshape(-1)
    out, at = [], 0
    for shape in SHAPES:
        size = int(np.prod(shape))
        out.append(w[at:at + size].reshape(shape))
        at += size
    return out


def bf16(w):
    """Round float32 values to BF16 (nearest, ties to even), returned as float32."""
    u = np.asarray(w, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def forward(w, x, mean=None, scale=None, bf16_operands=False):
    """Logits (n, 4) and the hidden activations; inputs standardised first.

    bf16_operands: the matrix-core inference's arithmetic. The inputs are
    standardised in float32 ((x - mean) * scale, two roundings) and rounded to BF16;
    each hidden layer's float32 value (bias added, ReLU) is rounded to BF16 before
    it multiplies the next layer's weights. Sums are exact here (float64)."""
    w1, b1, w2, b2, w3, b3 = unpack(w)
    if bf16_operands:
        x32 = np.asarray(x, dtype=np.float32)
        if mean is not None:
            x32 = (x32 - np.asarray(mean, dtype=np.float32)) * np.asarray(scale, dtype=np.float32)
        x = bf16(x32).astype(np.float64)
        a1 = x @ w1 + b1
        h1 = bf16(np.maximum(a1, 0.0).astype(np.float32)).astype(np.float64)
        a2 = h1 @ w2 + b2
        h2 = bf16(np.maximum(a2, 0.0).astype(np.float32)).astype(np.float64)
        return h2 @ w3 + b3, (x, a1, h1, a2, h2)
    x = np.asarray(x, dtype=np.float64)
    if mean is not None:
        x = (x - np.asarray(mean, dtype=np.float64)) * np.asarray(scale, dtype=np.float64)
    a1 = x @ w1 + b1
    h1 = np.maximum(a1, 0.0)
    a2 = h1 @ w2 + b2
    h2 = np.maximum(a2, 0.0)
    return h2 @ w3 + b3, (x, a1, h1, a2, h2)


def loss_and_grads(w, x, labels, class_weights=None, mean=None, scale=None, bf16_operands=False):
    """Mean class-weighted softmax cross-entropy over the samples and its gradient (PARAMS,).

    bf16_operands: the matrix-core training's arithmetic. The forward pass is
    forward(bf16_operands=True); dL/dlogits and the masked deltas of both hidden
    layers are rounded (through float32) to BF16 where they multiply weights or
    activations, and the bias gradients sum them before that rounding."""
    if bf16_operands:
        return _loss_and_grads_bf16(w, x, labels, class_weights, mean, scale)
    z, (x, a1, h1, a2, h2) = forward(w, x, mean, scale)
    n = len(x)
    labels = np.asarray(labels, dtype=np.int64)
    zmax = z.max(axis=1, keepdims=True)
    logsum = np.log(np.exp(z - zmax).sum(axis=1, keepdims=True)) + zmax
    cw = np.ones(CLASSES) if class_weights is None else np.asarray(class_weights, dtype=np.float64)
    per = cw[labels]
    loss = float(np.mean(per * (logsum[:, 0] - z[np.arange(n), labels])))
    p = np.exp(z - logsum)
    dz = p.copy()
    dz[np.arange(n), labels] -= 1.0
    dz *= per[:, None] / n
    _, _, w2, _, w3, _ = unpack(w)
    d2 = (dz @ w3.T) * (a2 > 0)
    d1 = (d2 @ w2.T) * (a1 > 0)
    grads = [x.T @ d1, d1.sum(0), h1.T @ d2, d2.sum(0), h2.T @ dz, dz.sum(0)]
    return loss, np.concatenate([g.reshape(-1) for g in grads])


def features(params, stats, extra, extent, iteration, views_per_step=1):
    """The 12 inputs per Gaussian (n, 12) from raw parameters (n, 16), statistics
    (n, 4: summed gradient length, views, largest radius, -) and the extra
    statistics (n, 4: largest gradient length, summed depth, summed gradient x, y)."""
    p = np.asarray(params, dtype=np.float64)
    s = np.asarray(stats, dtype=np.float64)
    e = np.asarray(extra, dtype=np.float64)
    views = s[:, 1]
    seen = views > 0
    mean_grad = np.where(seen, s[:, 0] / np.where(seen, views, 1.0), 0.0)
    scales = np.exp(p[:, 4:7])
    big, small = scales.max(axis=1), scales.min(axis=1)
    colour = 0.5 + SH_C0 * p[:, 12:15]
    depth = np.where(seen, e[:, 1] / np.where(seen, views, 1.0), 1.0)
    pull = np.hypot(e[:, 2], e[:, 3])
    f = np.zeros((len(p), INPUTS))
    f[:, 0] = np.log1p(mean_grad / 2e-4)
    f[:, 1] = np.log1p(e[:, 0] / 2e-4)
    f[:, 2] = 1.0 / (1.0 + np.exp(-p[:, 3]))
    f[:, 3] = np.log(big / extent)
    f[:, 4] = np.log(small / extent)
    f[:, 5] = np.log(big / small)
    f[:, 6] = colour @ np.array([0.299, 0.587, 0.114])
    f[:, 7] = np.log1p(views)
    f[:, 8] = np.log1p(np.maximum(iteration - p[:, 15], 0.0) * views_per_step)
    f[:, 9] = np.log1p(s[:, 2])
    f[:, 10] = np.log(np.maximum(depth, 1e-3))
    f[:, 11] = np.where(s[:, 0] > 0, np.maximum(1.0 - pull / np.where(s[:, 0] > 0, s[:, 0], 1.0), 0.0), 0.0)
    return f


def extra_stats(splats, splat_grads, viewport, extra):
    """One view's update of the extra statistics (largest gradient length, summed
    depth, summed gradient x, summed gradient y) for the visible splats."""
    sp = np.asarray(splats, dtype=np.float64)
    g = np.asarray(splat_grads, dtype=np.float64)
    out = np.asarray(extra, dtype=np.float64).copy()
    vis = sp[:, 3] != 0
    gx = g[:, 0] * viewport[0] * 0.5
    gy = g[:, 1] * viewport[1] * 0.5
    length = np.hypot(gx, gy)
    out[vis, 0] = np.maximum(out[vis, 0], length[vis])
    out[vis, 1] += sp[vis, 2]
    out[vis, 2] += gx[vis]
    out[vis, 3] += gy[vis]
    return out


def _bf16_64(v):
    return bf16(np.asarray(v, dtype=np.float64).astype(np.float32)).astype(np.float64)


def _loss_and_grads_bf16(w, x, labels, class_weights, mean, scale):
    z, (xb, a1, h1, a2, h2) = forward(w, x, mean, scale, bf16_operands=True)
    n = len(xb)
    labels = np.asarray(labels, dtype=np.int64)
    zmax = z.max(axis=1, keepdims=True)
    logsum = np.log(np.exp(z - zmax).sum(axis=1, keepdims=True)) + zmax
    cw = np.ones(CLASSES) if class_weights is None else np.asarray(class_weights, dtype=np.float64)
    per = cw[labels]
    loss = float(np.mean(per * (logsum[:, 0] - z[np.arange(n), labels])))
    dz = np.exp(z - logsum)
    dz[np.arange(n), labels] -= 1.0
    dz *= per[:, None] / n
    _, _, w2, _, w3, _ = unpack(w)
    dzb = _bf16_64(dz)
    d2 = (dzb @ w3.T) * (a2 > 0)
    d2b = _bf16_64(d2)
    d1 = (d2b @ w2.T) * (a1 > 0)
    d1b = _bf16_64(d1)
    grads = [xb.T @ d1b, d1.sum(0), h1.T @ d2b, d2.sum(0), h2.T @ dzb, dz.sum(0)]
    return loss, np.concatenate([g.reshape(-1) for g in grads])

