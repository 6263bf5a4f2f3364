"""Reference model of the learned motion proposer, for gradient checks.

TEST INFRASTRUCTURE ONLY: the CPU checker for the GPU proposer kernels
(only tests/ import it).  The network has no counterpart in the reference
crate; this file restates the GPU design:

    planes [N, 2, 4H, 4W] u8   downsampled (anchor, neighbour) luma
    x0 [N, H, W, 128]          features: 81 mean absolute differences of
                               each block's 4x4 anchor pixels against the
                               neighbour displaced by (ddx, ddy) in [-4, 4]^2
                               (clamped), then both frames' pixels / 255
    a_0 = lrelu(x0 @ W0[0].T + b0)            1x1, 128 -> 32
    a_l = lrelu(conv3x3(a_{l-1}, W_l) + b_l)  l = 1..4, channels 32, 64, 32, 32
    lrelu(z) = z for z > 0, z / 16 otherwise
    pred = a_4 @ W5.T + b5                    (mv / 8)
    loss = mean over blocks and both components of huber(pred - label / 8)

conv3x3 is zero-padded, tap t = 3 * dy + dx reads the input at offset
(dy - 1, dx - 1); weights are [9][cout][cin] as in the parameter vector.

Two precisions:
* float64 throughout (the exact model the GPU approximates), pinned by
  central finite differences in tests/test_oracle_proposer.py;
* ``bf16=True``: the GPU's rounding points restated (bf16 weights, inputs,
  activations and layer gradients, fp32-or-better accumulation), so a GPU
  forward pass can be compared element by element.
"""
from __future__ import annotations

import numpy as np

LAYERS = ((1, 128, 32), (9, 32, 32), (9, 32, 64), (9, 64, 32), (9, 32, 32))
MV_SCALE = 8.0
HUBER = 0.25
SEARCH = 4
IGNORE = -32768  # label mvx marking a block left out of the loss
LEAK = 1.0 / 16.0


def lrelu(z):
    return np.where(z > 0, z, LEAK * z)


def lrelu_grad(a):
    """Derivative given the activation (a > 0 exactly where z > 0)."""
    return np.where(a > 0, 1.0, LEAK)


def param_count() -> int:
    return sum(t * ci * co + co for t, ci, co in LAYERS) + 66


def split(params):
    """Parameter vector -> [(W [taps, cout, cin], b [cout])] * 5, W5 [2, 32], b5 [2]."""
    p = np.asarray(params, np.float64)
    out, o = [], 0
    for t, ci, co in LAYERS:
        w = p[o:o + t * co * ci].reshape(t, co, ci)
        o += t * co * ci
        b = p[o:o + co]
        o += co
        out.append((w, b))
    w5 = p[o:o + 64].reshape(2, 32)
    b5 = p[o + 64:o + 66]
    assert o + 66 == p.size
    return out, w5, b5


def bf16_round(x):
    """Round to the nearest bfloat16 (ties to even), returned as float64."""
    f = np.asarray(x, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def features(planes):
    """planes [N, 2, 4H, 4W] u8 -> x0 [N, H, W, 128] (float64, unrounded)."""
    planes = np.asarray(planes)
    n, _, dh, dw = planes.shape
    h, w = dh // 4, dw // 4
    A = planes[:, 0].astype(np.int64)
    B = planes[:, 1].astype(np.int64)
    x0 = np.zeros((n, h, w, 128))
    blocks = lambda P: P.reshape(n, h, 4, w, 4).transpose(0, 1, 3, 2, 4).reshape(n, h, w, 16)
    ys, xs = np.arange(dh), np.arange(dw)
    c = 0
    for ddy in range(-SEARCH, SEARCH + 1):
        for ddx in range(-SEARCH, SEARCH + 1):
            Bs = B[:, np.clip(ys + ddy, 0, dh - 1)][:, :, np.clip(xs + ddx, 0, dw - 1)]
            x0[..., c] = blocks(np.abs(A - Bs)).sum(-1) / 1024.0
            c += 1
    x0[..., 81:97] = blocks(A) / 255.0
    x0[..., 97:113] = blocks(B) / 255.0
    return x0


def feat_bf16(planes):
    """The GPU's feature rounding: fp32 products, then bf16."""
    x = features(planes)
    x[..., :81] = bf16_round(x[..., :81].astype(np.float32))
    pix = np.asarray(planes)
    n, _, dh, dw = pix.shape
    h, w = dh // 4, dw // 4
    blocks = lambda P: P.reshape(n, h, 4, w, 4).transpose(0, 1, 3, 2, 4).reshape(n, h, w, 16)
    x[..., 81:97] = bf16_round(blocks(pix[:, 0]).astype(np.float32) * np.float32(1.0 / 255.0))
    x[..., 97:113] = bf16_round(blocks(pix[:, 1]).astype(np.float32) * np.float32(1.0 / 255.0))
    return x


def conv(x, w):
    """taps 9: zero-padded 3x3; taps 1: 1x1."""
    if w.shape[0] == 1:
        return x @ w[0].T
    return conv3x3(x, w)


def conv3x3(x, w):
    n, h, wd, _ = x.shape
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)))
    out = np.zeros((n, h, wd, w.shape[1]))
    for t in range(9):
        dy, dx = divmod(t, 3)
        out += xp[:, dy:dy + h, dx:dx + wd, :] @ w[t].T
    return out


def conv3x3_input_grad(g, w):
    """d conv3x3(x, w) / dx applied to g: the tap-flipped, transposed conv."""
    return conv3x3(g, np.flip(w, 0).transpose(0, 2, 1))


def conv3x3_weight_grad(g, x):
    n, h, wd, _ = x.shape
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)))
    dw = np.zeros((9, g.shape[-1], x.shape[-1]))
    for t in range(9):
        dy, dx = divmod(t, 3)
        xs = xp[:, dy:dy + h, dx:dx + wd, :].reshape(-1, x.shape[-1])
        dw[t] = g.reshape(-1, g.shape[-1]).T @ xs
    return dw


def forward(params, planes, bf16=False):
    """Returns (pred [N, H, W, 2], [x0, a0 .. a4])."""
    layers, w5, b5 = split(params)
    r = bf16_round if bf16 else (lambda v: v)
    a = [feat_bf16(planes) if bf16 else features(planes)]
    for (w, b) in layers:
        a.append(r(lrelu(conv(a[-1], r(w)) + b)))
    pred = a[-1] @ w5.T + b5
    return pred, a


def loss_and_grads(params, planes, labels, bf16=False):
    """Mean Huber loss and its gradient with respect to every parameter."""
    layers, w5, b5 = split(params)
    r = bf16_round if bf16 else (lambda v: v)
    pred, a = forward(params, planes, bf16)
    lab = np.asarray(labels)
    keep = (lab[..., :1] != IGNORE).astype(np.float64)
    t = lab.astype(np.float64) / MV_SCALE
    d = pred - t
    ad = np.abs(d)
    count = d.size
    loss = (np.where(ad < HUBER, 0.5 * d * d / HUBER, ad - 0.5 * HUBER) * keep).sum() / count
    g = np.where(ad < HUBER, d / HUBER, np.sign(d)) * keep / count
    grads = []
    top = len(layers)
    dw5 = g.reshape(-1, 2).T @ a[top].reshape(-1, 32)
    db5 = g.reshape(-1, 2).sum(0)
    gl = r((g @ w5) * lrelu_grad(a[top]))
    per_layer = []
    for l in range(top - 1, -1, -1):
        w, _ = layers[l]
        if w.shape[0] == 1:
            dw = (gl.reshape(-1, gl.shape[-1]).T @ a[l].reshape(-1, a[l].shape[-1]))[None]
        else:
            dw = conv3x3_weight_grad(gl, a[l])
        db = gl.reshape(-1, gl.shape[-1]).sum(0)
        per_layer.append((dw, db))
        if l > 0:
            gl = r(conv3x3_input_grad(gl, r(w)) * lrelu_grad(a[l]))
    for dw, db in reversed(per_layer):
        grads += [dw.reshape(-1), db]
    grads += [dw5.reshape(-1), db5]
    return loss, np.concatenate(grads)


def he_init(seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    parts = []
    for t, ci, co in LAYERS:
        parts.append(rng.normal(0.0, np.sqrt(2.0 / (t * ci)), t * co * ci))
        parts.append(np.zeros(co))
    parts.append(rng.normal(0.0, np.sqrt(1.0 / 32), 64))
    parts.append(np.zeros(2))
    return np.concatenate(parts)
