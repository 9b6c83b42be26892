"""CPU oracle of the learned Schur-preconditioner network (test infrastructure
only: imported by tests/, never by the product).

A numpy restatement of the 3D convolutional encoder-decoder that the library
runs on the matrix cores (no counterpart in the reference, whose Uzawa loop
uses the diagonal Q of UzawaPreconditionerTerm.cxx:156-239 only):

  level 0: conv 4->16, conv 16->16                      (skip0)
  level 1: avgpool, conv 16->32, conv 32->32            (skip1)
  level 2: avgpool, conv 32->64, conv 64->64            (skip2)
  level 3: avgpool, conv 64->128, conv 128->64
  level 2: nearest up + skip2, conv 64->64, conv 64->32
  level 1: nearest up + skip1, conv 32->32, conv 32->16
  level 0: nearest up + skip0, conv 16->16, conv 16->1 (no ReLU)

3x3x3 convolutions with zero padding, channels last [B, z, y, x, C].
Parameters: per layer W[27*cin, cout] (row = tap*cin + ci, tap = (dz+1)*9 +
(dy+1)*3 + (dx+1)) then bias[cout], layers concatenated.

``mimic=True`` rounds to BF16 (round to nearest even) exactly where the
device rounds: the packed input, every convolution operand (weights,
activations, gradients) and output except the last, pooled and summed
tensors.  ``mimic=False`` is exact float64 arithmetic, which the finite-
difference and brute-force checks in tests/ pin.
"""
from __future__ import annotations

import numpy as np

# (cin stored, cout, level, relu, cin carrying data)
SPEC = [
    (4, 16, 0, 1, 2), (16, 16, 0, 1, 16), (16, 32, 1, 1, 16), (32, 32, 1, 1, 32), (32, 64, 2, 1, 32),
    (64, 64, 2, 1, 64), (64, 128, 3, 1, 64), (128, 64, 3, 1, 128), (64, 64, 2, 1, 64), (64, 32, 2, 1, 64),
    (32, 32, 1, 1, 32), (32, 16, 1, 1, 32), (16, 16, 0, 1, 16), (16, 1, 0, 0, 16),
]

TAPS = [(t % 3 - 1, (t // 3) % 3 - 1, t // 9 - 1) for t in range(27)]  # (dx, dy, dz)


def layer_slices():
    out, o = [], 0
    for cin, cout, *_ in SPEC:
        nw = 27 * cin * cout
        out.append((slice(o, o + nw), slice(o + nw, o + nw + cout)))
        o += nw + cout
    return out, o


def num_params():
    return layer_slices()[1]


def bf16(x):
    """Round to BF16 (nearest even), returned as float64."""
    f = np.asarray(x, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _id(x):
    return np.asarray(x, np.float64)


def conv3(x, W, cin, cout, bias=None):
    """y[b,z,y,x,co] = bias[co] + sum_tap sum_ci W[tap*cin+ci, co] x[b, z+dz, y+dy, x+dx, ci]."""
    B, nz, ny, nx, c = x.shape
    assert c == cin
    xp = np.zeros((B, nz + 2, ny + 2, nx + 2, cin))
    xp[:, 1:-1, 1:-1, 1:-1] = x
    y = np.zeros((B, nz, ny, nx, cout))
    for t, (dx, dy, dz) in enumerate(TAPS):
        y += xp[:, 1 + dz:1 + dz + nz, 1 + dy:1 + dy + ny, 1 + dx:1 + dx + nx] @ W[t * cin:(t + 1) * cin]
    if bias is not None:
        y += bias
    return y


def conv3_wgrad(x, g, cin, cout):
    """dW[tap*cin+ci, co] = sum_vox x[vox + off(tap), ci] g[vox, co]."""
    B, nz, ny, nx, _ = x.shape
    xp = np.zeros((B, nz + 2, ny + 2, nx + 2, cin))
    xp[:, 1:-1, 1:-1, 1:-1] = x
    dW = np.zeros((27 * cin, cout))
    g2 = g.reshape(-1, g.shape[-1])[:, :cout]
    for t, (dx, dy, dz) in enumerate(TAPS):
        xs = xp[:, 1 + dz:1 + dz + nz, 1 + dy:1 + dy + ny, 1 + dx:1 + dx + nx].reshape(-1, cin)
        dW[t * cin:(t + 1) * cin] = xs.T @ g2
    return dW


def conv3_dgrad(g, W, cin, cout):
    """dx[vox', ci] = sum_tap sum_co W[tap*cin+ci, co] g[vox' - off(tap), co]."""
    B, nz, ny, nx, _ = g.shape
    gp = np.zeros((B, nz + 2, ny + 2, nx + 2, cout))
    gp[:, 1:-1, 1:-1, 1:-1] = g[..., :cout]
    dx = np.zeros((B, nz, ny, nx, cin))
    for t, (ox, oy, oz) in enumerate(TAPS):
        dx += gp[:, 1 - oz:1 - oz + nz, 1 - oy:1 - oy + ny, 1 - ox:1 - ox + nx] @ W[t * cin:(t + 1) * cin].T
    return dx


def pool(x, rnd):
    """2x2x2 mean; with BF16 rounding the sum is formed in float32 as on the device."""
    B, nz, ny, nx, C = x.shape
    xr = x.reshape(B, nz // 2, 2, ny // 2, 2, nx // 2, 2, C)
    if rnd is _id:
        return xr.sum(axis=(2, 4, 6)) * 0.125
    s = xr.astype(np.float32).sum(axis=(2, 4, 6), dtype=np.float32)
    return rnd(s.astype(np.float64) * 0.125)


def up(x):
    return x.repeat(2, axis=1).repeat(2, axis=2).repeat(2, axis=3)


def down_sum(g):
    B, nz, ny, nx, C = g.shape
    return g.reshape(B, nz // 2, 2, ny // 2, 2, nx // 2, 2, C).sum(axis=(2, 4, 6))


def forward(params, in2, mimic=True):
    """in2: [B, nz, ny, nx, 2]; returns (c [B, nz, ny, nx], cache)."""
    rnd = bf16 if mimic else _id
    sl, _ = layer_slices()
    P = np.asarray(params, np.float64)
    Ws = [rnd(P[a]).reshape(27 * SPEC[l][0], SPEC[l][1]) for l, (a, _) in enumerate(sl)]
    bs = [P[b] for (_, b) in sl]
    B, nz, ny, nx, _ = in2.shape
    x4 = np.zeros((B, nz, ny, nx, 4))
    x4[..., :2] = rnd(in2)
    act = [None] * 14

    def L(l, x, last=False):
        cin, cout, _, relu, _ = SPEC[l]
        y = conv3(x, Ws[l], cin, cout, bs[l])
        if relu:
            y = np.maximum(y, 0.0)
        return y if last else rnd(y)

    act[0] = L(0, x4)
    act[1] = L(1, act[0])
    p1 = pool(act[1], rnd)
    act[2] = L(2, p1)
    act[3] = L(3, act[2])
    p2 = pool(act[3], rnd)
    act[4] = L(4, p2)
    act[5] = L(5, act[4])
    p3 = pool(act[5], rnd)
    act[6] = L(6, p3)
    act[7] = L(7, act[6])
    u2 = rnd(up(act[7]) + act[5])
    act[8] = L(8, u2)
    act[9] = L(9, act[8])
    u1 = rnd(up(act[9]) + act[3])
    act[10] = L(10, u1)
    act[11] = L(11, act[10])
    u0 = rnd(up(act[11]) + act[1])
    act[12] = L(12, u0)
    c = L(13, act[12], last=True)[..., 0]
    cache = dict(x4=x4, act=act, p1=p1, p2=p2, p3=p3, u2=u2, u1=u1, u0=u0, Ws=Ws, mimic=mimic)
    return c, cache


def backward(params, cache, gout):
    """Gradient of sum(gout * c) with respect to the parameters (flat vector)."""
    mimic = cache["mimic"]
    rnd = bf16 if mimic else _id
    sl, n = layer_slices()
    grads = np.zeros(n)
    Ws, act = cache["Ws"], cache["act"]
    mask = lambda g, a: np.where(a > 0, g, 0.0)  # noqa: E731

    def bw(l, x, g, need_dx=True):
        cin, cout = SPEC[l][0], SPEC[l][1]
        a, b = sl[l]
        grads[a] = conv3_wgrad(x, g, cin, cout).ravel()
        grads[b] = g.reshape(-1, g.shape[-1])[:, :cout].sum(axis=0)
        return rnd(conv3_dgrad(g, Ws[l], cin, cout)) if need_dx else None

    g4 = np.zeros(gout.shape + (1,))
    g4[..., 0] = rnd(gout)
    g = mask(bw(13, act[12], g4), act[12])
    gu0 = bw(12, cache["u0"], g)
    g = mask(rnd(down_sum(gu0)), act[11])
    g = mask(bw(11, act[10], g), act[10])
    gu1 = bw(10, cache["u1"], g)
    g = mask(rnd(down_sum(gu1)), act[9])
    g = mask(bw(9, act[8], g), act[8])
    gu2 = bw(8, cache["u2"], g)
    g = mask(rnd(down_sum(gu2)), act[7])
    g = mask(bw(7, act[6], g), act[6])
    gp3 = bw(6, cache["p3"], g)
    g = mask(rnd(up(gp3) * 0.125 + gu2), act[5])
    g = mask(bw(5, act[4], g), act[4])
    gp2 = bw(4, cache["p2"], g)
    g = mask(rnd(up(gp2) * 0.125 + gu1), act[3])
    g = mask(bw(3, act[2], g), act[2])
    gp1 = bw(2, cache["p1"], g)
    g = mask(rnd(up(gp1) * 0.125 + gu0), act[1])
    g = mask(bw(1, act[0], g), act[0])
    bw(0, cache["x4"], g, need_dx=False)
    return grads


def loss(c, a, w, proj):
    """L = (1/B) sum_b w_b |P_b (c_b - a_b)|^2 and dL/dc."""
    B = c.shape[0]
    d = (c - a).reshape(B, -1).astype(np.float64)
    mean = np.where(np.asarray(proj) != 0, d.mean(axis=1), 0.0)[:, None]
    d = d - mean
    w = np.asarray(w, np.float64)[:, None]
    L = float((w * d * d).sum() / B)
    return L, (2.0 * w * d / B).reshape(c.shape)


def init_params(seed=0, last_scale=0.01):
    """He-normal weights (for oracle-only tests; the device draws its own)."""
    rng = np.random.default_rng(seed)
    sl, n = layer_slices()
    P = np.zeros(n)
    for l, (cin, cout, _, _, creal) in enumerate(SPEC):
        sd = np.sqrt(2.0 / (27 * creal)) * (last_scale if l == len(SPEC) - 1 else 1.0)
        W = rng.standard_normal((27, cin, cout)) * sd
        W[:, creal:, :] = 0.0
        P[sl[l][0]] = W.ravel()
        P[sl[l][1]] = 0.01 * rng.standard_normal(cout)
    return P
