"""CPU oracle of the learned Schur-preconditioner network (test infrastructure
only: imported by tests/, never by the product).

A numpy restatement of the 3D convolutional encoder-decoder that the library
runs on the matrix cores (no counterpart in the reference, whose Uzawa loop
uses the diagonal Q of UzawaPreconditionerTerm.cxx:156-239 only).  U-Net with
LEVELS grid levels and channels CHAN[level]:

  encoder level l: [avgpool,] conv C(l-1)->C(l), conv C(l)->C(l)    (skip l)
  bottom level L-1: avgpool, conv C(L-2)->C(L-1), conv C(L-1)->C(L-2)
  decoder level l: nearest up + skip l, conv C(l)->C(l), conv C(l)->C(l-1)
  (level 0 ends with conv C(0)->1, no ReLU); the input has 4 stored channels.

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

LEVELS = 5
CHAN = [16, 32, 48, 64, 96]


def enc_a(l):
    return 2 * l


def enc_b(l):
    return 2 * l + 1


def dec_a(l):
    return 2 * LEVELS + 2 * (LEVELS - 2 - l)


def dec_b(l):
    return dec_a(l) + 1


def _make_spec():
    """(cin stored, cout, level, relu, cin carrying data) per layer."""
    spec = []
    for lev in range(LEVELS):
        cin = 4 if lev == 0 else CHAN[lev - 1]
        spec.append((cin, CHAN[lev], lev, 1, 2 if lev == 0 else cin))
        cout = CHAN[LEVELS - 2] if lev == LEVELS - 1 else CHAN[lev]
        spec.append((CHAN[lev], cout, lev, 1, CHAN[lev]))
    for lev in range(LEVELS - 2, -1, -1):
        spec.append((CHAN[lev], CHAN[lev], lev, 1, CHAN[lev]))
        spec.append((CHAN[lev], CHAN[lev - 1], lev, 1, CHAN[lev]) if lev else (CHAN[0], 1, 0, 0, CHAN[0]))
    return spec


SPEC = _make_spec()
NL = len(SPEC)
GRID_MULTIPLE = 2 ** (LEVELS - 1)

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
    assert nz % GRID_MULTIPLE == 0 and ny % GRID_MULTIPLE == 0 and nx % GRID_MULTIPLE == 0
    x4 = np.zeros((B, nz, ny, nx, 4))
    x4[..., :2] = rnd(in2)
    act = [None] * NL
    pools = [None] * LEVELS
    upadd = [None] * LEVELS

    def L(l, x, last=False):
        cin, cout, _, relu, _ = SPEC[l]
        y = conv3(x, Ws[l], cin, cout, bs[l])
        if relu:
            y = np.maximum(y, 0.0)
        return y if last else rnd(y)

    for lev in range(LEVELS):
        x = x4
        if lev:
            pools[lev] = pool(act[enc_b(lev - 1)], rnd)
            x = pools[lev]
        act[enc_a(lev)] = L(enc_a(lev), x)
        act[enc_b(lev)] = L(enc_b(lev), act[enc_a(lev)])
    cur = enc_b(LEVELS - 1)
    c = None
    for lev in range(LEVELS - 2, -1, -1):
        upadd[lev] = rnd(up(act[cur]) + act[enc_b(lev)])
        act[dec_a(lev)] = L(dec_a(lev), upadd[lev])
        if lev:
            act[dec_b(lev)] = L(dec_b(lev), act[dec_a(lev)])
        else:
            c = L(dec_b(0), act[dec_a(0)], last=True)[..., 0]
        cur = dec_b(lev)
    cache = dict(x4=x4, act=act, pools=pools, upadd=upadd, Ws=Ws, mimic=mimic)
    return c, cache


def backward(params, cache, gout):
    """Gradient of sum(gout * c) with respect to the parameters (flat vector)."""
    mimic = cache["mimic"]
    rnd = bf16 if mimic else _id
    sl, n = layer_slices()
    grads = np.zeros(n)
    Ws, act, pools, upadd = cache["Ws"], cache["act"], cache["pools"], cache["upadd"]
    mask = lambda g, a: np.where(a > 0, g, 0.0)  # noqa: E731

    def bw(l, x, g, need_dx=True):
        cin, cout = SPEC[l][0], SPEC[l][1]
        a, b = sl[l]
        grads[a] = conv3_wgrad(x, g, cin, cout).ravel()
        grads[b] = g.reshape(-1, g.shape[-1])[:, :cout].sum(axis=0)
        return rnd(conv3_dgrad(g, Ws[l], cin, cout)) if need_dx else None

    g4 = np.zeros(gout.shape + (1,))
    g4[..., 0] = rnd(gout)
    gskip = [None] * LEVELS
    g = g4
    for lev in range(LEVELS - 1):
        ga = mask(bw(dec_b(lev), act[dec_a(lev)], g), act[dec_a(lev)])
        gskip[lev] = bw(dec_a(lev), upadd[lev], ga)
        src = enc_b(LEVELS - 1) if lev == LEVELS - 2 else dec_b(lev + 1)
        g = mask(rnd(down_sum(gskip[lev])), act[src])
    Lb = LEVELS - 1
    ga = mask(bw(enc_b(Lb), act[enc_a(Lb)], g), act[enc_a(Lb)])
    gpool = bw(enc_a(Lb), pools[Lb], ga)
    for lev in range(LEVELS - 2, -1, -1):
        g = mask(rnd(up(gpool) * 0.125 + gskip[lev]), act[enc_b(lev)])
        ga = mask(bw(enc_b(lev), act[enc_a(lev)], g), act[enc_a(lev)])
        gpool = bw(enc_a(lev), pools[lev] if lev else cache["x4"], ga, need_dx=lev > 0)
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
