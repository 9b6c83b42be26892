"""CPU pins of the network oracle (oracle/lpc_net.py): the convolution against a
brute-force loop, the adjoint identities of every linear block, and the
exact-arithmetic backward pass against central finite differences.  The
device kernels are then checked against this oracle in test_gpu_learned.py."""
import numpy as np
import pytest

from oracle import lpc_net as N


def test_bf16_rounding_known_values():
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: to even -> 1; 1 + 3*2^-8 -> 1 + 2^-6
    x = np.array([1.0, 1 + 2**-8, 1 + 3 * 2**-8, -3.140625, 65504.0, 1e-3])
    r = N.bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1 + 2**-6
    assert r[3] == -3.140625  # exactly representable
    assert abs(r[5] - 1e-3) / 1e-3 < 2**-8


def test_conv_matches_brute_force():
    rng = np.random.default_rng(1)
    B, nz, ny, nx, cin, cout = 2, 3, 4, 5, 2, 3
    x = rng.standard_normal((B, nz, ny, nx, cin))
    W = rng.standard_normal((27 * cin, cout))
    b = rng.standard_normal(cout)
    y = N.conv3(x, W, cin, cout, b)
    ref = np.zeros_like(y)
    for bb in range(B):
        for z in range(nz):
            for yy in range(ny):
                for xx in range(nx):
                    for co in range(cout):
                        s = b[co]
                        for dz in (-1, 0, 1):
                            for dy in (-1, 0, 1):
                                for dx in (-1, 0, 1):
                                    zz, y2, x2 = z + dz, yy + dy, xx + dx
                                    if 0 <= zz < nz and 0 <= y2 < ny and 0 <= x2 < nx:
                                        t = (dz + 1) * 9 + (dy + 1) * 3 + (dx + 1)
                                        for ci in range(cin):
                                            s += W[t * cin + ci, co] * x[bb, zz, y2, x2, ci]
                        ref[bb, z, yy, xx, co] = s
    assert np.abs(y - ref).max() < 1e-12


def test_linear_blocks_are_adjoint():
    rng = np.random.default_rng(2)
    B, nz, ny, nx, cin, cout = 2, 4, 6, 8, 3, 5
    x = rng.standard_normal((B, nz, ny, nx, cin))
    g = rng.standard_normal((B, nz, ny, nx, cout))
    W = rng.standard_normal((27 * cin, cout))
    lhs = np.sum(N.conv3(x, W, cin, cout) * g)
    assert abs(lhs - np.sum(x * N.conv3_dgrad(g, W, cin, cout))) < 1e-9 * abs(lhs)
    assert abs(lhs - np.sum(W * N.conv3_wgrad(x, g, cin, cout))) < 1e-9 * abs(lhs)
    c = rng.standard_normal((B, nz // 2, ny // 2, nx // 2, cin))
    assert abs(np.sum(N.up(c) * x) - np.sum(c * N.down_sum(x))) < 1e-9
    assert abs(np.sum(N.pool(x, N._id) * c) - np.sum(x * N.up(c) * 0.125)) < 1e-5


@pytest.mark.parametrize("shape", [(1, 16, 16, 16), (2, 16, 16, 32)])
def test_backward_matches_finite_differences(shape):
    rng = np.random.default_rng(3)
    B, nz, ny, nx = shape
    P = N.init_params(seed=4, last_scale=1.0)
    x = rng.standard_normal((B, nz, ny, nx, 2))
    gout = rng.standard_normal((B, nz, ny, nx))
    c, cache = N.forward(P, x, mimic=False)
    grad = N.backward(P, cache, gout)
    sl, n = N.layer_slices()
    idx = []
    for l in range(len(N.SPEC)):
        a, b = sl[l]
        w = np.arange(a.start, a.stop)
        if l == 0:  # only the two data channels of the first layer carry gradient
            cin, cout = N.SPEC[0][0], N.SPEC[0][1]
            w = w[((w - a.start) // cout) % cin < 2]
        idx += list(rng.choice(w, 3, replace=False)) + [b.start]
    h = 1e-6
    for i in idx:
        Pp, Pm = P.copy(), P.copy()
        Pp[i] += h
        Pm[i] -= h
        fd = (np.sum(N.forward(Pp, x, mimic=False)[0] * gout) - np.sum(N.forward(Pm, x, mimic=False)[0] * gout)) / (2 * h)
        assert abs(fd - grad[i]) <= 1e-5 * max(1.0, abs(fd)), (i, fd, grad[i])


def test_loss_gradient_matches_finite_differences():
    rng = np.random.default_rng(5)
    c = rng.standard_normal((3, 8, 8, 8))
    a = rng.standard_normal((3, 8, 8, 8))
    w = np.array([0.5, 2.0, 1.0])
    proj = np.array([1, 0, 1])
    L, g = N.loss(c, a, w, proj)
    # mean-projected samples are blind to constants
    c2 = c.copy()
    c2[0] += 3.0
    assert abs(N.loss(c2, a, w, proj)[0] - L) < 1e-9
    h = 1e-6
    for (b, k, j, i) in [(0, 1, 2, 3), (1, 7, 0, 4), (2, 3, 3, 3)]:
        cp, cm = c.copy(), c.copy()
        cp[b, k, j, i] += h
        cm[b, k, j, i] -= h
        fd = (N.loss(cp, a, w, proj)[0] - N.loss(cm, a, w, proj)[0]) / (2 * h)
        assert abs(fd - g[b, k, j, i]) < 1e-6


def test_parameter_count_matches_library():
    import gale_amd
    gale_amd.lib()
    assert gale_amd.LearnedPreconditioner.num_params() == N.num_params()
