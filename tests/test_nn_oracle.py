"""Pins of the network oracle (oracle/nn_ref.py) by finite differences and
by a direct loop convolution, on the CPU."""
import numpy as np

from oracle import nn_ref


def _direct_conv(x, w):
    N, H, W, C = x.shape
    Co = w.shape[0]
    out = np.zeros((N, H, W, Co))
    for n in range(N):
        for y in range(H):
            for xx in range(W):
                for co in range(Co):
                    s = 0.0
                    for ky in range(3):
                        for kx in range(3):
                            yy, xs = y + ky - 1, xx + kx - 1
                            if 0 <= yy < H and 0 <= xs < W:
                                for ci in range(C):
                                    s += w[co, (3 * ky + kx) * C + ci] * x[n, yy, xs, ci]
                    out[n, y, xx, co] = s
    return out


def test_conv_matches_direct_loops():
    rng = np.random.default_rng(0)
    x = rng.standard_normal((2, 5, 4, 3))
    w = rng.standard_normal((2, 27))
    assert np.abs(nn_ref.conv3x3(x, w) - _direct_conv(x, w)).max() < 1e-12


def test_conv_gradients_by_finite_differences():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((2, 4, 5, 3))
    w = rng.standard_normal((4, 27))
    b = rng.standard_normal(4)
    g = rng.standard_normal((2, 4, 5, 4))      # dL/dout for L = sum(g * out)
    L = lambda x_, w_, b_: float((nn_ref.conv3x3(x_, w_, b_) * g).sum())
    dw, db = nn_ref.conv3x3_bwd_weight(x, g)
    dx = nn_ref.conv3x3_bwd_data(g, w, 3)
    h = 1e-6
    for idx in [(0, 0), (1, 5), (3, 26), (2, 13)]:
        wp = w.copy(); wp[idx] += h
        wm = w.copy(); wm[idx] -= h
        assert abs((L(x, wp, b) - L(x, wm, b)) / (2 * h) - dw[idx]) < 1e-6
    for idx in [(0, 0, 0, 0), (1, 3, 4, 2), (0, 2, 1, 1)]:
        xp = x.copy(); xp[idx] += h
        xm = x.copy(); xm[idx] -= h
        assert abs((L(xp, w, b) - L(xm, w, b)) / (2 * h) - dx[idx]) < 1e-6
    bp = b.copy(); bp[1] += h
    bm = b.copy(); bm[1] -= h
    assert abs((L(x, w, bp) - L(x, w, bm)) / (2 * h) - db[1]) < 1e-6


def test_pool_up_adjoints():
    """<pool(x), y> = <x, pool_bwd(y)> and <up(x), y> = <x, up_bwd(y)>."""
    rng = np.random.default_rng(2)
    x = rng.standard_normal((2, 6, 4, 3))
    y = rng.standard_normal((2, 3, 2, 3))
    assert abs((nn_ref.pool2(x) * y).sum() - (x * nn_ref.pool2_bwd(y)).sum()) < 1e-12
    z = rng.standard_normal((2, 12, 8, 3))
    assert abs((nn_ref.up2(x) * z).sum() - (x * nn_ref.up2_bwd(z)).sum()) < 1e-12


def test_bf16_rounding():
    v = np.array([1.0, 1.00390625, 1.005859375, -3.14159], dtype=np.float32)
    r = nn_ref.bf16_round(v)
    assert r[0] == 1.0 and r[1] == 1.0  # tie to even
    assert r[2] == np.float32(1.0078125)
    assert abs(r[3] + 3.140625) < 1e-7
