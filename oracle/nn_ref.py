"""numpy oracle of the warm-start network's kernels (TEST INFRASTRUCTURE).

Restates in float64 what the HIP kernels compute (NHWC tensors, conv weights
[Cout][9*Cin] with k = (3*ky + kx)*Cin + ci, zero padding 1) so that tests can
check them on the same bf16-rounded inputs. There is no reference
counterpart (idpy has no learned component); the functions are pinned by
finite differences in tests/test_nn_oracle.py.
"""
import numpy as np


def bf16_round(a):
    """Round float32 values to bf16 precision (nearest even), as float32."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32)


def im2col(x):
    """x (N, H, W, C) -> (N*H*W, 9*C) with zero padding 1."""
    N, H, W, C = x.shape
    xp = np.zeros((N, H + 2, W + 2, C), dtype=np.float64)
    xp[:, 1:-1, 1:-1, :] = x
    cols = [xp[:, ky:ky + H, kx:kx + W, :] for ky in range(3) for kx in range(3)]
    return np.concatenate(cols, axis=3).reshape(N * H * W, 9 * C)


def conv3x3(x, w, b=None):
    """out (N, H, W, Cout) = conv3x3(x, w) + b (no activation)."""
    N, H, W, C = x.shape
    out = im2col(x) @ np.asarray(w, dtype=np.float64).T
    if b is not None:
        out = out + b
    return out.reshape(N, H, W, -1)


def conv3x3_bwd_weight(x, dy):
    """dL/dw given dL/dout = dy: (Cout, 9*Cin); and dL/db."""
    N, H, W, Co = dy.shape
    d = dy.reshape(-1, Co).astype(np.float64)
    return d.T @ im2col(x), d.sum(axis=0)


def conv3x3_bwd_data(dy, w, Cin):
    """dL/dx (N, H, W, Cin) given dL/dout = dy."""
    N, H, W, Co = dy.shape
    cols = dy.reshape(-1, Co).astype(np.float64) @ np.asarray(w, dtype=np.float64)  # (P, 9*Cin)
    cols = cols.reshape(N, H, W, 9, Cin)
    dxp = np.zeros((N, H + 2, W + 2, Cin))
    for t in range(9):
        ky, kx = divmod(t, 3)
        dxp[:, ky:ky + H, kx:kx + W, :] += cols[:, :, :, t, :]
    return dxp[:, 1:-1, 1:-1, :]


def pool2(x):
    N, H, W, C = x.shape
    return x.reshape(N, H // 2, 2, W // 2, 2, C).mean(axis=(2, 4))


def pool2_bwd(dy):
    return 0.25 * np.repeat(np.repeat(dy, 2, axis=1), 2, axis=2)


def up2(x):
    return np.repeat(np.repeat(x, 2, axis=1), 2, axis=2)


def up2_bwd(dy):
    N, H2, W2, C = dy.shape
    return dy.reshape(N, H2 // 2, 2, W2 // 2, 2, C).sum(axis=(2, 4))


# --------------------------------------------------------------------------
# Whole-network restatement of idmx.warmstart.UNet (forward and backward),
# rounding to bf16 at the same points as the device: every stored
# activation and gradient tensor, and the conv weights; accumulation in
# float64 here (FP32 on the device).
ARCH = [('c1a', 8, 16, 1), ('c1b', 16, 16, 1), ('c2a', 16, 32, 1), ('c2b', 32, 32, 1),
        ('c3a', 32, 64, 1), ('c3b', 64, 64, 1), ('c4a', 64, 64, 1), ('c4b', 64, 64, 1),
        ('d3a', 128, 32, 1), ('d3b', 32, 32, 1), ('d2a', 64, 16, 1), ('d2b', 16, 16, 1),
        ('d1a', 32, 16, 1), ('d1b', 16, 16, 1), ('head', 16, 1, 0)]

R = bf16_round


def unet_forward(x, ws):
    """x (N, S, S, 8) bf16 values; ws {name: (w, b)} float32. Returns the
    FP32-like head output (N, S, S, 1) and the stored activations."""
    a = {'x': R(x)}

    def conv(name, src):
        w, b = ws[name]
        relu = dict((n, r) for n, _, _, r in ARCH)[name]
        pre = conv3x3(a[src], R(w), b)
        if name == 'head':
            a['head32'] = pre
        a[name] = R(np.maximum(pre, 0) if relu else pre)

    conv('c1a', 'x'); conv('c1b', 'c1a'); a['p1'] = R(pool2(a['c1b']))
    conv('c2a', 'p1'); conv('c2b', 'c2a'); a['p2'] = R(pool2(a['c2b']))
    conv('c3a', 'p2'); conv('c3b', 'c3a'); a['p3'] = R(pool2(a['c3b']))
    conv('c4a', 'p3'); conv('c4b', 'c4a')
    a['cat3'] = np.concatenate([up2(a['c4b']), a['c3b']], axis=3)
    conv('d3a', 'cat3'); conv('d3b', 'd3a')
    a['cat2'] = np.concatenate([up2(a['d3b']), a['c2b']], axis=3)
    conv('d2a', 'cat2'); conv('d2b', 'd2a')
    a['cat1'] = np.concatenate([up2(a['d2b']), a['c1b']], axis=3)
    conv('d1a', 'cat1'); conv('d1b', 'd1a'); conv('head', 'd1b')
    return a['head32'], a


def unet_backward(a, ws, dhead):
    """Gradients {name: (dw, db)} for dL/d(head output) = dhead."""
    grads = {}
    relu = dict((n, r) for n, _, _, r in ARCH)
    cin = dict((n, c) for n, c, _, _ in ARCH)

    def back(name, src, dout, need=True):
        d = R(dout)
        if relu[name]:
            d = np.where(a[name] > 0, d, 0.0)
        grads[name] = conv3x3_bwd_weight(a[src], d)
        return R(conv3x3_bwd_data(d, R(ws[name][0]), cin[name])) if need else None

    d = back('head', 'd1b', dhead)
    d = back('d1b', 'd1a', d)
    dcat1 = back('d1a', 'cat1', d)
    d = R(up2_bwd(dcat1[..., :16]))
    d = back('d2b', 'd2a', d)
    dcat2 = back('d2a', 'cat2', d)
    d = R(up2_bwd(dcat2[..., :32]))
    d = back('d3b', 'd3a', d)
    dcat3 = back('d3a', 'cat3', d)
    d = R(up2_bwd(dcat3[..., :64]))
    d = back('c4b', 'c4a', d)
    d = back('c4a', 'p3', d)
    d = R(R(pool2_bwd(d)) + dcat3[..., 64:])
    d = back('c3b', 'c3a', d)
    d = back('c3a', 'p2', d)
    d = R(R(pool2_bwd(d)) + dcat2[..., 32:])
    d = back('c2b', 'c2a', d)
    d = back('c2a', 'p1', d)
    d = R(R(pool2_bwd(d)) + dcat1[..., 16:])
    d = back('c1b', 'c1a', d)
    back('c1a', 'x', d, need=False)
    return grads
