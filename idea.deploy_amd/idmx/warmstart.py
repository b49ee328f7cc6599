"""Learned warm start for the batched droplet sweep.

Two models, both fitted to equilibrated fields that the batched solver
produces for other (G, R, L) points:

  ProfileModel  the steady droplet at lattice resolution: liquid and vapour
                densities, interface width and the angle-averaged profile
                shape as least-squares polynomials in (G, 1/R); a start is
                that profile rendered on the lattice with the radius that
                gives the point its mass.
  UNet          a 4-level convolutional U-Net (about 210k parameters) run and
                trained by the HIP kernels of libidmx.so (BF16 MFMA
                convolutions with FP32 accumulation, FP32 master weights and
                Adam). It sees an S x S (S = 64) image of the whole L x L
                domain and predicts the correction to a base image: the
                profile model's start where that model has a droplet, else
                the cold-start tanh droplet.

Input channels (8, NHWC, bf16): x and y relative to the centre in units of
L, r / L, the base image sampled at the pixel centres, G (scaled), R / L,
L / 512, and a zero pad channel. The warm start is the base field (the
profile at lattice resolution, or the cold tanh) plus the net's correction
resized bilinearly to the lattice, shifted by a constant so that its mass
equals the cold start's (mass is conserved by the dynamics, so the relaxed
state is the one the cold start reaches). The sweep then writes populations
at rest under the Guo forcing (idmx_batch_force_rest): with f = w n the
start would carry the velocity F/(2n) at every interface, and the damped
modes that kick excites take nearly as many steps as a cold start.

No reference counterpart: idpy has no learned component.
"""
import ctypes as C

import numpy as np

from . import _lib
from .configs import N_G, N_L

S = 64
CIN = 8

# (name, Cin, Cout, relu)
ARCH = [
    ('c1a', 8, 16, 1), ('c1b', 16, 16, 1),
    ('c2a', 16, 32, 1), ('c2b', 32, 32, 1),
    ('c3a', 32, 64, 1), ('c3b', 64, 64, 1),
    ('c4a', 64, 64, 1), ('c4b', 64, 64, 1),
    ('d3a', 128, 32, 1), ('d3b', 32, 32, 1),
    ('d2a', 64, 16, 1), ('d2b', 16, 16, 1),
    ('d1a', 32, 16, 1), ('d1b', 16, 16, 1),
    ('head', 16, 1, 0),
]


def n_params():
    return sum(9 * ci * co + co for _, ci, co, _ in ARCH)


def input_image(G, R, L, W=2.0, base=None):
    """(S, S, CIN) float32 input channels of one sweep point and the base
    image (S, S) they contain: `base` (an S x S image) when given, else the
    cold-start tanh droplet at the pixel centres."""
    c = (np.arange(S) + 0.5) * L / S - 0.5  # lattice coordinate of pixel centres
    X, Y = np.meshgrid(c, c)
    cx = cy = L // 2
    dx, dy = (X - cx), (Y - cy)
    r = np.sqrt(dx ** 2 + dy ** 2)
    if base is None:
        base = 0.5 * (N_G + N_L) - 0.5 * (N_L - N_G) * np.tanh((r - R) / W)
    img = np.zeros((S, S, CIN), dtype=np.float32)
    img[..., 0] = dx / L
    img[..., 1] = dy / L
    img[..., 2] = r / L
    img[..., 3] = base
    img[..., 4] = G * 3.0 / 5.0 + 1.0   # G ~ -5/3 .. -1.4 -> about 0 .. 0.2
    img[..., 5] = R / L
    img[..., 6] = L / 512.0
    return img, base.astype(np.float32)


def target_image(n_eq, L):
    """Area-average of an L x L field down to S x S (L a multiple of S/2... or
    bilinear sampling at the pixel centres otherwise)."""
    if L % S == 0:
        k = L // S
        return n_eq.reshape(S, k, S, k).mean(axis=(1, 3)).astype(np.float32)
    c = (np.arange(S) + 0.5) * L / S - 0.5
    x0 = np.clip(np.floor(c).astype(int), 0, L - 1)
    x1 = np.clip(x0 + 1, 0, L - 1)
    a = (c - np.floor(c))
    rows = n_eq[x0, :] * (1 - a)[:, None] + n_eq[x1, :] * a[:, None]
    return (rows[:, x0] * (1 - a)[None, :] + rows[:, x1] * a[None, :]).astype(np.float32)


def _bf16(a):
    """float32 -> bf16 bit patterns (round to nearest even), as uint16."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return u.astype(np.uint16)


class _Buf:
    """Raw device buffer owned by a Tenet."""

    def __init__(self, tenet, nbytes):
        self.tenet = tenet
        p = C.c_void_p()
        _lib.call('idmx_alloc', tenet.handle, max(int(nbytes), 1), C.byref(p))
        self.ptr = p
        self.nbytes = int(nbytes)
        _lib.call('idmx_memset', tenet.handle, p, 0, self.nbytes)

    def h2d(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes == self.nbytes, (a.nbytes, self.nbytes)
        _lib.call('idmx_h2d', self.tenet.handle, self.ptr, a.ctypes.data, a.nbytes)

    def d2h(self, dtype, shape):
        out = np.empty(shape, dtype=dtype)
        assert out.nbytes == self.nbytes
        _lib.call('idmx_d2h', self.tenet.handle, out.ctypes.data, self.ptr, out.nbytes)
        return out

    def free(self):
        if self.ptr is not None and self.tenet.handle is not None:
            _lib.call('idmx_free', self.tenet.handle, self.ptr)
        self.ptr = None


class UNet:
    """The warm-start network on the device."""

    def __init__(self, tenet, batch, seed=0):
        self.t, self.N = tenet, int(batch)
        h = tenet.handle
        self.h = h
        rng = np.random.default_rng(seed)
        self.layers = {}
        for name, ci, co, relu in ARCH:
            K = 9 * ci
            w = (rng.standard_normal((co, K)) * np.sqrt(2.0 / K)).astype(np.float32)
            if name == 'head':
                w *= 0.1
            L = {'ci': ci, 'co': co, 'relu': relu,
                 'w': _Buf(tenet, w.nbytes), 'wb': _Buf(tenet, w.size * 2),
                 'g': _Buf(tenet, w.nbytes), 'm': _Buf(tenet, w.nbytes), 'v': _Buf(tenet, w.nbytes),
                 'b': _Buf(tenet, co * 4), 'gb': _Buf(tenet, co * 4),
                 'mb': _Buf(tenet, co * 4), 'vb': _Buf(tenet, co * 4),
                 'wd': _Buf(tenet, w.size * 2)}
            L['w'].h2d(w)
            L['wb'].h2d(_bf16(w))
            self.layers[name] = L
        self.step = 0
        self._alloc_acts()

    # activation buffers (bf16) per resolution
    def _alloc_acts(self):
        N, t = self.N, self.t
        sz = lambda s, c: N * s * s * c * 2
        self.a = {}
        shapes = {'x': (S, CIN), 'c1a': (S, 16), 'c1b': (S, 16), 'p1': (S // 2, 16),
                  'c2a': (S // 2, 32), 'c2b': (S // 2, 32), 'p2': (S // 4, 32),
                  'c3a': (S // 4, 64), 'c3b': (S // 4, 64), 'p3': (S // 8, 64),
                  'c4a': (S // 8, 64), 'c4b': (S // 8, 64),
                  'cat3': (S // 4, 128), 'd3a': (S // 4, 32), 'd3b': (S // 4, 32),
                  'cat2': (S // 2, 64), 'd2a': (S // 2, 16), 'd2b': (S // 2, 16),
                  'cat1': (S, 32), 'd1a': (S, 16), 'd1b': (S, 16), 'head': (S, 1)}
        self.shapes = shapes
        for k, (s, c) in shapes.items():
            self.a[k] = _Buf(t, sz(s, c))
            self.a['d' + k] = _Buf(t, sz(s, c))
        self.a['tmp'] = _Buf(t, sz(S, 32))
        self.out32 = _Buf(t, N * S * S * 4)
        self.tgt = _Buf(t, N * S * S * 4)

    def free(self):
        for L in self.layers.values():
            for k, v in L.items():
                if isinstance(v, _Buf):
                    v.free()
        for v in self.a.values():
            v.free()
        self.out32.free()
        self.tgt.free()

    # -- primitives ------------------------------------------------------
    def _conv(self, name, src, dst, s, out32=None):
        L = self.layers[name]
        _lib.call('idmx_nn_conv3x3_fwd', self.h, self.a[src].ptr, L['wb'].ptr, L['b'].ptr,
                  self.a[dst].ptr, out32.ptr if out32 is not None else None,
                  self.N, s, s, L['ci'], L['co'], L['relu'])

    def _conv_bwd(self, name, src, dout_key, din_key, s, need_din=True):
        """dout_key holds dL/d(output); applies the ReLU mask in place, adds
        dL/dW, dL/db, and writes dL/d(input) into din_key."""
        L = self.layers[name]
        n = self.N * s * s * L['co']
        if L['relu']:
            _lib.call('idmx_nn_relu_bwd', self.h, self.a[name].ptr, self.a[dout_key].ptr,
                      self.a[dout_key].ptr, n)
        _lib.call('idmx_nn_conv3x3_bwd_weight', self.h, self.a[src].ptr, self.a[dout_key].ptr,
                  L['g'].ptr, L['gb'].ptr, self.N, s, s, L['ci'], L['co'])
        if need_din:
            _lib.call('idmx_nn_conv3x3_bwd_data', self.h, self.a[dout_key].ptr, L['w'].ptr,
                      L['wd'].ptr, self.a[din_key].ptr, self.N, s, s, L['ci'], L['co'])

    def _pool(self, src, dst, s, c, bwd=0):
        _lib.call('idmx_nn_pool2', self.h, self.a[src].ptr, self.a[dst].ptr, self.N, s, s, c, bwd)

    def _up(self, src, dst, s, c, cy, c0, bwd=0):
        _lib.call('idmx_nn_up2', self.h, self.a[src].ptr, self.a[dst].ptr, self.N, s, s, c, cy,
                  c0, bwd)

    def _copy(self, src, dst, P, cx, cx0, cy, cy0, c):
        _lib.call('idmx_nn_copy_channels', self.h, self.a[src].ptr, self.a[dst].ptr, P, cx, cx0,
                  cy, cy0, c)

    def _add(self, a, b, y, n):
        _lib.call('idmx_nn_add', self.h, self.a[a].ptr, self.a[b].ptr, self.a[y].ptr, n)

    # -- forward / backward -------------------------------------------------
    def set_input(self, x_bf16):
        self.a['x'].h2d(x_bf16)

    def forward(self):
        N, s = self.N, S
        self._conv('c1a', 'x', 'c1a', s)
        self._conv('c1b', 'c1a', 'c1b', s)
        self._pool('c1b', 'p1', s, 16)
        self._conv('c2a', 'p1', 'c2a', s // 2)
        self._conv('c2b', 'c2a', 'c2b', s // 2)
        self._pool('c2b', 'p2', s // 2, 32)
        self._conv('c3a', 'p2', 'c3a', s // 4)
        self._conv('c3b', 'c3a', 'c3b', s // 4)
        self._pool('c3b', 'p3', s // 4, 64)
        self._conv('c4a', 'p3', 'c4a', s // 8)
        self._conv('c4b', 'c4a', 'c4b', s // 8)
        self._up('c4b', 'cat3', s // 8, 64, 128, 0)
        self._copy('c3b', 'cat3', N * (s // 4) ** 2, 64, 0, 128, 64, 64)
        self._conv('d3a', 'cat3', 'd3a', s // 4)
        self._conv('d3b', 'd3a', 'd3b', s // 4)
        self._up('d3b', 'cat2', s // 4, 32, 64, 0)
        self._copy('c2b', 'cat2', N * (s // 2) ** 2, 32, 0, 64, 32, 32)
        self._conv('d2a', 'cat2', 'd2a', s // 2)
        self._conv('d2b', 'd2a', 'd2b', s // 2)
        self._up('d2b', 'cat1', s // 2, 16, 32, 0)
        self._copy('c1b', 'cat1', N * s * s, 16, 0, 32, 16, 16)
        self._conv('d1a', 'cat1', 'd1a', s)
        self._conv('d1b', 'd1a', 'd1b', s)
        self._conv('head', 'd1b', 'head', s, out32=self.out32)

    def loss_and_grad(self):
        loss = C.c_double()
        _lib.call('idmx_nn_mse', self.h, self.out32.ptr, self.tgt.ptr, self.a['dhead'].ptr,
                  self.N * S * S, C.byref(loss))
        return loss.value

    def backward(self):
        N, s = self.N, S
        self._conv_bwd('head', 'd1b', 'dhead', 'dd1b', s)
        self._conv_bwd('d1b', 'd1a', 'dd1b', 'dd1a', s)
        self._conv_bwd('d1a', 'cat1', 'dd1a', 'dcat1', s)
        self._up('dcat1', 'dd2b', s // 2, 16, 32, 0, bwd=1)
        self._conv_bwd('d2b', 'd2a', 'dd2b', 'dd2a', s // 2)
        self._conv_bwd('d2a', 'cat2', 'dd2a', 'dcat2', s // 2)
        self._up('dcat2', 'dd3b', s // 4, 32, 64, 0, bwd=1)
        self._conv_bwd('d3b', 'd3a', 'dd3b', 'dd3a', s // 4)
        self._conv_bwd('d3a', 'cat3', 'dd3a', 'dcat3', s // 4)
        self._up('dcat3', 'dc4b', s // 8, 64, 128, 0, bwd=1)
        self._conv_bwd('c4b', 'c4a', 'dc4b', 'dc4a', s // 8)
        self._conv_bwd('c4a', 'p3', 'dc4a', 'dp3', s // 8)
        # c3b feeds the pool and the skip into cat3
        self._pool('dp3', 'dc3b', s // 4, 64, bwd=1)
        P3 = N * (s // 4) ** 2
        self._copy('dcat3', 'tmp', P3, 128, 64, 64, 0, 64)
        self._add('dc3b', 'tmp', 'dc3b', P3 * 64)
        self._conv_bwd('c3b', 'c3a', 'dc3b', 'dc3a', s // 4)
        self._conv_bwd('c3a', 'p2', 'dc3a', 'dp2', s // 4)
        self._pool('dp2', 'dc2b', s // 2, 32, bwd=1)
        P2 = N * (s // 2) ** 2
        self._copy('dcat2', 'tmp', P2, 64, 32, 32, 0, 32)
        self._add('dc2b', 'tmp', 'dc2b', P2 * 32)
        self._conv_bwd('c2b', 'c2a', 'dc2b', 'dc2a', s // 2)
        self._conv_bwd('c2a', 'p1', 'dc2a', 'dp1', s // 2)
        self._pool('dp1', 'dc1b', s, 16, bwd=1)
        P1 = N * s * s
        self._copy('dcat1', 'tmp', P1, 32, 16, 16, 0, 16)
        self._add('dc1b', 'tmp', 'dc1b', P1 * 16)
        self._conv_bwd('c1b', 'c1a', 'dc1b', 'dc1a', s)
        self._conv_bwd('c1a', 'x', 'dc1a', 'dx', s, need_din=False)

    def adam(self, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        self.step += 1
        for L in self.layers.values():
            nw = L['co'] * 9 * L['ci']
            _lib.call('idmx_nn_adam', self.h, L['w'].ptr, L['g'].ptr, L['m'].ptr, L['v'].ptr,
                      L['wb'].ptr, nw, lr, b1, b2, eps, self.step)
            _lib.call('idmx_nn_adam', self.h, L['b'].ptr, L['gb'].ptr, L['mb'].ptr, L['vb'].ptr,
                      None, L['co'], lr, b1, b2, eps, self.step)

    # -- host helpers -------------------------------------------------------
    def get_weights(self):
        return {k: (L['w'].d2h(np.float32, (L['co'], 9 * L['ci'])),
                    L['b'].d2h(np.float32, (L['co'],))) for k, L in self.layers.items()}

    def set_weights(self, ws):
        for k, (w, b) in ws.items():
            L = self.layers[k]
            L['w'].h2d(np.asarray(w, dtype=np.float32))
            L['wb'].h2d(_bf16(w))
            L['b'].h2d(np.asarray(b, dtype=np.float32))

    def predict_images(self, x_f32):
        """x_f32: (N, S, S, CIN) -> (N, S, S) FP32 outputs."""
        self.set_input(_bf16(x_f32))
        self.forward()
        return self.out32.d2h(np.float32, (self.N, S, S))


def train(tenet, inputs, targets, epochs=200, batch=32, lr=2e-3, seed=0, log=None):
    """inputs (M, S, S, CIN) float32, targets (M, S, S) float32 residuals.
    Returns the trained UNet and the loss history (one value per epoch)."""
    if not (np.isfinite(inputs).all() and np.isfinite(targets).all()):
        raise ValueError('train: non-finite inputs or targets')
    rng = np.random.default_rng(seed)
    net = UNet(tenet, batch, seed=seed)
    M = inputs.shape[0]
    hist = []
    for ep in range(epochs):
        order = rng.permutation(M)
        tot, cnt = 0.0, 0
        for i in range(0, M - batch + 1, batch):
            idx = order[i:i + batch]
            net.set_input(_bf16(inputs[idx]))
            net.tgt.h2d(np.ascontiguousarray(targets[idx], dtype=np.float32))
            net.forward()
            tot += net.loss_and_grad()
            cnt += 1
            net.backward()
            net.adam(lr=lr * (0.5 ** (ep / max(1, epochs / 3))))
        hist.append(tot / max(cnt, 1))
        if log is not None and (ep % 10 == 0 or ep == epochs - 1):
            log('epoch %d loss %.3e' % (ep, hist[-1]))
    return net, hist


def training_set(points, fields, profile=None):
    """(inputs, targets) for train(): the converged, finite fields among
    `fields` (None entries are skipped), targets the field's S x S image minus
    the base image (the profile model's start for that point when it has
    one, else the cold tanh). Returns (X, Y, indices used)."""
    from .configs import droplet
    keep = [i for i, f in enumerate(fields) if f is not None and np.isfinite(f).all()]
    bases = [None] * len(points)
    if profile is not None:
        masses = [droplet((points[i].L, points[i].L), points[i].R, 2.0).sum() for i in keep]
        for i, b in zip(keep, profile.predict([points[i] for i in keep], masses)):
            bases[i] = b
    X, Y = [], []
    for i in keep:
        p = points[i]
        bimg = None if bases[i] is None else target_image(bases[i].reshape(p.L, p.L), p.L)
        x, base = input_image(p.G, p.R, p.L, base=bimg)
        X.append(x)
        Y.append(target_image(np.asarray(fields[i]).reshape(p.L, p.L), p.L) - base)
    return np.stack(X), np.stack(Y).astype(np.float32), keep


def predict_fields(tenet, net, points, cold_mass=None, profile=None):
    """Warm-start density fields (float64, L*L) for the sweep points. With a
    ProfileModel: its lattice-resolution start plus the net's correction
    resized to the lattice (where the model has no droplet for a point, the
    cold tanh image plus the correction, resized); without: the cold image
    plus the correction. Each field is shifted by a constant to the cold
    start's mass."""
    from .configs import droplet
    masses = [droplet((p.L, p.L), p.R, 2.0).sum() if cold_mass is None else cold_mass[i]
              for i, p in enumerate(points)]
    lat = profile.predict(points, masses) if profile is not None else [None] * len(points)
    out = []
    B = net.N
    for i in range(0, len(points), B):
        chunk = points[i:i + B]
        xs = np.zeros((B, S, S, CIN), dtype=np.float32)
        bases = []
        for j, p in enumerate(chunk):
            lj = lat[i + j]
            bimg = None if lj is None else target_image(lj.reshape(p.L, p.L), p.L)
            xs[j], base = input_image(p.G, p.R, p.L, base=bimg)
            bases.append(base)
        imgs = net.predict_images(xs)
        for j, p in enumerate(chunk):
            lj = lat[i + j]
            if profile is not None and profile.is_uniform(p, masses[i + j]):
                out.append(lj)          # already steady: no correction
                continue
            img = imgs[j] if lj is not None else bases[j] + imgs[j]
            img = np.ascontiguousarray(img, dtype=np.float32)
            dimg = _Buf(tenet, img.nbytes)
            dimg.h2d(img)
            dn = _Buf(tenet, p.L * p.L * 8)
            _lib.call('idmx_nn_resize_to_lattice', tenet.handle, dimg.ptr, S, dn.ptr, p.L, 0.0)
            n = dn.d2h(np.float64, (p.L * p.L,))
            dimg.free()
            dn.free()
            if lj is not None:
                n = n + lj
            n = n + (masses[i + j] - n.sum()) / n.size
            out.append(n)
    return out


def _centre_dist(L):
    """Distance of every site of an L x L domain from its centre (L/2, L/2),
    as the flat (y * L + x) array."""
    c = L // 2
    x = np.arange(L) - c
    return np.sqrt(x[None, :] ** 2 + x[:, None] ** 2).ravel()


class ProfileModel:
    """Lattice-resolution droplet profile learned from equilibrated fields.

    Each training field (a droplet centred on the site (L/2, L/2)) is reduced
    to its liquid and vapour densities n_l, n_g, the centre R and width W of
    n = n_g + (n_l - n_g) (1 - tanh((r - R) / W)) / 2 (a linear fit of
    atanh(1 - 2 phi) against r over the sites with 0.05 < phi < 0.95, phi the
    normalised density), and the angle average of the remainder, divided by
    n_l - n_g, in bins of s = (r - R) / W over [-S_MAX, S_MAX]. Least squares
    then fit n_l, n_g, W and every bin as polynomials in g = (G + 1.6) / 0.25
    and rho = 10 / R. A start for (G, L, mass) is that profile rendered on
    the lattice, with R chosen so that the field carries the given mass.

    The model sees the steady state at lattice resolution, which the U-Net's
    S x S image cannot (one pixel covers L / S sites, several interface
    widths at L = 256).

    Training fields that relaxed to a uniform gas (near the weak end of the G
    range a droplet of the sweep's mass evaporates) train a logistic
    classifier of droplet vs uniform on (G, mean density, 1/L); a point it
    calls uniform with probability above P_UNIFORM starts as the exactly
    uniform field, which is already steady."""

    P_UNIFORM = 0.95

    S_MAX = 6.0
    NB = 121

    def __init__(self):
        self.coef = self.bin_coef = self.cls_coef = None
        self.n_used = self.n_uniform = 0

    @staticmethod
    def _cls_feat(G, nu, L):
        g = (G + 1.6) / 0.25
        return np.array([1.0, g, nu, g * nu, nu * nu, g * g, 64.0 / L])

    @staticmethod
    def _logistic(X, y, ridge=1e-3, iters=50):
        """Logistic regression by Newton steps (IRLS) with a small ridge."""
        w = np.zeros(X.shape[1])
        for _ in range(iters):
            p = 1.0 / (1.0 + np.exp(-np.clip(X @ w, -40, 40)))
            H = X.T @ (X * (p * (1 - p))[:, None]) + ridge * np.eye(X.shape[1])
            step = np.linalg.solve(H, X.T @ (p - y) + ridge * w)
            w -= step
            if np.abs(step).max() < 1e-10:
                break
        return w

    def p_uniform(self, G, nu, L):
        """Probability that a point's steady state is a uniform gas (0 when
        the training fields held too few of either kind)."""
        if self.cls_coef is None:
            return 0.0
        return float(1.0 / (1.0 + np.exp(-np.clip(self._cls_feat(G, nu, L) @ self.cls_coef,
                                                   -40, 40))))

    def is_uniform(self, p, mass):
        return self.p_uniform(p.G, mass / (p.L * p.L), p.L) > self.P_UNIFORM

    @staticmethod
    def _feat(G, R):
        g, rho = (G + 1.6) / 0.25, 10.0 / R
        return np.array([1.0, g, g * g, g ** 3, rho, g * rho, rho * rho])

    @classmethod
    def describe(cls, n, L):
        """(n_l, n_g, R, W, residual bins) of one field, or None when it is
        not a droplet centred on the domain's centre site."""
        n = np.asarray(n, dtype=np.float64).ravel()
        r = _centre_dist(L)
        c = L // 2
        n_l, n_g = n[c * L + c], n[0]
        if not n_l > n_g + 0.05:
            return None
        phi = (n - n_g) / (n_l - n_g)
        band = (phi > 0.05) & (phi < 0.95)
        if band.sum() < 8:
            return None
        a, b = np.polyfit(r[band], np.arctanh(1 - 2 * phi[band]), 1)
        if not a > 0:
            return None
        W, R = 1.0 / a, -b / a
        if R < 3 * W or R > L / 2 - 3 * W:
            return None
        inner, outer = r < R - 4 * W, r > R + 6 * W
        if inner.any():
            n_l = float(n[inner].mean())
        if outer.any():
            n_g = float(n[outer].mean())
        s = (r - R) / W
        rest = (n - (n_g + (n_l - n_g) * 0.5 * (1 - np.tanh(s)))) / (n_l - n_g)
        k = np.round((s + cls.S_MAX) / (2 * cls.S_MAX) * (cls.NB - 1)).astype(int)
        ok = (k >= 0) & (k < cls.NB)
        cnt = np.bincount(k[ok], minlength=cls.NB)
        tot = np.bincount(k[ok], rest[ok], minlength=cls.NB)
        bins = np.where(cnt > 0, tot / np.maximum(cnt, 1), np.nan)
        return n_l, n_g, R, W, bins

    def fit(self, points, fields):
        """points: SweepPoint-like (G, L); fields: their equilibrated
        densities (None or non-droplet entries are skipped). Returns the
        number of fields used."""
        rows, cls_x, cls_y = [], [], []
        for p, f in zip(points, fields):
            if f is None or not np.isfinite(f).all():
                continue
            f = np.asarray(f, dtype=np.float64).ravel()
            d = self.describe(f, p.L)
            if d is not None:
                rows.append((p.G,) + d)
                cls_x.append(self._cls_feat(p.G, f.mean(), p.L))
                cls_y.append(0.0)
            elif np.ptp(f) <= 1e-6 * abs(f.mean()):
                cls_x.append(self._cls_feat(p.G, f.mean(), p.L))
                cls_y.append(1.0)
        self.cls_coef = None
        self.n_uniform = int(sum(cls_y))
        if self.n_uniform >= 4 and len(cls_y) - self.n_uniform >= 4:
            self.cls_coef = self._logistic(np.array(cls_x), np.array(cls_y))
        if len(rows) < 12:
            raise ValueError('ProfileModel.fit: %d usable fields, need 12' % len(rows))
        X = np.stack([self._feat(r[0], r[3]) for r in rows])
        Y = np.array([[r[1], r[2], r[4]] for r in rows])
        self.coef = np.linalg.lstsq(X, Y, rcond=None)[0]
        B = np.stack([r[5] for r in rows])
        self.bin_coef = np.zeros((X.shape[1], self.NB))
        for j in range(self.NB):
            m = np.isfinite(B[:, j])
            if m.sum() >= 2 * X.shape[1]:
                self.bin_coef[:, j] = np.linalg.lstsq(X[m], B[m, j], rcond=None)[0]
        self.G_range = (min(r[0] for r in rows), max(r[0] for r in rows))
        self.n_used = len(rows)
        return self.n_used

    def render(self, G, L, R):
        """The learned profile at centre radius R on an L x L lattice (flat)."""
        if self.coef is None:
            raise RuntimeError('ProfileModel: fit() first')
        f = self._feat(G, R)
        n_l, n_g, W = f @ self.coef
        s = (_centre_dist(L) - R) / W
        grid = np.linspace(-self.S_MAX, self.S_MAX, self.NB)
        rest = np.interp(s, grid, f @ self.bin_coef, left=0.0, right=0.0)
        return n_g + (n_l - n_g) * (0.5 * (1 - np.tanh(s)) + rest)

    def predict(self, points, masses):
        """Start densities of `points` whose totals are `masses`: the uniform
        field where the classifier calls the point uniform; else the profile,
        its radius by Newton steps on the rendered mass, then a uniform shift
        to the exact mass. None where the mass leaves no droplet of at least
        three interface widths (the caller starts those otherwise)."""
        if self.coef is None:
            raise RuntimeError('ProfileModel: fit() first')
        out = []
        for p, m in zip(points, masses):
            if self.is_uniform(p, m):
                out.append(np.full(p.L * p.L, m / (p.L * p.L)))
                continue
            R = float(p.R)
            field = None
            for _ in range(30):
                f = self._feat(p.G, R)
                n_l, n_g, W = f @ self.coef
                if not (n_l > n_g > 0 and W > 0):
                    break
                field = self.render(p.G, p.L, R)
                dm = m - field.sum()
                R_new = R + dm / (2 * np.pi * R * (n_l - n_g))
                if not 3 * W <= R_new <= p.L / 2 - 3 * W:
                    field = None
                    break
                if abs(R_new - R) < 1e-9:
                    R = R_new
                    break
                R = R_new
            if field is not None:
                field = self.render(p.G, p.L, R)
                field = field + (m - field.sum()) / field.size
            out.append(field)
        return out
