"""Pins of the oracle's initial-interface restatements (no GPU): numpy
closed forms of the same profiles, written from the geometry (distance to
the centre L//2, the slab's two tanh fronts), on odd and even sizes."""
import numpy as np
import pytest

from oracle import cpuref


def _pos(dims):
    g = np.arange(int(np.prod(dims)))
    p, r = [], g
    for L in dims:
        p.append(r % L)
        r = r // L
    return p


@pytest.mark.parametrize('dims', [(17, 12), (9, 10, 7), (16, 16)])
def test_radial(dims):
    n, u = cpuref.init_radial(dims, 0.1, 2.0, 3.3, 1.7)
    p = _pos(dims)
    r = np.sqrt(sum((L // 2 - x) ** 2 for L, x in zip(dims, p)).astype(float))
    want = 1.05 - 0.95 * np.tanh((r - 3.3) / 1.7)
    assert np.abs(n - want).max() < 1e-15 and not u.any()
    # the centre site sits on the liquid side, the corner on the gas side
    c = sum((L // 2) * int(np.prod(dims[:k])) for k, L in enumerate(dims))
    assert n[c] > 1.9 and n[0] < 0.2


@pytest.mark.parametrize('dims,d,full', [((20, 11), 0, True), ((20, 11), 1, False),
                                         ((8, 9, 13), 2, True), ((7, 12, 5), 1, True)])
def test_flat(dims, d, full):
    w = 6.0
    n, u = cpuref.init_flat(dims, 0.1, 2.0, w, d, full)
    x = _pos(dims)[d].astype(float)
    xm = dims[d] // 2
    prof = np.tanh(x - (xm - w / 2)) - np.tanh(x - (xm + w / 2))
    dn = 1.9 if full else -1.9
    assert np.abs(n - (1.05 + 0.5 * dn * (prof - 1.0))).max() < 1e-15 and not u.any()
    # the slab is centred on L//2 (symmetric under x -> 2 xm - x away from the wrap)
    line = n.reshape(dims[::-1]).swapaxes(len(dims) - 1 - d, -1)[(0,) * (len(dims) - 1)]
    k = np.arange(1, xm)
    k = k[2 * xm - k < dims[d]]
    assert np.allclose(line[xm - k], line[xm + k], atol=1e-14)


@pytest.mark.parametrize('dims,d', [((15, 14, 9), 2), ((11, 12, 13), 0), ((18, 13), 1)])
def test_cylinder(dims, d):
    n, u = cpuref.init_cylinder(dims, 0.1, 2.0, 3.6, d, True)
    p = _pos(dims)
    r2 = sum((dims[k] // 2 - p[k]) ** 2 for k in range(len(dims)) if k != d)
    want = 1.05 - 0.95 * np.tanh(np.sqrt(r2.astype(float)) - 3.6)
    assert np.abs(n - want).max() < 1e-15 and not u.any()


@pytest.mark.parametrize('dims,d,full', [((13, 10), 0, True), ((6, 9, 11), 2, False)])
def test_single_flat(dims, d, full):
    n, u = cpuref.init_single_flat(dims, 0.1, 2.0, d, full)
    x = _pos(dims)[d]
    dn = 1.9 if full else -1.9
    want = 0.1 + dn * (np.tanh((x - dims[d] // 2).astype(float)) + 1.0) * 0.5
    assert np.abs(n - want).max() < 1e-15 and not u.any()
