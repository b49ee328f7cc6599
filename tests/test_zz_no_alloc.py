"""idmx_max_abs_diff_f64_noalloc (partial results in a static device array,
no allocation) against idmx_max_abs_diff_f64 (the context's scratch buffer)
through the C ABI: two calls in a row, the second on a smaller input, so a
stale partial result of the first call would show; a NaN difference; n = 0."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _noalloc(a, b, n):
    from idmx import _lib
    out = C.c_double()
    _lib.call('idmx_max_abs_diff_f64_noalloc', None, a.ptr, b.ptr, int(n), C.byref(out))
    return out.value


def test_noalloc_matches_context_reduction(tenet):
    from idmx.IdpyCode import IdpyMemory
    rng = np.random.default_rng(23)
    # first call: a large input whose maximum sits in the last partial
    # result; second call: a smaller input with a smaller maximum
    n1, n2 = 3_000_017, 70_001
    a1 = rng.standard_normal(n1)
    b1 = a1 + 1e-3 * rng.standard_normal(n1)
    b1[-5] = a1[-5] + 50.0
    a2 = rng.standard_normal(n2)
    b2 = a2 + 1e-6 * rng.standard_normal(n2)
    for a, b in ((a1, b1), (a2, b2)):
        da, db = IdpyMemory.OnDevice(a, tenet), IdpyMemory.OnDevice(b, tenet)
        tenet.Sync()
        got = _noalloc(da, db, a.size)
        want = IdpyMemory.MaxAbsDiff(da, db)
        print('n', a.size, 'noalloc', got, 'context', want, flush=True)
        assert got == want == np.abs(a - b).max()
        da.Free()
        db.Free()
    # the second call must not see the first call's partials: a third, tiny
    # input (one workgroup)
    a3 = np.array([1.0, 2.0, 3.0])
    b3 = np.array([1.5, 2.0, 2.0])
    da, db = IdpyMemory.OnDevice(a3, tenet), IdpyMemory.OnDevice(b3, tenet)
    tenet.Sync()
    assert _noalloc(da, db, 3) == 1.0
    b3[1] = np.nan
    db.H2D(b3)
    tenet.Sync()
    assert _noalloc(da, db, 3) == np.inf
    assert _noalloc(da, db, 0) == -np.inf
    da.Free()
    db.Free()


def test_noalloc_rejects_bad_arguments(tenet):
    from idmx import _lib
    out = C.c_double()
    with pytest.raises(RuntimeError):
        _lib.call('idmx_max_abs_diff_f64_noalloc', None, None, None, 10, C.byref(out))
