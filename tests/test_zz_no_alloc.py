"""edh_vv_noalloc (partials in a static device array) against edh_vv
(partials in context buffers), through the C ABI: the same reduction on the
same device vectors, so the two results are bit-identical.  Two calls in a
row, the second on a smaller vector (the array keeps the first call's
partials, which the second must not read)."""
import ctypes

import numpy as np
import pytest

from edlib_amd import _lib

pytestmark = pytest.mark.gpu


def _dev(L, ctx, a):
    p = ctypes.c_void_p()
    _lib.check(L.edh_alloc(ctx, max(a.nbytes, 8), ctypes.byref(p)))
    if a.nbytes:
        _lib.check(L.edh_memcpy(ctx, p, a.ctypes.data_as(ctypes.c_void_p), a.nbytes, _lib.EDH_COPY_H2D))
    return p


def test_vv_noalloc_matches_vv():
    L = _lib.lib()
    ctx = ctypes.c_void_p()
    _lib.check(L.edh_create(ctypes.byref(ctx), 0))
    try:
        rng = np.random.default_rng(20261202)
        out = _dev(L, ctx, np.zeros(1))
        for n in (5_000_003, 1_000, 0):
            a = rng.uniform(-1, 1, n)
            b = rng.uniform(-1, 1, n)
            da, db = _dev(L, ctx, a), _dev(L, ctx, b)
            ref = ctypes.c_double()
            _lib.check(L.edh_vv(ctx, da, db, n, ctypes.byref(ref), _lib.EDH_MEM_DEVICE))
            _lib.check(L.edh_vv_noalloc(ctx, da, db, n, out))
            got = np.zeros(1)
            _lib.check(L.edh_memcpy(ctx, got.ctypes.data_as(ctypes.c_void_p), out, 8, _lib.EDH_COPY_D2H))
            _lib.check(L.edh_synchronize(ctx))
            print("n=%d vv=%.17g noalloc=%.17g" % (n, ref.value, got[0]), flush=True)
            assert got[0] == ref.value  # bit for bit: the same kernels and grid
            assert abs(got[0] - float(np.dot(a, b))) <= 1e-10 * max(1.0, np.abs(a).dot(np.abs(b)))
            for p in (da, db):
                _lib.check(L.edh_free(ctx, p))
        _lib.check(L.edh_free(ctx, out))
        # argument errors are reported, not launched
        assert L.edh_vv_noalloc(ctx, None, None, 10, None) == -1
    finally:
        L.edh_destroy(ctx)
