"""The allocation-free X^H Y (vp_vec_dots_noalloc) against vp_vec_dots through
the C ABI: two calls in a row (the second smaller, so stale partial sums of
the first would show), identical results, and a clear refusal of a call
whose scratch need exceeds the fixed array."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _dots(fn, ctx, X, Y, n):
    from voxphoton_amd import _lib as L
    k, m = len(X), len(Y)
    xs = (ctypes.c_void_p * k)(*[a.ptr for a in X])
    ys = (ctypes.c_void_p * m)(*[a.ptr for a in Y])
    out = np.full(k * m, np.nan, dtype=np.complex128)
    rc = fn(ctx.handle, xs, k, ys, m, int(n), L.ptr(out))
    return rc, out.reshape(k, m, order='F')


def _vectors(ctx, rng, count, n):
    from voxphoton_amd import vie
    out, host = [], []
    for _ in range(count):
        v = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        d = vie.DeviceArray(ctx, v.nbytes)
        d.upload(v)
        out.append(d)
        host.append(v)
    return out, host


def test_noalloc_dots_equal_allocating_dots(vp_ctx):
    from voxphoton_amd import _lib as L
    rng = np.random.default_rng(2024)
    # first call: 6 x 5 vectors of 300001 entries; second: 3 x 2 of 70001
    for (k, m, n) in ((6, 5, 300001), (3, 2, 70001)):
        print(f'noalloc dots: k={k} m={m} n={n}', flush=True)
        X, Xh = _vectors(vp_ctx, rng, k, n)
        Y, Yh = _vectors(vp_ctx, rng, m, n)
        rc0, ref = _dots(L.lib.vp_vec_dots, vp_ctx, X, Y, n)
        rc1, got = _dots(L.lib.vp_vec_dots_noalloc, vp_ctx, X, Y, n)
        assert rc0 == 0 and rc1 == 0, L.lib.vp_last_error()
        assert np.array_equal(got, ref)          # same kernels, same summation order
        dense = np.array(Xh).conj() @ np.array(Yh).T
        assert np.abs(got - dense).max() <= 1e-12 * np.abs(dense).max()


def test_noalloc_dots_refuse_oversize(vp_ctx):
    from voxphoton_amd import _lib as L
    rng = np.random.default_rng(7)
    n = 1 << 20                                    # 1024 blocks of partial sums per pair
    X, _ = _vectors(vp_ctx, rng, 1, n)
    # 40 x 40 pairs x 1024 blocks x 16 bytes = 26 MB of partial sums > 16 MiB
    rc, _ = _dots(L.lib.vp_vec_dots_noalloc, vp_ctx, X * 40, X * 40, n)
    assert rc == L.VP_ERR_UNSUPPORTED
    assert b'exceed' in L.lib.vp_last_error()
    # the library still serves the smaller call afterwards
    rc, got = _dots(L.lib.vp_vec_dots_noalloc, vp_ctx, X, X, n)
    rc0, ref = _dots(L.lib.vp_vec_dots, vp_ctx, X, X, n)
    assert rc == 0 and rc0 == 0 and np.array_equal(got, ref)
