"""alto_update_gram_noalloc through the C ABI against alto_update_gram: equal
bit for bit, on two calls in a row (the second smaller, so stale partials of
the first would show), in every storage precision; sizes beyond the scratch
array are rejected before any launch."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _A():
    import alto_amd
    return alto_amd


def _factor(A, U, prec):
    if prec == 0:
        return A.DeviceArray.from_numpy(U)
    if prec == 1:
        return A.DeviceArray.from_numpy(U.astype(np.float32))
    return A.DeviceArray.from_numpy(A.to_bf16_bits(U))


@pytest.mark.parametrize("prec", [0, 1, 2])
@pytest.mark.parametrize("R,I1,I2", [(16, 70000, 3000), (64, 131072, 777), (100, 20000, 513), (7, 5000, 1)])
def test_noalloc_gram_equals_update_gram(prec, R, I1, I2):
    A = _A()
    L = A.lib()
    rng = np.random.default_rng(R * 7 + prec)
    for I in (I1, I2):
        U = rng.random((I, R)) - 0.4
        d = _factor(A, U, prec)
        g_ref = A.DeviceArray((R, R), np.float64)
        g_new = A.DeviceArray((R, R), np.float64)
        A.check(L.alto_update_gram(d.ptr, I, R, prec, g_ref.ptr, None), "update_gram")
        A.check(L.alto_update_gram_noalloc(d.ptr, I, R, prec, g_new.ptr, None), "update_gram_noalloc")
        A.check(L.alto_device_synchronize(), "sync")
        np.testing.assert_array_equal(g_new.numpy(), g_ref.numpy())


def test_noalloc_gram_rejects_sizes_beyond_scratch():
    A = _A()
    L = A.lib()
    R, I = 128, 40000  # 157 blocks of 128 KiB partials > 16 MiB
    d = A.DeviceArray((I, R), np.float64)
    g = A.DeviceArray((R, R), np.float64)
    st = L.alto_update_gram_noalloc(d.ptr, I, R, 0, g.ptr, None)
    assert st == -4  # ALTO_ERR_UNSUPPORTED
    assert b"exceed" in L.alto_last_error()
    assert L.alto_update_gram_noalloc(d.ptr, I, 129, 0, g.ptr, None) == -1
    assert L.alto_update_gram_noalloc(None, I, R, 0, g.ptr, None) == -1
