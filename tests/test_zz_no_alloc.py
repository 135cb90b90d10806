"""sht_kernel_call_static (allocation-free staging through a fixed device
array) against sht_kernel_call with the same host operands, through the C
ABI: bit-identical results on consecutive calls, the second smaller; sizes
above the array are rejected."""
import ctypes as C

import numpy as np
import pytest

import shtensor_amd as sht
from shtensor_amd import _lib

pytestmark = pytest.mark.gpu


def _static_call(k, a, b, c, nb):
    p = lambda x: C.c_void_p(x.ctypes.data)
    return _lib.lib().sht_kernel_call_static(C.c_void_p(k.ptr), p(a), p(b), p(c), C.c_int64(nb), C.c_void_p(None))


@pytest.mark.parametrize("expr,sa,sb,sc,dtype,alpha,beta", [
    ("ijk,kl->ijl", [16, 16, 16], [16, 16], [16, 16, 16], np.float64, 1.0, 0.0),
    ("ik,kj->ij", [33, 7], [7, 21], [33, 21], np.float64, 0.75, 1.5),
    ("jik,kmj->im", [6, 5, 8], [8, 4, 6], [5, 4], np.float32, 1.0, 1.0),
])
def test_static_call_matches_call(expr, sa, sb, sc, dtype, alpha, beta):
    print("static call", expr, np.dtype(dtype).name, alpha, beta, flush=True)
    rng = np.random.default_rng(17)
    k = sht.Kernel(expr, sa, sb, sc, alpha, beta, sht.XMM, dtype)
    for nb in (57, 9):  # second call smaller: stale staging contents must not leak
        a = rng.uniform(-1, 1, nb * int(np.prod(sa))).astype(dtype)
        b = rng.uniform(-1, 1, nb * int(np.prod(sb))).astype(dtype)
        c0 = rng.uniform(-1, 1, nb * int(np.prod(sc))).astype(dtype)
        want = c0.copy()
        k.call(a, b, want, nb)
        got = c0.copy()
        assert _static_call(k, a, b, got, nb) == _lib.OK, _lib.lib().sht_last_error().decode()
        assert np.array_equal(got.view(np.uint8), want.view(np.uint8))
    k.close()


def test_static_call_rejects_oversize():
    k = sht.Kernel("ik,kj->ij", [64, 64], [64, 64], [64, 64])
    nb = (64 << 20) // (3 * 64 * 64 * 8) + 1  # A + B + C of nb ops exceed 64 MiB
    a = np.zeros(nb * 4096)
    b = np.zeros(nb * 4096)
    c = np.zeros(nb * 4096)
    assert _static_call(k, a, b, c, nb) == _lib.ERR_INVALID
    assert "64 MiB" in _lib.lib().sht_last_error().decode() or "67108864" in _lib.lib().sht_last_error().decode()
    assert _static_call(k, a, b, c, nb - 1) == _lib.OK
    k.close()
