"""GPU parity of the block-product Kernel (Kernel<T>::call replacement) with
the oracle's XMM chain: bitwise for every expression of
test/test_kernel.cpp, both dtypes, generic and hipRTC-specialised paths."""
import os
import zlib

import numpy as np
import pytest

import shtensor_amd as sht
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

CASES = [
    ("ik, kj -> ij", [8, 5], [5, 6], [8, 6]),
    ("jik, kmj -> im", [6, 5, 8], [8, 4, 6], [5, 4]),
    ("ikl, lkmj -> mij", [6, 5, 7], [7, 5, 6, 6], [6, 6, 6]),
    ("ijml, lmk -> jki", [5, 6, 8, 9], [9, 8, 7], [6, 7, 5]),
    ("ij, jk -> ik", [2, 2], [2, 2], [2, 2]),
    ("ij, jk -> ik", [5, 5], [5, 5], [5, 5]),
    ("ij, jk -> ik", [8, 8], [8, 8], [8, 8]),
    ("ij, jk -> ik", [10, 10], [10, 10], [10, 10]),
    ("ij, jk -> ki", [4, 4], [4, 4], [4, 4]),
    ("ikl, ljk -> ij", [8, 8, 8], [8, 8, 8], [8, 8]),
    ("kil, ljk -> ij", [8, 8, 8], [8, 8, 8], [8, 8]),
    ("ikl, ljk -> ji", [5, 5, 5], [5, 5, 5], [5, 5]),
    ("ik, kj -> ij", [16, 16], [16, 16], [16, 16]),
    ("ik,kjl->ilj", [3, 4], [4, 5, 2], [3, 2, 5]),
    ("ijk,kl->ijl", [16, 16, 16], [16, 16], [16, 16, 16]),
]


def run_gpu(expr, sa, sb, sc, alpha, beta, a, b, c, nb, dtype):
    k = sht.Kernel(expr, sa, sb, sc, alpha, beta, sht.XMM, dtype)
    da, db, dc = sht.DeviceArray(a), sht.DeviceArray(b), sht.DeviceArray(c)
    k.call(da, db, dc, nb)
    sht.call("sht_device_synchronize")
    out = dc.numpy()
    k.close()
    return out


@pytest.mark.parametrize("expr,sa,sb,sc", CASES)
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("alpha,beta", [(1.0, 0.0), (0.75, 1.5)])
def test_kernel_bitwise_vs_oracle(expr, sa, sb, sc, dtype, alpha, beta):
    print("kernel", expr, sa, sb, sc, np.dtype(dtype).name, alpha, beta, flush=True)
    rng = np.random.default_rng(zlib.crc32(repr((expr, sa)).encode()))
    nb = 7
    a = rng.uniform(-1, 1, nb * int(np.prod(sa))).astype(dtype)
    b = rng.uniform(-1, 1, nb * int(np.prod(sb))).astype(dtype)
    c = rng.uniform(-1, 1, nb * int(np.prod(sc))).astype(dtype)
    want = c.copy()
    orc.kernel_xmm(expr, sa, sb, sc, alpha, beta, a, b, want, nb)
    got = run_gpu(expr, sa, sb, sc, alpha, beta, a, b, c, nb, dtype)
    assert np.array_equal(got.view(np.uint8), want.view(np.uint8))


# FP64 MFMA form (M >= 16, K >= 4): partial row/column tiles, K % 4 tails,
# K above the LDS-staged panel limit (256), permuted maps on every operand.
MFMA_CASES = [
    ("ik,kj->ij", [16, 4], [4, 16], [16, 16]),
    ("ik,kj->ij", [33, 7], [7, 21], [33, 21]),
    ("ik,kj->ij", [100, 300], [300, 5], [100, 5]),
    ("ik,kj->ij", [17, 258], [258, 40], [17, 40]),
    ("ijk,kl->ijl", [16, 16, 16], [16, 16], [16, 16, 16]),
    ("ijkl,klm->ijm", [5, 7, 9, 11], [9, 11, 13], [5, 7, 13]),
    ("kil,ljk->ij", [6, 20, 5], [5, 9, 6], [20, 9]),
    ("ik,kjl->ilj", [40, 6], [6, 5, 3], [40, 3, 5]),
    ("jik,kmj->im", [6, 24, 8], [8, 19, 6], [24, 19]),
    # two row groups x two column tiles per workgroup (contiguous and
    # permuted maps, partial tiles)
    ("ik,kj->ij", [96, 24], [24, 40], [96, 40]),
    ("kil,ljk->ij", [6, 80, 5], [5, 35, 6], [80, 35]),
]


@pytest.mark.parametrize("expr,sa,sb,sc", MFMA_CASES)
@pytest.mark.parametrize("alpha,beta", [(1.0, 0.0), (0.75, 1.5), (-1.0, -2.0)])
def test_kernel_mfma_form_bitwise(expr, sa, sb, sc, alpha, beta):
    print("kernel mfma", expr, sa, sb, sc, alpha, beta, flush=True)
    rng = np.random.default_rng(zlib.crc32(repr((expr, sa, alpha)).encode()))
    nb = 5
    a = rng.uniform(-1, 1, nb * int(np.prod(sa)))
    b = rng.uniform(-1, 1, nb * int(np.prod(sb)))
    c = rng.uniform(-1, 1, nb * int(np.prod(sc)))
    k = sht.Kernel(expr, sa, sb, sc, alpha, beta)
    assert "FP64 MFMA" in k.get_info()
    k.close()
    want = c.copy()
    orc.kernel_xmm(expr, sa, sb, sc, alpha, beta, a, b, want, nb)
    got = run_gpu(expr, sa, sb, sc, alpha, beta, a, b, c, nb, np.float64)
    assert np.array_equal(got.view(np.uint8), want.view(np.uint8))


def test_kernel_mfma_signed_zero_chain():
    # a chain that stays -0.0 (beta = -1 on C = +0, zero products of negative
    # B entries): zero padding of the K % 4 tail would turn it into +0.0
    sa, sb, sc = [16, 5], [5, 16], [16, 16]
    a = np.zeros(80)
    b = -np.ones(80)
    c = np.zeros(256)
    want = c.copy()
    orc.kernel_xmm("ik,kj->ij", sa, sb, sc, 1.0, -1.0, a, b, want, 1)
    assert np.signbit(want).all()
    got = run_gpu("ik,kj->ij", sa, sb, sc, 1.0, -1.0, a, b, c, 1, np.float64)
    assert np.array_equal(got.view(np.uint8), want.view(np.uint8))


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_kernel_host_pointers(dtype):
    # the reference's caller passes std::vector data (test/test_kernel.cpp:143-150)
    expr, sa, sb, sc = "ijk,kl->ijl", [16, 16, 16], [16, 16], [16, 16, 16]
    rng = np.random.default_rng(9)
    nb = 11
    a = rng.uniform(-1, 1, nb * 4096).astype(dtype)
    b = rng.uniform(-1, 1, nb * 256).astype(dtype)
    c = rng.uniform(-1, 1, nb * 4096).astype(dtype)
    want = c.copy()
    orc.kernel_xmm(expr, sa, sb, sc, 0.5, 1.0, a, b, want, nb)
    k = sht.Kernel(expr, sa, sb, sc, 0.5, 1.0, sht.XMM, dtype)
    k.call(a, b, c, nb)  # host memory in and out; returns when C is written
    assert np.array_equal(c.view(np.uint8), want.view(np.uint8))
    # mixed: device A, host B and C
    c2 = want.copy()
    orc.kernel_xmm(expr, sa, sb, sc, 0.5, 1.0, a, b, want, nb)
    da = sht.DeviceArray(a)
    k.call(da, b, c2, nb)
    assert np.array_equal(c2.view(np.uint8), want.view(np.uint8))
    k.close()


@pytest.mark.parametrize("expr,sa,sb,sc", CASES[:5])
def test_kernel_generic_path_bitwise(expr, sa, sb, sc, monkeypatch):
    monkeypatch.setenv("SHT_JIT", "0")
    rng = np.random.default_rng(5)
    nb = 3
    a = rng.uniform(-1, 1, nb * int(np.prod(sa)))
    b = rng.uniform(-1, 1, nb * int(np.prod(sb)))
    c = rng.uniform(-1, 1, nb * int(np.prod(sc)))
    want = c.copy()
    orc.kernel_xmm(expr, sa, sb, sc, 1.0, 1.0, a, b, want, nb)
    k = sht.Kernel(expr, sa, sb, sc, 1.0, 1.0)
    assert "generic" in k.get_info()
    da, db, dc = sht.DeviceArray(a), sht.DeviceArray(b), sht.DeviceArray(c)
    k.call(da, db, dc, nb)
    sht.call("sht_device_synchronize")
    assert np.array_equal(dc.numpy(), want)


def test_kernel_beta_identity():
    # test/test_kernel.cpp:241-285 (alpha = 0, beta = 2)
    sa, sb, sc = [5, 6, 8, 9], [9, 8, 7], [6, 7, 5]
    rng = np.random.default_rng(1)
    a = rng.uniform(-1, 1, int(np.prod(sa))).astype(np.float32)
    b = rng.uniform(-1, 1, int(np.prod(sb))).astype(np.float32)
    c = rng.uniform(-1, 1, int(np.prod(sc))).astype(np.float32)
    got = run_gpu("ijml, lmk -> jki", sa, sb, sc, 0.0, 2.0, a, b, c, 1, np.float32)
    assert np.array_equal(got, 2 * c)
    assert abs(got.sum() - 2 * c.sum()) < 1e-6 * max(1.0, abs(c.sum()))


def test_kernel_nb_ops_zero_noop_and_many_ops():
    rng = np.random.default_rng(2)
    sa, sb, sc = [16, 16, 16], [16, 16], [16, 16, 16]
    nb = 300
    a = rng.uniform(-1, 1, nb * 4096)
    b = rng.uniform(-1, 1, nb * 256)
    c = rng.uniform(-1, 1, nb * 4096)
    k = sht.Kernel("ijk,kl->ijl", sa, sb, sc, 1.0, 0.0)
    da, db, dc = sht.DeviceArray(a), sht.DeviceArray(b), sht.DeviceArray(c)
    k.call(da, db, dc, 0)
    sht.call("sht_device_synchronize")
    assert np.array_equal(dc.numpy(), c)
    k.call(da, db, dc, nb)
    sht.call("sht_device_synchronize")
    want = c.copy()
    orc.kernel_xmm("ijk,kl->ijl", sa, sb, sc, 1.0, 0.0, a, b, want, nb)
    assert np.array_equal(dc.numpy(), want)


def test_kernel_errors_and_info():
    with pytest.raises(sht.ShtError, match="Duplicate tensor indices"):
        sht.Kernel("ii,ik->ik", [2, 2], [2, 2], [2, 2])
    with pytest.raises(sht.ShtError, match="Sizes for dimensions 1 in tensor A"):
        sht.Kernel("ij,jk->ik", [2, 3], [2, 4], [2, 4])
    with pytest.raises(sht.ShtError, match="Unknown kernel type"):
        sht.Kernel("ij,jk->ik", [2, 3], [3, 4], [2, 4], kernel_type=7)
    k = sht.Kernel("jik, kmj -> im", [6, 5, 8], [8, 4, 6], [5, 4], kernel_type=sht.LAPACK)
    info = k.get_info()
    assert "Expression: jik, kmj -> im" in info
    assert "map_row: 1\nmap_col: 0,2\n" in info
    assert "scatter row: 0,6,12,18,24\n" in info


def test_diag_mfma_rate():
    """The FP64 MFMA loop probe runs and reports a rate in a plausible range
    (well above zero, not above the matrix cores' 78.6 TFLOP/s peak by more
    than clock slack)."""
    r = sht.diag_mfma_f64_rate()
    assert 5.0 < r < 90.0, r


def test_diag_step_rate():
    """The grouped kernel's step-shape probe: FP64 within the FP64 peak,
    FP32 within the FP32 peak, wider sweeps no slower than narrow ones by
    much, and out-of-range arguments rejected."""
    r64 = {ns: sht.diag_step_rate(sht.PREC_FP64, ns) for ns in (1, 4, 8)}
    r32 = {ns: sht.diag_step_rate(sht.PREC_FP32, ns) for ns in (1, 4, 8)}
    assert all(2.0 < v < 90.0 for v in r64.values()), r64
    assert all(2.0 < v < 180.0 for v in r32.values()), r32
    assert r64[8] > 0.8 * r64[1] and r32[8] > 0.8 * r32[1], (r64, r32)
    for bad in ((sht.PREC_FP64, 0), (sht.PREC_FP64, 9), (sht.PREC_BF16X3, 4)):
        with pytest.raises(sht.ShtError):
            sht.diag_step_rate(*bad)


@pytest.mark.parametrize("expr,sa,sb,sc", [CASES[0], CASES[1], CASES[2], CASES[9], CASES[14]])
def test_lapack_typed_kernel_vs_lapack_oracle(expr, sa, sb, sc):
    """A kernel created with KernelType::LAPACK (one HIP backend: the same
    kernels) against the oracle's reshape + netlib dgemm restatement of the
    reference's LAPACK variant (src/Kernel/KernelImpl.cpp:142-208): the two
    orders of summation agree to 1e-13 normwise."""
    rng = np.random.default_rng(zlib.crc32(expr.encode()))
    nb = 6
    a = rng.uniform(-1, 1, nb * int(np.prod(sa)))
    b = rng.uniform(-1, 1, nb * int(np.prod(sb)))
    c = rng.uniform(-1, 1, nb * int(np.prod(sc)))
    want = c.copy()
    orc.kernel_lapack(expr, sa, sb, sc, 0.75, 1.5, a, b, want, nb)
    k = sht.Kernel(expr, sa, sb, sc, 0.75, 1.5, sht.LAPACK, np.float64)
    got = c.copy()
    k.call(a, b, got, nb)
    k.close()
    assert np.linalg.norm(got - want) <= 1e-13 * np.linalg.norm(want)
