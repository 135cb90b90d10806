"""Pins the oracle's block-product kernels (XMM chain and LAPACK semantics,
src/Kernel/KernelImpl.cpp) with closed forms, exact integer arithmetic,
numpy einsum in Fortran order (the reference's intended oracle,
test/test_kernel.cpp:44-116) and hand-built rounding cases."""
import numpy as np
import pytest

from oracle import oracle as orc

# test/test_kernel.cpp:306-336 plus permuted-output cases
KERNEL_CASES = [
    ("ik, kj -> ij", [8, 5], [5, 6], [8, 6]),
    ("jik, kmj -> im", [6, 5, 8], [8, 4, 6], [5, 4]),
    ("ikl, lkmj -> mij", [6, 5, 7], [7, 5, 6, 6], [6, 6, 6]),
    ("ijml, lmk -> jki", [5, 6, 8, 9], [9, 8, 7], [6, 7, 5]),
    ("ij, jk -> ik", [2, 2], [2, 2], [2, 2]),
    ("ij, jk -> ik", [10, 10], [10, 10], [10, 10]),
    ("ij, jk -> ki", [4, 4], [4, 4], [4, 4]),
    ("ik,kjl->ilj", [3, 4], [4, 5, 2], [3, 2, 5]),
    ("kil, ljk -> ij", [8, 8, 8], [8, 8, 8], [8, 8]),
]


def einsum_f(expr, a, sa, b, sb, nb_ops):
    e = expr.replace(" ", "")
    A = a.reshape([nb_ops] + sa[::-1]).transpose([0] + list(range(len(sa), 0, -1)))
    B = b.reshape([nb_ops] + sb[::-1]).transpose([0] + list(range(len(sb), 0, -1)))
    lhs, out = e.split("->")
    ea, eb = lhs.split(",")
    res = np.einsum("Z%s,Z%s->Z%s" % (ea, eb, out), A, B)
    return np.stack([r.ravel(order="F") for r in res])


@pytest.mark.parametrize("expr,sa,sb,sc", KERNEL_CASES)
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_xmm_integer_exact(expr, sa, sb, sc, dtype):
    # |v| <= 8 integers: every partial sum is exact, so any order agrees bitwise
    rng = np.random.default_rng(7)
    nb = 3
    a = rng.integers(-8, 9, nb * int(np.prod(sa))).astype(dtype)
    b = rng.integers(-8, 9, nb * int(np.prod(sb))).astype(dtype)
    c = np.full(nb * int(np.prod(sc)), 123, dtype=dtype)
    orc.kernel_xmm(expr, sa, sb, sc, 1.0, 0.0, a, b, c, nb)
    want = einsum_f(expr, a.astype(np.float64), sa, b.astype(np.float64), sb, nb).ravel()
    assert np.array_equal(c.astype(np.float64), want)


@pytest.mark.parametrize("expr,sa,sb,sc", KERNEL_CASES)
def test_xmm_vs_lapack_vs_einsum(expr, sa, sb, sc):
    rng = np.random.default_rng(11)
    nb = 4
    a = rng.uniform(-1, 1, nb * int(np.prod(sa)))
    b = rng.uniform(-1, 1, nb * int(np.prod(sb)))
    c0 = rng.uniform(-1, 1, nb * int(np.prod(sc)))
    cx, cl = c0.copy(), c0.copy()
    orc.kernel_xmm(expr, sa, sb, sc, 0.5, 1.5, a, b, cx, nb)
    orc.kernel_lapack(expr, sa, sb, sc, 0.5, 1.5, a, b, cl, nb)
    want = 0.5 * einsum_f(expr, a, sa, b, sb, nb).ravel() + 1.5 * c0
    assert np.allclose(cx, want, rtol=1e-13, atol=1e-13)
    assert np.allclose(cl, want, rtol=1e-13, atol=1e-13)


def test_beta_identity():
    # test/test_kernel.cpp:241-285: alpha=0, beta=2 scales C exactly
    sa, sb, sc = [5, 6, 8, 9], [9, 8, 7], [6, 7, 5]
    rng = np.random.default_rng(3)
    a = rng.uniform(-1, 1, int(np.prod(sa))).astype(np.float32)
    b = rng.uniform(-1, 1, int(np.prod(sb))).astype(np.float32)
    c = rng.uniform(-1, 1, int(np.prod(sc))).astype(np.float32)
    c2 = c.copy()
    orc.kernel_xmm("ijml, lmk -> jki", sa, sb, sc, 0.0, 2.0, a, b, c2, 1)
    assert np.array_equal(c2, 2 * c)


def test_fused_multiply_add_pinned():
    # a*b = 1 + 2^-29 + 2^-60 is not a double: the fused chain keeps 2^-60,
    # netlib dgemm (separate multiply and add) loses it.
    a = np.array([1 + 2.0 ** -30])
    b = np.array([1 + 2.0 ** -30])
    cx = np.array([-1.0])
    cl = np.array([-1.0])
    orc.kernel_xmm("ik,kj->ij", [1, 1], [1, 1], [1, 1], 1.0, 1.0, a, b, cx, 1)
    orc.kernel_lapack("ik,kj->ij", [1, 1], [1, 1], [1, 1], 1.0, 1.0, a, b, cl, 1)
    assert cx[0] == 2.0 ** -29 + 2.0 ** -60
    assert cl[0] == 2.0 ** -29


def test_k_order_pinned():
    # products [1, 1e16, -1e16] summed in ascending k give 0, descending give 1
    a = np.array([1.0, 1e16, -1e16])
    b = np.ones(3)
    c = np.zeros(1)
    orc.kernel_xmm("ik,kj->ij", [1, 3], [3, 1], [1, 1], 1.0, 0.0, a, b, c, 1)
    assert c[0] == 0.0
    c = np.zeros(1)
    orc.kernel_xmm("ik,kj->ij", [1, 3], [3, 1], [1, 1], 1.0, 0.0, a[::-1].copy(), b, c, 1)
    assert c[0] == 1.0


def test_alpha_applied_to_a_before_fma():
    # alpha*A is rounded before the fma: alpha=3, a=1/3 -> 3*(1/3) rounds to 1
    a = np.array([1.0 / 3.0])
    b = np.array([1.0])
    c = np.array([-1.0])
    orc.kernel_xmm("ik,kj->ij", [1, 1], [1, 1], [1, 1], 3.0, 1.0, a, b, c, 1)
    assert c[0] == 3.0 * (1.0 / 3.0) - 1.0


def test_beta_zero_does_not_read_c():
    a = np.ones(4)
    b = np.ones(4)
    c = np.full(4, np.nan)
    orc.kernel_xmm("ik,kj->ij", [2, 2], [2, 2], [2, 2], 1.0, 0.0, a, b, c, 1)
    assert np.array_equal(c, np.full(4, 2.0))
    c = np.full(4, np.nan)
    orc.kernel_xmm("ik,kj->ij", [2, 2], [2, 2], [2, 2], 1.0, -0.0, a, b, c, 1)
    assert np.isnan(c).all()  # -0.0 has a set byte: C is read


def test_nb_ops_zero_is_noop():
    a = np.ones(4)
    b = np.ones(4)
    c = np.full(4, 5.0)
    orc.kernel_xmm("ik,kj->ij", [2, 2], [2, 2], [2, 2], 1.0, 0.0, a, b, c, 0)
    assert np.array_equal(c, np.full(4, 5.0))


def test_expression_error():
    with pytest.raises(orc.OracleError, match="Duplicate tensor indices"):
        orc.kernel_xmm("ii,ik->ik", [2, 2], [2, 2], [2, 2], 1.0, 0.0, np.ones(4), np.ones(4), np.ones(4), 1)
