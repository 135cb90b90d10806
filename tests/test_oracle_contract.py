"""Pins the oracle's storage format, task list / C pattern rule, block-sparse
contraction values and synthetic generator."""
import itertools

import numpy as np
import pytest

from oracle import oracle as orc


def test_reserve_all_360_blocks():
    # test/test_tensor.cpp:39-73: 4x5x6x3 blocks, dims follow the block sizes
    bs = [[4, 5, 3, 4], [5, 5, 8, 9, 2], [5, 6, 3, 7, 7, 9], [4, 3, 2]]
    T = orc.BlockSparse(bs)
    grid = np.array(list(itertools.product(*[range(len(b)) for b in bs]))).T
    T.reserve(grid[:, ::-1].copy())  # any order: reserve sorts
    assert len(T.indices) == 360
    assert np.array_equal(T.indices, np.arange(360))
    vols = [int(np.prod(T.dims(i))) for i in range(360)]
    assert np.array_equal(T.offsets, np.concatenate([[0], np.cumsum(vols)[:-1]]))
    assert len(T.data) == sum(vols)
    for p in range(360):
        c = T.coords(p)
        assert T.dims(p) == [bs[d][c[d]] for d in range(4)]
    # rows: per-mode-0 count prefix (BaseTensor.cpp:243-312)
    assert T.rows.tolist() == [0, 90, 180, 270]


def test_reserve_sorts_mode0_fastest():
    T = orc.BlockSparse([[2, 3], [4, 5, 6]])
    T.reserve([[1, 0, 1], [2, 1, 0]])  # ids 1+2*2=5, 0+2*1=2, 1+0=1
    assert T.indices.tolist() == [1, 2, 5]
    assert T.offsets.tolist() == [0, 3 * 4, 3 * 4 + 2 * 5]


def test_reserve_errors():
    T = orc.BlockSparse([[2, 3], [4, 5]])
    with pytest.raises(orc.OracleError, match="duplicate"):
        T.reserve([[0, 0], [1, 1]])
    with pytest.raises(orc.OracleError, match="out of range"):
        T.reserve([[0, 2], [1, 1]])


def random_bst(rng, block_sizes, p, integer=False):
    T = orc.BlockSparse(block_sizes)
    total = int(np.prod([len(b) for b in block_sizes]))
    ids = np.nonzero(rng.random(total) < p)[0]
    T.reserve_ids(ids)
    if integer:
        T.data[:] = rng.integers(-4, 5, len(T.data))
    else:
        T.data[:] = rng.uniform(-1, 1, len(T.data))
    return T


def brute_tasks(expr, A, B, Cblk):
    """All (a, b) pairs with equal contracted block coords, by brute force."""
    e = expr.replace(" ", "")
    lhs, out = e.split("->")
    ea, eb = lhs.split(",")
    res = []
    for pa in range(len(A.indices)):
        ca = dict(zip(ea, A.coords(int(A.indices[pa]))))
        for pb in range(len(B.indices)):
            cb = dict(zip(eb, B.coords(int(B.indices[pb]))))
            if all(ca[x] == cb[x] for x in ea if x in eb):
                allc = {**ca, **cb}
                cid = sum(allc[x] * int(s) for x, s in zip(out, Cblk.bstride))
                kk = tuple(ca[x] for x in ea if x in eb)
                res.append((cid, kk[::-1], pa, pb))
    res.sort()
    return res


CONTRACT_CASES = [
    ("ijk,kl->ijl", "ijk", "kl", "ijl"),
    ("ijkl,klm->ijm", "ijkl", "klm", "ijm"),
    ("ik,kj->ij", "ik", "kj", "ij"),
    ("kil,ljk->ij", "kil", "ljk", "ij"),
    ("ik,kjl->ilj", "ik", "kjl", "ilj"),
    ("akbl,lcka->bc", "akbl", "lcka", "bc"),
]


def make_case(rng, ea, eb, ec, p=0.5, integer=False):
    letters = sorted(set(ea + eb))
    bs = {x: rng.integers(1, 5, rng.integers(2, 4)).tolist() for x in letters}
    A = random_bst(rng, [bs[x] for x in ea], p, integer)
    B = random_bst(rng, [bs[x] for x in eb], p, integer)
    Cp = orc.BlockSparse([bs[x] for x in ec])
    return A, B, Cp


@pytest.mark.parametrize("expr,ea,eb,ec", CONTRACT_CASES)
def test_tasks_match_brute_force(expr, ea, eb, ec):
    rng = np.random.default_rng(5)
    A, B, Cp = make_case(rng, ea, eb, ec)
    t, kk, cids = orc.tasks(expr, A, B, Cp)
    want = brute_tasks(expr, A, B, Cp)
    assert len(t) == len(want)
    got_c = cids[t[:, 2]]
    assert [(int(c), int(a), int(b)) for c, (a, b, _) in zip(got_c, t)] == [(w[0], w[2], w[3]) for w in want]
    assert cids.tolist() == sorted(set(w[0] for w in want))
    # invariant: task count = sum over K of nA(K) * nB(K)
    assert len(t) == len(want)


@pytest.mark.parametrize("expr,ea,eb,ec", CONTRACT_CASES)
def test_contract_integer_exact_vs_dense(expr, ea, eb, ec):
    rng = np.random.default_rng(9)
    A, B, Cp = make_case(rng, ea, eb, ec, integer=True)
    _, _, cids = orc.tasks(expr, A, B, Cp)
    Cp.reserve_ids(cids)
    orc.contract(expr, 1.0, A, B, 0.0, Cp, mode=0)
    want = np.einsum(expr.replace(" ", ""), A.dense(), B.dense())
    assert np.array_equal(Cp.dense(), want)


@pytest.mark.parametrize("expr,ea,eb,ec", CONTRACT_CASES)
def test_contract_float_chain_vs_longdouble(expr, ea, eb, ec):
    rng = np.random.default_rng(13)
    A, B, Cp = make_case(rng, ea, eb, ec)
    _, _, cids = orc.tasks(expr, A, B, Cp)
    Cp.reserve_ids(cids)
    Cp.data[:] = rng.uniform(-1, 1, len(Cp.data))
    C1 = orc.BlockSparse(Cp.block_sizes).reserve_ids(cids)
    C1.data[:] = Cp.data
    orc.contract(expr, 0.7, A, B, 1.3, Cp, mode=0)
    orc.contract(expr, 0.7, A, B, 1.3, C1, mode=1)
    assert np.allclose(Cp.data, C1.data, rtol=1e-13, atol=1e-14)


@pytest.mark.parametrize("expr,ea,eb,ec", CONTRACT_CASES)
def test_contract_chain_is_kernel_chain_per_task(expr, ea, eb, ec):
    """Float data: the contraction's cross-task rule is exactly the block
    Kernel's chain (kernel_xmm) applied task by task in key order -- the
    first task with the contraction's beta, every later one with beta = 1
    (acc = 1*C is exact, so the fma chain simply continues).  Pins the order
    of summation across blocks, which integer data and long-double
    comparisons cannot see."""
    rng = np.random.default_rng(17)
    A, B, Cp = make_case(rng, ea, eb, ec, p=0.7)
    t, _, cids = orc.tasks(expr, A, B, Cp)
    C = orc.BlockSparse(Cp.block_sizes).reserve_ids(cids)
    c0 = rng.uniform(-1, 1, len(C.data))
    C.data[:] = c0
    W = orc.BlockSparse(Cp.block_sizes).reserve_ids(cids)
    W.data[:] = c0
    alpha, beta = 0.7, -1.3
    orc.contract(expr, alpha, A, B, beta, C, mode=0)
    ends_c = np.r_[W.offsets[1:], len(W.data)]
    seen = set()
    for pa, pb, pc in t:  # tasks sorted by (C block, contracted key)
        da = A.dims(A.indices[pa])
        db = B.dims(B.indices[pb])
        dc = W.dims(W.indices[pc])
        a = A.block(pa).ravel(order="F")
        b = B.block(pb).ravel(order="F")
        cblk = W.data[W.offsets[pc]:ends_c[pc]].copy()
        orc.kernel_xmm(expr, da, db, dc, alpha, 1.0 if pc in seen else beta, a, b, cblk, 1)
        W.data[W.offsets[pc]:ends_c[pc]] = cblk
        seen.add(pc)
    assert len(seen) == len(cids)  # (every C block of the product pattern has a task)
    assert np.array_equal(C.data.view(np.uint64), W.data.view(np.uint64))


def test_kept_pattern_drops_outside_tasks():
    rng = np.random.default_rng(21)
    A, B, Cp = make_case(rng, "ik", "kj", "ij", p=0.8, integer=True)
    _, _, cids = orc.tasks("ik,kj->ij", A, B, Cp)
    keep = cids[::2]
    Ck = orc.BlockSparse(Cp.block_sizes).reserve_ids(keep)
    Ck.data[:] = 1.0
    t, _, c2 = orc.tasks("ik,kj->ij", A, B, Ck)
    assert np.array_equal(c2, keep)
    orc.contract("ik,kj->ij", 1.0, A, B, 2.0, Ck, mode=0)
    full = np.einsum("ik,kj->ij", A.dense(), B.dense())
    mask = Ck.dense() != 0
    Cd = orc.BlockSparse(Cp.block_sizes).reserve_ids(keep)
    Cd.data[:] = 1.0
    expect = full + 2.0 * Cd.dense()
    assert np.array_equal(Ck.dense()[mask], expect[mask])


def test_threaded_equals_serial():
    rng = np.random.default_rng(17)
    A, B, Cp = make_case(rng, "ijkl", "klm", "ijm", p=0.6)
    _, _, cids = orc.tasks("ijkl,klm->ijm", A, B, Cp)
    C1 = orc.BlockSparse(Cp.block_sizes).reserve_ids(cids)
    C4 = orc.BlockSparse(Cp.block_sizes).reserve_ids(cids)
    orc.contract("ijkl,klm->ijm", 1.0, A, B, 0.0, C1, nthreads=1)
    orc.contract("ijkl,klm->ijm", 1.0, A, B, 0.0, C4, nthreads=4)
    assert np.array_equal(C1.data, C4.data)


def test_subset_equals_full():
    rng = np.random.default_rng(19)
    A, B, Cp = make_case(rng, "ijk", "kl", "ijl", p=0.6)
    _, _, cids = orc.tasks("ijk,kl->ijl", A, B, Cp)
    Cf = orc.BlockSparse(Cp.block_sizes).reserve_ids(cids)
    Cs = orc.BlockSparse(Cp.block_sizes).reserve_ids(cids)
    orc.contract("ijk,kl->ijl", 1.0, A, B, 0.0, Cf)
    sub = np.arange(0, len(cids), 3)
    orc.contract("ijk,kl->ijl", 1.0, A, B, 0.0, Cs, subset=sub)
    for p in sub:
        assert np.array_equal(Cs.block(p), Cf.block(p))


def test_block_norms():
    rng = np.random.default_rng(2)
    T = random_bst(rng, [[2, 3], [4, 1, 2]], 1.0)
    T.data[:] *= 0.1  # |v| < 1: the reference's int accumulator would give 0
    fro = orc.block_norms(T, 0)
    mx = orc.block_norms(T, 1)
    for p in range(len(T.indices)):
        assert np.isclose(fro[p], np.linalg.norm(T.block(p)), rtol=1e-15)
        assert mx[p] == np.abs(T.block(p)).max()


def test_splitmix64_known_answer():
    # SplitMix64 (Steele, Lea, Flood 2014) seeded with 0: first two outputs
    assert orc.lib().orc_splitmix64(0) == 0xE220A8397B1DCDAF
    assert orc.lib().orc_splitmix64(0x9E3779B97F4A7C15) == 0x6E789E6AA1B965F4


def test_generator_properties():
    choices = [5, 7, 9, 11, 13, 15, 21, 28]
    bs = orc.gen_block_sizes(1001, 0, 600, choices)
    assert bs.sum() == 600 and set(bs[:-1].tolist()) <= set(choices) and 1 <= bs[-1] <= 28
    assert not np.array_equal(bs, orc.gen_block_sizes(1001, 1, 600, choices))
    ids = orc.gen_presence(1000, 0, [16, 16, 16], 0.3)
    assert np.all(np.diff(ids) > 0)
    assert abs(len(ids) / 4096 - 0.3) < 0.03
    T = orc.BlockSparse([[16] * 16] * 3).reserve_ids(ids)
    orc.gen_values(1000, 0, T, 0.0)
    assert T.data.min() >= -1 and T.data.max() < 1 and abs(T.data.mean()) < 0.01
    T2 = orc.BlockSparse([[16] * 16] * 3).reserve_ids(ids)
    orc.gen_values(1000, 0, T2, 0.3)
    for p in range(0, len(ids), 97):
        c = T.coords(int(ids[p]))
        f = np.exp(-0.3 * abs(c[0] - c[1]))
        assert np.allclose(T2.block(p), T.block(p) * f, rtol=4e-16, atol=0)
