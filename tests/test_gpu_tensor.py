"""GPU tensor storage: reserve/format parity with the oracle, the
reference's test_tensor.cpp scenario, norms, filter/compress, generator."""
import itertools

import numpy as np
import pytest

import shtensor_amd as sht
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = sht.Context(2 << 30)
    yield c
    c.close()


def test_reserve_all_and_iterate(ctx):
    # test/test_tensor.cpp:39-73
    bs = [[4, 5, 3, 4], [5, 5, 8, 9, 2], [5, 6, 3, 7, 7, 9], [4, 3, 2]]
    T = sht.Tensor(ctx, bs, np.float32)
    T.reserve_all()
    assert T.get_nb_nzblocks_global() == 360
    assert T.get_occupation() == 1.0
    ref = orc.BlockSparse(bs, np.float32)
    grid = np.array(list(itertools.product(*[range(len(b)) for b in bs]))).T
    ref.reserve(grid)
    info = T.get_sparse_info()
    assert np.array_equal(info["indices"], ref.indices)
    assert np.array_equal(info["offsets"], ref.offsets)
    assert np.array_equal(info["rows"], ref.rows)
    for p, bid, blk in T.blocks():
        c = T.coords(bid)
        assert list(blk.shape) == [bs[d][c[d]] for d in range(4)]
        assert not blk.any()  # zero-filled (BaseTensor.cpp:343)


def test_reserve_subset_unsorted_and_errors(ctx):
    T = sht.Tensor(ctx, [[4, 5, 3, 4], [5, 5, 8, 9, 2], [5, 6, 3, 7, 7, 9]])
    idx = [[0, 0, 0, 1, 2, 2, 3, 3, 3], [0, 0, 1, 0, 2, 3, 2, 3, 4], [0, 1, 1, 0, 1, 2, 0, 1, 2]]
    T.reserve(idx)
    ref = orc.BlockSparse(T.block_sizes).reserve(idx)
    info = T.get_sparse_info()
    assert np.array_equal(info["indices"], ref.indices)
    assert np.array_equal(info["offsets"], ref.offsets)
    with pytest.raises(sht.ShtError, match="same size"):
        T.reserve([[0, 1], [0], [0, 1]])
    with pytest.raises(sht.ShtError, match="duplicate"):
        T.reserve([[0, 0], [0, 0], [1, 1]])
    with pytest.raises(sht.ShtError, match="out of range"):
        T.reserve([[4], [0], [0]])


def test_upload_download_and_norms(ctx):
    rng = np.random.default_rng(3)
    bs = [[3, 2, 4], [5, 1], [2, 2, 2]]
    ref = orc.BlockSparse(bs)
    ref.reserve_ids(np.nonzero(rng.random(18) < 0.6)[0])
    ref.data[:] = rng.uniform(-1, 1, len(ref.data)) * 0.3
    T = sht.Tensor(ctx, bs)
    T.reserve_ids(ref.indices)
    T.upload(ref.data)
    assert np.array_equal(T.download(), ref.data)
    assert np.allclose(T.block_norms(sht.FROBENIUS), orc.block_norms(ref, 0), rtol=1e-14, atol=0)
    assert np.array_equal(T.block_norms(sht.MAXABS), orc.block_norms(ref, 1))


def test_filter_and_compress(ctx):
    # test/test_tensor.cpp:75-120: even block ids hold 0, odd hold 1
    bs = [[4, 5, 3, 4], [5, 5, 8, 9, 2], [5, 6, 3, 7, 7, 9], [4, 3, 2]]
    T = sht.Tensor(ctx, bs, np.float32)
    T.reserve_all()
    data = T.download()
    info = T.get_sparse_info()
    for bid, off in zip(info["indices"], info["offsets"]):
        n = int(np.prod(T.block_dims(int(bid))))
        data[off:off + n] = 0.0 if bid % 2 == 0 else 1.0
    T.upload(data)
    T.filter(1e-6, sht.FROBENIUS, compress=False)
    ids = T.get_sparse_info()["indices"]
    absid = np.where(ids >= 0, ids, -(ids + 1))
    assert np.array_equal(ids < 0, absid % 2 == 0)
    kept = int(np.sum(ids >= 0))
    T.compress()
    info = T.get_sparse_info()
    assert (info["indices"] >= 0).all() and T.get_nb_nzblocks_local() == kept
    # compaction keeps every surviving block's values, offsets are re-packed
    vols = [int(np.prod(T.block_dims(int(b)))) for b in info["indices"]]
    assert np.array_equal(info["offsets"], np.concatenate([[0], np.cumsum(vols)[:-1]]))
    assert (T.download() == 1.0).all()


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_compress_keeps_block_values(ctx, dtype):
    # BaseTensor::compress (BaseTensor.cpp:253-354), fixed: every surviving
    # block keeps its own values at its re-packed offset (one gather launch)
    rng = np.random.default_rng(11)
    bs = [[3, 7, 2, 5], [4, 1, 6], [2, 9, 3]]
    T = sht.Tensor(ctx, bs, dtype)
    T.reserve_all()
    info = T.get_sparse_info()
    data = rng.uniform(-1, 1, T.nelem()).astype(dtype)
    small = rng.random(len(info["indices"])) < 0.4
    before = {}
    for q, (bid, off) in enumerate(zip(info["indices"], info["offsets"])):
        n = int(np.prod(T.block_dims(int(bid))))
        if small[q]:
            data[off:off + n] *= dtype(1e-9)
        else:
            before[int(bid)] = data[off:off + n].copy()
    T.upload(data)
    T.filter(1e-6, sht.FROBENIUS, compress=True)
    info = T.get_sparse_info()
    assert sorted(before) == list(info["indices"])
    out = T.download()
    for bid, off in zip(info["indices"], info["offsets"]):
        b = before[int(bid)]
        assert np.array_equal(out[off:off + len(b)], b)


def test_filtered_tensor_rejected_by_plan(ctx):
    # filter without compress leaves negated ids: a plan must refuse the
    # tensor before any kernel reads its block coordinates
    bs = [[4, 4], [4, 4]]
    A = sht.Tensor(ctx, bs)
    A.reserve_all()
    A.upload(np.r_[np.zeros(16), np.ones(48)])
    A.filter(1e-6, sht.FROBENIUS, compress=False)
    assert (A.get_sparse_info()["indices"] < 0).any()
    B = sht.Tensor(ctx, bs)
    B.reserve_all()
    C = sht.Tensor(ctx, bs)
    with pytest.raises(sht.ShtError, match="compress the tensor first"):
        sht.Plan(ctx, "ik,kj->ij", A, B, C)
    A.compress()
    P = sht.Plan(ctx, "ik,kj->ij", A, B, C)
    P.close()


def test_filter_small_values_fixed_norm(ctx):
    # the reference's int accumulator (BlockSpan.h:147) would drop |v| < 1 blocks
    T = sht.Tensor(ctx, [[2, 2], [2]])
    T.reserve_all()
    T.upload(np.full(T.nelem(), 0.1))
    T.filter(1e-3, sht.FROBENIUS, compress=True)
    assert T.get_nb_nzblocks_local() == 2


@pytest.mark.parametrize("lam", [0.0, 0.3])
def test_generator_matches_oracle(ctx, lam):
    bs = [orc.gen_block_sizes(1001, ax, 120, [5, 7, 9, 11, 13, 15, 21, 28]).tolist() for ax in range(3)]
    T = sht.Tensor(ctx, bs)
    T.reserve_random(1001, 0, 0.3)
    ref = orc.BlockSparse(bs)
    ref.reserve_ids(orc.gen_presence(1001, 0, [len(b) for b in bs], 0.3))
    orc.gen_values(1001, 0, ref, lam)
    T.fill_random(1001, 0, lam)
    info = T.get_sparse_info()
    assert np.array_equal(info["indices"], ref.indices)
    assert np.array_equal(T.download(), ref.data)
