"""Serving loop: requests of different orders, sizes, sparsities and
precisions, submitted together, merged into shared launches; each result is
bit-identical to the same contraction run alone; admission by footprint."""
import numpy as np
import pytest

import shtensor_amd as sht
from oracle import oracle as orc
from gpu_util import make_case, random_bst, to_device

pytestmark = pytest.mark.gpu

MIX = [("ik,kj->ij", "ik", "kj", "ij"), ("ijk,kl->ijl", "ijk", "kl", "ijl"),
       ("ijkl,klm->ijm", "ijkl", "klm", "ijm"), ("kil,ljk->ij", "kil", "ljk", "ij")]


@pytest.fixture(scope="module")
def ctx():
    c = sht.Context(4 << 30)
    yield c
    c.close()


def build_requests(ctx, n, seed):
    rng = np.random.default_rng(seed)
    reqs = []
    for r in range(n):
        expr, ea, eb, ec = MIX[r % len(MIX)]
        A, B, Cp = make_case(rng, ea, eb, ec, p=float(rng.uniform(0.2, 0.8)), bmin=2, bmax=12)
        prec = [sht.PREC_FP64, sht.PREC_FP64_EXACT, sht.PREC_FP32][r % 3]
        reqs.append((expr, A, B, Cp, prec))
    return reqs


def test_mix_equals_alone(ctx):
    reqs = build_requests(ctx, 12, 5)
    alone, dev = [], []
    for expr, A, B, Cp, prec in reqs:
        dA, dB = to_device(ctx, A), to_device(ctx, B)
        dC = sht.Tensor(ctx, Cp.block_sizes)
        sht.contract(ctx, expr, 1.0, dA, dB, 0.0, dC, prec)
        alone.append((dC.get_indices(), dC.download()))
        dev.append((expr, dA, dB, Cp, prec))
    srv = sht.Server(ctx, 2 << 30, max_batch=16)
    outs = []
    ids = []
    for expr, dA, dB, Cp, prec in dev:
        dC = sht.Tensor(ctx, Cp.block_sizes)
        outs.append(dC)
        ids.append(srv.submit(expr, 1.0, dA, dB, 0.0, dC, prec))
    lat = [srv.wait(i) for i in ids]
    st = srv.stats()
    srv.close()
    assert st["done"] == len(reqs)
    assert st["max_merged"] >= 2  # requests shared launches
    assert all(x >= 0 for x in lat)
    for (ind, data), dC in zip(alone, outs):
        assert np.array_equal(dC.get_indices(), ind)
        assert np.array_equal(dC.download(), data)


def test_mix_matches_oracle(ctx):
    reqs = build_requests(ctx, 6, 9)
    srv = sht.Server(ctx, 2 << 30)
    pend = []
    for expr, A, B, Cp, prec in reqs:
        dA, dB = to_device(ctx, A), to_device(ctx, B)
        dC = sht.Tensor(ctx, Cp.block_sizes)
        pend.append((srv.submit(expr, 1.0, dA, dB, 0.0, dC, prec), expr, A, B, Cp, prec, dC))
    for rid, expr, A, B, Cp, prec, dC in pend:
        srv.wait(rid)
        _, _, cids = orc.tasks(expr, A, B, Cp)
        R = orc.BlockSparse(Cp.block_sizes).reserve_ids(cids)
        orc.contract(expr, 1.0, A, B, 0.0, R)
        got = dC.download()
        if prec == sht.PREC_FP32:
            assert np.linalg.norm(got - R.data) <= 1e-6 * np.linalg.norm(R.data)
        else:
            assert np.array_equal(got, R.data)
    srv.close()


def test_wide_blocks_merged(ctx):
    """Requests with C blocks wider than 32 columns (four-wave 64 x 64
    tiles) and narrow ones share launches; each result equals the same
    contraction run alone."""
    rng = np.random.default_rng(13)
    reqs = []
    for r in range(4):
        bs_i = [int(x) for x in rng.integers(20, 90, 3)]
        bs_k = [int(x) for x in rng.integers(3, 40, 3)]
        bs_j = [int(x) for x in rng.integers(33, 80, 3)] if r % 2 == 0 else [int(x) for x in rng.integers(4, 30, 3)]
        A = random_bst(rng, [bs_i, bs_k], 0.7)
        B = random_bst(rng, [bs_k, bs_j], 0.7)
        reqs.append((A, B, orc.BlockSparse([bs_i, bs_j]), [sht.PREC_FP64, sht.PREC_FP32][(r // 2) % 2]))
    alone, dev = [], []
    for A, B, Cp, prec in reqs:
        dA, dB = to_device(ctx, A), to_device(ctx, B)
        dC = sht.Tensor(ctx, Cp.block_sizes)
        sht.contract(ctx, "ik,kj->ij", 1.0, dA, dB, 0.0, dC, prec)
        alone.append(dC.download())
        dev.append((dA, dB, Cp, prec))
    srv = sht.Server(ctx, 2 << 30, max_batch=8)
    outs, ids = [], []
    for dA, dB, Cp, prec in dev:
        dC = sht.Tensor(ctx, Cp.block_sizes)
        outs.append(dC)
        ids.append(srv.submit("ik,kj->ij", 1.0, dA, dB, 0.0, dC, prec))
    for i in ids:
        srv.wait(i)
    srv.close()
    for data, dC in zip(alone, outs):
        assert np.array_equal(dC.download(), data)


def test_admission_rejects_oversized(ctx):
    rng = np.random.default_rng(3)
    A, B, Cp = make_case(rng, "ik", "kj", "ij", p=0.9, bmin=8, bmax=16)
    dA, dB = to_device(ctx, A), to_device(ctx, B)
    dC = sht.Tensor(ctx, Cp.block_sizes)
    srv = sht.Server(ctx, 1024)  # smaller than A alone
    with pytest.raises(sht.ShtError, match="admission budget") as ei:
        srv.submit("ik,kj->ij", 1.0, dA, dB, 0.0, dC)
    assert ei.value.code == 7
    srv.close()


def test_admission_serializes_under_tight_budget(ctx):
    reqs = build_requests(ctx, 5, 17)
    dev = []
    biggest = 0
    for expr, A, B, Cp, prec in reqs:
        dA, dB = to_device(ctx, A), to_device(ctx, B)
        biggest = max(biggest, 8 * (dA.nelem() + dB.nelem()))
        dev.append((expr, dA, dB, Cp, prec))
    srv = sht.Server(ctx, biggest)  # only about one request fits at a time
    ids = [srv.submit(e, 1.0, a, b, 0.0, sht.Tensor(ctx, cp.block_sizes), p) for e, a, b, cp, p in dev]
    for i in ids:
        srv.wait(i)
    assert srv.stats()["done"] == len(ids)
    srv.close()


def test_failed_request_reports_error(ctx):
    A = sht.Tensor(ctx, [[2, 3], [4]])
    A.reserve_all()
    B = sht.Tensor(ctx, [[3, 1], [5]])  # contracted block sizes differ from A's
    B.reserve_all()
    C = sht.Tensor(ctx, [[2, 3], [5]])
    srv = sht.Server(ctx, 1 << 30)
    rid = srv.submit("ik,kj->ij", 1.0, A, B, 0.0, C)
    with pytest.raises(sht.ShtError):
        srv.wait(rid)
    srv.close()
