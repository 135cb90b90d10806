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


def dense_pair(ctx, nk, bk=512, bi=512, ni=4, seed=0):
    """Dense ik,kj->ij operands with ni x nk blocks of bi x bk (A) and nk x ni
    blocks (B), uniform data: a request whose FP64_EXACT execution runs for
    tens of milliseconds."""
    A = sht.Tensor(ctx, [[bi] * ni, [bk] * nk])
    B = sht.Tensor(ctx, [[bk] * nk, [bi] * ni])
    A.reserve_all()
    B.reserve_all()
    A.fill_random(seed, 0)
    B.fill_random(seed, 1)
    return A, B


def test_short_requests_overlap_a_long_one():
    """Continuous batching: short requests submitted while a long one runs
    are admitted, launched on other streams and finish before it; each
    result equals the same contraction run alone."""
    ctx = sht.Context(6 << 30)
    try:
        LA, LB = dense_pair(ctx, nk=256)
        shorts = build_requests(ctx, 6, 23)
        alone, dev = [], []
        for expr, A, B, Cp, _ in shorts:
            dA, dB = to_device(ctx, A), to_device(ctx, B)
            dC = sht.Tensor(ctx, Cp.block_sizes)
            sht.contract(ctx, expr, 1.0, dA, dB, 0.0, dC, sht.PREC_FP64)
            alone.append(dC.download())
            dev.append((expr, dA, dB, Cp))
        srv = sht.Server(ctx, 5 << 30)
        LC = sht.Tensor(ctx, [[512] * 4, [512] * 4])
        long_id = srv.submit("ik,kj->ij", 1.0, LA, LB, 0.0, LC, sht.PREC_FP64_EXACT)
        import time
        time.sleep(0.03)  # the long request is planned and on the device
        outs, ids = [], []
        for expr, dA, dB, Cp in dev:
            dC = sht.Tensor(ctx, Cp.block_sizes)
            outs.append(dC)
            ids.append(srv.submit(expr, 1.0, dA, dB, 0.0, dC, sht.PREC_FP64))
        for i in ids:
            srv.wait(i)
        long_running = not srv.poll(long_id)
        srv.wait(long_id)
        st = srv.stats()
        srv.close()
        assert long_running, "the short requests waited for the long one"
        assert st["batches"] >= 2
        for data, dC in zip(alone, outs):
            assert np.array_equal(dC.download(), data)
    finally:
        ctx.close()


def test_cancel_waiting_request(ctx):
    reqs = build_requests(ctx, 3, 31)
    dev = []
    for expr, A, B, Cp, prec in reqs:
        dev.append((expr, to_device(ctx, A), to_device(ctx, B), Cp, prec))
    foot = max(8 * (a.nelem() + b.nelem()) for _, a, b, _, _ in dev)
    LA, LB = dense_pair(ctx, nk=32)
    LC = sht.Tensor(ctx, [[512] * 4, [512] * 4])
    # the budget holds the long request alone: the others wait for memory
    srv = sht.Server(ctx, 8 * (LA.nelem() + LB.nelem()) + foot // 2)
    long_id = srv.submit("ik,kj->ij", 1.0, LA, LB, 0.0, LC, sht.PREC_FP64_EXACT)
    ids = [srv.submit(e, 1.0, a, b, 0.0, sht.Tensor(ctx, cp.block_sizes), p) for e, a, b, cp, p in dev]
    srv.cancel(ids[1])
    with pytest.raises(sht.ShtError) as ei:
        srv.wait(ids[1])
    assert ei.value.code == sht._lib.ERR_CANCELLED
    srv.wait(long_id)
    srv.wait(ids[0])
    srv.wait(ids[2])
    with pytest.raises(sht.ShtError):
        srv.cancel(ids[0])  # finished requests are unknown once waited for
    assert srv.counters()["cancelled"] == 1
    assert srv.stats()["done"] == 4
    srv.close()


def test_accumulating_requests_on_one_c_keep_order(ctx):
    """Two requests accumulating into the same C (beta = 1) and a request
    reading the C another one writes run in submission order: the results
    equal the same contractions run one after the other alone."""
    rng = np.random.default_rng(41)
    A1, B1, Cp = make_case(rng, "ik", "kj", "ij", p=0.8, bmin=3, bmax=9)
    A2 = random_bst(rng, A1.block_sizes, 0.8)
    B2 = random_bst(rng, B1.block_sizes, 0.8)
    _, _, cids = orc.tasks("ik,kj->ij", A1, B1, Cp)
    C0 = orc.BlockSparse(Cp.block_sizes).reserve_ids(cids)
    C0.data[:] = rng.uniform(-1, 1, len(C0.data))
    E = random_bst(rng, [Cp.block_sizes[1], Cp.block_sizes[1]], 0.8)
    dA1, dB1, dA2, dB2, dE = (to_device(ctx, x) for x in (A1, B1, A2, B2, E))
    # alone, one after the other
    ref = to_device(ctx, C0)
    sht.contract(ctx, "ik,kj->ij", 1.0, dA1, dB1, 1.0, ref, sht.PREC_FP64)
    sht.contract(ctx, "ik,kj->ij", 0.5, dA2, dB2, 1.0, ref, sht.PREC_FP64)
    Dref = sht.Tensor(ctx, [Cp.block_sizes[0], Cp.block_sizes[1]])
    sht.contract(ctx, "ik,kj->ij", 1.0, ref, dE, 0.0, Dref, sht.PREC_FP64)
    # served, all submitted before any completes
    C = to_device(ctx, C0)
    D = sht.Tensor(ctx, [Cp.block_sizes[0], Cp.block_sizes[1]])
    srv = sht.Server(ctx, 2 << 30)
    r1 = srv.submit("ik,kj->ij", 1.0, dA1, dB1, 1.0, C, sht.PREC_FP64)
    r2 = srv.submit("ik,kj->ij", 0.5, dA2, dB2, 1.0, C, sht.PREC_FP64)
    r3 = srv.submit("ik,kj->ij", 1.0, C, dE, 0.0, D, sht.PREC_FP64)
    for r in (r1, r2, r3):
        srv.wait(r)
    srv.close()
    assert np.array_equal(C.download(), ref.download())
    assert np.array_equal(D.get_indices(), Dref.get_indices())
    assert np.array_equal(D.download(), Dref.download())


def test_plan_out_of_arena_is_requeued():
    """A request whose plan does not fit the arena while another request
    holds memory goes back to the queue and runs once that one finishes.
    The long request (A stored k-major, so its plan holds a permuted staging
    copy of A for as long as it runs) leaves too little arena for the second
    plan until it completes."""
    import time
    rng = np.random.default_rng(5)
    A, B, Cp = make_case(rng, "ijk", "kl", "ijl", p=0.9, bmin=6, bmax=12, nmin=6, nmax=8)
    ctx = sht.Context(4 << 30)
    try:
        LA = sht.Tensor(ctx, [[512] * 128, [512] * 4])  # (k, i) blocks
        LB = sht.Tensor(ctx, [[512] * 128, [512] * 4])
        for T, tid in ((LA, 0), (LB, 1)):
            T.reserve_all()
            T.fill_random(3, tid)
        LC = sht.Tensor(ctx, [[512] * 4, [512] * 4])
        dA, dB = to_device(ctx, A), to_device(ctx, B)
        # memory the long plan holds beyond its (kept) C, and what the short
        # request needs (its C + plan)
        P1 = sht.Plan(ctx, "ki,kj->ij", LA, LB, LC, sht.PREC_FP64_EXACT)
        f1 = ctx.arena_info()["free"]
        P1.close()
        hold = ctx.arena_info()["free"] - f1
        f3 = ctx.arena_info()["free"]
        Cprobe = sht.Tensor(ctx, Cp.block_sizes)
        P2 = sht.Plan(ctx, "ijk,kl->ijl", dA, dB, Cprobe, sht.PREC_FP64)
        need = f3 - ctx.arena_info()["free"]
        P2.execute()
        sht.call("sht_device_synchronize")
        want = Cprobe.download()
        P2.close()
        Cprobe.close()
        assert hold >= need, (hold, need)
        filler = ctx.allocate(ctx.arena_info()["free"] - (hold + need // 2))
        srv = sht.Server(ctx, 64 << 30)
        long_id = srv.submit("ki,kj->ij", 1.0, LA, LB, 0.0, LC, sht.PREC_FP64_EXACT)
        time.sleep(0.03)  # the long request holds its plan and runs
        dC = sht.Tensor(ctx, Cp.block_sizes)
        rid = srv.submit("ijk,kl->ijl", 1.0, dA, dB, 0.0, dC, sht.PREC_FP64)
        srv.wait(rid)
        srv.wait(long_id)
        requeued = srv.counters()["requeued"]
        srv.close()
        ctx.free(filler)
        assert np.array_equal(dC.download(), want)
        assert requeued >= 1
    finally:
        ctx.close()


@pytest.mark.parametrize("prec", [sht.PREC_FP64, sht.PREC_FP32])
@pytest.mark.parametrize("beta", [0.0, 0.5])
def test_a_stationary_requests_merged(ctx, prec, beta):
    """Requests whose plans pick the grouped A-stationary schedule (A much
    larger than B, long contracted blocks) share one launch of that kernel
    (sweep, column and unit records rebased to the arena base); every result
    equals the same contraction run alone, bit for bit, with beta = 0 (first
    sweeps start from zeros) and beta != 0 (C scaled first)."""
    rng = np.random.default_rng(41 + int(beta * 10) + int(prec))
    expr = "ijkl,klm->ijm"
    reqs = []
    for r in range(4):
        bs = {"i": rng.integers(10, 40, 3).tolist(), "j": rng.integers(2, 9, 3).tolist(),
              "k": rng.integers(8, 13, 3).tolist(), "l": rng.integers(8, 13, 2).tolist(),
              "m": rng.integers(5, 20, 2).tolist()}
        A = random_bst(rng, [bs[x] for x in "ijkl"], 0.7)
        B = random_bst(rng, [bs[x] for x in "klm"], 0.5)
        Cp = orc.BlockSparse([bs[x] for x in "ijm"])
        _, _, cids = orc.tasks(expr, A, B, Cp)
        Ck = orc.BlockSparse(Cp.block_sizes).reserve_ids(cids)
        Ck.data[:] = rng.uniform(-1, 1, len(Ck.data))
        reqs.append((A, B, Ck))
    alone, dev = [], []
    for A, B, Ck in reqs:
        dA, dB, dC = to_device(ctx, A), to_device(ctx, B), to_device(ctx, Ck)
        plan = sht.Plan(ctx, expr, dA, dB, dC, prec)
        assert plan.get_schedule() == 2  # the default picks the grouped A-stationary form
        plan.execute(0.75, beta)
        sht.call("sht_device_synchronize")
        plan.close()
        alone.append(dC.download())
        dev.append((dA, dB, Ck))
    srv = sht.Server(ctx, 2 << 30, max_batch=16)
    outs, ids = [], []
    for dA, dB, Ck in dev:
        dC = to_device(ctx, Ck)
        outs.append(dC)
        ids.append(srv.submit(expr, 0.75, dA, dB, beta, dC, prec))
    for i in ids:
        srv.wait(i)
    st = srv.stats()
    srv.close()
    assert st["done"] == len(reqs)
    assert st["max_merged"] >= 2, st
    for ref, dC in zip(alone, outs):
        assert np.array_equal(dC.download(), ref)
