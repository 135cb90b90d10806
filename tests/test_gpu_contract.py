"""GPU parity of the block-sparse contraction with the oracle.

Task list and C block pattern: bit-exact.  Values: FP64_EXACT bitwise vs the
oracle's fma chain; FP64 (MFMA) bitwise as well (stated contract 1e-12
normwise); FP32 and three-term split-BF16 within 1e-6; integer-valued operands bitwise in every form (all
products and partial sums are exact)."""
import numpy as np
import pytest

import shtensor_amd as sht
from oracle import oracle as orc
from gpu_util import make_case, normwise, random_bst, to_device

pytestmark = pytest.mark.gpu

CASES = [
    ("ijk,kl->ijl", "ijk", "kl", "ijl"),
    ("ijkl,klm->ijm", "ijkl", "klm", "ijm"),
    ("ik,kj->ij", "ik", "kj", "ij"),
    ("kil,ljk->ij", "kil", "ljk", "ij"),
    ("ik,kjl->ilj", "ik", "kjl", "ilj"),
    ("akbl,lcka->bc", "akbl", "lcka", "bc"),
    ("ijk,km->ijm", "ijk", "km", "ijm"),
]
# suffix: none = C-stationary (schedule 0), _asN = A-stationary schedule N
# (1: one-wave units, 2: four units of a block row per workgroup, 3: narrow
# sweeps at three waves per SIMD)
PRECS = {"fp64_exact": sht.PREC_FP64_EXACT, "fp64_exact_as1": sht.PREC_FP64_EXACT, "fp64": sht.PREC_FP64,
         "fp64_as1": sht.PREC_FP64, "fp64_as2": sht.PREC_FP64, "fp64_as3": sht.PREC_FP64, "fp32": sht.PREC_FP32,
         "fp32_as1": sht.PREC_FP32, "fp32_as2": sht.PREC_FP32, "fp32_as3": sht.PREC_FP32, "bf16x3": sht.PREC_BF16X3,
         "bf16x3_as1": sht.PREC_BF16X3, "bf16x3_as2": sht.PREC_BF16X3}
TOL = {k: (0.0 if k.startswith("fp64") else 1e-6) for k in PRECS}
# v_mfma_f64_16x16x4f64 accumulates its four products as a sequential fused
# multiply-add chain in k order, so the FP64 MFMA forms equal the fma chain
# bit for bit (asserted here; the stated contract of SHT_PREC_FP64 is 1e-12).
EXACT = tuple(k for k in PRECS if k.startswith("fp64"))


def make_plan(ctx, expr, dA, dB, dC, prec):
    plan = sht.Plan(ctx, expr, dA, dB, dC, PRECS[prec])
    sched = int(prec[-1]) if "_as" in prec else 0
    plan.set_schedule(sched)
    assert plan.get_schedule() == sched
    return plan


@pytest.fixture(scope="module")
def ctx():
    c = sht.Context(4 << 30)
    yield c
    c.close()


def oracle_result(expr, A, B, Cp, alpha=1.0, beta=0.0, c0=None, mode=0):
    t, kk, cids = orc.tasks(expr, A, B, Cp)
    R = orc.BlockSparse(Cp.block_sizes).reserve_ids(cids if len(Cp.indices) == 0 else Cp.indices)
    if c0 is not None:
        R.data[:] = c0
    orc.contract(expr, alpha, A, B, beta, R, mode=mode)
    return t, kk, cids, R


@pytest.mark.parametrize("prec", list(PRECS))
@pytest.mark.parametrize("expr,ea,eb,ec", CASES)
def test_contract_vs_oracle(ctx, expr, ea, eb, ec, prec):
    print("contract", expr, prec, flush=True)
    rng = np.random.default_rng(len(expr) * 7 + len(prec))
    A, B, Cp = make_case(rng, ea, eb, ec, p=0.6, bmin=2, bmax=9)
    t, kk, cids, R = oracle_result(expr, A, B, Cp)
    dA, dB = to_device(ctx, A), to_device(ctx, B)
    dC = sht.Tensor(ctx, Cp.block_sizes)
    plan = make_plan(ctx, expr, dA, dB, dC, prec)
    got = plan.tasks()
    assert np.array_equal(got["a"], t[:, 0]) and np.array_equal(got["b"], t[:, 1])
    assert np.array_equal(got["c"], t[:, 2]) and np.array_equal(got["key"], kk)
    assert np.array_equal(dC.get_sparse_info()["indices"], cids)
    assert np.array_equal(dC.get_sparse_info()["offsets"], R.offsets)
    plan.execute(1.0, 0.0)
    sht.call("sht_device_synchronize")
    out = dC.download()
    if prec in EXACT:
        assert np.array_equal(out, R.data)
    else:
        assert normwise(out, R.data) <= TOL[prec], normwise(out, R.data)
    plan.close()


@pytest.mark.parametrize("prec", list(PRECS))
@pytest.mark.parametrize("expr,ea,eb,ec", CASES[:4])
def test_contract_integer_exact_every_form(ctx, expr, ea, eb, ec, prec):
    rng = np.random.default_rng(99)
    A, B, Cp = make_case(rng, ea, eb, ec, p=0.7, integer=True, bmin=2, bmax=7)
    _, _, _, R = oracle_result(expr, A, B, Cp)
    dA, dB = to_device(ctx, A), to_device(ctx, B)
    dC = sht.Tensor(ctx, Cp.block_sizes)
    plan = make_plan(ctx, expr, dA, dB, dC, prec)
    plan.execute(1.0, 0.0)
    sht.call("sht_device_synchronize")
    assert np.array_equal(dC.download(), R.data)
    plan.close()


@pytest.mark.parametrize("prec", ["fp64_exact", "fp64_exact_as1", "fp64", "fp64_as1", "fp64_as2", "fp64_as3"])
def test_alpha_beta_kept_pattern(ctx, prec):
    rng = np.random.default_rng(7)
    expr = "ijkl,klm->ijm"
    A, B, Cp = make_case(rng, "ijkl", "klm", "ijm", p=0.6, bmin=2, bmax=9)
    _, _, cids = orc.tasks(expr, A, B, Cp)
    keep = cids[::2]
    Ck = orc.BlockSparse(Cp.block_sizes).reserve_ids(keep)
    c0 = rng.uniform(-1, 1, len(Ck.data))
    Ck.data[:] = c0
    _, _, _, R = oracle_result(expr, A, B, Ck, alpha=0.5, beta=-1.25, c0=c0)
    dA, dB = to_device(ctx, A), to_device(ctx, B)
    dC = to_device(ctx, Ck)
    plan = make_plan(ctx, expr, dA, dB, dC, prec)
    plan.execute(0.5, -1.25)
    sht.call("sht_device_synchronize")
    plan.close()
    assert np.array_equal(dC.get_sparse_info()["indices"], keep)
    out = dC.download()
    if prec in EXACT:
        assert np.array_equal(out, R.data)
    else:
        assert normwise(out, R.data) <= 1e-12


@pytest.mark.parametrize("extra", [False, True])
@pytest.mark.parametrize("prec", ["fp64", "fp64_as1", "fp64_as2", "fp64_as3", "fp64_exact_as1", "fp32_as2",
                                  "bf16x3_as2"])
def test_beta_zero_does_not_read_c(ctx, prec, extra):
    """beta == 0: C's previous contents (NaN here) are never read.  With
    extra C blocks that no task reaches, those blocks become zero; without,
    the A-stationary schedules start every column's first sweep from zeros
    instead of a C = 0 pass."""
    print("beta0", prec, extra, flush=True)
    rng = np.random.default_rng(11)
    expr = "ijkl,klm->ijm"
    A, B, Cp = make_case(rng, "ijkl", "klm", "ijm", p=0.6, bmin=2, bmax=9)
    _, _, cids = orc.tasks(expr, A, B, Cp)
    ids = cids
    if extra:
        total = int(np.prod([len(b) for b in Cp.block_sizes]))
        free = np.setdiff1d(np.arange(total), cids)
        assert len(free) > 0
        ids = np.union1d(cids, free[:: max(1, len(free) // 5)])
    Ck = orc.BlockSparse(Cp.block_sizes).reserve_ids(ids)
    c0 = np.full(len(Ck.data), np.nan)
    Ck.data[:] = c0
    _, _, _, R = oracle_result(expr, A, B, Ck, alpha=1.0, beta=0.0, c0=c0)
    assert not np.isnan(R.data).any()
    dA, dB = to_device(ctx, A), to_device(ctx, B)
    dC = to_device(ctx, Ck)
    plan = make_plan(ctx, expr, dA, dB, dC, prec)
    plan.execute(1.0, 0.0)
    sht.call("sht_device_synchronize")
    plan.close()
    out = dC.download()
    assert not np.isnan(out).any()
    if prec in EXACT:
        assert np.array_equal(out, R.data)
    else:
        assert normwise(out, R.data) <= TOL[prec]


def test_alpha_zero_scales_only(ctx):
    rng = np.random.default_rng(8)
    A, B, Cp = make_case(rng, "ik", "kj", "ij", p=0.8)
    _, _, cids = orc.tasks("ik,kj->ij", A, B, Cp)
    Ck = orc.BlockSparse(Cp.block_sizes).reserve_ids(cids)
    Ck.data[:] = rng.uniform(-1, 1, len(Ck.data))
    dA, dB, dC = to_device(ctx, A), to_device(ctx, B), to_device(ctx, Ck)
    sht.contract(ctx, "ik,kj->ij", 0.0, dA, dB, 3.0, dC, sht.PREC_FP64)
    assert np.array_equal(dC.download(), 3.0 * Ck.data)


@pytest.mark.parametrize("prec", ["fp64", "fp64_as1", "fp64_as2", "fp64_as3", "fp64_exact", "fp64_exact_as1"])
@pytest.mark.parametrize("case", ["neg_zero", "neg_zero_alpha", "inf_last_column"])
def test_padded_k_keeps_signed_zero_and_inf(ctx, prec, case):
    """Contracted lengths that are not multiples of 4 leave padded positions
    in the last step of every MFMA (the A stream re-reads column K-1 there).
    Their products must not change any chain: a C element whose reference
    chain ends at -0 stays -0, and an inf in column K-1 stays inf rather
    than becoming NaN (inf * 0).  Compared bit for bit with the oracle."""
    rng = np.random.default_rng(91)
    expr = "ik,kj->ij"
    bs_i, bs_k, bs_j = [16, 40, 7, 33], [5, 13, 6, 21], [16, 20, 33, 3]
    A = random_bst(rng, [bs_i, bs_k], 0.9)
    B = random_bst(rng, [bs_k, bs_j], 0.9)
    Cp = orc.BlockSparse([bs_i, bs_j])
    _, _, cids = orc.tasks(expr, A, B, Cp)
    Ck = orc.BlockSparse(Cp.block_sizes).reserve_ids(cids)
    alpha, beta = 1.0, 1.0
    if case in ("neg_zero", "neg_zero_alpha"):
        # every product is -0 and C starts at -0: the chains end at -0
        B.data[:] = rng.uniform(-1.0, -0.5, len(B.data))
        A.data[:] = 0.0 if case == "neg_zero" else -0.0
        alpha = 1.0 if case == "neg_zero" else -0.5
        c0 = np.full(len(Ck.data), -0.0)
    else:
        # the last contracted column of every A block is +inf, B > 0
        B.data[:] = rng.uniform(0.5, 1.0, len(B.data))
        for p in range(len(A.indices)):
            ik = np.unravel_index(int(A.indices[p]), (len(bs_i), len(bs_k)), order="F")
            m, k = bs_i[ik[0]], bs_k[ik[1]]
            A.data[A.offsets[p] + (k - 1) * m:A.offsets[p] + k * m] = np.inf
        beta = 0.5
        c0 = rng.uniform(-1, 1, len(Ck.data))
    _, _, _, R = oracle_result(expr, A, B, Ck, alpha=alpha, beta=beta, c0=c0)
    if case != "inf_last_column":
        assert np.signbit(R.data).all() and not R.data.any()
    else:
        assert np.isinf(R.data).any() and not np.isnan(R.data).any()
    Ck.data[:] = c0
    dA, dB, dC = to_device(ctx, A), to_device(ctx, B), to_device(ctx, Ck)
    plan = make_plan(ctx, expr, dA, dB, dC, prec)
    plan.execute(alpha, beta)
    sht.call("sht_device_synchronize")
    plan.close()
    out = dC.download()
    assert np.array_equal(out.view(np.int64), R.data.view(np.int64))


def test_empty_operand(ctx):
    A = orc.BlockSparse([[2, 3], [4]])
    A.reserve_ids([])
    B = orc.BlockSparse([[4], [5, 1]])
    B.reserve_ids([0, 1])
    B.data[:] = 1.0
    dA, dB = to_device(ctx, A), to_device(ctx, B)
    dC = sht.Tensor(ctx, [[2, 3], [5, 1]])
    sht.contract(ctx, "ik,kj->ij", 1.0, dA, dB, 0.0, dC)
    assert dC.get_nb_nzblocks_local() == 0


def test_expression_and_block_size_errors(ctx):
    A = sht.Tensor(ctx, [[2, 3], [4]])
    A.reserve_all()
    B = sht.Tensor(ctx, [[4], [5]])
    B.reserve_all()
    C = sht.Tensor(ctx, [[2, 3], [5]])
    with pytest.raises(sht.ShtError, match="Duplicate tensor indices"):
        sht.contract(ctx, "ii,kj->ij", 1.0, A, B, 0.0, C)
    B2 = sht.Tensor(ctx, [[3, 1], [5]])
    B2.reserve_all()
    with pytest.raises(sht.ShtError, match="Sizes for dimensions|Block sizes"):
        sht.contract(ctx, "ik,kj->ij", 1.0, A, B2, 0.0, C)
    B3 = sht.Tensor(ctx, [[4], [5, 2]])
    B3.reserve_all()
    with pytest.raises(sht.ShtError):
        sht.contract(ctx, "ik,kj->ij", 1.0, A, B3, 0.0, C)


@pytest.mark.parametrize("prec", ["fp64_exact", "fp64_exact_as1", "fp64", "fp64_as1", "fp64_as2", "fp64_as3",
                                  "fp32_as1", "fp32_as3", "bf16x3", "bf16x3_as1"])
def test_graph_replay_matches_eager(ctx, prec):
    rng = np.random.default_rng(12)
    expr = "ijk,kl->ijl"
    A, B, Cp = make_case(rng, "ijk", "kl", "ijl", p=0.6, bmin=3, bmax=12)
    dA, dB = to_device(ctx, A), to_device(ctx, B)
    dC = sht.Tensor(ctx, Cp.block_sizes)
    plan = make_plan(ctx, expr, dA, dB, dC, prec)
    plan.execute(1.0, 0.0)
    sht.call("sht_device_synchronize")
    eager = dC.download()
    s = sht.Stream()
    for rep in range(3):
        dC.upload(np.full(dC.nelem(), 7.0))
        plan.graph_replay(1.0, 0.0, 2, s)
        s.synchronize()
        assert np.array_equal(dC.download(), eager), rep
    s.close()
    plan.close()


@pytest.mark.parametrize("prec", ["fp64", "fp64_exact", "fp32"])
def test_concurrent_tile_classes(ctx, prec):
    """A C-stationary plan with several tile classes runs them on up to four
    streams (fork and join by events): eager on a created stream, eager on
    the null stream and graph replays (with eager work between them) all give
    the oracle's result, FP64 bit for bit."""
    rng = np.random.default_rng(131)
    expr = "ik,kj->ij"
    bs_i = [16, 50, 64, 100, 7, 130]
    bs_k = [3, 16, 37, 64]
    bs_j = [33, 48, 64, 80, 130, 20, 9, 16, 5]
    A = random_bst(rng, [bs_i, bs_k], 0.8)
    B = random_bst(rng, [bs_k, bs_j], 0.8)
    Cp = orc.BlockSparse([bs_i, bs_j])
    _, _, cids = orc.tasks(expr, A, B, Cp)
    Ck = orc.BlockSparse(Cp.block_sizes).reserve_ids(cids)
    c0 = rng.uniform(-1, 1, len(Ck.data))
    _, _, _, R = oracle_result(expr, A, B, Ck, alpha=0.75, beta=1.5, c0=c0, mode=0)
    Ck.data[:] = c0
    dA, dB, dC = to_device(ctx, A), to_device(ctx, B), to_device(ctx, Ck)
    plan = make_plan(ctx, expr, dA, dB, dC, prec)
    classes = plan.execute_timed(0.0, 1.0)  # (C = 1 * C: values unchanged)
    assert len(classes) >= 3, classes  # (concurrency needs two or more)
    check = (lambda x: np.array_equal(x, R.data)) if prec in EXACT else (lambda x: normwise(x, R.data) <= TOL[prec])
    s = sht.Stream()
    # (the first execution is a graph capture: the plan creates its side
    # streams before capturing)
    for run in range(4):
        dC.upload(c0)
        if run == 1:
            plan.execute(0.75, 1.5, s)
            s.synchronize()
        elif run == 2:
            plan.execute(0.75, 1.5)
            sht.call("sht_device_synchronize")
        else:
            plan.graph_replay(0.75, 1.5, 1, s)
            s.synchronize()
        assert check(dC.download()), run
    s.close()
    plan.close()


def test_config0_full_size(ctx):
    """ijk,kl->ijl, every mode 256 in 16 blocks of 16, p=0.3 (seeded)."""
    seed = 1000
    bs = [[16] * 16] * 3
    bsB = [[16] * 16] * 2
    A = orc.BlockSparse(bs).reserve_ids(orc.gen_presence(seed, 0, [16, 16, 16], 0.3))
    B = orc.BlockSparse(bsB).reserve_ids(orc.gen_presence(seed, 1, [16, 16], 0.3))
    orc.gen_values(seed, 0, A, 0.0)
    orc.gen_values(seed, 1, B, 0.0)
    Cp = orc.BlockSparse(bs)
    t, kk, cids, R = oracle_result("ijk,kl->ijl", A, B, Cp)
    dA = sht.Tensor(ctx, bs)
    dA.reserve_random(seed, 0, 0.3)
    dA.fill_random(seed, 0)
    dB = sht.Tensor(ctx, bsB)
    dB.reserve_random(seed, 1, 0.3)
    dB.fill_random(seed, 1)
    for prec in ("fp64_exact", "fp64_exact_as1", "fp64", "fp64_as1", "fp64_as3"):
        dC = sht.Tensor(ctx, bs)
        plan = make_plan(ctx, "ijk,kl->ijl", dA, dB, dC, prec)
        got = plan.tasks()
        assert len(got["a"]) == len(t)
        assert np.array_equal(got["a"], t[:, 0]) and np.array_equal(got["b"], t[:, 1])
        assert np.array_equal(dC.get_sparse_info()["indices"], cids)
        plan.execute()
        sht.call("sht_device_synchronize")
        out = dC.download()
        if prec in EXACT:
            assert np.array_equal(out, R.data)
        else:
            assert normwise(out, R.data) <= 1e-12
        plan.close()
        dC.close()


def test_autotune_keeps_results(ctx):
    rng = np.random.default_rng(31)
    expr = "ijkl,klm->ijm"
    A, B, Cp = make_case(rng, "ijkl", "klm", "ijm", p=0.5, bmin=3, bmax=12)
    _, _, _, R = oracle_result(expr, A, B, Cp)
    dA, dB = to_device(ctx, A), to_device(ctx, B)
    dC = sht.Tensor(ctx, Cp.block_sizes)
    plan = sht.Plan(ctx, expr, dA, dB, dC, sht.PREC_FP64)
    times = plan.autotune(1.0, 0.0, reps=2)
    assert set(times) == {0, 1, 2, 3} and min(times.values()) > 0 and plan.get_schedule() in (0, 1, 2, 3)
    plan.execute(1.0, 0.0)
    sht.call("sht_device_synchronize")
    assert normwise(dC.download(), R.data) <= 1e-12
    plan.close()


@pytest.mark.parametrize("prec", ["fp64", "fp32", "bf16x3"])
@pytest.mark.parametrize("alpha,beta", [(1.0, 0.0), (0.5, -1.25)])
def test_wide_blocks_c_stationary(ctx, alpha, beta, prec):
    """C blocks of 33..130 columns run on the four-wave 64 x 64 tiles (A and
    B chunks shared through LDS), with partial tiles in both directions and
    contracted lengths that are not multiples of 16: FP64 bit for bit equal
    to the oracle's fma chain, FP32 and split-BF16 within 1e-6."""
    rng = np.random.default_rng(78)
    expr = "ik,kj->ij"
    bs_i = [16, 50, 64, 100, 7]
    bs_k = [3, 16, 37, 64]
    bs_j = [33, 48, 64, 80, 130, 20]
    A = random_bst(rng, [bs_i, bs_k], 0.8)
    B = random_bst(rng, [bs_k, bs_j], 0.8)
    Cp = orc.BlockSparse([bs_i, bs_j])
    _, _, cids = orc.tasks(expr, A, B, Cp)
    Ck = orc.BlockSparse(Cp.block_sizes).reserve_ids(cids)
    c0 = rng.uniform(-1, 1, len(Ck.data))
    _, _, _, R = oracle_result(expr, A, B, Ck, alpha=alpha, beta=beta, c0=c0)
    Ck.data[:] = c0
    dA, dB = to_device(ctx, A), to_device(ctx, B)
    dC = to_device(ctx, Ck)
    plan = make_plan(ctx, expr, dA, dB, dC, prec)
    plan.execute(alpha, beta)
    sht.call("sht_device_synchronize")
    plan.close()
    if prec in EXACT:
        assert np.array_equal(dC.download(), R.data)
    else:
        assert normwise(dC.download(), R.data) <= TOL[prec]


@pytest.mark.parametrize("prec,names", [("fp64", ("mfma64_n16_r128", "mfma64_n32_r128")),
                                        ("fp32", ("mfma32_n16w_r128", "mfma32_n32w_r128"))])
def test_large_launch_128_row_tiles(ctx, prec, names):
    """A launch of more than 64 * 1024 64-row tiles runs FP64 and FP32 C
    blocks of more than 96 rows on 128-row tiles (then 64- and 32-row tiles
    for the rest): 38 400 C blocks of 20..230 rows and 4 or 20 columns, FP64
    bit for bit equal to the oracle, FP32 within its tolerance."""
    rng = np.random.default_rng(79)
    expr = "ik,kj->ij"
    bs_i = [100, 230, 40, 20] * 32
    bs_k = [8, 16, 5, 3]
    bs_j = [4, 20] * 150
    A = random_bst(rng, [bs_i, bs_k], 1.0)
    B = random_bst(rng, [bs_k, bs_j], 1.0)
    Cp = orc.BlockSparse([bs_i, bs_j])
    _, _, _, R = oracle_result(expr, A, B, Cp)
    dA, dB = to_device(ctx, A), to_device(ctx, B)
    dC = sht.Tensor(ctx, Cp.block_sizes)
    plan = make_plan(ctx, expr, dA, dB, dC, prec)
    launches = [k for k, _ in plan.execute_timed(1.0, 0.0)]
    assert all(n in launches for n in names), launches
    sht.call("sht_device_synchronize")
    plan.close()
    if prec in EXACT:
        assert np.array_equal(dC.download(), R.data)
    else:
        assert normwise(dC.download(), R.data) <= TOL[prec]
    dC.close()


@pytest.mark.parametrize("prec,name", [("fp64", "mfma64_n32s"), ("fp32", "mfma32_n32s"), ("bf16x3", "bf16x3_n32s")])
def test_alike_operands_square_order(ctx, prec, name):
    """A and B of the same size, and more than 32 four-wave tiles per SIMD:
    the tiles are dealt to the XCD queues in squares of 16 x 16 C blocks
    (about 10 000 C blocks of 32 x 32 in 100 x 100 block rows and columns:
    100 = 6 x 16 + 4, so the last square of every row and column of squares
    is partial, and the eight queues are padded); FP64 bit for bit equal to
    the oracle, FP32 and split-BF16 within their tolerance."""
    rng = np.random.default_rng(83)
    expr = "ik,kj->ij"
    bs = [32] * 100
    A = random_bst(rng, [bs, bs], 0.25)
    B = random_bst(rng, [bs, bs], 0.25)
    Cp = orc.BlockSparse([bs, bs])
    _, _, _, R = oracle_result(expr, A, B, Cp)
    dA, dB = to_device(ctx, A), to_device(ctx, B)
    dC = sht.Tensor(ctx, Cp.block_sizes)
    plan = make_plan(ctx, expr, dA, dB, dC, prec)
    launches = [k for k, _ in plan.execute_timed(1.0, 0.0)]
    assert name in launches, launches
    sht.call("sht_device_synchronize")
    plan.close()
    assert np.array_equal(dC.get_indices(), R.indices)
    if prec in EXACT:
        assert np.array_equal(dC.download(), R.data)
    else:
        assert normwise(dC.download(), R.data) <= TOL[prec]
    dC.close()


@pytest.mark.parametrize("prec,scheds", [(sht.PREC_FP64, (0, 1, 2, 3)), (sht.PREC_FP32, (0, 1, 2, 3)),
                                         (sht.PREC_BF16X3, (0, 1, 2))])
def test_schedules_agree_on_uneven_block_rows(ctx, prec, scheds):
    """Block rows of 1..7 32-row chunks (so the grouped schedule pads its
    workgroups with 0..3 ghost units), panels of up to 17 16-column slots
    (so sweeps are cut at 8 and at 4 slots) and C blocks up to 70 columns wide
    (C-stationary 64-column tiles plus a remainder, 32 x 32 four-wave tiles
    for 17-32 columns): every schedule gives the oracle's result, FP64 bit
    for bit, FP32 and split-BF16 within 1e-6."""
    rng = np.random.default_rng(77)
    expr = "ik,kj->ij"
    bs_i = [5, 32, 33, 64, 97, 130, 161, 200, 224, 17]
    bs_k = [7, 16, 40]
    bs_j = [3, 16, 17, 30, 9, 28, 25, 11, 20, 40, 70]
    A = random_bst(rng, [bs_i, bs_k], 0.8)
    B = random_bst(rng, [bs_k, bs_j], 0.9)
    Cp = orc.BlockSparse([bs_i, bs_j])
    _, _, _, R = oracle_result(expr, A, B, Cp)
    dA, dB = to_device(ctx, A), to_device(ctx, B)
    outs = {}
    for sc in scheds:
        dC = sht.Tensor(ctx, Cp.block_sizes)
        plan = sht.Plan(ctx, expr, dA, dB, dC, prec)
        plan.set_schedule(sc)
        plan.execute(1.0, 0.0)
        sht.call("sht_device_synchronize")
        outs[sc] = dC.download()
        plan.close()
        if prec == sht.PREC_FP64:
            assert np.array_equal(outs[sc], R.data), sc
        else:
            assert normwise(outs[sc], R.data) <= 1e-6, sc


@pytest.mark.parametrize("prec", [sht.PREC_FP64, sht.PREC_FP64_EXACT, sht.PREC_FP32, sht.PREC_BF16X3])
def test_tile_variants_and_autotune_ex(ctx, prec):
    """Every C-stationary tile variant (automatic, without / with 128-row
    tiles, cost order, uniform 64 x 64, one persistent launch) gives the
    oracle's result (FP64 forms bit for bit); the autotuner times every
    (schedule, variant) and the variant it keeps gives the same result
    again.  The fma chain has no persistent variant."""
    rng = np.random.default_rng(202)
    expr = "ik,kj->ij"
    bs_i = [5, 32, 33, 64, 97, 130, 161, 17]
    bs_k = [7, 16, 40, 3]
    bs_j = [3, 16, 17, 30, 9, 70, 40]
    A = random_bst(rng, [bs_i, bs_k], 0.8)
    B = random_bst(rng, [bs_k, bs_j], 0.8)
    Cp = orc.BlockSparse([bs_i, bs_j])
    _, _, _, R = oracle_result(expr, A, B, Cp)
    dA, dB = to_device(ctx, A), to_device(ctx, B)
    dC = sht.Tensor(ctx, Cp.block_sizes)
    plan = sht.Plan(ctx, expr, dA, dB, dC, prec)
    plan.set_schedule(0)

    def check():
        plan.execute(1.0, 0.0)
        sht.call("sht_device_synchronize")
        out = dC.download()
        if prec in (sht.PREC_FP64, sht.PREC_FP64_EXACT):
            assert np.array_equal(out, R.data)
        else:
            assert normwise(out, R.data) <= 1e-6

    exact = prec == sht.PREC_FP64_EXACT
    for v in range(5 if exact else 6):
        plan.set_variant(v)
        assert plan.get_variant() == v
        check()
    if exact:
        with pytest.raises(sht.ShtError, match="persistent variant"):
            plan.set_variant(5)
        assert plan.get_variant() == 4
    cands = plan.autotune_ex(1.0, 0.0, reps=1)
    names = {(s, v) for s, v, _ in cands}
    assert (0, "uniform_64x64") in names and (0, "auto") in names
    assert ((0, "persistent") in names) == (not exact)
    assert all(ms > 0 for _, _, ms in cands)
    check()
    plan.close()


@pytest.mark.parametrize("prec", [sht.PREC_FP64, sht.PREC_FP32, sht.PREC_BF16X3])
def test_persistent_launch_many_tiles(ctx, prec):
    """The persistent variant over more tiles than the device keeps
    workgroups (every workgroup walks its XCD queue through several tiles):
    4 096 C blocks of 32 x 32 (ik,kj->ij, L = 2048, p = 0.25) give the same C
    bit for bit as the per-class launches of variant 0 (the same tile class,
    32 x 32, in both), the FP64 form the oracle's; then a mix of widths
    (ijk,km->ijm, uneven blocks from 5 to 40: 64 x 64, 64 x 32, 32 x 32 and
    64 x 16 tiles in one launch) within the form's tolerance, FP64 bit for
    bit."""
    seed, L, b = 2026, 2048, 32
    bs = [[b] * (L // b)] * 2
    A = orc.BlockSparse(bs).reserve_ids(orc.gen_presence(seed, 0, [L // b] * 2, 0.25))
    B = orc.BlockSparse(bs).reserve_ids(orc.gen_presence(seed, 1, [L // b] * 2, 0.25))
    orc.gen_values(seed, 0, A, 0.1)
    orc.gen_values(seed, 1, B, 0.1)
    rng = np.random.default_rng(77)
    cases = [("ik,kj->ij", A, B, orc.BlockSparse(bs))]
    szs = {x: rng.choice([5, 9, 16, 17, 28, 33, 40], size=rng.integers(6, 10)).tolist() for x in "ijkm"}
    A2 = random_bst(rng, [szs["i"], szs["j"], szs["k"]], 0.3)
    B2 = random_bst(rng, [szs["k"], szs["m"]], 0.5)
    cases.append(("ijk,km->ijm", A2, B2, orc.BlockSparse([szs["i"], szs["j"], szs["m"]])))
    for n, (expr, a, bb, Cp) in enumerate(cases):
        _, _, _, R = oracle_result(expr, a, bb, Cp)
        dA, dB = to_device(ctx, a), to_device(ctx, bb)
        outs = {}
        for v in (0, 5):
            dC = sht.Tensor(ctx, Cp.block_sizes)
            plan = sht.Plan(ctx, expr, dA, dB, dC, prec)
            plan.set_schedule(0)
            plan.set_variant(v)
            if v == 5:
                assert plan.execute_timed(1.0, 0.0)[0][0].endswith("_persist")
            print("persistent launch", expr, prec, "variant", v, flush=True)
            plan.execute(1.0, 0.0)
            sht.call("sht_device_synchronize")
            outs[v] = dC.download()
            assert np.array_equal(dC.get_indices(), R.indices)
            plan.close()
            dC.close()
        if prec == sht.PREC_FP64:
            assert np.array_equal(outs[5], R.data)
        else:
            assert normwise(outs[5], R.data) <= 1e-6
        if n == 0:
            assert np.array_equal(outs[5], outs[0])
        dA.close()
        dB.close()


C1_FORMS = [(sht.PREC_FP64, (0, 1, 2, 3)), (sht.PREC_FP64_EXACT, (0, 1)), (sht.PREC_FP32, (0, 1, 2, 3)),
            (sht.PREC_BF16X3, (0, 1, 2))]


@pytest.fixture(scope="module")
def c1_recipe():
    """Config 1's recipe at L = 180: ijkl,klm->ijm, uneven blocks drawn from
    {5,7,9,11,13,15,21,28}, p = 0.1, values decaying by exp(-0.3|b0-b1|),
    seed 1001 -- generated on the device and by the oracle from the same
    counters; the oracle's contraction of it."""
    seed, L, choices = 1001, 180, [5, 7, 9, 11, 13, 15, 21, 28]
    bs = [orc.gen_block_sizes(seed, ax, L, choices).tolist() for ax in range(5)]
    bsA, bsB, bsC = [bs[0], bs[1], bs[2], bs[3]], [bs[2], bs[3], bs[4]], [bs[0], bs[1], bs[4]]
    A = orc.BlockSparse(bsA).reserve_ids(orc.gen_presence(seed, 0, [len(b) for b in bsA], 0.1))
    B = orc.BlockSparse(bsB).reserve_ids(orc.gen_presence(seed, 1, [len(b) for b in bsB], 0.1))
    orc.gen_values(seed, 0, A, 0.3)
    orc.gen_values(seed, 1, B, 0.3)
    _, _, _, R = oracle_result("ijkl,klm->ijm", A, B, orc.BlockSparse(bsC))
    return seed, bsA, bsB, bsC, R


@pytest.mark.parametrize("prec,scheds", C1_FORMS)
def test_config1_recipe_every_schedule(ctx, c1_recipe, prec, scheds):
    """Every precision form through every schedule on config 1's block
    recipe: FP64 forms bit for bit, FP32 and split-BF16 within 1e-6
    normwise and max-blockwise."""
    seed, bsA, bsB, bsC, R = c1_recipe
    dA = sht.Tensor(ctx, bsA)
    dA.reserve_random(seed, 0, 0.1)
    dA.fill_random(seed, 0, 0.3)
    dB = sht.Tensor(ctx, bsB)
    dB.reserve_random(seed, 1, 0.1)
    dB.fill_random(seed, 1, 0.3)
    ends = np.r_[R.offsets[1:], len(R.data)]
    for sc in scheds:
        print("config 1 recipe", prec, "schedule", sc, flush=True)
        dC = sht.Tensor(ctx, bsC)
        plan = sht.Plan(ctx, "ijkl,klm->ijm", dA, dB, dC, prec)
        plan.set_schedule(sc)
        plan.execute(1.0, 0.0)
        sht.call("sht_device_synchronize")
        assert np.array_equal(dC.get_indices(), R.indices)
        out = dC.download()
        if prec in (sht.PREC_FP64, sht.PREC_FP64_EXACT):
            assert np.array_equal(out, R.data), sc
        else:
            assert normwise(out, R.data) <= 1e-6, sc
            blk = max(np.linalg.norm(out[a:b] - R.data[a:b]) / np.linalg.norm(R.data[a:b])
                      for a, b in zip(R.offsets, ends) if np.linalg.norm(R.data[a:b]) > 0)
            assert blk <= 1e-6, (sc, blk)
        plan.close()
        dC.close()
