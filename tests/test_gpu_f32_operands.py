"""FP32 operand tensors (Tensor<float,N>, Tensor.h:19-296 with T = float)
for the FP32 and split-BF16 forms: A and B stored in FP32 in HBM, C in FP64.
Both forms round each operand value to FP32 before the block products, so
reading FP32 tensors must give results equal bit for bit to reading the FP64
tensors they were rounded from -- in every schedule, tile variant and served
batch -- and within 1e-6 of the FP64 oracle."""
import numpy as np
import pytest

import shtensor_amd as sht
from oracle import oracle as orc
from gpu_util import normwise, random_bst, to_device

pytestmark = pytest.mark.gpu

FORMS = {"fp32": (sht.PREC_FP32, (0, 1, 2, 3)), "bf16x3": (sht.PREC_BF16X3, (0, 1, 2))}


@pytest.fixture(scope="module")
def ctx():
    c = sht.Context(4 << 30)
    yield c
    c.close()


def to_device32(ctx, T):
    """FP32 device tensor with T's block ids and T's values rounded to FP32."""
    D = sht.Tensor(ctx, [list(map(int, b)) for b in T.block_sizes], dtype=np.float32)
    D.reserve_ids(T.indices)
    if len(T.data):
        D.upload(T.data.astype(np.float32))
    return D


def run(ctx, expr, dA, dB, bsC, prec, sched, variant=None, alpha=1.0, beta=0.0, c0=None, cids=None):
    dC = sht.Tensor(ctx, bsC)
    if c0 is not None:
        dC.reserve_ids(cids)
        dC.upload(c0)
    plan = sht.Plan(ctx, expr, dA, dB, dC, prec)
    plan.set_schedule(sched)
    if variant is not None:
        plan.set_variant(variant)
    plan.execute(alpha, beta)
    sht.call("sht_device_synchronize")
    out, ids = dC.download(), dC.get_indices()
    plan.close()
    dC.close()
    return out, ids


def blockwise(out, R):
    ends = np.r_[R.offsets[1:], len(R.data)]
    return max((np.linalg.norm(out[a:b] - R.data[a:b]) / np.linalg.norm(R.data[a:b])
                for a, b in zip(R.offsets, ends) if np.linalg.norm(R.data[a:b]) > 0), default=0.0)


@pytest.mark.parametrize("form", list(FORMS))
def test_config1_recipe_fp32_operands(ctx, form):
    """Config 1's recipe at L = 180 (ijkl,klm->ijm, blocks from
    {5,...,28}, p = 0.1, decaying values), generated once as FP64 and once as
    FP32 tensors from the same counters (the FP32 values are the FP64 values
    rounded): every schedule gives the same C bit for bit from either, within
    1e-6 normwise and max-blockwise of the oracle."""
    prec, scheds = FORMS[form]
    seed, L, choices = 1001, 180, [5, 7, 9, 11, 13, 15, 21, 28]
    bs = [orc.gen_block_sizes(seed, ax, L, choices).tolist() for ax in range(5)]
    bsA, bsB, bsC = [bs[0], bs[1], bs[2], bs[3]], [bs[2], bs[3], bs[4]], [bs[0], bs[1], bs[4]]
    A = orc.BlockSparse(bsA).reserve_ids(orc.gen_presence(seed, 0, [len(b) for b in bsA], 0.1))
    B = orc.BlockSparse(bsB).reserve_ids(orc.gen_presence(seed, 1, [len(b) for b in bsB], 0.1))
    orc.gen_values(seed, 0, A, 0.3)
    orc.gen_values(seed, 1, B, 0.3)
    _, _, cids = orc.tasks("ijkl,klm->ijm", A, B, orc.BlockSparse(bsC))
    R = orc.BlockSparse(bsC).reserve_ids(cids)
    orc.contract("ijkl,klm->ijm", 1.0, A, B, 0.0, R)
    dev = {}
    for dt in (np.float64, np.float32):
        dA = sht.Tensor(ctx, bsA, dtype=dt)
        dA.reserve_random(seed, 0, 0.1)
        dA.fill_random(seed, 0, 0.3)
        dB = sht.Tensor(ctx, bsB, dtype=dt)
        dB.reserve_random(seed, 1, 0.1)
        dB.fill_random(seed, 1, 0.3)
        dev[dt] = (dA, dB)
    a32 = dev[np.float32][0].download()
    assert a32.dtype == np.float32 and np.array_equal(a32, A.data.astype(np.float32))
    for sc in scheds:
        print("config 1 recipe, FP32 operands", form, "schedule", sc, flush=True)
        o64, i64 = run(ctx, "ijkl,klm->ijm", *dev[np.float64], bsC, prec, sc)
        o32, i32 = run(ctx, "ijkl,klm->ijm", *dev[np.float32], bsC, prec, sc)
        assert np.array_equal(i32, R.indices) and np.array_equal(i64, R.indices)
        assert np.array_equal(o32, o64), sc
        assert normwise(o32, R.data) <= 1e-6, sc
        assert blockwise(o32, R) <= 1e-6, sc
    for d in dev.values():
        for t in d:
            t.close()


# C-stationary tile classes: blocks wider than 32 columns (64 x 64 tiles),
# 17-32 columns (64 x 32 / 32 x 32) and at most 16 (64 x 16, one-wave 32 x
# 16 for short last row tiles); permuted expressions stage FP32 copies.
CASES = [
    ("ik,kj->ij", {"i": [64, 64, 40], "k": [64, 48, 64], "j": [64, 40, 64]}),
    ("ik,kj->ij", {"i": [32, 20, 32], "k": [32, 32, 9], "j": [24, 32, 17]}),
    ("ik,kj->ij", {"i": [150, 33, 16], "k": [16, 13, 16], "j": [16, 11, 7]}),
    ("kil,ljk->ij", {"i": [19, 40], "j": [9, 33], "k": [5, 12], "l": [7, 3]}),
    ("ijkl,klm->ijm", {"i": [12, 30], "j": [5, 7, 3], "k": [9, 11], "l": [8, 13], "m": [21, 5, 28]}),
]


@pytest.mark.parametrize("form", list(FORMS))
@pytest.mark.parametrize("case", range(len(CASES)))
def test_fp32_operands_every_schedule_and_variant(ctx, form, case):
    """Random FP64 data and its FP32 rounding: equal results bit for bit in
    every schedule and every C-stationary tile variant, alpha and beta
    applied, and within 1e-6 of the oracle."""
    prec, scheds = FORMS[form]
    expr, sizes = CASES[case]
    ea, rest = expr.split(",")
    eb, ec = rest.split("->")
    rng = np.random.default_rng(700 + case)
    A = random_bst(rng, [sizes[x] for x in ea], 0.7)
    B = random_bst(rng, [sizes[x] for x in eb], 0.7)
    A.data[:] = A.data.astype(np.float32)  # the oracle sees the stored FP32 values
    B.data[:] = B.data.astype(np.float32)
    bsC = [sizes[x] for x in ec]
    _, _, cids = orc.tasks(expr, A, B, orc.BlockSparse(bsC))
    R = orc.BlockSparse(bsC).reserve_ids(cids)
    c0 = rng.uniform(-1, 1, len(R.data))
    R.data[:] = c0
    orc.contract(expr, 0.5, A, B, 1.25, R)
    d64 = (to_device(ctx, A), to_device(ctx, B))
    d32 = (to_device32(ctx, A), to_device32(ctx, B))
    runs = [(s, None) for s in scheds] + [(0, v) for v in range(6)]
    for sc, var in runs:
        print(form, expr, "schedule", sc, "variant", var, flush=True)
        o64, _ = run(ctx, expr, *d64, bsC, prec, sc, var, 0.5, 1.25, c0, cids)
        o32, _ = run(ctx, expr, *d32, bsC, prec, sc, var, 0.5, 1.25, c0, cids)
        assert np.array_equal(o32, o64), (sc, var)
        assert normwise(o32, R.data) <= 1e-6, (sc, var)
    for t in d64 + d32:
        t.close()


def test_fp32_operands_rejected_where_unsupported(ctx):
    """FP32 operands need the FP32 or split-BF16 form; C stays FP64; A and B
    share one type."""
    bs = [[4, 4], [4, 4]]
    a32 = sht.Tensor(ctx, bs, dtype=np.float32)
    a32.reserve_all()
    b32 = sht.Tensor(ctx, bs, dtype=np.float32)
    b32.reserve_all()
    b64 = sht.Tensor(ctx, bs)
    b64.reserve_all()
    for prec in (sht.PREC_FP64, sht.PREC_FP64_EXACT):
        with pytest.raises(sht.ShtError, match="FLOAT32 operands"):
            sht.Plan(ctx, "ik,kj->ij", a32, b32, sht.Tensor(ctx, bs), prec)
    with pytest.raises(sht.ShtError, match="same float type"):
        sht.Plan(ctx, "ik,kj->ij", a32, b64, sht.Tensor(ctx, bs), sht.PREC_FP32)
    with pytest.raises(sht.ShtError, match="C must be FLOAT64"):
        sht.Plan(ctx, "ik,kj->ij", a32, b32, sht.Tensor(ctx, bs, dtype=np.float32), sht.PREC_FP32)


@pytest.mark.parametrize("form", list(FORMS))
def test_served_fp32_operands_equal_alone(ctx, form):
    """Requests with FP32 operands merged into shared launches (grouped
    A-stationary and C-stationary) equal the same contractions run alone,
    bit for bit."""
    prec = FORMS[form][0]
    rng = np.random.default_rng(77 + prec)
    reqs = []
    for r in range(4):
        if r % 2 == 0:  # picks the grouped A-stationary schedule (A >> B, long k)
            expr = "ijkl,klm->ijm"
            bs = {"i": rng.integers(10, 40, 3).tolist(), "j": rng.integers(2, 9, 3).tolist(),
                  "k": rng.integers(8, 13, 3).tolist(), "l": rng.integers(8, 13, 2).tolist(),
                  "m": rng.integers(5, 20, 2).tolist()}
        else:  # C-stationary
            expr = "ik,kj->ij"
            bs = {x: rng.integers(8, 40, 3).tolist() for x in "ikj"}
        ea, rest = expr.split(",")
        eb, ec = rest.split("->")
        A = random_bst(rng, [bs[x] for x in ea], 0.7)
        B = random_bst(rng, [bs[x] for x in eb], 0.5)
        reqs.append((expr, to_device32(ctx, A), to_device32(ctx, B), [bs[x] for x in ec]))
    alone = []
    for expr, dA, dB, bsC in reqs:
        dC = sht.Tensor(ctx, bsC)
        plan = sht.Plan(ctx, expr, dA, dB, dC, prec)
        plan.execute(0.75, 0.0)
        sht.call("sht_device_synchronize")
        plan.close()
        alone.append(dC.download())
        dC.close()
    srv = sht.Server(ctx, 2 << 30, max_batch=16)
    outs, ids = [], []
    for expr, dA, dB, bsC in reqs:
        dC = sht.Tensor(ctx, bsC)
        outs.append(dC)
        ids.append(srv.submit(expr, 0.75, dA, dB, 0.0, dC, prec))
    for i in ids:
        srv.wait(i)
    st = srv.stats()
    srv.close()
    assert st["done"] == len(reqs)
    for ref, dC in zip(alone, outs):
        assert np.array_equal(dC.download(), ref)


def test_served_fp32_operands_counted_at_four_bytes(ctx):
    """The admission footprint counts FP32 operand tensors at 4 bytes per
    element: a budget that holds A and B in FP32 but not in FP64 admits the
    request, and its result equals the same contraction run alone."""
    rng = np.random.default_rng(91)
    sizes = {"i": [24, 16, 20], "k": [32, 8, 16], "j": [12, 28]}
    A = random_bst(rng, [sizes["i"], sizes["k"]], 0.8)
    B = random_bst(rng, [sizes["k"], sizes["j"]], 0.8)
    dA, dB = to_device32(ctx, A), to_device32(ctx, B)
    op_bytes = 4 * (dA.nelem() + dB.nelem())
    bsC = [sizes["i"], sizes["j"]]
    dC = sht.Tensor(ctx, bsC)
    plan = sht.Plan(ctx, "ik,kj->ij", dA, dB, dC, sht.PREC_FP32)
    plan.execute(1.0, 0.0)
    sht.call("sht_device_synchronize")
    plan.close()
    alone = dC.download()
    srv = sht.Server(ctx, int(1.5 * op_bytes))  # below 8 bytes per operand element
    dC2 = sht.Tensor(ctx, bsC)
    rid = srv.submit("ik,kj->ij", 1.0, dA, dB, 0.0, dC2, sht.PREC_FP32)
    srv.wait(rid)
    srv.close()
    assert np.array_equal(dC2.download(), alone)
