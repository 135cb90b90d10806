"""FP32 staging of FP64 operands: an FP32 or split-BF16 plan over FLOAT64
tensors may convert A and B once per execution into FP32 copies that its
C-stationary tiles read (sht_plan_set_fp32_staging), instead of rounding the
FP64 values at every use.  Both forms round every operand to FP32 before the
block products, so C must be equal bit for bit with and without the staging
-- and to the same plan over FLOAT32 tensors -- in every tile variant, with
ragged and permuted blocks, alpha and beta, and graph replay; and within 1e-6
of the FP64 oracle."""
import numpy as np
import pytest

import shtensor_amd as sht
from oracle import oracle as orc
from gpu_util import normwise, random_bst, to_device

pytestmark = pytest.mark.gpu

FORMS = {"fp32": sht.PREC_FP32, "bf16x3": sht.PREC_BF16X3}


@pytest.fixture(scope="module")
def ctx():
    c = sht.Context(4 << 30)
    yield c
    c.close()


def to_device32(ctx, T):
    D = sht.Tensor(ctx, [list(map(int, b)) for b in T.block_sizes], dtype=np.float32)
    D.reserve_ids(T.indices)
    if len(T.data):
        D.upload(T.data.astype(np.float32))
    return D


def run(ctx, expr, dA, dB, bsC, prec, staging, variant=0, alpha=1.0, beta=0.0, c0=None, cids=None, replays=0):
    dC = sht.Tensor(ctx, bsC)
    if c0 is not None:
        dC.reserve_ids(cids)
        dC.upload(c0)
    plan = sht.Plan(ctx, expr, dA, dB, dC, prec)
    plan.set_schedule(0)
    plan.set_variant(variant)
    if staging is not None:
        plan.set_fp32_staging(staging)
        assert plan.get_fp32_staging() == staging
    if replays:
        s = sht.Stream()
        plan.graph_replay(alpha, beta, replays, s)
        s.synchronize()
        s.close()
    else:
        plan.execute(alpha, beta)
    sht.call("sht_device_synchronize")
    out, ids = dC.download(), dC.get_indices()
    plan.close()
    dC.close()
    return out, ids


# every C-stationary FP32 / split-BF16 tile class, ragged contracted lengths,
# permuted layouts (staged permutation, then the FP32 copy)
CASES = [
    ("ik,kj->ij", {"i": [64, 64, 40], "k": [64, 48, 64], "j": [64, 40, 64]}),
    ("ik,kj->ij", {"i": [32, 20, 32], "k": [32, 32, 9], "j": [24, 32, 17]}),
    ("ik,kj->ij", {"i": [150, 33, 16], "k": [16, 13, 16], "j": [16, 11, 7]}),
    ("kil,ljk->ij", {"i": [19, 40], "j": [9, 33], "k": [5, 12], "l": [7, 3]}),
    ("ijkl,klm->ijm", {"i": [12, 30], "j": [5, 7, 3], "k": [9, 11], "l": [8, 13], "m": [21, 5, 28]}),
]


@pytest.mark.parametrize("form", list(FORMS))
@pytest.mark.parametrize("case", range(len(CASES)))
def test_staging_equals_unstaged_and_fp32_tensors(ctx, case, form):
    """Staging on and off over FP64 tensors, and the FP32 tensors of the same
    values: bitwise equal in every tile variant (alpha 0.5, beta 1.25 onto a
    random C), within 1e-6 of the oracle; alpha = 0 leaves beta * C."""
    prec = FORMS[form]
    expr, sizes = CASES[case]
    ea, rest = expr.split(",")
    eb, ec = rest.split("->")
    rng = np.random.default_rng(1100 + case)
    A = random_bst(rng, [sizes[x] for x in ea], 0.7)
    B = random_bst(rng, [sizes[x] for x in eb], 0.7)
    bsC = [sizes[x] for x in ec]
    _, _, cids = orc.tasks(expr, A, B, orc.BlockSparse(bsC))
    R = orc.BlockSparse(bsC).reserve_ids(cids)
    c0 = rng.uniform(-1, 1, len(R.data))
    R.data[:] = c0
    orc.contract(expr, 0.5, A, B, 1.25, R)
    d64 = (to_device(ctx, A), to_device(ctx, B))
    d32 = (to_device32(ctx, A), to_device32(ctx, B))
    for var in range(6):
        print("fp32 staging", form, expr, "variant", var, flush=True)
        o0, i0 = run(ctx, expr, *d64, bsC, prec, False, var, 0.5, 1.25, c0, cids)
        o1, i1 = run(ctx, expr, *d64, bsC, prec, True, var, 0.5, 1.25, c0, cids)
        o2, _ = run(ctx, expr, *d32, bsC, prec, None, var, 0.5, 1.25, c0, cids)
        assert np.array_equal(i1, R.indices) and np.array_equal(i0, i1)
        assert np.array_equal(o1, o0) and np.array_equal(o2, o1), var
        assert normwise(o1, R.data) <= 1e-6, var
    o1, _ = run(ctx, expr, *d64, bsC, prec, True, 0, 0.0, 1.25, c0, cids)
    assert np.array_equal(o1, c0 * 1.25)
    for t in d64 + d32:
        t.close()


@pytest.mark.parametrize("form", list(FORMS))
def test_staging_default_and_graph_replay(ctx, form):
    """A 32-block product (ik,kj->ij, L = 2048, p = 0.25): each operand
    element is read about sixteen times, so a new plan over FP64 tensors
    stages by default; C equals the unstaged plan bit for bit, also replayed
    as a graph, and the oracle within 1e-6."""
    seed, L, b = 2025, 2048, 32
    bs = [[b] * (L // b)] * 2
    A = orc.BlockSparse(bs).reserve_ids(orc.gen_presence(seed, 0, [L // b] * 2, 0.25))
    B = orc.BlockSparse(bs).reserve_ids(orc.gen_presence(seed, 1, [L // b] * 2, 0.25))
    orc.gen_values(seed, 0, A, 0.1)
    orc.gen_values(seed, 1, B, 0.1)
    _, _, cids = orc.tasks("ik,kj->ij", A, B, orc.BlockSparse(bs))
    R = orc.BlockSparse(bs).reserve_ids(cids)
    orc.contract("ik,kj->ij", 1.0, A, B, 0.0, R)
    dA, dB = to_device(ctx, A), to_device(ctx, B)
    dC = sht.Tensor(ctx, bs)
    plan = sht.Plan(ctx, "ik,kj->ij", dA, dB, dC, FORMS[form])
    assert plan.get_schedule() == 0 and plan.get_fp32_staging()
    plan.close()
    dC.close()
    o0, _ = run(ctx, "ik,kj->ij", dA, dB, bs, FORMS[form], False)
    o1, i1 = run(ctx, "ik,kj->ij", dA, dB, bs, FORMS[form], True)
    og, _ = run(ctx, "ik,kj->ij", dA, dB, bs, FORMS[form], True, replays=3)
    assert np.array_equal(i1, R.indices)
    assert np.array_equal(o1, o0) and np.array_equal(og, o1)
    assert normwise(o1, R.data) <= 1e-6
    dA.close()
    dB.close()


def test_staging_rejected_where_it_does_not_apply(ctx):
    rng = np.random.default_rng(5)
    A = random_bst(rng, [[8, 8], [8, 8]], 1.0)
    B = random_bst(rng, [[8, 8], [8, 8]], 1.0)
    d64 = (to_device(ctx, A), to_device(ctx, B))
    d32 = (to_device32(ctx, A), to_device32(ctx, B))
    C = sht.Tensor(ctx, [[8, 8], [8, 8]])
    for prec, ops in ((sht.PREC_FP64, d64), (sht.PREC_FP64_EXACT, d64), (sht.PREC_FP32, d32),
                      (sht.PREC_BF16X3, d32)):
        plan = sht.Plan(ctx, "ik,kj->ij", *ops, C, prec)
        assert not plan.get_fp32_staging()
        with pytest.raises(sht.ShtError, match="FLOAT64 operands"):
            plan.set_fp32_staging(True)
        plan.set_fp32_staging(False)
        plan.close()
    for t in d64 + d32 + (C,):
        t.close()
