"""Helpers shared by the GPU tests: move oracle block-sparse tensors to the
device and compare results."""
import numpy as np

import shtensor_amd as sht
from oracle import oracle as orc


def to_device(ctx, T):
    """Device Tensor with the same block ids and data as oracle tensor T."""
    D = sht.Tensor(ctx, [list(map(int, b)) for b in T.block_sizes])
    D.reserve_ids(T.indices)
    info = D.get_sparse_info()
    assert np.array_equal(info["indices"], T.indices)
    assert np.array_equal(info["offsets"], T.offsets)  # storage format parity
    if len(T.data):
        D.upload(T.data)
    return D


def normwise(x, ref):
    n = np.linalg.norm(ref)
    return float(np.linalg.norm(x - ref) / (n if n > 0 else 1.0))


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


def make_case(rng, ea, eb, ec, p=0.5, integer=False, bmin=1, bmax=5, nmin=2, nmax=4):
    letters = sorted(set(ea + eb))
    bs = {x: rng.integers(bmin, bmax, rng.integers(nmin, nmax)).tolist() for x in letters}
    A = random_bst(rng, [bs[x] for x in ea], p, integer)
    B = random_bst(rng, [bs[x] for x in eb], p, integer)
    Cp = orc.BlockSparse([bs[x] for x in ec])
    return A, B, Cp
