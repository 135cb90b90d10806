"""ctypes wrapper of the oracle (liboracle.so): the CPU restatement of
shtensor's block-sparse contraction semantics.

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the shtensor_amd package.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

I32P = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
I64P = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
F64P = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
F32P = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")


class Bst(C.Structure):
    _fields_ = [("N", C.c_int), ("nblk", C.c_void_p), ("bsizes", C.c_void_p), ("nnz", C.c_int64),
                ("indices", C.c_void_p), ("offsets", C.c_void_p), ("data", C.c_void_p)]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (make -C oracle)")
        L = C.CDLL(path)
        L.orc_contract_info.restype = C.c_int
        L.orc_contract_info.argtypes = [C.c_char_p, I32P, C.c_int, I32P, C.c_int, I32P, C.c_int, I32P,
                                        C.c_char_p, C.c_int]
        L.orc_scatter.restype = C.c_int
        L.orc_scatter.argtypes = [I32P, C.c_int, I32P, C.c_int, I32P]
        for name, fp in (("orc_kernel_xmm_f64", F64P), ("orc_kernel_lapack_f64", F64P)):
            f = getattr(L, name)
            f.restype = C.c_int
            f.argtypes = [C.c_char_p, I32P, C.c_int, I32P, C.c_int, I32P, C.c_int, C.c_double, C.c_double,
                          fp, fp, fp, C.c_int64, C.c_char_p, C.c_int]
        L.orc_kernel_xmm_f32.restype = C.c_int
        L.orc_kernel_xmm_f32.argtypes = [C.c_char_p, I32P, C.c_int, I32P, C.c_int, I32P, C.c_int, C.c_float,
                                         C.c_float, F32P, F32P, F32P, C.c_int64, C.c_char_p, C.c_int]
        L.orc_reshape_f64.restype = None
        L.orc_reshape_f64.argtypes = [F64P, I32P, C.c_int, I32P, F64P]
        L.orc_reserve.restype = C.c_int64
        L.orc_reserve.argtypes = [C.c_int, I32P, I32P, I32P, C.c_int64, I64P, I64P, I64P, C.c_char_p, C.c_int]
        L.orc_tasks.restype = C.c_int64
        L.orc_tasks.argtypes = [C.c_char_p, C.POINTER(Bst), C.POINTER(Bst), C.POINTER(Bst), C.c_void_p,
                                C.c_int64, C.c_void_p, C.POINTER(C.c_int64), C.c_void_p, C.c_char_p, C.c_int]
        L.orc_contract_f64.restype = C.c_int64
        L.orc_contract_f64.argtypes = [C.c_char_p, C.c_double, C.POINTER(Bst), C.POINTER(Bst), C.c_double,
                                       C.POINTER(Bst), C.c_int, C.c_void_p, C.c_int64, C.c_int, C.c_char_p,
                                       C.c_int]
        L.orc_contract_f64_t.restype = C.c_int64
        L.orc_contract_f64_t.argtypes = [C.c_char_p, C.c_double, C.POINTER(Bst), C.POINTER(Bst), C.c_double,
                                         C.POINTER(Bst), C.c_int, C.c_void_p, C.c_int64, C.c_int, C.c_char_p,
                                         C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.orc_block_norms_f64.restype = None
        L.orc_block_norms_f64.argtypes = [C.POINTER(Bst), C.c_int, F64P]
        L.orc_splitmix64.restype = C.c_uint64
        L.orc_splitmix64.argtypes = [C.c_uint64]
        L.orc_gen_block_sizes.restype = C.c_int
        L.orc_gen_block_sizes.argtypes = [C.c_uint64, C.c_int, C.c_int, I32P, C.c_int, I32P, C.c_int]
        L.orc_gen_presence.restype = C.c_int64
        L.orc_gen_presence.argtypes = [C.c_uint64, C.c_int, C.c_int, I32P, C.c_double, C.c_void_p, C.c_int64]
        L.orc_gen_values_f64.restype = None
        L.orc_gen_values_f64.argtypes = [C.c_uint64, C.c_int, C.POINTER(Bst), C.c_double, C.c_void_p, C.c_int64]
        _LIB = L
    return _LIB


class OracleError(RuntimeError):
    pass


def _i32(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.int32))


def contract_info(expr, sa, sb, sc):
    """Maps (a_row, a_col, b_row, b_col, c_row, c_col) or raises OracleError."""
    out = np.zeros(6 * 17, dtype=np.int32)
    err = C.create_string_buffer(512)
    sa, sb, sc = _i32(sa), _i32(sb), _i32(sc)
    ok = lib().orc_contract_info(expr.encode(), sa, len(sa), sb, len(sb), sc, len(sc), out, err, 512)
    if not ok:
        raise OracleError(err.value.decode())
    names = ("a_row", "a_col", "b_row", "b_col", "c_row", "c_col")
    return {n: [int(v) for v in out[r * 17 + 1:r * 17 + 1 + out[r * 17]]] for r, n in enumerate(names)}


def scatter(map_, sizes):
    map_, sizes = _i32(map_), _i32(sizes)
    n = int(np.prod([sizes[a] for a in map_])) if len(map_) else 1
    out = np.zeros(max(n, 1), dtype=np.int32)
    k = lib().orc_scatter(map_, len(map_), sizes, len(sizes), out)
    return out[:k]


def kernel_xmm(expr, sa, sb, sc, alpha, beta, a, b, c, nb_ops):
    """XMM-chain semantics on packed same-shape ops; c updated in place."""
    sa, sb, sc = _i32(sa), _i32(sb), _i32(sc)
    err = C.create_string_buffer(512)
    if c.dtype == np.float64:
        ok = lib().orc_kernel_xmm_f64(expr.encode(), sa, len(sa), sb, len(sb), sc, len(sc), alpha, beta, a, b,
                                      c, nb_ops, err, 512)
    else:
        ok = lib().orc_kernel_xmm_f32(expr.encode(), sa, len(sa), sb, len(sb), sc, len(sc), alpha, beta, a, b,
                                      c, nb_ops, err, 512)
    if not ok:
        raise OracleError(err.value.decode())


def kernel_lapack(expr, sa, sb, sc, alpha, beta, a, b, c, nb_ops):
    sa, sb, sc = _i32(sa), _i32(sb), _i32(sc)
    err = C.create_string_buffer(512)
    ok = lib().orc_kernel_lapack_f64(expr.encode(), sa, len(sa), sb, len(sb), sc, len(sc), alpha, beta, a, b, c,
                                     nb_ops, err, 512)
    if not ok:
        raise OracleError(err.value.decode())


def reshape(x, sizes, order):
    sizes, order = _i32(sizes), _i32(order)
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros_like(x)
    lib().orc_reshape_f64(x, sizes, len(sizes), order, out)
    return out


class BlockSparse:
    """Host block-sparse tensor in the reference storage format."""

    def __init__(self, block_sizes, dtype=np.float64):
        self.block_sizes = [np.asarray(b, dtype=np.int32) for b in block_sizes]
        self.N = len(block_sizes)
        self.nblk = _i32([len(b) for b in block_sizes])
        self.bsizes = _i32(np.concatenate(self.block_sizes))
        self.dtype = dtype
        self.indices = np.zeros(0, dtype=np.int64)
        self.offsets = np.zeros(0, dtype=np.int64)
        self.rows = np.zeros(len(block_sizes[0]), dtype=np.int64)
        self.data = np.zeros(0, dtype=dtype)
        self.bstride = np.cumprod([1] + [len(b) for b in block_sizes[:-1]]).astype(np.int64)

    def reserve(self, idx_per_mode):
        idx = _i32(np.asarray(idx_per_mode, dtype=np.int32).reshape(self.N, -1))
        nb = idx.shape[1]
        ind = np.zeros(nb, dtype=np.int64)
        off = np.zeros(nb, dtype=np.int64)
        rows = np.zeros(self.N and len(self.block_sizes[0]), dtype=np.int64)
        err = C.create_string_buffer(512)
        tot = lib().orc_reserve(self.N, self.nblk, self.bsizes, idx.ravel(), nb, ind, off, rows, err, 512)
        if tot < 0:
            raise OracleError(err.value.decode())
        self.indices, self.offsets, self.rows = ind, off, rows
        self.data = np.zeros(int(tot), dtype=self.dtype)
        return self

    def reserve_ids(self, ids):
        ids = np.asarray(ids, dtype=np.int64)
        coords = [(ids // s) % n for s, n in zip(self.bstride, self.nblk)]
        return self.reserve(np.stack(coords) if len(ids) else np.zeros((self.N, 0), np.int32))

    def coords(self, bid):
        return [int((bid // s) % n) for s, n in zip(self.bstride, self.nblk)]

    def dims(self, bid):
        return [int(self.block_sizes[d][c]) for d, c in enumerate(self.coords(bid))]

    def block(self, pos):
        bid = int(self.indices[pos])
        dims = self.dims(bid)
        n = int(np.prod(dims))
        o = int(self.offsets[pos])
        return self.data[o:o + n].reshape(dims, order="F")

    def struct(self):
        self._keep = (self.nblk, self.bsizes, self.indices, self.offsets, self.data)
        return Bst(self.N, self.nblk.ctypes.data, self.bsizes.ctypes.data, len(self.indices),
                   self.indices.ctypes.data, self.offsets.ctypes.data, self.data.ctypes.data)

    def dense(self):
        shape = [int(np.sum(b)) for b in self.block_sizes]
        starts = [np.concatenate([[0], np.cumsum(b)[:-1]]) for b in self.block_sizes]
        out = np.zeros(shape, dtype=self.data.dtype)
        for p in range(len(self.indices)):
            c = self.coords(int(self.indices[p]))
            sl = tuple(slice(int(starts[d][c[d]]), int(starts[d][c[d]] + self.block_sizes[d][c[d]]))
                       for d in range(self.N))
            out[sl] = self.block(p)
        return out


def tasks(expr, A, B, Cpat):
    """(tasks[nt,3] positions (a,b,c), kkeys[nt], C pattern ids).

    Cpat is a BlockSparse with C's block sizes; when it holds no blocks the
    pattern is derived from the product, otherwise it is kept."""
    sa, sb = A.struct(), B.struct()
    sc = Cpat.struct()
    err = C.create_string_buffer(512)
    nc = C.c_int64(0)
    nt = lib().orc_tasks(expr.encode(), C.byref(sa), C.byref(sb), C.byref(sc), None, 0, None, C.byref(nc), None,
                         err, 512)
    if nt < 0:
        raise OracleError(err.value.decode())
    t = np.zeros((max(nt, 1), 3), dtype=np.int64)
    kk = np.zeros(max(nt, 1), dtype=np.int64)
    cid = np.zeros(max(nc.value, 1), dtype=np.int64)
    nt2 = lib().orc_tasks(expr.encode(), C.byref(sa), C.byref(sb), C.byref(sc), t.ctypes.data, nt,
                          cid.ctypes.data, C.byref(nc), kk.ctypes.data, err, 512)
    assert nt2 == nt
    return t[:nt], kk[:nt], cid[:nc.value]


def contract(expr, alpha, A, B, beta, Cten, mode=0, subset=None, nthreads=1, timing=None):
    """Compute C in place (pattern must be reserved).  Returns task count.
    timing (a dict, optional) receives "plan_s" (task list + C pattern) and
    "compute_s" (block products) wall times."""
    sa, sb, sc = A.struct(), B.struct(), Cten.struct()
    err = C.create_string_buffer(512)
    sub = None
    nsub = 0
    if subset is not None:
        subset = np.ascontiguousarray(subset, dtype=np.int64)
        sub, nsub = subset.ctypes.data, len(subset)
    ps, cs = C.c_double(0.0), C.c_double(0.0)
    r = lib().orc_contract_f64_t(expr.encode(), alpha, C.byref(sa), C.byref(sb), beta, C.byref(sc), mode, sub, nsub,
                                 nthreads, err, 512, C.byref(ps), C.byref(cs))
    if r < 0:
        raise OracleError(err.value.decode())
    if timing is not None:
        timing["plan_s"] = ps.value
        timing["compute_s"] = cs.value
    return r


def block_norms(T, norm=0):
    out = np.zeros(max(len(T.indices), 1), dtype=np.float64)
    s = T.struct()
    lib().orc_block_norms_f64(C.byref(s), norm, out)
    return out[:len(T.indices)]


def gen_block_sizes(seed, axis, length, choices):
    choices = _i32(choices)
    out = np.zeros(length, dtype=np.int32)
    n = lib().orc_gen_block_sizes(seed, axis, length, choices, len(choices), out, length)
    return out[:n]


def gen_presence(seed, tid, nblk, p):
    nblk = _i32(nblk)
    n = lib().orc_gen_presence(seed, tid, len(nblk), nblk, p, None, 0)
    out = np.zeros(max(n, 1), dtype=np.int64)
    lib().orc_gen_presence(seed, tid, len(nblk), nblk, p, out.ctypes.data, n)
    return out[:n]


def gen_values(seed, tid, T, lam, positions=None):
    s = T.struct()
    if positions is None:
        lib().orc_gen_values_f64(seed, tid, C.byref(s), lam, None, len(T.indices))
    else:
        positions = np.ascontiguousarray(positions, dtype=np.int64)
        lib().orc_gen_values_f64(seed, tid, C.byref(s), lam, positions.ctypes.data, len(positions))
