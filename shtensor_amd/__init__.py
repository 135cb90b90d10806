"""shtensor_amd — MI355X-native block-sparse tensor contraction.

Host-side mirror of shtensor's interfaces over the C ABI (include/sht.h):

* ``Kernel``  — ``Shtensor::Kernel<T>`` (src/Kernel/Kernel.h:62-93): construct
  from an einsum expression and block sizes, ``call(a, b, c, nb_ops)`` on
  device buffers, ``get_info()``.
* ``Tensor``  — ``Shtensor::Tensor<T,N>`` (src/Tensor/Tensor.h:19-296):
  ``reserve``, ``reserve_all``, ``get_sparse_info``, block access, ``filter``,
  ``compress``, counts and occupation.
* ``Context`` — ``Shtensor::Context`` + ``MemoryPool`` (src/Shmem): one GPU
  and its HBM arena.
* ``contract`` / ``Plan`` — the block-sparse contraction driver.

Every call goes to libshtensor_amd.so; nothing here computes on the CPU.
Errors raise ``ShtError`` carrying the library's message (the reference's
``std::runtime_error`` text where it has one).
"""
import ctypes as C
import weakref

import numpy as np

from . import _lib
from ._lib import (FLOAT32, FLOAT64, FROBENIUS, LAPACK, MAXABS, PREC_BF16X3, PREC_FP32, PREC_FP64,
                   PREC_FP64_EXACT, XMM, ShtError, call, check, lib)

__all__ = ["Context", "DeviceArray", "Kernel", "Tensor", "Plan", "Server", "contract", "contract_info", "ShtError",
           "FLOAT32", "FLOAT64", "LAPACK", "XMM", "FROBENIUS", "MAXABS", "PREC_FP64", "PREC_FP64_EXACT",
           "PREC_FP32", "PREC_BF16X3", "gen_block_sizes", "Stream"]

_DT = {np.dtype(np.float64): FLOAT64, np.dtype(np.float32): FLOAT32}
_NP = {FLOAT64: np.float64, FLOAT32: np.float32}


def _ints(v):
    a = np.ascontiguousarray(np.asarray(list(v), dtype=np.int32))
    return a, a.ctypes.data_as(C.POINTER(C.c_int))


def contract_info(expr, sizes_a, sizes_b, sizes_c):
    """compute_contract_info (src/Basics/ContractInfo.h:68-346): the six maps."""
    a, pa = _ints(sizes_a)
    b, pb = _ints(sizes_b)
    c, pc = _ints(sizes_c)
    out = np.zeros(6 * 17, dtype=np.int32)
    call("sht_contract_info", expr.encode(), pa, len(a), pb, len(b), pc, len(c),
         out.ctypes.data_as(C.POINTER(C.c_int)))
    names = ("a_row", "a_col", "b_row", "b_col", "c_row", "c_col")
    return {n: [int(x) for x in out[r * 17 + 1:r * 17 + 1 + out[r * 17]]] for r, n in enumerate(names)}


def gen_block_sizes(seed, axis, length, choices):
    ch, pch = _ints(choices)
    out = np.zeros(max(length, 1), dtype=np.int32)
    n = C.c_int(0)
    call("sht_gen_block_sizes", seed, axis, length, pch, len(ch), out.ctypes.data_as(C.POINTER(C.c_int)), len(out),
         C.byref(n))
    return out[:n.value].tolist()


class Stream:
    """A HIP stream (non-blocking, default priority)."""

    def __init__(self):
        p = C.c_void_p()
        call("sht_stream_create", C.byref(p))
        self.ptr = p.value

    def synchronize(self):
        call("sht_stream_synchronize", C.c_void_p(self.ptr))

    def close(self):
        if self.ptr:
            call("sht_stream_destroy", C.c_void_p(self.ptr))
            self.ptr = None

    __del__ = close


def _sptr(stream):
    return C.c_void_p(stream.ptr if isinstance(stream, Stream) else stream)


class DeviceArray:
    """Plain device allocation (hipMalloc) holding a numpy-shaped array."""

    def __init__(self, host=None, shape=None, dtype=np.float64):
        if host is not None:
            host = np.ascontiguousarray(host)
            self.shape, self.dtype = host.shape, host.dtype
        else:
            self.shape, self.dtype = tuple(shape), np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        call("sht_dev_alloc", max(self.nbytes, 1), C.byref(p))
        self.ptr = p.value
        if host is not None:
            self.upload(host)

    def upload(self, host):
        host = np.ascontiguousarray(host, dtype=self.dtype)
        assert host.nbytes == self.nbytes
        call("sht_copy_h2d", C.c_void_p(self.ptr), host.ctypes.data_as(C.c_void_p), self.nbytes)

    def numpy(self):
        out = np.empty(self.shape, dtype=self.dtype)
        call("sht_copy_d2h", out.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr), self.nbytes)
        return out

    def free(self):
        if self.ptr:
            call("sht_dev_free", C.c_void_p(self.ptr))
            self.ptr = None

    __del__ = free


class Context:
    """One GPU + its HBM arena (Context.cpp:9-48, MemoryPool)."""

    def __init__(self, arena_bytes=1 << 30, device=0, host=False):
        p = C.c_void_p()
        if host:
            call("sht_ctx_create_host", arena_bytes, C.byref(p))
        else:
            call("sht_ctx_create", arena_bytes, device, C.byref(p))
        self.ptr = p.value
        self._children = weakref.WeakSet()  # tensors and plans: closed before the context

    def arena_info(self):
        t, f, l = C.c_size_t(), C.c_size_t(), C.c_size_t()
        call("sht_ctx_arena_info", C.c_void_p(self.ptr), C.byref(t), C.byref(f), C.byref(l))
        return {"total": t.value, "free": f.value, "largest_free": l.value}

    def allocate(self, nbytes):
        p = C.c_void_p()
        call("sht_arena_alloc", C.c_void_p(self.ptr), nbytes, C.byref(p))
        return p.value

    def reallocate(self, ptr, nbytes):
        p = C.c_void_p()
        call("sht_arena_realloc", C.c_void_p(self.ptr), C.c_void_p(ptr), nbytes, C.byref(p))
        return p.value

    def free(self, ptr):
        call("sht_arena_free", C.c_void_p(self.ptr), C.c_void_p(ptr))

    def size_of(self, ptr):
        s = C.c_size_t()
        call("sht_arena_size", C.c_void_p(self.ptr), C.c_void_p(ptr), C.byref(s))
        return s.value

    def close(self):
        if self.ptr:
            for kind in (Server, Plan, Tensor):
                for child in list(self._children):
                    if isinstance(child, kind):
                        child.close()
            call("sht_ctx_destroy", C.c_void_p(self.ptr))
            self.ptr = None

    __del__ = close

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class Kernel:
    """Shtensor::Kernel<T> (Kernel.h:62-93) on the GPU."""

    def __init__(self, expr, sizes_a, sizes_b, sizes_c, alpha=1.0, beta=0.0, kernel_type=XMM, dtype=np.float64):
        self.dtype = np.dtype(dtype)
        a, pa = _ints(sizes_a)
        b, pb = _ints(sizes_b)
        c, pc = _ints(sizes_c)
        p = C.c_void_p()
        call("sht_kernel_create", expr.encode(), pa, len(a), pb, len(b), pc, len(c), float(alpha), float(beta),
             _DT[self.dtype], kernel_type, C.byref(p))
        self.ptr = p.value

    def call(self, a, b, c, nb_ops, stream=None):
        """a, b, c: DeviceArray, raw device pointers, or contiguous numpy
        arrays of the kernel's dtype (host memory: the library stages them
        through device buffers and returns after C is written back)."""
        def ptr(x):
            if isinstance(x, DeviceArray):
                return C.c_void_p(x.ptr)
            if isinstance(x, np.ndarray):
                assert x.dtype == self.dtype and x.flags["C_CONTIGUOUS"], "host operand dtype/layout"
                return C.c_void_p(x.ctypes.data)
            return C.c_void_p(x)
        call("sht_kernel_call", C.c_void_p(self.ptr), ptr(a), ptr(b), ptr(c), int(nb_ops), _sptr(stream))

    def get_info(self):
        buf = C.create_string_buffer(1 << 16)
        call("sht_kernel_info", C.c_void_p(self.ptr), buf, len(buf))
        return buf.value.decode()

    def close(self):
        if getattr(self, "ptr", None):
            call("sht_kernel_destroy", C.c_void_p(self.ptr))
            self.ptr = None

    __del__ = close


class Tensor:
    """Shtensor::Tensor<T,N> (Tensor.h:19-296) with data in the HBM arena."""

    def __init__(self, ctx, block_sizes, dtype=np.float64):
        self.ctx = ctx
        self.dtype = np.dtype(dtype)
        self.block_sizes = [list(map(int, b)) for b in block_sizes]
        self.N = len(block_sizes)
        self._bs = [np.ascontiguousarray(np.asarray(b, dtype=np.int32)) for b in self.block_sizes]
        arr = (C.POINTER(C.c_int) * self.N)(*[b.ctypes.data_as(C.POINTER(C.c_int)) for b in self._bs])
        nb, pnb = _ints([len(b) for b in self.block_sizes])
        p = C.c_void_p()
        call("sht_tensor_create", C.c_void_p(ctx.ptr), _DT[self.dtype], self.N, arr, pnb, C.byref(p))
        self.ptr = p.value
        ctx._children.add(self)
        self.nblk = [len(b) for b in self.block_sizes]
        self.bstride = np.cumprod([1] + self.nblk[:-1]).astype(np.int64)

    # --- reservation (BaseTensor.cpp:97-355)
    def reserve(self, idx_per_mode):
        idx = [np.ascontiguousarray(np.asarray(x, dtype=np.int32)) for x in idx_per_mode]
        if len(idx) != self.N or len({len(x) for x in idx}) > 1:
            raise ShtError(_lib.ERR_INVALID, "reserve: Indices do not have the same size in each dimension")
        arr = (C.POINTER(C.c_int) * self.N)(*[x.ctypes.data_as(C.POINTER(C.c_int)) for x in idx])
        call("sht_tensor_reserve", C.c_void_p(self.ptr), arr, len(idx[0]))

    def reserve_ids(self, ids):
        ids = np.ascontiguousarray(np.asarray(ids, dtype=np.int64))
        call("sht_tensor_reserve_ids", C.c_void_p(self.ptr), ids.ctypes.data_as(C.POINTER(C.c_int64)), len(ids))

    def reserve_all(self):
        call("sht_tensor_reserve_all", C.c_void_p(self.ptr))

    def reserve_random(self, seed, tid, p):
        call("sht_gen_reserve", C.c_void_p(self.ptr), seed, tid, float(p))

    def fill_random(self, seed, tid, lam=0.0, stream=None):
        call("sht_gen_fill", C.c_void_p(self.ptr), seed, tid, float(lam), _sptr(stream))

    # --- sparse info / counts
    def get_sparse_info(self):
        ind, off, rows = C.POINTER(C.c_int64)(), C.POINTER(C.c_int64)(), C.POINTER(C.c_int64)()
        n = C.c_int64()
        call("sht_tensor_sparse_info", C.c_void_p(self.ptr), C.byref(ind), C.byref(off), C.byref(rows), C.byref(n))
        k = n.value
        cp = lambda p, m: np.ctypeslib.as_array(p, shape=(m,)).copy() if m else np.zeros(0, np.int64)
        return {"indices": cp(ind, k), "offsets": cp(off, k), "rows": cp(rows, self.nblk[0])}

    def get_indices(self):
        return self.get_sparse_info()["indices"]

    def get_nb_nzblocks_local(self):
        l, g = C.c_int64(), C.c_int64()
        call("sht_tensor_nb_nzblocks", C.c_void_p(self.ptr), C.byref(l), C.byref(g))
        return l.value

    def get_nb_nzblocks_global(self):
        l, g = C.c_int64(), C.c_int64()
        call("sht_tensor_nb_nzblocks", C.c_void_p(self.ptr), C.byref(l), C.byref(g))
        return g.value

    def get_occupation(self):
        o = C.c_double()
        call("sht_tensor_occupation", C.c_void_p(self.ptr), C.byref(o))
        return o.value

    def nelem(self):
        p, n = C.c_void_p(), C.c_int64()
        call("sht_tensor_data", C.c_void_p(self.ptr), C.byref(p), C.byref(n))
        return n.value

    def data_ptr(self):
        p, n = C.c_void_p(), C.c_int64()
        call("sht_tensor_data", C.c_void_p(self.ptr), C.byref(p), C.byref(n))
        return p.value

    # --- data movement
    def upload(self, host):
        host = np.ascontiguousarray(host, dtype=self.dtype)
        call("sht_tensor_upload", C.c_void_p(self.ptr), host.ctypes.data_as(C.c_void_p), len(host))

    def download(self):
        out = np.zeros(self.nelem(), dtype=self.dtype)
        call("sht_tensor_download", C.c_void_p(self.ptr), out.ctypes.data_as(C.c_void_p), len(out))
        return out

    # --- blocks (Tensor.h:38-52; dims from the block id)
    def coords(self, bid):
        return [int((bid // s) % n) for s, n in zip(self.bstride, self.nblk)]

    def block_dims(self, bid):
        return [self.block_sizes[d][c] for d, c in enumerate(self.coords(bid))]

    def blocks(self, data=None):
        """Yield (position, block id, column-major block array) from a download."""
        info = self.get_sparse_info()
        data = self.download() if data is None else data
        for p, (bid, off) in enumerate(zip(info["indices"], info["offsets"])):
            b = int(bid) if bid >= 0 else -int(bid) - 1
            dims = self.block_dims(b)
            n = int(np.prod(dims))
            yield p, int(bid), data[int(off):int(off) + n].reshape(dims, order="F")

    def block_norms(self, norm=FROBENIUS):
        n = self.get_nb_nzblocks_local()
        out = DeviceArray(shape=(max(n, 1),), dtype=np.float64)
        call("sht_tensor_block_norms", C.c_void_p(self.ptr), norm, C.c_void_p(out.ptr), None)
        call("sht_device_synchronize")
        return out.numpy()[:n]

    def converted(self, dtype, stream=None):
        """New tensor of the same block pattern with this tensor's values in
        `dtype` (sht_tensor_convert; FP64 -> FP32 rounds to nearest)."""
        out = Tensor(self.ctx, self.block_sizes, dtype=dtype)
        call("sht_tensor_convert", C.c_void_p(self.ptr), C.c_void_p(out.ptr), _sptr(stream))
        return out

    def filter(self, eps, norm=FROBENIUS, compress=True):
        call("sht_tensor_filter", C.c_void_p(self.ptr), float(eps), norm, 1 if compress else 0)

    def compress(self):
        call("sht_tensor_compress", C.c_void_p(self.ptr))

    def close(self):
        if getattr(self, "ptr", None):
            call("sht_tensor_destroy", C.c_void_p(self.ptr))
            self.ptr = None

    __del__ = close


class Plan:
    """A built contraction: GPU task list, C pattern and schedule."""

    def __init__(self, ctx, expr, A, B, Cten, precision=PREC_FP64):
        p = C.c_void_p()
        call("sht_plan_create", C.c_void_p(ctx.ptr), expr.encode(), C.c_void_p(A.ptr), C.c_void_p(B.ptr),
             C.c_void_p(Cten.ptr), precision, C.byref(p))
        self.ptr = p.value
        self._keep = (A, B, Cten)
        ctx._children.add(self)

    def execute(self, alpha=1.0, beta=0.0, stream=None):
        call("sht_plan_execute", C.c_void_p(self.ptr), float(alpha), float(beta), _sptr(stream))

    def graph_replay(self, alpha, beta, replays, stream):
        call("sht_plan_graph_replay", C.c_void_p(self.ptr), float(alpha), float(beta), int(replays), _sptr(stream))

    def execute_timed(self, alpha=1.0, beta=0.0, stream=None):
        """One execution with events around every grouped launch:
        list of (kernel class name, ms)."""
        kc = (C.c_int * 64)()
        ms = (C.c_float * 64)()
        n = C.c_int()
        call("sht_plan_execute_timed", C.c_void_p(self.ptr), float(alpha), float(beta), _sptr(stream), kc, ms, 64,
             C.byref(n))
        return [(lib().sht_kclass_name(kc[i]).decode(), float(ms[i])) for i in range(min(n.value, 64))]

    def set_schedule(self, schedule):
        """0: C-stationary tiles; 1: A-stationary, one-wave units; 2: A-stationary,
        four units of a block row per workgroup; 3: A-stationary, narrow sweeps
        (FP64 / FP32)."""
        call("sht_plan_set_schedule", C.c_void_p(self.ptr), int(schedule))

    def get_schedule(self):
        v = C.c_int()
        call("sht_plan_get_schedule", C.c_void_p(self.ptr), C.byref(v))
        return v.value

    def autotune(self, alpha=1.0, beta=0.0, reps=2, stream=None):
        """Time every applicable schedule, keep the fastest; {schedule: ms}."""
        ms = (C.c_float * 4)(-1.0, -1.0, -1.0, -1.0)
        call("sht_plan_autotune", C.c_void_p(self.ptr), float(alpha), float(beta), int(reps), _sptr(stream), ms)
        return {i: float(ms[i]) for i in range(4) if ms[i] >= 0}

    def tasks(self):
        ptrs = [C.POINTER(C.c_int64)() for _ in range(4)]
        n = C.c_int64()
        call("sht_plan_tasks", C.c_void_p(self.ptr), *[C.byref(x) for x in ptrs], C.byref(n))
        k = n.value
        arr = [np.ctypeslib.as_array(x, shape=(k,)).copy() if k else np.zeros(0, np.int64) for x in ptrs]
        return {"a": arr[0], "b": arr[1], "c": arr[2], "key": arr[3]}

    def autotune_ex(self, alpha=1.0, beta=0.0, reps=2, stream=None):
        """Time every (schedule, C-stationary tile variant) and keep the
        fastest; returns [(schedule, variant name, ms), ...]."""
        cap = 16
        n = C.c_int()
        sc, va = (C.c_int * cap)(), (C.c_int * cap)()
        ms = (C.c_float * cap)()
        call("sht_plan_autotune_ex", C.c_void_p(self.ptr), float(alpha), float(beta), int(reps), _sptr(stream), cap,
             C.byref(n), sc, va, ms)
        name = lambda v: lib().sht_cs_variant_name(v).decode()
        return [(sc[i], name(va[i]), ms[i]) for i in range(min(n.value, cap))]

    def set_variant(self, variant):
        call("sht_plan_set_variant", C.c_void_p(self.ptr), int(variant))

    def get_variant(self):
        v = C.c_int()
        call("sht_plan_get_variant", C.c_void_p(self.ptr), C.byref(v))
        return v.value

    def set_fp32_staging(self, on):
        """FP32 / split-BF16 plans over FP64 operands: convert A and B to
        FP32 once per execution (True) or round them at every use (False);
        bitwise the same."""
        call("sht_plan_set_fp32_staging", C.c_void_p(self.ptr), 1 if on else 0)

    def get_fp32_staging(self):
        v = C.c_int()
        call("sht_plan_get_fp32_staging", C.c_void_p(self.ptr), C.byref(v))
        return bool(v.value)

    def build_ms(self):
        """Host wall time of the build steps (ms): task list + C pattern,
        C-stationary tiles, A-stationary schedule (-1 if not built)."""
        t, c, a = C.c_double(), C.c_double(), C.c_double()
        call("sht_plan_build_ms", C.c_void_p(self.ptr), C.byref(t), C.byref(c), C.byref(a))
        return {"tasks": t.value, "cstat": c.value, "astat": a.value}

    def stats(self):
        f, b = C.c_double(), C.c_double()
        nt, nc = C.c_int64(), C.c_int64()
        call("sht_plan_stats", C.c_void_p(self.ptr), C.byref(f), C.byref(b), C.byref(nt), C.byref(nc))
        return {"flops": f.value, "bytes": b.value, "ntasks": nt.value, "ncblocks": nc.value}

    def close(self):
        if getattr(self, "ptr", None):
            call("sht_plan_destroy", C.c_void_p(self.ptr))
            self.ptr = None

    __del__ = close


class Server:
    """Continuous-batching request loop over one context (see sht_server_*)."""

    def __init__(self, ctx, budget_bytes, max_batch=64):
        p = C.c_void_p()
        call("sht_server_create", C.c_void_p(ctx.ptr), int(budget_bytes), int(max_batch), C.byref(p))
        self.ptr = p.value
        self._keep = {}
        ctx._children.add(self)

    def submit(self, expr, alpha, A, B, beta, Cten, precision=PREC_FP64):
        rid = C.c_int64()
        call("sht_server_submit", C.c_void_p(self.ptr), expr.encode(), float(alpha), C.c_void_p(A.ptr),
             C.c_void_p(B.ptr), float(beta), C.c_void_p(Cten.ptr), precision, C.byref(rid))
        self._keep[rid.value] = (A, B, Cten)
        return rid.value

    def wait(self, rid):
        """Latency in ms (raises ShtError if the request failed)."""
        lat = C.c_double()
        try:
            call("sht_server_wait", C.c_void_p(self.ptr), int(rid), C.byref(lat))
        finally:
            self._keep.pop(rid, None)
        return lat.value

    def poll(self, rid):
        d = C.c_int()
        call("sht_server_poll", C.c_void_p(self.ptr), int(rid), C.byref(d))
        return bool(d.value)

    def cancel(self, rid):
        """Cancel a request not yet launched (its wait then raises ShtError
        with code ERR_CANCELLED); raises ShtError(ERR_STATE) once launched."""
        call("sht_server_cancel", C.c_void_p(self.ptr), int(rid))

    def counters(self):
        nc, nr, iu = C.c_int64(), C.c_int64(), C.c_size_t()
        call("sht_server_counters", C.c_void_p(self.ptr), C.byref(nc), C.byref(nr), C.byref(iu))
        return {"cancelled": nc.value, "requeued": nr.value, "in_use": iu.value}

    def stats(self):
        v = [C.c_int64() for _ in range(4)]
        call("sht_server_stats", C.c_void_p(self.ptr), *[C.byref(x) for x in v])
        return dict(zip(("done", "batches", "max_merged", "launches"), [x.value for x in v]))

    def close(self):
        if getattr(self, "ptr", None):
            call("sht_server_destroy", C.c_void_p(self.ptr))
            self.ptr = None

    __del__ = close


def diag_mfma_f64_rate(stream=None):
    """TFLOP/s of a register-resident FP64 MFMA loop on this device (the
    matrix-core rate the contraction kernels can be compared with)."""
    v = C.c_double()
    call("sht_diag_mfma_f64_rate", _sptr(stream), C.byref(v))
    return v.value


def diag_step_rate(precision, slots, stream=None):
    """TFLOP/s of the grouped A-stationary kernel's step shape (A load, LDS
    panel fragments, 2 x slots MFMAs, a barrier every two steps) with every
    operand in cache: that loop's ceiling (PREC_FP64 or PREC_FP32)."""
    v = C.c_double()
    call("sht_diag_step_rate", int(precision), int(slots), _sptr(stream), C.byref(v))
    return v.value


def contract(ctx, expr, alpha, A, B, beta, Cten, precision=PREC_FP64, stream=None):
    """C = beta*C + alpha*contract(expr, A, B) on the GPU (synchronous)."""
    call("sht_contract", C.c_void_p(ctx.ptr), expr.encode(), float(alpha), C.c_void_p(A.ptr), C.c_void_p(B.ptr),
         float(beta), C.c_void_p(Cten.ptr), precision, _sptr(stream))
