"""ctypes binding of libshtensor_amd.so (the C ABI declared in include/sht.h).

The library must exist: there is no CPU fallback for any entry point.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libshtensor_amd.so")

OK, ERR_INVALID, ERR_EXPR, ERR_OOM, ERR_HIP, ERR_STATE, ERR_UNSUPPORTED, ERR_ADMISSION, ERR_CANCELLED = range(9)
FLOAT32, FLOAT64 = 0, 1
LAPACK, XMM = 0, 1
FROBENIUS, MAXABS = 0, 1
PREC_FP64, PREC_FP64_EXACT, PREC_FP32, PREC_BF16X3 = 0, 1, 2, 3

_P = C.c_void_p
_I = C.c_int
_I64 = C.c_int64
_SZ = C.c_size_t
_D = C.c_double
_PI = C.POINTER(C.c_int)
_PI64 = C.POINTER(C.c_int64)
_PPI = C.POINTER(C.POINTER(C.c_int))
_PP = C.POINTER(C.c_void_p)

# name: (argtypes)   — every function returns sht_status (int)
SIGNATURES = {
    "sht_ctx_create": [_SZ, _I, _PP],
    "sht_ctx_create_host": [_SZ, _PP],
    "sht_ctx_destroy": [_P],
    "sht_contract_info": [C.c_char_p, _PI, _I, _PI, _I, _PI, _I, _PI],
    "sht_ctx_arena_info": [_P, C.POINTER(_SZ), C.POINTER(_SZ), C.POINTER(_SZ)],
    "sht_arena_alloc": [_P, _SZ, _PP],
    "sht_arena_realloc": [_P, _P, _SZ, _PP],
    "sht_arena_free": [_P, _P],
    "sht_arena_size": [_P, _P, C.POINTER(_SZ)],
    "sht_dev_alloc": [_SZ, _PP],
    "sht_dev_free": [_P],
    "sht_copy_h2d": [_P, _P, _SZ],
    "sht_copy_d2h": [_P, _P, _SZ],
    "sht_device_synchronize": [],
    "sht_stream_create": [_PP],
    "sht_stream_destroy": [_P],
    "sht_stream_synchronize": [_P],
    "sht_diag_mfma_f64_rate": [_P, C.POINTER(C.c_double)],
    "sht_diag_step_rate": [C.c_int, C.c_int, _P, C.POINTER(C.c_double)],
    "sht_kernel_create": [C.c_char_p, _PI, _I, _PI, _I, _PI, _I, _D, _D, _I, _I, _PP],
    "sht_kernel_call": [_P, _P, _P, _P, _I64, _P],
    "sht_kernel_call_static": [_P, _P, _P, _P, _I64, _P],
    "sht_plan_build_ms": [_P, C.POINTER(_D), C.POINTER(_D), C.POINTER(_D)],
    "sht_server_cancel": [_P, _I64],
    "sht_plan_autotune_ex": [_P, _D, _D, _I, _P, _I, _PI, _PI, _PI, C.POINTER(C.c_float)],
    "sht_plan_set_variant": [_P, _I],
    "sht_plan_get_variant": [_P, _PI],
    "sht_plan_set_fp32_staging": [_P, _I],
    "sht_plan_get_fp32_staging": [_P, _PI],
    "sht_server_counters": [_P, _PI64, _PI64, C.POINTER(_SZ)],
    "sht_kernel_info": [_P, C.c_char_p, _SZ],
    "sht_kernel_destroy": [_P],
    "sht_kernel_prewarm": [C.c_char_p, _PI, _I, _PI, _I, _PI, _I, _D, _D, _I],
    "sht_tensor_create": [_P, _I, _I, _PPI, _PI, _PP],
    "sht_tensor_destroy": [_P],
    "sht_tensor_reserve": [_P, _PPI, _I64],
    "sht_tensor_reserve_ids": [_P, _PI64, _I64],
    "sht_tensor_reserve_all": [_P],
    "sht_tensor_sparse_info": [_P, C.POINTER(_PI64), C.POINTER(_PI64), C.POINTER(_PI64), _PI64],
    "sht_tensor_data": [_P, _PP, _PI64],
    "sht_tensor_nb_nzblocks": [_P, _PI64, _PI64],
    "sht_tensor_occupation": [_P, C.POINTER(_D)],
    "sht_tensor_upload": [_P, _P, _I64],
    "sht_tensor_download": [_P, _P, _I64],
    "sht_tensor_convert": [_P, _P, _P],
    "sht_tensor_block_norms": [_P, _I, _P, _P],
    "sht_tensor_filter": [_P, _D, _I, _I],
    "sht_tensor_compress": [_P],
    "sht_gen_block_sizes": [C.c_uint64, _I, _I, _PI, _I, _PI, _I, _PI],
    "sht_gen_reserve": [_P, C.c_uint64, _I, _D],
    "sht_gen_fill": [_P, C.c_uint64, _I, _D, _P],
    "sht_contract": [_P, C.c_char_p, _D, _P, _P, _D, _P, _I, _P],
    "sht_plan_create": [_P, C.c_char_p, _P, _P, _P, _I, _PP],
    "sht_plan_execute": [_P, _D, _D, _P],
    "sht_plan_graph_replay": [_P, _D, _D, _I, _P],
    "sht_plan_execute_timed": [_P, _D, _D, _P, _PI, C.POINTER(C.c_float), _I, _PI],
    "sht_plan_set_schedule": [_P, _I],
    "sht_plan_get_schedule": [_P, _PI],
    "sht_plan_autotune": [_P, _D, _D, _I, _P, C.POINTER(C.c_float)],
    "sht_plan_tasks": [_P, C.POINTER(_PI64), C.POINTER(_PI64), C.POINTER(_PI64), C.POINTER(_PI64), _PI64],
    "sht_plan_stats": [_P, C.POINTER(_D), C.POINTER(_D), _PI64, _PI64],
    "sht_plan_destroy": [_P],
    "sht_server_create": [_P, _SZ, _I, _PP],
    "sht_server_submit": [_P, C.c_char_p, _D, _P, _P, _D, _P, _I, _PI64],
    "sht_server_wait": [_P, _I64, C.POINTER(_D)],
    "sht_server_poll": [_P, _I64, _PI],
    "sht_server_stats": [_P, _PI64, _PI64, _PI64, _PI64],
    "sht_server_destroy": [_P],
}

_lib = None


class ShtError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libshtensor_amd.so is not built: run `make -C shtensor_amd/csrc` "
                               "(or __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = C.c_int
        L.sht_last_error.restype = C.c_char_p
        L.sht_last_error.argtypes = []
        L.sht_version.restype = C.c_char_p
        L.sht_version.argtypes = []
        L.sht_kclass_name.restype = C.c_char_p
        L.sht_kclass_name.argtypes = [C.c_int]
        L.sht_cs_variant_name.restype = C.c_char_p
        L.sht_cs_variant_name.argtypes = [C.c_int]
        _lib = L
    return _lib


def check(status):
    if status != OK:
        raise ShtError(status, lib().sht_last_error().decode(errors="replace"))


def call(name, *args):
    check(getattr(lib(), name)(*args))
