"""CPU checks of the product library: it loads, exports every entry point
include/sht.h declares, and its host logic (expression analysis, arena,
generator) matches the golden vectors / oracle without touching a GPU."""
import ctypes as C
import json
import os
import re
import subprocess

import numpy as np
import pytest

import shtensor_amd as sht
from shtensor_amd import _lib
from oracle import oracle as orc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "sht.h")
GOLD = json.load(open(os.path.join(ROOT, "tests", "golden", "contract_info.json")))["cases"]


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sht_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r" T (sht_[a-z0-9_]+)$", out, flags=re.M))
    missing = [s for s in declared_symbols() if s not in exported]
    assert not missing, missing
    L = _lib.lib()
    for s in declared_symbols():
        assert hasattr(L, s)
    assert b"gfx950" in L.sht_version()


def test_python_binding_covers_header():
    bound = set(_lib.SIGNATURES) | {"sht_last_error", "sht_version", "sht_kclass_name", "sht_cs_variant_name"}
    assert set(declared_symbols()) <= bound


DIVERGENT = {"ij,jk->il"}


@pytest.mark.parametrize("case", GOLD, ids=[c["expr"] or "<empty>" for c in GOLD])
def test_product_contract_info_matches_reference(case):
    expr = case["expr"]
    if expr in DIVERGENT:
        with pytest.raises(sht.ShtError, match="Dimension mismatch"):
            sht.contract_info(expr, case["sa"], case["sb"], case["sc"])
        return
    if not case["ok"]:
        with pytest.raises(sht.ShtError) as ei:
            sht.contract_info(expr, case["sa"], case["sb"], case["sc"])
        assert str(ei.value) == case["err"]
        assert ei.value.code == _lib.ERR_EXPR
        return
    m = sht.contract_info(expr, case["sa"], case["sb"], case["sc"])
    for k in ("a_row", "a_col", "b_row", "b_col", "c_row", "c_col"):
        assert m[k] == case[k], k


def test_arena_accounting_mempool_semantics():
    # test/test_mempool.cpp:15-114 re-derived for out-of-band headers and
    # 256-byte granularity: free memory counts payload only.
    A = 256
    total = 1024 * 8 + A
    ctx = sht.Context(total, host=True)
    try:
        assert ctx.arena_info()["free"] == total
        p0 = ctx.allocate(100 * 8)   # 800 -> 1024
        p1 = ctx.allocate(100 * 4)   # 400 -> 512
        assert ctx.arena_info()["free"] == total - 1024 - 512
        ctx.free(p0)
        assert ctx.arena_info()["free"] == total - 512
        p2 = ctx.allocate(100 * 8)
        assert p2 == p0  # first fit picks the freed chunk again (test_mempool.cpp:62-65)
        ctx.free(p2)
        ctx.free(p1)
        info = ctx.arena_info()
        assert info["free"] == total and info["largest_free"] == total  # fully coalesced
        q = ctx.allocate(total)
        assert ctx.arena_info()["free"] == 0
        q = ctx.reallocate(q, 256)          # shrink in place
        assert q == p0 and ctx.arena_info()["free"] == total - 256
        q2 = ctx.reallocate(q, 512)         # grow in place into the free successor
        assert q2 == q and ctx.arena_info()["free"] == total - 512
        q3 = ctx.reallocate(q2, total)      # grow to the whole pool
        assert q3 == q and ctx.arena_info()["free"] == 0
        with pytest.raises(sht.ShtError, match="Out of memory") as ei:
            ctx.allocate(1)
        assert ei.value.code == _lib.ERR_OOM
        ctx.free(q3)
        # move path: a live successor blocks in-place growth
        a = ctx.allocate(256)
        b = ctx.allocate(256)
        a2 = ctx.reallocate(a, 1024)
        assert a2 != a and ctx.size_of(a2) == 1024
        ctx.free(b)
        ctx.free(a2)
        assert ctx.arena_info()["free"] == total
        with pytest.raises(sht.ShtError):
            ctx.free(a2)  # double free is rejected
    finally:
        ctx.close()


def test_arena_first_fit_fragmentation():
    ctx = sht.Context(16 * 256, host=True)
    try:
        ps = [ctx.allocate(256) for _ in range(16)]
        for p in ps[::2]:
            ctx.free(p)
        info = ctx.arena_info()
        assert info["free"] == 8 * 256 and info["largest_free"] == 256
        with pytest.raises(sht.ShtError):
            ctx.allocate(512)
        ctx.free(ps[1])  # coalesces 0,1,2
        assert ctx.arena_info()["largest_free"] == 3 * 256
        assert ctx.allocate(700) == ps[0]
    finally:
        ctx.close()


def test_tensor_needs_device_context():
    ctx = sht.Context(1 << 20, host=True)
    try:
        with pytest.raises(sht.ShtError, match="device context"):
            sht.Tensor(ctx, [[2, 2], [3]])
    finally:
        ctx.close()


@pytest.mark.parametrize("axis", [0, 1, 3])
def test_generator_block_sizes_match_oracle(axis):
    choices = [5, 7, 9, 11, 13, 15, 21, 28]
    assert sht.gen_block_sizes(1001, axis, 600, choices) == orc.gen_block_sizes(1001, axis, 600, choices).tolist()
