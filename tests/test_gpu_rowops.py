"""The row-format join (pgpu_hj_*) through the C ABI on a context with an HBM
budget: the output multiset is byte-exact against a numpy equi-join on two
queries in a row (the second reuses the context's pooled streams, pinned
words and device blocks), a probe whose output outgrows its first buffer
takes the regrowth path, and blocks kept for reuse never stand in the way of
an allocation within the budget."""
import ctypes as C

import numpy as np
import pytest

import pythia_amd as pa
from pythia_amd import _lib

pytestmark = pytest.mark.gpu


class Proj(C.Structure):
    _fields_ = [("side", C.c_uint32), ("src_offset", C.c_uint32), ("bytes", C.c_uint32)]


class HjDesc(C.Structure):
    _fields_ = [("build_row_bytes", C.c_uint32), ("build_key_offset", C.c_uint32), ("build_key_bytes", C.c_uint32),
                ("probe_row_bytes", C.c_uint32), ("probe_key_offset", C.c_uint32), ("probe_key_bytes", C.c_uint32),
                ("nproj", C.c_uint32), ("proj", Proj * 32), ("slots", C.c_uint32)]


def _fn(name, args):
    f = getattr(_lib.lib(), name)
    f.restype = C.c_int
    f.argtypes = args
    return f


def _ok(rc, what):
    assert rc == 0, f"{what}: {_lib.lib().pgpu_last_error()}"


ROW = np.dtype([("k", np.int32), ("v", np.int32)])


def _expected(B, P):
    """Sorted (build v | probe v << 32) words of the equi-join, duplicates included."""
    order = np.argsort(B["k"], kind="stable")
    bk, bv = B["k"][order], B["v"][order]
    lo = np.searchsorted(bk, P["k"], "left")
    cnt = np.searchsorted(bk, P["k"], "right") - lo
    total = int(cnt.sum())
    pidx = np.repeat(np.arange(len(P)), cnt)
    bidx = lo[pidx] + (np.arange(total) - np.repeat(np.cumsum(cnt) - cnt, cnt))
    words = bv[bidx].astype(np.uint32).astype(np.uint64) | (P["v"][pidx].astype(np.uint32).astype(np.uint64) << 32)
    return np.sort(words)


def test_hj_repeated_queries_reuse_pools_within_budget():
    create = _fn("pgpu_hj_create", [C.c_void_p, C.POINTER(HjDesc), C.POINTER(C.c_void_p)])
    destroy = _fn("pgpu_hj_destroy", [C.c_void_p])
    build = _fn("pgpu_hj_build_device", [C.c_void_p, C.c_void_p, C.c_uint64])
    probe = _fn("pgpu_hj_probe_device", [C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint64, C.POINTER(C.c_uint64)])
    output = _fn("pgpu_hj_output", [C.c_void_p, C.c_uint32, C.POINTER(C.c_void_p), C.POINTER(C.c_uint64)])
    d2h = _fn("pgpu_d2h", [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64])

    budget = 512 << 20
    ctx = pa.Context(0, hbm_budget=budget)
    rng = np.random.default_rng(7)
    nb, npr = 200_000, 1_500_000
    B = np.empty(nb, ROW)
    B["k"] = rng.integers(0, 150_000, nb)  # about 1.3 build rows per key
    B["v"] = rng.integers(-2**31, 2**31 - 1, nb)
    P = np.empty(npr, ROW)
    P["k"] = rng.integers(-1000, 160_000, npr)  # some probe keys match nothing
    P["v"] = rng.integers(-2**31, 2**31 - 1, npr)
    small = 1000
    exp = _expected(B, P)
    exp_small = _expected(B, P[:small])
    assert len(exp) > npr  # the output outgrows the probe-sized first buffer
    db, dp = ctx.upload(B), ctx.upload(P)
    inputs, _, _ = ctx.hbm_usage()

    d = HjDesc()
    d.build_row_bytes = d.probe_row_bytes = 8
    d.build_key_offset = d.probe_key_offset = 0
    d.build_key_bytes = d.probe_key_bytes = 4
    d.nproj = 2
    d.proj[0] = Proj(0, 4, 4)  # B$1
    d.proj[1] = Proj(1, 4, 4)  # P$1
    d.slots = 2

    def rows_of(h, slot, n):
        out_ptr, n_rows = C.c_void_p(), C.c_uint64()
        _ok(output(h, slot, C.byref(out_ptr), C.byref(n_rows)), "pgpu_hj_output")
        assert n_rows.value == n
        host = np.empty(n, np.uint64)
        _ok(d2h(ctx.handle, host.ctypes.data, out_ptr, 8 * n), "pgpu_d2h")
        return np.sort(host)

    for _ in range(2):
        h = C.c_void_p()
        _ok(create(ctx.handle, C.byref(d), C.byref(h)), "pgpu_hj_create")
        _ok(build(h, db.ptr, nb), "pgpu_hj_build_device")
        n_out = C.c_uint64()
        _ok(probe(h, 0, dp.ptr, npr, C.byref(n_out)), "pgpu_hj_probe_device")
        assert n_out.value == len(exp)
        assert np.array_equal(rows_of(h, 0, n_out.value), exp)
        _ok(probe(h, 1, dp.ptr, small, C.byref(n_out)), "pgpu_hj_probe_device")
        assert np.array_equal(rows_of(h, 1, n_out.value), exp_small)
        _ok(destroy(h), "pgpu_hj_destroy")
        used, _, _ = ctx.hbm_usage()
        assert used == inputs  # blocks kept for reuse are not in use

    # nearly the whole remaining budget: the context drops its kept blocks
    big = ctx.alloc(budget - inputs - (1 << 20))
    big.free()
    db.free()
    dp.free()
    ctx.close()
