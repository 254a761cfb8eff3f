"""pgpu_join_count_sum_noalloc (the join that allocates nothing: its table
and partials live in the library's per-device scratch array) against the
existing entry point pgpu_join_count_sum on the same inputs, through the C
ABI: two calls in a row (the second smaller, so stale scratch contents from
the first would show), both key widths, with and without probe payloads;
builds beyond the scratch table are refused."""
import ctypes as C

import numpy as np
import pytest

import pythia_amd as pa
from oracle import oracle as O
from pythia_amd import _lib

pytestmark = pytest.mark.gpu

LIB = C.CDLL(_lib.LIBPGPU)
LIB.pgpu_last_error.restype = C.c_char_p
LIB.pgpu_join_count_sum_noalloc.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p, C.c_uint64,
                                            C.c_void_p, C.c_void_p, C.c_uint64, C.POINTER(_lib.JoinResult)]
LIB.pgpu_join_count_sum.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p, C.c_void_p,
                                    C.c_uint64, C.POINTER(_lib.JoinResult)]
LIB.pgpu_join_plan_create.argtypes = [C.c_void_p, C.POINTER(_lib.JoinCfg), C.c_uint64, C.c_uint64,
                                      C.POINTER(C.c_void_p)]
LIB.pgpu_join_plan_destroy.argtypes = [C.c_void_p]
MAXB = {4: 1 << 20, 8: 1 << 19}


def _inputs(ctx, kb, nR, nS, seed):
    rng = np.random.default_rng(seed)
    dt = np.int32 if kb == 4 else np.int64
    rk = rng.permutation(nR).astype(dt) * 3 - nR
    rk[: min(3, nR)] = np.iinfo(dt).min  # the minimum as a real key (and duplicated)
    rp = rng.integers(-2**30, 2**30, nR).astype(dt)
    sk = (rng.integers(0, nR + nR // 8 + 1, nS) * 3 - nR).astype(dt)
    sk[:5] = np.iinfo(dt).min
    sp = rng.integers(-2**30, 2**30, nS).astype(dt)
    return [ctx.upload(np.ascontiguousarray(a)) for a in (rk, rp, sk, sp)]


def _noalloc(ctx, kb, hv, bufs, nR, nS):
    r = _lib.JoinResult()
    rc = LIB.pgpu_join_count_sum_noalloc(ctx.handle, kb, hv, bufs[0].ptr, bufs[1].ptr, nR, bufs[2].ptr,
                                         bufs[3].ptr if hv else None, nS, C.byref(r))
    assert rc == 0, LIB.pgpu_last_error()
    return (r.matches, r.sum_build, r.sum_probe)


def _existing(ctx, kb, hv, bufs, nR, nS):
    cfg = _lib.JoinCfg(kb, 0, hv, _lib.ALGO_GLOBAL, 0)
    plan = C.c_void_p()
    assert LIB.pgpu_join_plan_create(ctx.handle, C.byref(cfg), max(nR, 1), max(nS, 1), C.byref(plan)) == 0
    r = _lib.JoinResult()
    rc = LIB.pgpu_join_count_sum(plan, bufs[0].ptr, bufs[1].ptr, nR, bufs[2].ptr, bufs[3].ptr if hv else None, nS,
                                 C.byref(r))
    assert rc == 0, LIB.pgpu_last_error()
    LIB.pgpu_join_plan_destroy(plan)
    return (r.matches, r.sum_build, r.sum_probe)


@pytest.mark.parametrize("kb", [4, 8])
@pytest.mark.parametrize("hv", [0, 1])
def test_noalloc_matches_existing_entry_point_twice(ctx, kb, hv):
    for nR, nS, seed in ((MAXB[kb], 3 * MAXB[kb], 1), (4099, 20011, 2)):
        bufs = _inputs(ctx, kb, nR, nS, seed)
        got = _noalloc(ctx, kb, hv, bufs, nR, nS)
        assert got == _existing(ctx, kb, hv, bufs, nR, nS)
        assert got[0] > nS // 2
        for b in bufs:
            b.free()


def test_noalloc_c1_shape_against_closed_form(ctx):
    """The C1 shape (2^20 int32 build rows, the largest the scratch table
    holds) against the oracle's closed form."""
    nR, nS = 1 << 20, 1 << 22
    bufs = [ctx.alloc(4 * n) for n in (nR, nR, nS, nS)]
    for b, col, n in zip(bufs, ["C1_RKEY", "C1_RPAY", "C1_SFK", "C1_SPAY"], [nR, nR, nS, nS]):
        ctx.gen(col, n, b, n_build=nR)
    ctx.sync()
    ref = O.c1_closed(nR, nS)
    assert _noalloc(ctx, 4, 1, bufs, nR, nS) == (ref.count, ref.sum_build, ref.sum_probe)
    for b in bufs:
        b.free()


def test_noalloc_refuses_builds_beyond_its_table(ctx):
    r = _lib.JoinResult()
    for kb in (4, 8):
        n = MAXB[kb] + 1  # real buffers of that size: nothing is read out of bounds either way
        bufs = [ctx.alloc(kb * n) for _ in range(3)]
        rc = LIB.pgpu_join_count_sum_noalloc(ctx.handle, kb, 0, bufs[0].ptr, bufs[1].ptr, n, bufs[2].ptr, None, 1,
                                             C.byref(r))
        assert rc == -1
        assert b"exceed" in LIB.pgpu_last_error()
        for b in bufs:
            b.free()
    # empty sides: nothing joins
    assert LIB.pgpu_join_count_sum_noalloc(ctx.handle, 8, 0, None, None, 0, None, None, 0, C.byref(r)) == 0
    assert (r.matches, r.sum_build, r.sum_probe) == (0, 0, 0)
