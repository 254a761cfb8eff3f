"""GPU parity of the fused join + count/sum path (libpgpu.so through its C ABI).

Every expected value comes from the oracle: the chained-bucket restatement of
HashJoinOp (oracle.hashjoin_count_sum) on small inputs, and the generators'
closed forms (exact int64 sums) up to the full C2 shape.  Integer results
must be bit-exact.
"""
import os
import zlib

import numpy as np
import pytest

import pythia_amd as pa
from oracle import oracle as O

pytestmark = pytest.mark.gpu
THREADS = min(16, os.cpu_count() or 1)


def _dev(ctx, arr):
    return ctx.upload(np.ascontiguousarray(arr))


@pytest.mark.parametrize("column,dtype,kind", [
    ("C1_RKEY", np.int32, "c1"), ("C1_RPAY", np.int32, "c1"), ("C1_SFK", np.int32, "c1"),
    ("C1_SPAY", np.int32, "c1"), ("C2_RKEY", np.int64, "c2"), ("C2_RPAY", np.int64, "c2"),
    ("C2_SFK_UNIFORM", np.int64, "c2"), ("C2_SFK_ZIPF", np.int64, "c2z"),
    ("C3_D_ORDERKEY", np.int32, "c3"), ("C3_D_CUSTKEY", np.int32, "c3"), ("C3_F_ORDERKEY", np.int32, "c3"),
    ("C3_F_PRICE", np.float64, "c3")])
def test_device_generators_match_host(ctx, column, dtype, kind):
    nR, nS = 1 << 14, 1 << 16
    if kind == "c1":
        rk, rp, sk, sp = O.gen_c1(nR, nS)
        exp = {"C1_RKEY": rk, "C1_RPAY": rp, "C1_SFK": sk, "C1_SPAY": sp}[column]
    elif kind.startswith("c2"):
        z = 1.0 if kind == "c2z" else 0.0
        rk, rp, sk = O.gen_c2(nR, nS, z)
        exp = {"C2_RKEY": rk, "C2_RPAY": rp}.get(column, sk)
    else:
        dk, dc, fk, fp = O.gen_c3(nR, nS, norders=nR)
        exp = {"C3_D_ORDERKEY": dk, "C3_D_CUSTKEY": dc, "C3_F_ORDERKEY": fk, "C3_F_PRICE": fp}[column]
    n = len(exp)
    out = ctx.alloc(n * np.dtype(dtype).itemsize)
    cdf = _dev(ctx, O.zipf_cdf(nR, 1.0)) if kind == "c2z" else None
    ctx.gen(column, n, out, n_build=nR, cdf=cdf)
    got = out.download(dtype, n)
    assert np.array_equal(got, exp)
    out.free()
    if cdf:
        cdf.free()


def _c2_device(ctx, nR, nS, zipf=0.0):
    rk, rp, sk = ctx.alloc(nR * 8), ctx.alloc(nR * 8), ctx.alloc(nS * 8)
    ctx.gen("C2_RKEY", nR, rk)
    ctx.gen("C2_RPAY", nR, rp)
    cdf = None
    if zipf:
        cdf = _dev(ctx, O.zipf_cdf(nR, zipf))
        ctx.gen("C2_SFK_ZIPF", nS, sk, n_build=nR, cdf=cdf)
    else:
        ctx.gen("C2_SFK_UNIFORM", nS, sk, n_build=nR)
    ctx.sync()
    if cdf:
        cdf.free()
    return rk, rp, sk


@pytest.mark.parametrize("bits", [2, 5, 10, 14])
@pytest.mark.parametrize("zipf", [0.0, 1.0])
def test_c2_radix_matches_closed_form(ctx, bits, zipf):
    nR, nS = 1 << 16, 1 << 21
    rk, rp, sk = _c2_device(ctx, nR, nS, zipf)
    plan = pa.JoinPlan(ctx, 8, nR, nS, radix_bits=bits, algo=pa.ALGO_RADIX)
    r = plan.count_sum(rk, rp, nR, sk, None, nS)
    ref = O.c2_closed(nR, nS, zipf, threads=THREADS)
    assert (r.matches, r.sum_build) == (ref.count, ref.sum_build)
    assert r.radix_bits == bits
    if bits == 2:
        assert r.lds_passes > 1  # 16K build rows per partition: LDS pieces
    plan.close()
    for b in (rk, rp, sk):
        b.free()


@pytest.mark.parametrize("variant", [0, 1, 2, 3, 4, 5, 7, 8])
@pytest.mark.parametrize("bits", [4, 9])
def test_join_variants_match_closed_form(ctx, variant, bits):
    """Every LDS join shape gives the same count and sum (4 bits: 4K build
    rows per partition, several LDS pieces per partition)."""
    nR, nS = 1 << 16, 1 << 20
    rk, rp, sk = _c2_device(ctx, nR, nS, 0.0)
    plan = pa.JoinPlan(ctx, 8, nR, nS, radix_bits=bits, algo=pa.ALGO_RADIX, join_variant=variant)
    r = plan.count_sum(rk, rp, nR, sk, None, nS)
    ref = O.c2_closed(nR, nS, 0.0, threads=THREADS)
    assert (r.matches, r.sum_build) == (ref.count, ref.sum_build)
    plan.close()
    for b in (rk, rp, sk):
        b.free()


def test_c2_global_matches_closed_form(ctx):
    nR, nS = 1 << 18, 1 << 21
    rk, rp, sk = _c2_device(ctx, nR, nS, 0.5)
    plan = pa.JoinPlan(ctx, 8, nR, nS, algo=pa.ALGO_GLOBAL)
    r = plan.count_sum(rk, rp, nR, sk, None, nS)
    ref = O.c2_closed(nR, nS, 0.5, threads=THREADS)
    assert (r.matches, r.sum_build) == (ref.count, ref.sum_build)
    plan.close()
    for b in (rk, rp, sk):
        b.free()


@pytest.mark.parametrize("algo", [pa.ALGO_RADIX, pa.ALGO_GLOBAL, pa.ALGO_RADIX1, pa.ALGO_AUTO])
def test_c1_int32_with_probe_payload(ctx, algo):
    nR, nS = 1 << 20, 1 << 24  # the C1 shape
    bufs = [ctx.alloc(nR * 4), ctx.alloc(nR * 4), ctx.alloc(nS * 4), ctx.alloc(nS * 4)]
    for b, col, n in zip(bufs, ["C1_RKEY", "C1_RPAY", "C1_SFK", "C1_SPAY"], [nR, nR, nS, nS]):
        ctx.gen(col, n, b, n_build=nR)
    plan = pa.JoinPlan(ctx, 4, nR, nS, probe_payload=True, algo=algo)
    r = plan.count_sum(bufs[0], bufs[1], nR, bufs[2], bufs[3], nS)
    ref = O.c1_closed(nR, nS, threads=THREADS)
    assert (r.matches, r.sum_build, r.sum_probe) == (ref.count, ref.sum_build, ref.sum_probe)
    plan.close()
    for b in bufs:
        b.free()


@pytest.mark.parametrize("kb", [4, 8])
@pytest.mark.parametrize("algo,bits", [(pa.ALGO_RADIX, 3), (pa.ALGO_RADIX, 9), (pa.ALGO_GLOBAL, 0),
                                       (pa.ALGO_RADIX1, 0)])
def test_random_keys_vs_oracle_join(ctx, kb, algo, bits):
    """Duplicates on both sides, negative keys, the minimum integer (the
    table's empty marker) as a real key, and non-matching probes."""
    rng = np.random.default_rng(kb * 100 + bits)
    dt = np.int32 if kb == 4 else np.int64
    lo = np.iinfo(dt).min
    nR, nS = 9000, 50000
    bk = rng.integers(-3000, 3000, nR).astype(dt)
    bk[:7] = lo
    bp = rng.integers(np.iinfo(dt).min // 4, np.iinfo(dt).max // 4, nR).astype(dt)
    pk = rng.integers(-3500, 3500, nS).astype(dt)
    pk[:11] = lo
    pp = rng.integers(np.iinfo(dt).min // 4, np.iinfo(dt).max // 4, nS).astype(dt)
    ct = "int" if kb == 4 else "long"
    bs, ps = O.schema((ct,), (ct,)), O.schema((ct,), (ct,))
    d = O.join_desc(bs, ps, 0, 0, ["B$1", "P$1"], buckets=1 << 10, tuplesperbucket=4, threads=4)
    csb, _ = O.hashjoin_count_sum(d, O.rows(bs, bk, bp), O.rows(ps, pk, pp), 0)
    csp, _ = O.hashjoin_count_sum(d, O.rows(bs, bk, bp), O.rows(ps, pk, pp), 1)
    plan = pa.JoinPlan(ctx, kb, nR, nS, radix_bits=bits, probe_payload=True, algo=algo)
    bufs = [_dev(ctx, a) for a in (bk, bp, pk, pp)]
    r = plan.count_sum(bufs[0], bufs[1], nR, bufs[2], bufs[3], nS)
    assert r.matches == csb.count
    mask = (1 << (8 * kb)) - 1  # the reference accumulates in the column type
    assert r.sum_build & mask == csb.isum & mask
    assert r.sum_probe & mask == csp.isum & mask
    # exact 64-bit sums from numpy
    bsum = {}
    for k, v in zip(bk.tolist(), bp.tolist()):
        c, s = bsum.get(k, (0, 0))
        bsum[k] = (c + 1, s + v)
    ex_b = sum(bsum.get(k, (0, 0))[1] for k in pk.tolist()) % 2**64
    ex_p = sum(bsum.get(k, (0, 0))[0] * v for k, v in zip(pk.tolist(), pp.tolist())) % 2**64
    assert (r.sum_build, r.sum_probe) == (ex_b, ex_p)
    plan.close()
    for b in bufs:
        b.free()


@pytest.mark.parametrize("kb", [4, 8])
def test_one_pass_unique_keys_vs_exact(ctx, kb):
    """The one-pass route (ALGO_RADIX1; ALGO_AUTO takes it for builds of
    2^16 to 2^22 rows) with unique build keys, the minimum integer as a real
    key on both sides and probe keys that miss, called twice with different
    sizes on one plan."""
    rng = np.random.default_rng(77 + kb)
    dt = np.int32 if kb == 4 else np.int64
    lo = np.iinfo(dt).min
    plan = pa.JoinPlan(ctx, kb, 1 << 18, 1 << 21, probe_payload=True, algo=pa.ALGO_AUTO)
    for nR, nS in ((1 << 18, 1 << 21), (100000, 700001)):
        bk = rng.permutation(np.arange(-2 * nR, 2 * nR, 2, dtype=np.int64))[:nR].astype(dt)
        bk[7] = lo
        pk = rng.choice(bk, nS).astype(dt)
        pk[:33] = lo
        pk[33:333] += 1  # odd: no match
        bp = rng.integers(np.iinfo(dt).min // 4, np.iinfo(dt).max // 4, nR).astype(dt)
        pp = rng.integers(np.iinfo(dt).min // 4, np.iinfo(dt).max // 4, nS).astype(dt)
        bsum = dict(zip(bk.tolist(), bp.tolist()))
        hits = [(bsum[k], v) for k, v in zip(pk.tolist(), pp.tolist()) if k in bsum]
        bufs = [_dev(ctx, a) for a in (bk, bp, pk, pp)]
        r = plan.count_sum(bufs[0], bufs[1], nR, bufs[2], bufs[3], nS)
        assert r.matches == len(hits)
        assert r.sum_build == sum(b for b, _ in hits) % 2**64
        assert r.sum_probe == sum(v for _, v in hits) % 2**64
        assert r.radix_bits == int(np.ceil(np.log2(nR / 4096)))  # one pass at ~4096 build rows per partition
        for b in bufs:
            b.free()
    plan.close()


@pytest.mark.parametrize("nR,nS", [(0, 1000), (1000, 0), (0, 0), (1, 1)])
def test_empty_and_tiny_inputs(ctx, nR, nS):
    bk = np.arange(nR, dtype=np.int64)
    pk = np.zeros(nS, dtype=np.int64)
    plan = pa.JoinPlan(ctx, 8, max(nR, 1), max(nS, 1), algo=pa.ALGO_RADIX)
    bufs = [_dev(ctx, a) for a in (bk, bk * 3, pk)]
    r = plan.count_sum(bufs[0], bufs[1], nR, bufs[2], None, nS)
    exp = nS if (nR >= 1 and nS >= 1) else 0
    assert r.matches == exp and r.sum_build == 0
    plan.close()
    for b in bufs:
        b.free()


def test_plan_rejects_oversize_and_bad_config(ctx):
    with pytest.raises(pa.PgpuError):
        pa.JoinPlan(ctx, 3, 10, 10)
    with pytest.raises(pa.PgpuError):
        pa.JoinPlan(ctx, 8, 10, 10, radix_bits=15)
    plan = pa.JoinPlan(ctx, 8, 10, 10)
    b = ctx.alloc(8 * 100)
    with pytest.raises(pa.PgpuError):
        plan.count_sum(b, b, 11, b, None, 5)
    plan.close()
    b.free()


def test_hbm_budget_is_enforced():
    ctx = pa.Context(0, hbm_budget=64 << 20)
    b = ctx.alloc(32 << 20)
    with pytest.raises(pa.PgpuError) as e:
        ctx.alloc(40 << 20)
    assert e.value.code == -3
    with pytest.raises(pa.PgpuError):
        pa.JoinPlan(ctx, 8, 1 << 22, 1 << 22)
    b.free()
    used, budget, peak = ctx.hbm_usage()
    assert used == 0 and budget == 64 << 20 and peak >= 32 << 20
    ctx.close()


@pytest.mark.parametrize("zipf", [0.0, 0.5, 1.0])
def test_c2_full_shape(ctx, zipf):
    """C2 at its full size: R 2^24 x (int64, int64), S 2^28 int64 uniform or
    Zipf s = 0.5 / 1.0, 14 radix bits; count and int64 sum bit-exact against
    the closed form."""
    nR, nS = 1 << 24, 1 << 28
    rk, rp, sk = _c2_device(ctx, nR, nS, zipf)
    plan = pa.JoinPlan(ctx, 8, nR, nS, radix_bits=14, algo=pa.ALGO_RADIX)
    r = plan.count_sum(rk, rp, nR, sk, None, nS)
    ref = O.c2_closed(nR, nS, zipf, threads=THREADS)
    assert (r.matches, r.sum_build) == (ref.count, ref.sum_build)
    assert r.lds_passes == 1
    plan.close()
    for b in (rk, rp, sk):
        b.free()


@pytest.mark.parametrize("nS,bits,zipf", [((1 << 23) + 12345, 14, 0.0), ((5 << 20) + 1, 6, 1.0), (3 << 22, 9, 0.5)])
def test_pass1_tiles_ragged(ctx, nS, bits, zipf):
    """Pass 1 runs one workgroup per tile at offsets of its own (block prefix
    + the tile's offset in its block): probe sides whose blocks hold several
    tiles, a partial last tile and a partial last block, uniform and skewed
    keys.  Count and sum bit-exact against the closed form."""
    nR = 1 << 16
    rk, rp, sk = _c2_device(ctx, nR, nS, zipf)
    plan = pa.JoinPlan(ctx, 8, nR, nS, radix_bits=bits, algo=pa.ALGO_RADIX)
    r = plan.count_sum(rk, rp, nR, sk, None, nS)
    ref = O.c2_closed(nR, nS, zipf, threads=THREADS)
    assert (r.matches, r.sum_build) == (ref.count, ref.sum_build)
    plan.close()
    for b in (rk, rp, sk):
        b.free()


def test_pass1_tiles_with_payloads(ctx):
    """Per-tile pass 1 for 8-byte (int32 key + payload) and 16-byte (int64
    key + payload) probe rows: 4096- and 8192-row tiles, several per block."""
    nR, nS = 1 << 20, (1 << 24) + 777  # the C1 shape, a ragged probe side
    bufs = [ctx.alloc(nR * 4), ctx.alloc(nR * 4), ctx.alloc(nS * 4), ctx.alloc(nS * 4)]
    for b, col, n in zip(bufs, ["C1_RKEY", "C1_RPAY", "C1_SFK", "C1_SPAY"], [nR, nR, nS, nS]):
        ctx.gen(col, n, b, n_build=nR)
    plan = pa.JoinPlan(ctx, 4, nR, nS, probe_payload=True, radix_bits=8, algo=pa.ALGO_RADIX)
    r = plan.count_sum(bufs[0], bufs[1], nR, bufs[2], bufs[3], nS)
    ref = O.c1_closed(nR, nS, threads=THREADS)
    assert (r.matches, r.sum_build, r.sum_probe) == (ref.count, ref.sum_build, ref.sum_probe)
    plan.close()
    for b in bufs:
        b.free()
    # int64 keys with a probe payload: every probe key is a build key (the
    # build keys are a permutation of 0..nR-1), so sum_probe is the payload sum
    nR, nS = 1 << 16, (1 << 23) + 77
    rk, rp, sk = _c2_device(ctx, nR, nS)
    pv = np.random.default_rng(5).integers(-(1 << 40), 1 << 40, nS, dtype=np.int64)
    sp = _dev(ctx, pv)
    plan = pa.JoinPlan(ctx, 8, nR, nS, probe_payload=True, radix_bits=5, algo=pa.ALGO_RADIX)
    r = plan.count_sum(rk, rp, nR, sk, sp, nS)
    ref = O.c2_closed(nR, nS, 0.0, threads=THREADS)
    assert (r.matches, r.sum_build) == (ref.count, ref.sum_build) and r.matches == nS
    assert r.sum_probe == int(pv.astype(object).sum()) % 2**64
    plan.close()
    for b in (rk, rp, sk, sp):
        b.free()


def _c3_device(ctx, nD, nF, norders):
    dk, dc, fk, fp = ctx.alloc(nD * 4), ctx.alloc(nD * 4), ctx.alloc(nF * 4), ctx.alloc(nF * 8)
    ctx.gen("C3_D_ORDERKEY", nD, dk, n_build=nD)
    ctx.gen("C3_D_CUSTKEY", nD, dc)
    ctx.gen("C3_F_ORDERKEY", nF, fk, n_build=norders)
    ctx.gen("C3_F_PRICE", nF, fp)
    ctx.sync()
    return dk, dc, fk, fp


@pytest.mark.parametrize("nD,nF", [(20000, 100000), (15_000_000, 60_000_000)])
def test_c3_groupjoin_matches_exact_cents(ctx, nD, nF):
    """sum(price) group by custkey over F join D on orderkey: every group
    within 1e-12 relative of the exact integer-cent sum."""
    bufs = _c3_device(ctx, nD, nF, nD)
    gj = pa.GroupJoin(ctx, nD, nF, 1_500_000)
    ng, joined = gj.run(*bufs[:2], nD, *bufs[2:], nF)
    _, cents = O.c3_closed(nD, nF, norders=nD, threads=THREADS)
    assert joined == nF
    assert ng == np.count_nonzero(cents)
    k, s, c = gj.fetch(ng)
    exact = cents[k] / 100.0
    assert np.all(np.abs(s - exact) <= 1e-12 * np.abs(exact))
    assert len(np.unique(k)) == ng
    assert int(c.sum()) == nF
    assert gj.path() == pa.GJ_PARTITIONED
    gj.close()
    for b in bufs:
        b.free()


def test_groupjoin_rows_pinned_and_pageable(ctx):
    """The group join's result as the aggregate's 12-byte (int32 key, FP64
    sum or int64 count) rows: written by the packing kernel straight into
    pinned memory from pgpu_host_alloc (16-byte aligned places), packed in
    HBM and copied into pinned memory at an unaligned place and into
    pageable memory; every row range gives the fetched columns' bytes,
    including ranges whose length leaves a partial 16-byte tail."""
    import ctypes as C

    nD, nF = 20000, 100000
    bufs = _c3_device(ctx, nD, nF, nD)
    gj = pa.GroupJoin(ctx, nD, nF, 1_500_000)
    ng, _ = gj.run(*bufs[:2], nD, *bufs[2:], nF)
    k, s, c = gj.fetch(ng)
    row = np.dtype({"names": ["k", "a"], "formats": [np.int32, np.uint64], "offsets": [0, 4], "itemsize": 12})
    pinned = ctx.host_alloc(ng * 12 + 64)
    try:
        for acc, a in ((pa.GJ_ROW_SUM, s.view(np.uint64)), (pa.GJ_ROW_COUNT, c)):
            want = np.zeros(ng, row)
            want["k"], want["a"] = k, a
            want = want.tobytes()
            for first, n, off in ((0, ng, 0), (0, 1, 0), (0, 2, 16), (0, 3, 32), (1, 255, 0), (7, 257, 48),
                                  (4, ng - 4, 0), (5, ng - 5, 4), (3, 300, 20)):
                gj.rows(acc, first, n, pinned + off)
                got = bytes((C.c_char * (n * 12)).from_address(pinned + off))
                assert got == want[first * 12:(first + n) * 12], (acc, first, n, off)
            page = np.zeros(ng * 12, np.uint8)
            gj.rows(acc, 0, ng, page.ctypes.data)
            assert page.tobytes() == want
    finally:
        ctx.host_free(pinned)
        gj.close()
        for b in bufs:
            b.free()


def test_groupjoin_duplicates_and_sentinel_keys(ctx):
    """Duplicate dimension keys (each match contributes), the INT32_MIN
    marker as a real key and group, unmatched fact rows."""
    rng = np.random.default_rng(11)
    lo = np.iinfo(np.int32).min
    dk = rng.integers(-50, 50, 300).astype(np.int32)
    dk[:3] = lo
    dg = rng.integers(-5, 5, 300).astype(np.int32)
    dg[5:9] = lo
    fk = rng.integers(-60, 60, 5000).astype(np.int32)
    fk[:4] = lo
    fv = rng.random(5000)
    exp = {}
    for k, v in zip(fk.tolist(), fv.tolist()):
        for kk, gg in zip(dk.tolist(), dg.tolist()):
            if kk == k:
                exp[gg] = exp.get(gg, 0.0) + v
    bufs = [_dev(ctx, a) for a in (dk, dg, fk, fv)]
    gj = pa.GroupJoin(ctx, 300, 5000, 100)
    ng, joined = gj.run(bufs[0], bufs[1], 300, bufs[2], bufs[3], 5000)
    k, s, c = gj.fetch(ng)
    assert gj.path() == pa.GJ_GENERAL  # duplicate dimension keys
    assert dict(zip(k.tolist(), s.tolist())).keys() == exp.keys()
    for kk, ss in zip(k.tolist(), s.tolist()):
        assert abs(ss - exp[kk]) <= 1e-12 * max(1.0, abs(exp[kk]))
    gj.close()
    for b in bufs:
        b.free()


I32_MIN, I32_MAX = np.iinfo(np.int32).min, np.iinfo(np.int32).max


def _gj_oracle(dk, dg, fk, fv):
    """sum(F.val) group by D.grp over F join D on key through the oracle's
    HashJoinOp + GenericAggregate restatement: (joined rows, {grp: sum})."""
    ds, fs = O.schema(("int",), ("int",)), O.schema(("int",), ("dec",))
    d = O.join_desc(ds, fs, 0, 0, ["B$1", "P$1"], buckets=65536, tuplesperbucket=4)
    joined, _ = O.hashjoin(d, O.rows(ds, dk, dg), O.rows(fs, fk, fv))
    out = O.aggregate(O.agg_desc(O.join_outschema(d), [0], 1, buckets=65536), joined)
    return len(joined), dict(zip(out["c0"].tolist(), out["c1"].tolist()))


def _gj_device(ctx, dk, dg, fk, fv, max_groups):
    bufs = [_dev(ctx, a) for a in (dk, dg, fk, fv)]
    gj = pa.GroupJoin(ctx, len(dk), len(fk), max_groups)
    try:
        ng, joined = gj.run(bufs[0], bufs[1], len(dk), bufs[2], bufs[3], len(fk))
        k, s, c = gj.fetch(ng)
        return gj.path(), joined, k, s, c
    finally:
        gj.close()
        for b in bufs:
            b.free()


def _gj_case(name, rng):
    """(dim keys, dim groups, fact keys, fact values, max_groups, expected path)."""
    if name == "gaps_misses_low":  # keys and groups at the bottom of the int32 range
        dk = (I32_MIN + rng.choice(100_000, 50_000, replace=False)).astype(np.int32)
        dg = (I32_MIN + rng.integers(0, 3000, dk.size)).astype(np.int32)
        fk = (I32_MIN + rng.integers(0, 110_000, 300_000)).astype(np.int32)
        return dk, dg, fk, rng.random(fk.size) + 0.5, 3000, pa.GJ_PARTITIONED
    if name == "gaps_misses_high":  # keys and groups at the top, misses below the key range
        dk = (I32_MAX - rng.choice(100_000, 60_000, replace=False)).astype(np.int32)
        dg = (I32_MAX - rng.integers(0, 50_000, dk.size)).astype(np.int32)
        fk = (I32_MAX - rng.integers(0, 120_000, 300_000)).astype(np.int32)
        return dk, dg, fk, rng.random(fk.size) + 0.5, 50_000, pa.GJ_PARTITIONED
    if name == "skew":  # one key holds 60 % of the fact rows; 7 groups: multi-item partitions
        dk = rng.permutation(5000).astype(np.int32)
        dg = (dk % 7).astype(np.int32)
        fk = rng.integers(0, 5000, 400_000).astype(np.int32)
        fk[rng.random(fk.size) < 0.6] = 17
        return dk, dg, fk, rng.random(fk.size) + 0.5, 7, pa.GJ_PARTITIONED
    if name == "widest_dense":  # 2^25 key range, 2^22 group range: every partition count at its maximum
        dk = np.sort(rng.choice(1 << 25, 100_000, replace=False)).astype(np.int32)
        dk[0], dk[-1] = 0, (1 << 25) - 1
        dg = rng.integers(0, 1 << 22, dk.size).astype(np.int32)
        dg[0], dg[-1] = 0, (1 << 22) - 1
        fk = rng.choice(dk, 400_000).astype(np.int32)
        miss = rng.random(fk.size) < 0.1
        fk[miss] = rng.integers(-5, (1 << 25) + 5, int(miss.sum()))
        return dk, dg, fk, rng.random(fk.size) + 0.5, 100_000, pa.GJ_PARTITIONED
    if name == "wide_keys":  # key range beyond 2^25: general path
        dk = np.unique(rng.integers(I32_MIN, I32_MAX, 60_000)).astype(np.int32)[:50_000]
        dg = rng.integers(0, 1000, dk.size).astype(np.int32)
        fk = rng.choice(dk, 200_000).astype(np.int32)
        return dk, dg, fk, rng.random(fk.size) + 0.5, 1000, pa.GJ_GENERAL
    if name == "wide_groups":  # group range beyond 2^22: general path
        dk = rng.permutation(20_000).astype(np.int32)
        dg = rng.integers(I32_MIN + 1, I32_MAX, dk.size).astype(np.int32)
        fk = rng.integers(0, 20_000, 200_000).astype(np.int32)
        return dk, dg, fk, rng.random(fk.size) + 0.5, 20_000, pa.GJ_GENERAL
    if name == "tiny":
        dk, dg = np.array([5], np.int32), np.array([I32_MIN], np.int32)
        fk = np.array([5, 4, 5, 6, 5], np.int32)
        return dk, dg, fk, np.array([1.0, 2.0, 3.0, 4.0, 5.0]), 1, pa.GJ_PARTITIONED
    raise ValueError(name)


@pytest.mark.parametrize("name", ["gaps_misses_low", "gaps_misses_high", "skew", "widest_dense", "wide_keys",
                                  "wide_groups", "tiny"])
def test_groupjoin_vs_oracle(ctx, name):
    """Group join against the oracle's hashjoin -> aggregate_sum: the key /
    group range edges, misses, skewed partitions split into several work
    items, the largest partition counts, and the general path's cases."""
    rng = np.random.default_rng(zlib.crc32(name.encode()))
    dk, dg, fk, fv, max_groups, path = _gj_case(name, rng)
    exp_joined, exp = _gj_oracle(dk, dg, fk, fv)
    got_path, joined, k, s, c = _gj_device(ctx, dk, dg, fk, fv, max_groups)
    assert got_path == path
    assert joined == exp_joined and int(c.sum()) == exp_joined
    assert len(np.unique(k)) == len(k)
    assert set(k.tolist()) == set(exp)
    for kk, ss in zip(k.tolist(), s.tolist()):
        assert abs(ss - exp[kk]) <= 1e-12 * abs(exp[kk])


@pytest.mark.parametrize("dense", [True, False])
def test_groupjoin_more_groups_than_planned_is_an_error(ctx, dense):
    """More result groups than max_groups: a clear overflow error on both
    paths, no write past the outputs."""
    rng = np.random.default_rng(5)
    dk = rng.permutation(10_000).astype(np.int32) if dense else \
        np.unique(rng.integers(I32_MIN, I32_MAX, 12_000)).astype(np.int32)[:10_000]
    dg = np.arange(dk.size, dtype=np.int32)
    fk = rng.choice(dk, 50_000).astype(np.int32)
    bufs = [_dev(ctx, a) for a in (dk, dg, fk, rng.random(fk.size))]
    gj = pa.GroupJoin(ctx, dk.size, fk.size, 100)
    with pytest.raises(pa.PgpuError):
        gj.run(bufs[0], bufs[1], dk.size, bufs[2], bufs[3], fk.size)
    gj.close()
    for b in bufs:
        b.free()


@pytest.mark.parametrize("nD,nF", [(0, 1000), (1000, 0)])
def test_groupjoin_empty_sides(ctx, nD, nF):
    dk = np.arange(max(nD, 1), dtype=np.int32)
    fk = np.arange(max(nF, 1), dtype=np.int32)
    bufs = [_dev(ctx, a) for a in (dk, dk, fk, fk.astype(np.float64))]
    gj = pa.GroupJoin(ctx, max(nD, 1), max(nF, 1), 10)
    assert gj.run(bufs[0], bufs[1], nD, bufs[2], bufs[3], nF) == (0, 0)
    gj.close()
    for b in bufs:
        b.free()
