"""Pins the CPU oracle (oracle/oracle.c) before it is trusted as a checker.

Sources of truth, none of them the oracle's own formulas:
  * the reference's unit-test known answers: unit_tests/testhash.cpp,
    unit_tests/testhashtable.cpp, unit_tests/queryhashjoin.cpp,
    unit_tests/queryaggsum_global.cpp, unit_tests/querypartition.cpp;
  * brute force (nested loops / numpy group-by) on small inputs;
  * closed forms of the seeded generators (include/pgpu_gen.h).
"""
import collections
import ctypes as C
import math
import os

import numpy as np
import pytest

from oracle import oracle as O


def test_getlogarithm_is_ceil_log2():
    # unit_tests/testhash.cpp:58-82
    rng = np.random.default_rng(1)
    for a in rng.integers(1, 1 << 10, size=2000):
        assert O.lib().orc_getlogarithm(int(a)) == math.ceil(math.log2(int(a)))
    assert O.lib().orc_getlogarithm(1) == 0


def test_generate_multipass_split():
    # unit_tests/testhash.cpp:53-66: ParameterizedModuloValueHasher(0,1024,2).generate(3) -> 8, 8, 16
    b, s = O.generate(1024, 2, 3)
    assert b == [8, 8, 16]
    # MSD-first: the pass masks are disjoint and tile bits [skip, skip+10)
    masks = [(bb - 1) << ss for bb, ss in zip(b, s)]
    assert masks[0] & masks[1] == 0 and masks[1] & masks[2] == 0 and masks[0] & masks[2] == 0
    assert masks[0] | masks[1] | masks[2] == (1023 << 2)
    assert s[0] > s[1] > s[2] == 2
    b2, s2 = O.generate(1 << 14, 0, 2)  # the 2-pass radix split of 14 bits: 7 high then 7 low
    assert b2 == [128, 128] and s2 == [7, 0]
    b3, s3 = O.generate(1 << 13, 0, 2)
    assert b3 == [64, 128] and s3 == [7, 0]


def test_modulo_bounds_and_values():
    s = O.schema(("long",))
    h = O.hasher(O.H_MODULO, 1000, s, 0)  # rounds up to 1024 buckets
    assert h.buckets == 1024
    rng = np.random.default_rng(2)
    for v in rng.integers(-10**12, 10**12, size=500):
        got = O.lib().orc_hash_value(C.byref(h), int(v))
        assert got < 1024 and got == int(v) % 1024  # two's complement low bits == python mod 2^k


def test_int32_key_sign_extension():
    s = O.schema(("int",))
    h = O.hasher(O.H_MODULO, 16, s, 0)
    r = O.rows(s, np.array([-1, -17, 5], np.int32))
    got = [O.lib().orc_hash_tuple(C.byref(h), C.c_void_p(r.ctypes.data + i * 4)) for i in range(3)]
    assert got == [15, 15, 5]


def test_alwayszero():
    s = O.schema(("long",))
    h = O.hasher(O.H_ALWAYSZERO, 123, s, 0)
    assert h.buckets == 1
    assert O.lib().orc_hash_value(C.byref(h), 987654321) == 0


def test_exactrange_bucket_boundaries():
    # unit_tests/testhash.cpp:103-125
    s = O.schema(("long",))
    rng = np.random.default_rng(3)
    for _ in range(300):
        mx = int(rng.integers(2, 2**31 - 1)) & ~1
        mn = int(rng.integers(0, mx))
        h = O.hasher(O.H_EXACTRANGE, 80, s, 0, rmin=mn, rmax=mx)
        assert h.buckets == 80
        assert O.lib().orc_hash_value(C.byref(h), mx) <= 79
        for j in range(1, 80):
            m = O.lib().orc_exactrange_minimum(C.byref(h), j)
            assert O.lib().orc_hash_value(C.byref(h), m - 1) == j - 1
            assert O.lib().orc_hash_value(C.byref(h), m) == j


def test_knuth_and_parammodulo_closed_form():
    s = O.schema(("long",))
    hk = O.hasher(O.H_KNUTH, 256, s, 0, rmin=7, skipbits=3)
    hp = O.hasher(O.H_PARAMMODULO, 256, s, 0, rmin=7, skipbits=3)
    for v in [0, 1, 12345, 2**40 + 3, 999999937]:
        assert O.lib().orc_hash_value(C.byref(hk), v) == ((v * 2654435761) % 2**64 & (255 << 3)) >> 3
        assert O.lib().orc_hash_value(C.byref(hp), v) == (((v - 7) % 2**64) & (255 << 3)) >> 3


def _fnv_bytes(data: bytes, k: int) -> int:
    # FNV-1a 64 over *signed* chars, then the (h >> k) ^ h fold (hash.h:157-189)
    h = 14695981039346656037
    for b in data:
        sb = b - 256 if b >= 128 else b
        h ^= sb % 2**64
        h = (h * 1099511628211) % 2**64  # FNV prime = 2^40 + 0x1B3
    hs = h - 2**64 if h >= 2**63 else h
    return ((hs >> k) % 2**64 ^ h) & ((1 << k) - 1)


def test_bytes_hash_fnv():
    s = O.schema(("int",), ("char", 8))
    h = O.hasher(O.H_BYTES, 1 << 12, s, 0, fieldmax=1)
    data = O.rows(s, np.array([7, -3], np.int32), np.array([b"ab\xffcd", b"\x80zz"], dtype="S8"))
    for i in range(2):
        raw = data[i:i + 1].tobytes()[:12]
        got = O.lib().orc_hash_tuple(C.byref(h), C.c_void_p(data.ctypes.data + i * 12))
        assert got == _fnv_bytes(raw, 12)


def test_hashtable_chain_iteration():
    # unit_tests/testhashtable.cpp:44-85: one bucket of two int slots; every
    # inserted value is iterated exactly once through long overflow chains
    L = O.lib()
    for n in [0, 1, 2, 3, 1001, 9999]:
        t = L.orc_ht_create(1, 8, 4)
        for i in range(n):
            p = L.orc_ht_allocate(t, 0)
            C.c_int32.from_address(p).value = i
        buf = (C.c_void_p * max(n, 1))()
        m = L.orc_ht_bucket_tuples(t, 0, buf, max(n, 1))
        assert m == n
        vals = [C.c_int32.from_address(buf[i]).value for i in range(m)]
        assert vals == list(range(n))  # insertion (chain) order
        assert L.orc_ht_spills(t) == max(0, (n + 1) // 2 - 1)
        L.orc_ht_destroy(t)


def test_queryhashjoin_known_answer():
    # unit_tests/queryhashjoin.cpp: build (long i, long i), probe (long i,
    # dec i+0.1), i = 1..20, modulo 16 buckets, tuplesperbucket 2, 4 threads,
    # projection [B$1, P$1]; every key exactly once and P$1 == v + 0.1
    bs = O.schema(("long",), ("long",))
    ps = O.schema(("long",), ("dec",))
    i = np.arange(1, 21)
    b = O.rows(bs, i, i)
    p = O.rows(ps, i, i + 0.1)
    d = O.join_desc(bs, ps, 0, 0, ["B$1", "P$1"], buckets=16, tuplesperbucket=2, threads=4)
    out, st = O.hashjoin(d, b, p)
    assert len(out) == 20
    assert sorted(out["c0"].tolist()) == list(range(1, 21))
    assert np.all(out["c1"] == out["c0"] + 0.1)
    assert st.spills == 0  # keys 17..20 share buckets 1..4 with 1..4: two per bucket fit two slots
    d1 = O.join_desc(bs, ps, 0, 0, ["B$1", "P$1"], buckets=16, tuplesperbucket=1, threads=1)
    out1, st1 = O.hashjoin(d1, b, p)
    assert len(out1) == 20 and st1.spills == 4  # one slot: buckets 1..4 each spill once


def _brute_join(bkeys, bcols, pkeys, pcols, proj):
    out = []
    for i, bk in enumerate(bkeys):
        for j, pk in enumerate(pkeys):
            if int(bk) == int(pk):
                out.append(tuple(bcols[c][i] if side == "B" else pcols[c][j] for side, c in proj))
    return sorted(out)


@pytest.mark.parametrize("threads", [1, 3, 8])
@pytest.mark.parametrize("tpb", [1, 4])
def test_join_matches_nested_loops(threads, tpb):
    rng = np.random.default_rng(threads * 10 + tpb)
    nb, np_ = 300, 700
    bs = O.schema(("int",), ("long",), ("dec",))
    ps = O.schema(("long",), ("int",))
    bk = rng.integers(-40, 40, nb).astype(np.int32)  # duplicates, negative keys
    bl = rng.integers(-10**12, 10**12, nb)
    bd = rng.random(nb)
    pk = rng.integers(-50, 50, np_)
    pi = rng.integers(-1000, 1000, np_).astype(np.int32)
    d = O.join_desc(bs, ps, 0, 0, ["P$1", "B$2", "B$1", "P$0"], buckets=32, tuplesperbucket=tpb, threads=threads)
    out, _ = O.hashjoin(d, O.rows(bs, bk, bl, bd), O.rows(ps, pk, pi))
    got = sorted(zip(out["c0"].tolist(), out["c1"].tolist(), out["c2"].tolist(), out["c3"].tolist()))
    exp = _brute_join(bk, [bk, bl, bd], pk, [pk, pi], [("P", 1), ("B", 2), ("B", 1), ("P", 0)])
    assert got == [tuple(e) for e in exp]


def test_join_count_sum_int32_wraps():
    # AggregateSum accumulates an int column in int (aggregatesum.cpp:81-107)
    bs = O.schema(("int",), ("int",))
    ps = O.schema(("int",))
    n = 1000
    keys = np.arange(n, dtype=np.int32)
    pay = np.full(n, 2**31 - 7, dtype=np.int32)
    d = O.join_desc(bs, ps, 0, 0, ["B$1"], buckets=64, tuplesperbucket=4, threads=4)
    cs, _ = O.hashjoin_count_sum(d, O.rows(bs, keys, pay), O.rows(ps, np.repeat(keys, 3)), 0)
    assert cs.count == 3 * n
    exact = 3 * n * (2**31 - 7)
    wrapped = ((exact + 2**31) % 2**32) - 2**31
    assert cs.isum == wrapped


@pytest.mark.parametrize("threads,buckets", [(t, b) for b in (4, 4096) for t in (1, 2, 4, 8, 80)])
def test_queryaggsum_global_known_answer(threads, buckets):
    # unit_tests/queryaggsum_global.cpp: group i holds 1..i, sum = i(i+1)/2,
    # each group appears once in global mode
    s = O.schema(("long",), ("dec",))
    g = np.concatenate([np.full(i, i) for i in range(1, 21)])
    v = np.concatenate([np.arange(1, i + 1) for i in range(1, 21)]).astype(np.float64)
    d = O.agg_desc(s, [0], 1, buckets=buckets, global_=True, threads=threads)
    out = O.aggregate(d, O.rows(s, g, v))
    assert sorted(out["c0"].tolist()) == list(range(1, 21))
    for k, sm in zip(out["c0"], out["c1"]):
        assert round(sm) == k * (k + 1) // 2


def test_threadlocal_then_final_matches_groupby():
    # ThreadLocal mode emits one partial per (thread, group); a final
    # aggregate over the partials equals a numpy group-by
    rng = np.random.default_rng(7)
    s = O.schema(("int",), ("int",))
    g = rng.integers(0, 50, 5000).astype(np.int32)
    v = rng.integers(-2**31, 2**31, 5000).astype(np.int32)
    part = O.aggregate(O.agg_desc(s, [0], 1, buckets=16, threads=5), O.rows(s, g, v))
    assert len(part) >= 50 and len(part) <= 5 * 50
    final = O.aggregate(O.agg_desc(s, [0], 1, buckets=16, threads=1), O.rows(s, part["c0"], part["c1"]))
    got = {int(a): int(b) for a, b in zip(final["c0"], final["c1"])}
    exp = collections.defaultdict(int)
    for a, b in zip(g.tolist(), v.tolist()):
        exp[a] += b
    assert got == {k: ((x + 2**31) % 2**32) - 2**31 for k, x in exp.items()}


def test_aggregate_count_and_no_groupby():
    s = O.schema(("long",), ("long",))
    r = O.rows(s, np.arange(1000), np.arange(1000) * 3)
    out = O.aggregate(O.agg_desc(s, [], -1, threads=1), r)
    assert len(out) == 1 and out["c0"][0] == 1000
    out = O.aggregate(O.agg_desc(s, [], 1, threads=1), r)
    assert out["c0"][0] == 3 * 999 * 1000 // 2


def test_composite_key_bytes_hash():
    # queryagg_compositekey: AggregateCount over a composite key hashed with `bytes`
    rng = np.random.default_rng(9)
    s = O.schema(("int",), ("int",), ("long",))
    a = rng.integers(0, 5, 2000).astype(np.int32)
    b = rng.integers(0, 7, 2000).astype(np.int32)
    out = O.aggregate(O.agg_desc(s, [0, 1], -1, hashfn=O.H_BYTES, buckets=64, global_=True, threads=4),
                      O.rows(s, a, b, np.zeros(2000, np.int64)))
    got = {(int(x), int(y)): int(c) for x, y, c in zip(out["c0"], out["c1"], out["c2"])}
    assert got == dict(collections.Counter(zip(a.tolist(), b.tolist())))


@pytest.mark.parametrize("threads", [1, 3, 7, 15])
def test_querypartition_invariants(threads):
    # unit_tests/querypartition.cpp: 2 x 128K keys in [1, 128K]; multiset
    # preserved, each key in exactly one partition, partitions range-ordered
    n = 128 * 1024
    rng = np.random.default_rng(threads)
    s = O.schema(("long",), ("long",))
    vals = rng.integers(1, n + 1, 2 * n)
    out, counts = O.partition(s, 0, 1, n, threads, O.rows(s, vals, vals ^ 0xABCDEF))
    assert int(counts.sum()) == 2 * n
    assert np.array_equal(np.sort(out["c0"]), np.sort(vals))
    assert np.all(out["c1"] == out["c0"] ^ 0xABCDEF)
    bounds = np.concatenate([[0], np.cumsum(counts.astype(np.int64))])
    prev_max = 0
    for p in range(threads):
        seg = out["c0"][bounds[p]:bounds[p + 1]]
        if len(seg):
            assert seg.min() > prev_max
            prev_max = seg.max()


# ---------------------------------------------------------------- generators
def test_c1_generator_is_permutation_and_closed_form():
    nR, nS = 1 << 12, 1 << 15
    rk, rp, sk, sp = O.gen_c1(nR, nS)
    assert np.array_equal(np.sort(rk), np.arange(nR))
    assert sk.min() >= 0 and sk.max() < nR
    cs = O.c1_closed(nR, nS)
    # independent check: numpy direct-address join
    inv = np.empty(nR, np.int64)
    inv[rk] = np.arange(nR)
    assert cs.count == nS
    assert cs.sum_build == int(rp[inv[sk]].astype(np.int64).sum()) % 2**64
    assert cs.sum_probe == int(sp.astype(np.int64).sum()) % 2**64


def test_c1_oracle_join_equals_closed_form():
    nR, nS = 1 << 12, 1 << 15
    rk, rp, sk, sp = O.gen_c1(nR, nS)
    bs, ps = O.schema(("int",), ("int",)), O.schema(("int",), ("int",))
    d = O.join_desc(bs, ps, 0, 0, ["B$1", "P$1"], buckets=1 << 8, tuplesperbucket=16, threads=4)
    cs_b, _ = O.hashjoin_count_sum(d, O.rows(bs, rk, rp), O.rows(ps, sk, sp), 0)
    cs_p, _ = O.hashjoin_count_sum(d, O.rows(bs, rk, rp), O.rows(ps, sk, sp), 1)
    ref = O.c1_closed(nR, nS)
    assert cs_b.count == ref.count == nS
    assert cs_b.isum % 2**32 == ref.sum_build % 2**32  # int32 accumulator (low 32 bits)
    assert cs_p.isum % 2**32 == ref.sum_probe % 2**32


@pytest.mark.parametrize("zipf", [0.0, 0.5, 1.0])
def test_c2_generator_and_closed_form(zipf):
    nR, nS = 1 << 12, 1 << 16
    rk, rp, sk = O.gen_c2(nR, nS, zipf)
    assert len(np.unique(rk)) == nR  # bijective mix: unique keys
    assert rp.min() >= 0 and rp.max() < 2**31
    assert np.isin(sk, rk).all()
    ref = O.c2_closed(nR, nS, zipf)
    pos = {int(k): i for i, k in enumerate(rk)}
    rows = np.array([pos[int(k)] for k in sk])
    assert ref.count == nS and ref.sum_build == int(rp[rows].sum())
    if zipf == 1.0:
        # the hottest rank's share is 1 / H_n
        hn = sum(1.0 / r for r in range(1, nR + 1))
        top = np.bincount(rows).max() / nS
        assert abs(top - 1 / hn) < 0.01
    # the chained-bucket restatement agrees with the closed form
    secs, cs = O.baseline_join(8, rk, rp, sk, 1 << 10, 4, 4)
    assert cs.count == ref.count and cs.sum_build == ref.sum_build


@pytest.mark.parametrize("join,n", [(0, 1000), (7, 5000), (1023, 1 << 14)])
def test_c4_generator_and_closed_form(join, n):
    """Build keys are a permutation of 0..n-1, about 10% of probe keys fall
    in the disjoint miss range [n, n + 2^30), and the closed form agrees with
    the chained-bucket restatement of the join and with a direct-address join."""
    m = 8 * n
    rk, rp, sk, sp = O.gen_c4(join, n, m)
    assert np.array_equal(np.sort(rk), np.arange(n))
    miss = sk >= n
    assert (sk[miss] < n + 2**30).all() and (sk >= 0).all()
    assert abs(miss.mean() - 0.1) < 4 * np.sqrt(0.09 / m) + 1e-3
    ref = O.c4_closed(join, n, m)
    inv = np.empty(n, np.int64)
    inv[rk] = np.arange(n)
    hit = ~miss
    assert ref.count == int(hit.sum())
    assert ref.sum_build == int(rp[inv[sk[hit]]].astype(np.int64).sum()) % 2**64
    assert ref.sum_probe == int(sp[hit].astype(np.int64).sum()) % 2**64
    bs, ps = O.schema(("int",), ("int",)), O.schema(("int",), ("int",))
    d = O.join_desc(bs, ps, 0, 0, ["B$1", "P$1"], buckets=1 << 10, tuplesperbucket=4, threads=4)
    cs_b, _ = O.hashjoin_count_sum(d, O.rows(bs, rk, rp), O.rows(ps, sk, sp), 0)
    cs_p, _ = O.hashjoin_count_sum(d, O.rows(bs, rk, rp), O.rows(ps, sk, sp), 1)
    assert cs_b.count == ref.count
    assert cs_b.isum % 2**32 == ref.sum_build % 2**32
    assert cs_p.isum % 2**32 == ref.sum_probe % 2**32
    many = O.c4_closed_many(join, [n], [m], threads=2)
    assert (many[0].count, many[0].sum_build, many[0].sum_probe) == (ref.count, ref.sum_build, ref.sum_probe)


def test_zipf_cdf_shape():
    cdf = O.zipf_cdf(1000, 1.0)
    assert np.all(np.diff(cdf) > 0) and cdf[-1] == 1.0
    hn = sum(1.0 / r for r in range(1, 1001))
    assert abs(cdf[0] - 1 / hn) < 1e-12


def test_c3_closed_form_matches_oracle_join_aggregate():
    nD, nF, norders = 20000, 80000, 20000
    dk, dc, fk, fp = O.gen_c3(nD, nF, norders)
    assert np.array_equal(np.sort(dk), np.arange(1, nD + 1))
    joined, cents = O.c3_closed(nD, nF, norders)
    assert joined == nF  # every fact orderkey is in 1..norders == D's keys
    # join F with D on orderkey -> (custkey, price), then sum(price) group by custkey
    ds, fs = O.schema(("int",), ("int",)), O.schema(("int",), ("dec",))
    d = O.join_desc(ds, fs, 0, 0, ["B$1", "P$1"], buckets=1 << 12, tuplesperbucket=4, threads=4)
    jout, _ = O.hashjoin(d, O.rows(ds, dk, dc), O.rows(fs, fk, fp), cap=2 * nF)
    js = O.join_outschema(d)
    agg = O.aggregate(O.agg_desc(js, [0], 1, buckets=1 << 10, global_=True, threads=4), jout)
    assert len(agg) == np.count_nonzero(cents)
    for c, s in zip(agg["c0"], agg["c1"]):
        exact = cents[c] / 100.0
        assert abs(s - exact) <= 1e-12 * abs(exact)


def test_c5_closed_form_is_c2_family_with_offsets():
    """A served query reads rows first.. of a table's uniform FK column: with
    first = 0 it is the C2 closed form; with an offset it equals a direct
    computation over the generated table."""
    n, m = 1 << 12, 5000
    a, b = O.c5_closed(3, n, 0, m), O.c2_closed(n, m, sub=3)
    assert (a.count, a.sum_build) == (b.count, b.sum_build)
    rk, rp, sk = O.gen_c2(n, (1 << 13) + m, sub=3)
    pos = {int(k): i for i, k in enumerate(rk)}
    tail = sk[1 << 13:]
    c = O.c5_closed(3, n, 1 << 13, m, threads=3)
    assert c.count == m and c.sum_build == int(sum(int(rp[pos[int(k)]]) for k in tail)) % 2**64


def _ht_fill(nbuckets, tpb, nvalues):
    t = O.lib().orc_ht_create(nbuckets, 4 * tpb, 4)
    for v in range(nvalues):
        C.cast(O.lib().orc_ht_allocate(t, v % nbuckets), C.POINTER(C.c_int32))[0] = v
    return t


def _ht_values_once(t, nbuckets, nvalues):
    seen = np.zeros(nvalues, np.int64)
    out = (C.c_void_p * 64)()
    for b in range(nbuckets):
        n = O.lib().orc_ht_bucket_tuples(t, b, out, 64)
        for i in range(min(n, 64)):
            v = C.cast(out[i], C.POINTER(C.c_int32))[0]
            assert 0 <= v < nvalues and v % nbuckets == b
            seen[v] += 1
    return seen


@pytest.mark.parametrize("tpb", [1, 8])
def test_hashtable_serialize_round_trip(tmp_path, tpb):
    """unit_tests/testhashtable.cpp:87-182: a table of int tuples (value v in
    bucket v % buckets) serialized to a segment file (util/hashtable.cpp:
    99-121) and deserialized into a fresh table of the same geometry
    (:178-201): every value appears exactly once, in its bucket.  A table
    that spilled into overflow chunks is refused (the reference asserts
    spills == 0)."""
    rng = np.random.default_rng(9 + tpb)
    nbuckets = int(rng.integers(100, 10000)) if tpb == 1 else 256
    nvalues = nbuckets * tpb - (0 if tpb == 1 else 37)
    path = str(tmp_path / "ht.seg").encode()
    t = _ht_fill(nbuckets, tpb, nvalues)
    assert O.lib().orc_ht_spills(t) == 0
    assert O.lib().orc_ht_serialize(t, path) == 0
    O.lib().orc_ht_destroy(t)
    assert os.path.getsize(path) == nbuckets * (16 + 4 * tpb)  # BucketHeader + bucket space, per bucket
    u = O.lib().orc_ht_create(nbuckets, 4 * tpb, 4)
    assert O.lib().orc_ht_deserialize(u, path) == 0
    assert np.all(_ht_values_once(u, nbuckets, nvalues) == 1)
    O.lib().orc_ht_destroy(u)
    spilled = _ht_fill(16, 1, 40)
    assert O.lib().orc_ht_spills(spilled) > 0
    assert O.lib().orc_ht_serialize(spilled, path) == -1
    O.lib().orc_ht_destroy(spilled)
