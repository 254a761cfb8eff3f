"""The oracle against the compiled reference (oracle/_ref/refjoin, built by
oracle/ref.mk from the reference's own sources when they are present).

refjoin runs the reference's HashJoinOp -> AggregateSum / AggregateCount ->
MergeOp -> final aggregate on the generated inputs; the oracle's closed forms
and its chained-bucket restatement must give the same count and sums,
including the int32 accumulator wrap of AggregateSum as compiled."""
import json
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O

REF = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref", "refjoin")
pytestmark = pytest.mark.skipif(not os.path.exists(REF), reason="compiled reference not built (oracle/ref.mk)")


def refjoin(wl, lr, ls, threads, zipf, buckets, tpb, agg):
    out = subprocess.run([REF, wl, str(lr), str(ls), str(threads), str(zipf), str(buckets), str(tpb), agg],
                         stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=300, cwd="/tmp")
    assert out.returncode == 0, out.stderr
    return json.loads(out.stdout)


def i32(x):
    return int(np.int32(np.uint32(x & 0xFFFFFFFF)))


@pytest.mark.parametrize("threads", [1, 3, 8])
@pytest.mark.parametrize("tpb", [1, 4, 16])
def test_c1_reference_equals_oracle(threads, tpb):
    lr, ls = 12, 17
    ref = O.c1_closed(1 << lr, 1 << ls)
    assert refjoin("c1", lr, ls, threads, 0, 1 << 8, tpb, "count")["value"] == ref.count
    assert refjoin("c1", lr, ls, threads, 0, 1 << 8, tpb, "sum")["value"] == i32(ref.sum_build)
    assert refjoin("c1", lr, ls, threads, 0, 1 << 8, tpb, "probesum")["value"] == i32(ref.sum_probe)


@pytest.mark.parametrize("zipf", [0.0, 0.5, 1.0])
@pytest.mark.parametrize("threads", [2, 8])
def test_c2_reference_equals_oracle(zipf, threads):
    lr, ls = 14, 18
    ref = O.c2_closed(1 << lr, 1 << ls, zipf)
    assert refjoin("c2", lr, ls, threads, zipf, 1 << 12, 4, "count")["value"] == ref.count
    got = refjoin("c2", lr, ls, threads, zipf, 1 << 12, 4, "sum")["value"]
    assert got % 2**64 == ref.sum_build
    # and the oracle's restated chained-bucket join agrees too
    rk, rp, sk = O.gen_c2(1 << lr, 1 << ls, zipf)
    secs, cs = O.baseline_join(8, rk, rp, sk, 1 << 12, 4, threads)
    assert (cs.count, cs.sum_build) == (ref.count, ref.sum_build)


@pytest.mark.parametrize("threads", [1, 4])
def test_c4_reference_equals_oracle(threads):
    """The batched-join family: the reference's plan run once per join gives
    the closed forms' sums (int32 accumulator) for each join."""
    import sys

    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    import pythia_amd as pa

    first, count = 5, 4
    sizes = pa.sizes(4, count, 10, 22, first=first)
    ref = O.c4_closed_many(first, sizes, 8 * sizes, threads=4)
    for agg, field in (("sum", "sum_build"), ("probesum", "sum_probe")):
        out = subprocess.run([REF, "c4", str(first), str(count), str(threads), "4", agg], stdout=subprocess.PIPE,
                             stderr=subprocess.PIPE, text=True, timeout=300, cwd="/tmp")
        assert out.returncode == 0, out.stderr
        r = json.loads(out.stdout)
        assert r["tuples"] == int(9 * sizes.sum())
        assert r["values"] == [i32(getattr(c, field)) for c in ref]


@pytest.mark.parametrize("threads", [1, 4])
def test_c3_reference_equals_oracle(threads):
    """Join + group-by: the reference's hashjoin -> aggregate_sum(field 0)
    plan finds the closed form's groups, and its FP64 total matches the exact
    integer-cent total within rounding."""
    nD, nF = 20000, 100000
    out = subprocess.run([REF, "c3", str(nD), str(nF), str(threads)], stdout=subprocess.PIPE,
                         stderr=subprocess.PIPE, text=True, timeout=300, cwd="/tmp")
    assert out.returncode == 0, out.stderr
    r = json.loads(out.stdout)
    joined, cents = O.c3_closed(nD, nF, norders=nD)
    assert r["groups"] == int(np.count_nonzero(cents))
    exact = int(cents.sum()) / 100.0
    assert abs(r["total"] - exact) <= 1e-12 * exact


# The oracle's hash functions against the reference's own TupleHasher::create
# (hash.cpp:99-234; hash.h:129-479) on fixed vectors: values around zero, the
# extremes of both key widths, the range bounds and random values.
_I32 = [0, 1, -1, 2, 7, 8, 255, 256, 4095, 4096, 12345, -12345, -1000, -1001, -999, 99999, 100000, 100001,
        2**31 - 1, -2**31, -2**31 + 1, 1 << 20, (1 << 20) - 1]
_I64 = _I32 + [2**40 + 3, -(2**40), 2**63 - 1, -2**63, -2**63 + 1, 2**62, 0x123456789ABCDEF, -0x123456789ABCDEF]
_RNG = np.random.default_rng(77)
_I32 += [int(x) for x in _RNG.integers(-2**31, 2**31 - 1, 40)]
_I64 += [int(x) for x in _RNG.integers(-2**63, 2**63 - 1, 40, dtype=np.int64)]
_IN_RANGE = [int(x) for x in _RNG.integers(-1000, 100001, 40)]  # range hashes overflow (UB) far outside


def _ref_hashes(ty, node, values):
    out = subprocess.run([REF, "hash", ty, node] + [str(v) for v in values], stdout=subprocess.PIPE,
                         stderr=subprocess.PIPE, text=True, timeout=60, cwd="/tmp")
    assert out.returncode == 0, out.stderr
    return json.loads(out.stdout)


@pytest.mark.parametrize("ty", ["int", "long"])
@pytest.mark.parametrize("node,fn,buckets,rmin,rmax,skip,values", [
    ('{ fn = "modulo"; buckets = 1000; field = 0; }', O.H_MODULO, 1000, 0, 0, 0, None),
    ('{ fn = "parammodulo"; buckets = 256; field = 0; offset = 7; skipbits = 3; }', O.H_PARAMMODULO, 256, 7, 0, 3,
     None),
    ('{ fn = "knuth"; buckets = 256; field = 0; offset = 5; skipbits = 2; }', O.H_KNUTH, 256, 5, 0, 2, None),
    ('{ fn = "range"; buckets = 64; field = 0; range = [-1000, 100000]; }', O.H_RANGE, 64, -1000, 100000, 0,
     "in_range"),
    ('{ fn = "range"; buckets = 100; field = 0; range = [0, 1048575]; }', O.H_RANGE, 100, 0, 1048575, 0, "in_range"),
    ('{ fn = "exactrange"; buckets = 10; field = 0; range = [-1000, 100000]; }', O.H_EXACTRANGE, 10, -1000, 100000,
     0, "in_range"),
    ('{ fn = "willis"; buckets = 4096; field = 0; }', O.H_WILLIS, 4096, 0, 0, 0, None),
    ('{ fn = "willis"; buckets = 3; field = 0; }', O.H_WILLIS, 3, 0, 0, 0, None),
    ('{ fn = "tpchorderkey"; buckets = 1024; field = 0; }', O.H_TPCH, 1024, 0, 0, 0, None),
    ('{ fn = "bytes"; buckets = 4096; field = 0; }', O.H_BYTES, 4096, 0, 0, 0, None),
    ('{ fn = "alwayszero"; }', O.H_ALWAYSZERO, 1, 0, 0, 0, None),
])
def test_hashers_equal_the_reference(ty, node, fn, buckets, rmin, rmax, skip, values):
    if values == "in_range":
        vals = [-1000, -999, 0, 1, 50000, 99999, 100000] + _IN_RANGE
        vals = [v for v in vals if rmin <= v <= rmax]
    else:
        vals = _I32 if ty == "int" else _I64
    s = O.schema((ty,))
    h = O.hasher(fn, buckets, s, 0, rmin=rmin, rmax=rmax, skipbits=skip)
    ref = _ref_hashes(ty, node, vals)
    assert ref["buckets"] == h.buckets
    r = O.rows(s, np.array(vals, np.int32 if ty == "int" else np.int64))
    w = 4 if ty == "int" else 8
    got = [O.lib().orc_hash_tuple(__import__("ctypes").byref(h), r.ctypes.data + i * w) for i in range(len(vals))]
    assert got == ref["hashes"]


def test_tpchq1magic_equals_the_reference():
    """TpchQ1MagicByteHasher over the (returnflag, linestatus) char(1) pair:
    every byte pattern of the flags TPC-H uses and random 4-byte spans."""
    import ctypes

    vals = [0, 16, 65536, 2, 4294967295, 0x0046_0041, 0x004F_0052, 0x0046_004E, 0x004F_0041]
    vals += [int(x) for x in _RNG.integers(0, 2**32, 60)]
    ref = _ref_hashes("chars", '{ fn = "tpchq1magic"; buckets = 4; fieldrange = [0, 1]; }', vals)
    s = O.schema(("char", 1), ("char", 1))
    h = O.hasher(O.H_TPCHQ1, 4, s, 0, fieldmax=1)
    assert ref["buckets"] == h.buckets == 4
    buf = np.array(vals, np.uint32)
    got = [O.lib().orc_hash_tuple(ctypes.byref(h), buf.ctypes.data + 4 * i) for i in range(len(vals))]
    assert got == ref["hashes"]


@pytest.mark.parametrize("nb,tpb,nv", [(5000, 1, 5000), (256, 8, 2000), (1024, 3, 2900)])
def test_hashtable_segments_interchange_with_the_reference(tmp_path, nb, tpb, nv):
    """HashTable::serialize / deserialize (util/hashtable.cpp:99-201): a
    segment the reference writes is read by the oracle with every value
    once in its bucket, and one the oracle writes is read by the reference
    the same way; both are the same size."""
    import ctypes

    ref_path, orc_path = str(tmp_path / "ref.seg"), str(tmp_path / "orc.seg")
    out = subprocess.run([REF, "htwrite", str(nb), str(tpb), str(nv), ref_path], stdout=subprocess.PIPE,
                         stderr=subprocess.PIPE, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    u = O.lib().orc_ht_create(nb, 4 * tpb, 4)
    assert O.lib().orc_ht_deserialize(u, ref_path.encode()) == 0
    seen = np.zeros(nv, np.int64)
    ptrs = (ctypes.c_void_p * 64)()
    for b in range(nb):
        n = O.lib().orc_ht_bucket_tuples(u, b, ptrs, 64)
        for i in range(n):
            v = ctypes.cast(ptrs[i], ctypes.POINTER(ctypes.c_int32))[0]
            assert v % nb == b
            seen[v] += 1
    assert np.all(seen == 1)
    assert O.lib().orc_ht_serialize(u, orc_path.encode()) == 0
    O.lib().orc_ht_destroy(u)
    assert os.path.getsize(orc_path) == os.path.getsize(ref_path)
    out = subprocess.run([REF, "htread", str(nb), str(tpb), str(nv), orc_path], stdout=subprocess.PIPE,
                         stderr=subprocess.PIPE, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    assert json.loads(out.stdout) == {"once": nv, "bad": 0}
