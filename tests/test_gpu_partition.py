"""GPU parity of the range partitioner behind `gpupartition`.

Through the C ABI: the partitioned rows must equal, byte for byte, the
oracle's restatement of PartitionOp (oracle.partition: exactrange hash,
thread t owning input slice t, idxstart order), with the same per-bucket
counts; with sorting, each bucket equals the oracle's bucket stably sorted by
the sort column.  Through the plan engine: the invariants of the reference's
unit_tests/querypartition.cpp (every tuple once, one thread per value,
range-partitioned thread ids) on a scan -> gpupartition -> threadidprepend ->
merge plan.
"""
import textwrap

import numpy as np
import pytest

import pythia_amd as pa
from oracle import oracle as O
from pythia_amd import plan as P

pytestmark = pytest.mark.gpu

COLTYPE = {np.dtype(np.int32): pa.COL_INT32, np.dtype(np.int64): pa.COL_INT64, np.dtype(np.float64): pa.COL_F64}


def _slices(n, T):
    return [(n * t // T, n * (t + 1) // T) for t in range(T)]


def _gpu_partition(ctx, s, rows, attr, rmin, rmax, buckets, sort_attr=None):
    dt = s.dtype()
    kdt, koff = dt.fields[f"c{attr}"][:2]
    sort_off, sort_type = None, pa.COL_INT64
    if sort_attr is not None:
        sdt, sort_off = dt.fields[f"c{sort_attr}"][:2]
        sort_type = COLTYPE[sdt]
    p = pa.Partition(ctx, dt.itemsize, koff, kdt.itemsize, rmin, rmax, buckets, sort_off, sort_type)
    try:
        for t, (a, b) in enumerate(_slices(len(rows), buckets)):
            if b > a:
                p.append(t, rows[a:b])
        counts = p.finish()
        parts = [p.pull(b, dt) for b in range(buckets)]
    finally:
        p.close()
    return counts, parts


def _case(name, rng):
    """(schema, rows, attr, rmin, rmax)"""
    if name == "long_uniform":
        s = O.schema(("long",), ("long",))
        v = rng.integers(1, 100_001, 250_000)
        return s, O.rows(s, v, v ^ 0xABCDEF), 0, 1, 100_000
    if name == "negative_edges":  # keys below min and above max that C division still maps into a bucket
        s = O.schema(("long",), ("int",))
        v = rng.integers(-1285, 1002, 200_000)  # range [-1000, 1000], 7 buckets: bucket range 286
        return s, O.rows(s, v, np.arange(v.size)), 0, -1000, 1000
    if name == "odd_width_int_key":  # 14-byte rows, int key at byte offset 2
        s = O.schema(("char", 2), ("int",), ("dec",))
        v = rng.integers(0, 5000, 100_000)
        c = np.where(np.arange(v.size) % 2 == 0, b"az", b"bz").astype("S2")
        return s, O.rows(s, c, v, rng.random(v.size)), 1, 0, 4999
    raise ValueError(name)


@pytest.mark.parametrize("name,buckets", [("long_uniform", 1), ("long_uniform", 3), ("long_uniform", 16),
                                          ("long_uniform", 127), ("negative_edges", 7),
                                          ("odd_width_int_key", 5)])
def test_partition_matches_oracle(ctx, name, buckets):
    rng = np.random.default_rng(buckets * 7919 + len(name))
    s, rows, attr, rmin, rmax = _case(name, rng)
    exp, exp_counts = O.partition(s, attr, rmin, rmax, buckets, rows)
    counts, parts = _gpu_partition(ctx, s, rows, attr, rmin, rmax, buckets)
    assert counts.tolist() == exp_counts.tolist()
    assert np.concatenate(parts).tobytes() == exp.tobytes()


@pytest.mark.parametrize("sort_col", [0, 1, 2])
def test_partition_sorted_buckets(ctx, sort_col):
    """sort = "yes": each bucket ordered by the sort column (int, dec with
    negatives and signed zeros, long); ties keep input order."""
    rng = np.random.default_rng(17 + sort_col)
    s = O.schema(("int",), ("dec",), ("long",))
    n = 120_000
    k = rng.integers(0, 1000, n)
    d = np.round(rng.normal(0, 50, n), 1)
    d[rng.random(n) < 0.05] = -0.0
    d[rng.random(n) < 0.05] = 0.0
    x = rng.integers(-(1 << 40), 1 << 40, n)
    rows = O.rows(s, k, d, x)
    buckets = 9
    exp, exp_counts = O.partition(s, 0, 0, 999, buckets, rows)
    counts, parts = _gpu_partition(ctx, s, rows, 0, 0, 999, buckets, sort_attr=sort_col)
    assert counts.tolist() == exp_counts.tolist()
    at = 0
    for b in range(buckets):
        e = exp[at:at + int(exp_counts[b])]
        at += int(exp_counts[b])
        key = e[f"c{sort_col}"]
        want = e[np.argsort(key, kind="stable")]
        assert parts[b].tobytes() == want.tobytes(), f"bucket {b}"


def test_partition_out_of_range_key_is_an_error(ctx):
    s = O.schema(("long",), ("long",))
    rows = O.rows(s, np.array([1, 5, 200]), np.array([0, 0, 0]))  # 200 lies past the last of 4 buckets of [1, 100]
    p = pa.Partition(ctx, 16, 0, 8, 1, 100, 4)
    p.append(0, rows)
    with pytest.raises(pa.PgpuError):
        p.finish()
    p.close()


def test_partition_reset_reuse_and_empty(ctx):
    s = O.schema(("long",), ("long",))
    p = pa.Partition(ctx, 16, 0, 8, 0, 99, 4)
    assert p.finish().tolist() == [0, 0, 0, 0]
    assert len(p.pull(2, s.dtype())) == 0
    rng = np.random.default_rng(3)
    for n in (5000, 300):
        p.reset()
        v = rng.integers(0, 100, n)
        rows = O.rows(s, v, np.arange(n))
        p.append(1, rows)  # one source only; sources 0, 2, 3 stay empty
        counts = p.finish()
        got = np.concatenate([p.pull(b, s.dtype()) for b in range(4)])
        assert counts.sum() == n
        order = np.argsort((v // 25), kind="stable")
        assert got.tobytes() == rows[order].tobytes()
    p.close()


def _querypartition_plan(tmp_path, f, threads, tuples, buffsize, sort):
    tl = ", ".join(str(i) for i in range(threads))
    return textwrap.dedent(f"""
        path = "{tmp_path}/"; buffsize = {buffsize};
        treeroot = {{ name = "merge"; input = {{ name = "tid"; input = {{ name = "part"; input = {{ name = "scan"; }}; }}; }}; }};
        merge = {{ type = "merge"; threads = {threads}; }};
        tid = {{ type = "threadidprepend"; }};
        part = {{ type = "gpupartition"; attr = 0; maxtuples = {2 * tuples * threads}; range = [ 1, {tuples} ];
                 buckets = {threads}; sort = "{sort}"; }};
        scan = {{ type = "parallelscan"; filetype = "text"; files = ( "{f}" ); mapping = ( [ {tl} ] );
                  schema = ( "long", "long" ); }};
    """)


@pytest.mark.parametrize("threads,tuples,buffsize,sort", [(1, 4096, 20, "no"), (3, 131072, 4096, "no"),
                                                          (8, 131072, 4096, "yes"), (15, 16384, 20, "no")])
def test_querypartition_invariants(tmp_path, threads, tuples, buffsize, sort):
    """unit_tests/querypartition.cpp: 2 x TUPLES rows (v, v ^ 0xABCDEF), v in
    [1, TUPLES]; every row comes out once, all copies of v from the same
    thread, thread ids nondecreasing in v (range partitioned), and with
    sort = "yes" each thread's rows in nondecreasing v."""
    rng = np.random.default_rng(threads)
    v = rng.integers(1, tuples + 1, 2 * tuples)
    (tmp_path / "part.txt").write_text("".join(f"{x}|{x ^ 0xABCDEF}\n" for x in v.tolist()))
    out, meta = P.run(_querypartition_plan(tmp_path, "part.txt", threads, tuples, buffsize, sort))
    assert meta["schema"] == "int,long,long"
    tid, val, chk = out["c0"], out["c1"], out["c2"]
    assert len(out) == len(v)
    assert np.all((chk ^ 0xABCDEF) == val)
    assert np.all((tid >= 0) & (tid < threads))
    assert np.array_equal(np.bincount(val, minlength=tuples + 1), np.bincount(v, minlength=tuples + 1))
    src = np.full(tuples + 1, -1)
    src[val] = tid
    assert np.all(src[val] == tid), "a value appears in two partitions"
    seen = src[src >= 0]
    assert np.all(np.diff(seen) >= 0), "output is not range partitioned"
    br = (tuples - 1 + threads) // threads
    assert np.all(tid == (val - 1) // br)
    if sort == "yes":
        for t in range(threads):
            assert np.all(np.diff(val[tid == t]) >= 0)


def test_partition_plan_rejects_bad_config(tmp_path):
    (tmp_path / "p.txt").write_text("1|2\n")
    base = _querypartition_plan(tmp_path, "p.txt", 2, 10, 4096, "no")
    with pytest.raises(Exception):
        P.run(base.replace("buckets = 2;", "buckets = 128;"))
    with pytest.raises(Exception):
        P.run(base.replace('sort = "no";', ""))
    with pytest.raises(Exception):
        P.run(base.replace("maxtuples = 40;", ""))
