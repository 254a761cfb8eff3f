"""`gpuhashjoin` / `gpuaggregate_sum` / `gpuaggregate_count` through the plan
engine (Operator lifecycle, Page protocol, merge workers) on the GPU, checked
against the oracle's restatement of the reference operators.  Plans follow
the reference's unit tests (unit_tests/queryhashjoin.cpp, queryaggsum.cpp,
queryaggsum_global.cpp, queryagg_compositekey.cpp)."""
import textwrap

import numpy as np
import pytest

from oracle import oracle as O
from pythia_amd import plan as P

pytestmark = pytest.mark.gpu


def write_text(tmp_path, name, rows):
    (tmp_path / name).write_text("".join("|".join(str(x) for x in r) + "\n" for r in rows))
    return name


def write_bin(tmp_path, name, arr):
    (tmp_path / name).write_bytes(np.ascontiguousarray(arr).tobytes())
    return name


def join_plan(tmp_path, fb, fp, threads, bschema, pschema, projection, buffsize=16, groups=None, bmap="",
              pmap="", filetype="text", checker=True, extra="", bj=0, pj=0):
    groups = groups or f"[ {', '.join(str(i) for i in range(threads))} ]"
    join = 'name = "join"; build = { name = "b"; }; probe = { name = "p"; };'
    if checker:
        join = f'name = "chk"; input = {{ {join} }};'
    return textwrap.dedent(f"""
        path = "{tmp_path}/"; buffsize = {buffsize};
        treeroot = {{ name = "merge"; input = {{ {join} }}; }};
        merge = {{ type = "merge"; threads = {threads}; }};
        chk = {{ type = "checker_callstate"; }};
        join = {{ type = "gpuhashjoin"; hash = {{ fn = "modulo"; buckets = 16; }}; threadgroups = ( {groups} );
                 tuplesperbucket = 2; allocpolicy = "local"; buildjattr = {bj}; probejattr = {pj};
                 projection = ( {projection} ); {extra} }};
        b = {{ type = "parallelscan"; filetype = "{filetype}"; files = ( {fb} ); schema = ( {bschema} ); {bmap} }};
        p = {{ type = "parallelscan"; filetype = "{filetype}"; files = ( {fp} ); schema = ( {pschema} ); {pmap} }};
    """)


def test_queryhashjoin_known_answer(tmp_path):
    # unit_tests/queryhashjoin.cpp with type = "gpuhashjoin": 4 threads,
    # buffsize 16, every key 1..20 exactly once, P$1 == v + 0.1
    i = np.arange(1, 21)
    fb = write_text(tmp_path, "b.tbl", zip(i, i))
    fp = write_text(tmp_path, "p.tbl", [(v, f"{v + 0.1:.1f}") for v in i])
    out, meta = P.run(join_plan(tmp_path, f'"{fb}"', f'"{fp}"', 4, '"long", "long"', '"long", "dec"',
                                '"B$1", "P$1"'))
    assert meta["schema"] == "long,dec"
    assert sorted(out["c0"].tolist()) == list(range(1, 21))
    assert np.all(out["c1"] == out["c0"] + 0.1)


def rows_sorted(arr):
    return sorted(tuple(x.tolist() if hasattr(x, "tolist") else x for x in r) for r in arr)


@pytest.mark.parametrize("threads,buffsize,batch", [(1, 64, 1 << 20), (4, 40, 1 << 20), (3, 4096, 1000)])
def test_gpuhashjoin_matches_oracle(tmp_path, threads, buffsize, batch):
    """int build key vs long probe key (compared by value), duplicates,
    negative keys, char and dec payload columns, a projection mixing sides."""
    rng = np.random.default_rng(threads)
    nb, np_ = 3000, 20000
    bs = O.schema(("int",), ("char", 6), ("dec",))
    ps = O.schema(("long",), ("long",), ("int",))
    bk = rng.integers(-500, 500, nb).astype(np.int32)
    bc = np.array([b"k%04d" % (x % 10000) for x in rng.integers(0, 10000, nb)], dtype="S6")
    bd = rng.random(nb)
    pk = rng.integers(-600, 600, np_)
    pl = rng.integers(-2**62, 2**62, np_)
    pi = rng.integers(-1000, 1000, np_).astype(np.int32)
    brow, prow = O.rows(bs, bk, bc, bd), O.rows(ps, pk, pl, pi)
    proj = ["P$2", "B$1", "P$0", "B$2", "B$0", "P$1"]
    d = O.join_desc(bs, ps, 0, 0, proj, buckets=256, tuplesperbucket=4, threads=2)
    exp, _ = O.hashjoin(d, brow, prow, cap=len(prow) * 20)
    fb, fp = write_bin(tmp_path, "b.bin", brow), write_bin(tmp_path, "p.bin", prow)
    out, meta = P.run(join_plan(tmp_path, f'"{fb}"', f'"{fp}"', threads, '"int", "char(5)", "dec"',
                                '"long", "long", "int"', ", ".join(f'"{x}"' for x in proj), buffsize=buffsize,
                                filetype="binary", extra=f"gpu = {{ probebatch = {batch}L; }};"))
    assert meta["schema"] == "int,char(5),long,dec,int,long"
    assert len(out) == len(exp) > 0
    rb = meta["row_bytes"]
    raw_got, raw_want = out.tobytes(), exp.tobytes()
    got = sorted(raw_got[i * rb:(i + 1) * rb] for i in range(len(out)))
    want = sorted(raw_want[i * rb:(i + 1) * rb] for i in range(len(exp)))
    assert got == want  # byte-exact multiset of output rows


def test_threadgroups_see_only_their_build_side(tmp_path):
    """threadgroups [[0,1],[2,3]] with mapped scans: group g joins only its
    own build file (operators/join.cpp:198-239)."""
    b0 = write_text(tmp_path, "b0.tbl", [(k, 100 + k) for k in range(0, 50)])
    b1 = write_text(tmp_path, "b1.tbl", [(k, 200 + k) for k in range(0, 50)])
    p0 = write_text(tmp_path, "p0.tbl", [(k, 1) for k in range(0, 50)])
    p1 = write_text(tmp_path, "p1.tbl", [(k, 2) for k in range(0, 50)])
    out, _ = P.run(join_plan(tmp_path, f'"{b0}", "{b1}"', f'"{p0}", "{p1}"', 4, '"long", "long"', '"long", "long"',
                             '"B$1", "P$1"', groups="[ 0, 1 ], [ 2, 3 ]", bmap="mapping = ( ( 0, 1 ), ( 2, 3 ) );",
                             pmap="mapping = ( ( 0, 1 ), ( 2, 3 ) );"))
    got = sorted(zip(out["c0"].tolist(), out["c1"].tolist()))
    assert got == sorted([(100 + k, 1) for k in range(50)] + [(200 + k, 2) for k in range(50)])


def test_empty_probe_is_finished(tmp_path):
    fb = write_text(tmp_path, "b.tbl", [(1, 1)])
    (tmp_path / "p.tbl").write_text("")
    out, _ = P.run(join_plan(tmp_path, f'"{fb}"', '"p.tbl"', 2, '"long", "long"', '"long", "long"', '"B$1"'))
    assert len(out) == 0


def agg_plan(tmp_path, f, schema, kind, threads, buckets, fields="field = 0;", sumfield=1, global_=True,
             hashfield="field = 0;", fn="modulo", two_level=False, buffsize=16):
    g = f"global = true; threads = {threads};" if global_ else ""
    sf = f"sumfield = {sumfield};" if kind == "gpuaggregate_sum" else ""
    if two_level:
        tree = ('name = "final"; input = { name = "merge"; input = { name = "chk"; input = '
                '{ name = "agg"; input = { name = "scan"; }; }; }; };')
    else:
        tree = 'name = "merge"; input = { name = "chk"; input = { name = "agg"; input = { name = "scan"; }; }; };'
    return textwrap.dedent(f"""
        path = "{tmp_path}/"; buffsize = {buffsize};
        treeroot = {{ {tree} }};
        merge = {{ type = "merge"; threads = {threads}; }};
        chk = {{ type = "checker_callstate"; }};
        agg = {{ type = "{kind}"; {fields} {sf} {g} hash = {{ fn = "{fn}"; buckets = {buckets}; {hashfield} }}; }};
        final = {{ type = "{kind}"; {fields} sumfield = {len(fields.split(",")) if "fields" in fields else 1};
                  hash = {{ fn = "{fn}"; buckets = {buckets}; {hashfield} }}; }};
        scan = {{ type = "parallelscan"; filetype = "text"; files = ( "{f}" ); schema = ( {schema} ); }};
    """)


def createaggfile(tmp_path):
    # unit_tests/queryaggsum.cpp createaggfile: group i holds 1..i
    return write_text(tmp_path, "agg.tbl", [(i, k) for i in range(1, 21) for k in range(1, i + 1)])


@pytest.mark.parametrize("threads,buckets", [(t, b) for b in (4, 4096) for t in (1, 2, 4, 8)])
def test_queryaggsum_global_known_answer(tmp_path, threads, buckets):
    f = createaggfile(tmp_path)
    out, meta = P.run(agg_plan(tmp_path, f, '"long", "dec"', "gpuaggregate_sum", threads, buckets))
    assert meta["schema"] == "long,dec"
    assert sorted(out["c0"].tolist()) == list(range(1, 21))  # each group exactly once
    for k, s in zip(out["c0"], out["c1"]):
        assert round(s) == k * (k + 1) // 2


def test_threadlocal_partials_then_final_aggregate(tmp_path):
    # unit_tests/queryaggsum.cpp plan shape: partial aggregate per thread ->
    # merge(4) -> final aggregate; int sums wrap as the reference's int column
    rng = np.random.default_rng(3)
    g = rng.integers(0, 300, 20000)
    v = rng.integers(-2**31, 2**31, 20000).astype(np.int32)
    f = write_text(tmp_path, "a.tbl", zip(g, v))
    out, _ = P.run(agg_plan(tmp_path, f, '"long", "int"', "gpuaggregate_sum", 4, 64, global_=False,
                            two_level=True))
    exp = {}
    for a, b in zip(g.tolist(), v.tolist()):
        exp[a] = exp.get(a, 0) + b
    assert {int(a): int(b) for a, b in zip(out["c0"], out["c1"])} == \
        {k: ((x + 2**31) % 2**32) - 2**31 for k, x in exp.items()}


def test_composite_key_count(tmp_path):
    # unit_tests/queryagg_compositekey.cpp: aggregate_count, composite key, `bytes` hash
    rng = np.random.default_rng(5)
    a, b = rng.integers(0, 6, 5000), rng.integers(0, 9, 5000)
    f = write_text(tmp_path, "c.tbl", zip(a, b, np.zeros(5000, int)))
    plan = agg_plan(tmp_path, f, '"int", "long", "long"', "gpuaggregate_count", 3, 128, fields="fields = ( 0, 1 );",
                    hashfield="fieldrange = ( 0, 1 );", fn="bytes", buffsize=64)
    out, meta = P.run(plan)
    assert meta["schema"] == "int,long,long"
    got = {(int(x), int(y)): int(c) for x, y, c in zip(out["c0"], out["c1"], out["c2"])}
    exp = {}
    for x, y in zip(a.tolist(), b.tolist()):
        exp[(x, y)] = exp.get((x, y), 0) + 1
    assert got == exp


def test_join_then_groupby_sum_fp64(tmp_path):
    """The C3 query shape at small scale: F join D on orderkey, then
    sum(price) group by custkey; FP64 within 1e-12 relative of exact cents."""
    nD, nF = 20000, 100000
    dk, dc, fk, fp = O.gen_c3(nD, nF, norders=nD)
    joined, cents = O.c3_closed(nD, nF, norders=nD)
    ds, fs = O.schema(("int",), ("int",)), O.schema(("int",), ("dec",))
    fb = write_bin(tmp_path, "d.bin", O.rows(ds, dk, dc))
    fpr = write_bin(tmp_path, "f.bin", O.rows(fs, fk, fp))
    plan = textwrap.dedent(f"""
        path = "{tmp_path}/"; buffsize = 65536;
        treeroot = {{ name = "merge"; input = {{ name = "agg"; input = {{ name = "join";
            build = {{ name = "d"; }}; probe = {{ name = "f"; }}; }}; }}; }};
        merge = {{ type = "merge"; threads = 4; }};
        agg = {{ type = "gpuaggregate_sum"; field = 0; sumfield = 1; global = true; threads = 4;
                hash = {{ fn = "modulo"; buckets = 4096; field = 0; }}; }};
        join = {{ type = "gpuhashjoin"; hash = {{ fn = "modulo"; buckets = 65536; }}; threadgroups = ( [ 0, 1, 2, 3 ] );
                 tuplesperbucket = 4; allocpolicy = "local"; buildjattr = 0; probejattr = 0;
                 projection = ( "B$1", "P$1" ); }};
        d = {{ type = "parallelscan"; filetype = "binary"; files = ( "{fb}" ); schema = ( "int", "int" ); }};
        f = {{ type = "parallelscan"; filetype = "binary"; files = ( "{fpr}" ); schema = ( "int", "dec" ); }};
    """)
    out, meta = P.run(plan)
    assert meta["schema"] == "int,dec"
    assert len(out) == np.count_nonzero(cents)
    for c, s in zip(out["c0"], out["c1"]):
        exact = cents[c] / 100.0
        assert abs(s - exact) <= 1e-12 * abs(exact)


@pytest.mark.parametrize("threads,batch", [(1, 1 << 20), (4, 997)])
def test_resident_join_matches_oracle(tmp_path, threads, batch):
    """gpu { resident = true; }: the scanned relations are joined from their
    HBM copies; the output is the same byte-exact multiset as the reference
    join, on the first run and on a second run that reuses the copies."""
    rng = np.random.default_rng(40 + threads)
    nb, np_ = 3000, 20000
    bs = O.schema(("int",), ("char", 6), ("dec",))
    ps = O.schema(("long",), ("long",), ("int",))
    bk = rng.integers(-500, 500, nb).astype(np.int32)
    bc = np.array([b"k%04d" % (x % 10000) for x in rng.integers(0, 10000, nb)], dtype="S6")
    brow = O.rows(bs, bk, bc, rng.random(nb))
    prow = O.rows(ps, rng.integers(-600, 600, np_), rng.integers(-2**62, 2**62, np_),
                  rng.integers(-1000, 1000, np_).astype(np.int32))
    proj = ["P$2", "B$1", "P$0", "B$2", "B$0", "P$1"]
    d = O.join_desc(bs, ps, 0, 0, proj, buckets=256, tuplesperbucket=4, threads=2)
    exp, _ = O.hashjoin(d, brow, prow, cap=len(prow) * 20)
    fb, fp = write_bin(tmp_path, "b.bin", brow), write_bin(tmp_path, "p.bin", prow)
    plan = join_plan(tmp_path, f'"{fb}"', f'"{fp}"', threads, '"int", "char(5)", "dec"', '"long", "long", "int"',
                     ", ".join(f'"{x}"' for x in proj), buffsize=4096, filetype="binary", checker=False,
                     extra=f"gpu = {{ resident = true; probebatch = {batch}L; }};")
    try:
        for _ in range(2):
            out, meta = P.run(plan)
            rb = meta["row_bytes"]
            assert len(out) == len(exp) > 0
            got = sorted(out.tobytes()[i * rb:(i + 1) * rb] for i in range(len(out)))
            want = sorted(exp.tobytes()[i * rb:(i + 1) * rb] for i in range(len(exp)))
            assert got == want
    finally:
        P.clear_resident()


def _resident_agg_plan(agg, threads, extra_agg=""):
    groups = ", ".join(str(i) for i in range(threads))
    sumfield = "sumfield = 0;" if agg == "aggregate_sum" else ""
    return textwrap.dedent(f"""
        path = "./"; buffsize = 65536;
        treeroot = {{ name = "merge"; input = {{ name = "agg"; input = {{ name = "join";
                      build = {{ name = "r"; }}; probe = {{ name = "s"; }}; }}; }}; }};
        merge = {{ type = "merge"; threads = {threads}; }};
        agg = {{ type = "gpu{agg}"; fields = (); {sumfield} {extra_agg} }};
        join = {{ type = "gpuhashjoin"; hash = {{ fn = "modulo"; buckets = 4096; }};
                  threadgroups = ( [ {groups} ] ); tuplesperbucket = 4; allocpolicy = "local";
                  buildjattr = 0; probejattr = 0; projection = ( "B$1" );
                  gpu = {{ resident = true; probebatch = 30000L; }}; }};
        r = {{ type = "memscan"; table = "res_r"; schema = ( "int", "int" ); }};
        s = {{ type = "memscan"; table = "res_s"; schema = ( "int", "int" ); }};
    """)


@pytest.mark.parametrize("threads", [1, 4])
def test_resident_join_with_fused_aggregate(threads):
    """The queryhashjoin plan with a resident join: gpuaggregate_sum /
    gpuaggregate_count fold the join output in HBM.  The int32 sum wraps as
    the reference's CtInt accumulator (each thread's partial, then the
    consumer's sum, all mod 2^32); re-registering a table replaces its HBM
    copy."""
    rng = np.random.default_rng(threads)
    rdt = np.dtype({"names": ["k", "v"], "formats": [np.int32, np.int32], "offsets": [0, 4], "itemsize": 8})
    try:
        for rep in range(2):
            nR, nS = 5000 + rep, 200_000
            R = np.empty(nR, rdt)
            R["k"] = rng.permutation(nR).astype(np.int32) * 3
            R["v"] = rng.integers(-2**31, 2**31 - 1, nR).astype(np.int32)
            S = np.empty(nS, rdt)
            S["k"] = (rng.integers(0, nR + 500, nS) * 3).astype(np.int32)
            S["v"] = 0
            P.register_table("res_r", R)
            P.register_table("res_s", S)
            bv = dict(zip(R["k"].tolist(), R["v"].tolist()))
            matched = [bv[k] for k in S["k"].tolist() if k in bv]
            out, _ = P.run(_resident_agg_plan("aggregate_sum", threads))
            cnt, _ = P.run(_resident_agg_plan("aggregate_count", threads))
            assert int(cnt["c0"].sum()) == len(matched)
            got = int(out["c0"].astype(np.int64).sum()) & 0xFFFFFFFF
            assert got == sum(matched) & 0xFFFFFFFF
    finally:
        P.clear_tables()
        P.clear_resident()


@pytest.mark.parametrize("vtype", ["long", "dec"])
def test_resident_large_probe_and_single_group_sums(vtype):
    """A resident probe batch of 1.2M rows (several 256-row tiles per
    workgroup) whose output, with every build key twice, outgrows the first
    output buffer (the write pass runs again into a larger one); the
    single-group fold then runs its grid-stride loop over ~2.3M join rows.
    int64 sums wrap mod 2^64 exactly; FP64 sums within 1e-12 relative of the
    exact sum (positive values)."""
    rng = np.random.default_rng(11 if vtype == "long" else 12)
    vdt = np.int64 if vtype == "long" else np.float64
    rdt = np.dtype({"names": ["k", "v"], "formats": [np.int32, vdt], "offsets": [0, 4], "itemsize": 12})
    sdt = np.dtype({"names": ["k", "v"], "formats": [np.int32, np.int32], "offsets": [0, 4], "itemsize": 8})
    nR, nS = 100_000, 1_300_000
    R = np.empty(nR, rdt)
    R["k"] = (rng.permutation(nR) // 2 * 3).astype(np.int32)  # every key twice
    R["v"] = rng.integers(-2**62, 2**62, nR) if vtype == "long" else rng.random(nR) * 1e6
    S = np.empty(nS, sdt)
    S["k"] = (rng.integers(0, nR // 2 + 1000, nS) * 3).astype(np.int32)
    S["v"] = 0
    nkeys = nR // 2 + 1000
    mult = np.bincount(R["k"] // 3, minlength=nkeys)
    pk = S["k"] // 3
    try:
        P.register_table("res_r", R)
        P.register_table("res_s", S)
        plan = _resident_agg_plan("aggregate_sum", 2).replace("probebatch = 30000L", "probebatch = 1200000L")
        plan = plan.replace('r = { type = "memscan"; table = "res_r"; schema = ( "int", "int" ); };',
                            f'r = {{ type = "memscan"; table = "res_r"; schema = ( "int", "{vtype}" ); }};')
        assert f'"int", "{vtype}"' in plan
        cplan = _resident_agg_plan("aggregate_count", 2).replace("probebatch = 30000L", "probebatch = 1200000L")
        cplan = cplan.replace('table = "res_r"; schema = ( "int", "int" )', f'table = "res_r"; schema = ( "int", "{vtype}" )')
        cnt, _ = P.run(cplan)
        assert int(cnt["c0"].sum()) == int(mult[pk].sum()) > nS
        out, meta = P.run(plan)
        if vtype == "long":
            ksum = np.zeros(nkeys, np.uint64)
            np.add.at(ksum, R["k"] // 3, R["v"].view(np.uint64))
            with np.errstate(over="ignore"):
                want = int(ksum[pk].sum(dtype=np.uint64))
                got = int(out["c0"].astype(np.int64).view(np.uint64).sum(dtype=np.uint64))
            assert got == want
        else:
            ksum = np.zeros(nkeys, np.float64)
            np.add.at(ksum, R["k"] // 3, R["v"])
            want = float(np.sum(ksum[pk].astype(np.longdouble)))
            got = float(out["c0"].sum())
            assert abs(got - want) <= 1e-12 * abs(want)
    finally:
        P.clear_tables()
        P.clear_resident()


def test_host_pages_large_batches_stage_in_rounds():
    """Host-page join with a 5M-row probe batch (40 MB, more than one pinned
    staging round up) and 64 MB output pages (pulls of more than one round
    down): the output multiset is byte-exact against numpy."""
    rng = np.random.default_rng(21)
    dt = np.dtype({"names": ["k", "v"], "formats": [np.int32, np.int32], "offsets": [0, 4], "itemsize": 8})
    nR, nS = 100_000, 5_000_000
    R = np.empty(nR, dt)
    R["k"] = rng.permutation(nR).astype(np.int32)
    R["v"] = rng.integers(-2**31, 2**31 - 1, nR)
    S = np.empty(nS, dt)
    S["k"] = rng.integers(-5000, nR, nS).astype(np.int32)  # a few probe keys match nothing
    S["v"] = rng.integers(-2**31, 2**31 - 1, nS)
    bv = np.full(nR, 0, np.int64)
    bv[R["k"]] = R["v"]
    hit = S["k"] >= 0
    want = np.sort(bv[S["k"][hit]].astype(np.uint32).astype(np.uint64)
                   | (S["v"][hit].astype(np.uint32).astype(np.uint64) << 32))
    plan = textwrap.dedent("""
        path = "./"; buffsize = 67108864;
        treeroot = { name = "merge"; input = { name = "join"; build = { name = "r"; }; probe = { name = "s"; }; }; };
        merge = { type = "merge"; threads = 1; };
        join = { type = "gpuhashjoin"; hash = { fn = "modulo"; buckets = 4096; }; threadgroups = ( [ 0 ] );
                 tuplesperbucket = 4; allocpolicy = "local"; buildjattr = 0; probejattr = 0;
                 projection = ( "B$1", "P$1" ); gpu = { probebatch = 8388608L; }; };
        r = { type = "memscan"; table = "big_r"; schema = ( "int", "int" ); };
        s = { type = "memscan"; table = "big_s"; schema = ( "int", "int" ); };
    """)
    try:
        P.register_table("big_r", R)
        P.register_table("big_s", S)
        out, meta = P.run(plan)
        assert meta["row_bytes"] == 8 and len(out) == len(want)
        got = np.sort(np.frombuffer(out.tobytes(), np.uint64))
        assert np.array_equal(got, want)
    finally:
        P.clear_tables()


@pytest.mark.parametrize("global_", [False, True])
def test_resident_join_fused_groupby_sum(global_):
    """A resident join whose output is grouped where it lies in HBM:
    sum(B$2) group by B$1, int64 sums wrapping mod 2^64, against numpy.
    Thread-local aggregates emit per-thread partial groups (summed here);
    the global one emits every group once."""
    rng = np.random.default_rng(33 if global_ else 34)
    rdt = np.dtype({"names": ["k", "g", "v"], "formats": [np.int32, np.int32, np.int64], "offsets": [0, 4, 8],
                    "itemsize": 16})
    sdt = np.dtype({"names": ["k", "p"], "formats": [np.int32, np.int32], "offsets": [0, 4], "itemsize": 8})
    nR, nS, ng = 50_000, 400_000, 3000
    R = np.empty(nR, rdt)
    R["k"] = rng.permutation(nR).astype(np.int32)
    R["g"] = rng.integers(0, ng, nR)
    R["v"] = rng.integers(-2**62, 2**62, nR)
    S = np.empty(nS, sdt)
    S["k"] = rng.integers(0, nR + 2000, nS).astype(np.int32)
    S["p"] = 0
    hit = S["k"] < nR
    pos = np.empty(nR, np.int64)
    pos[R["k"]] = np.arange(nR)
    m = pos[S["k"][hit]]
    want = {}
    sums = np.zeros(ng, np.uint64)
    np.add.at(sums, R["g"][m], R["v"][m].view(np.uint64))
    present = np.bincount(R["g"][m], minlength=ng) > 0
    want = {int(g): int(np.int64(sums[g].view(np.int64))) for g in np.nonzero(present)[0]}
    agg = 'type = "gpuaggregate_sum"; field = 0; sumfield = 1; hash = { fn = "modulo"; buckets = 4096; field = 0; };'
    if global_:
        agg += " global = true; threads = 3;"
    plan = textwrap.dedent(f"""
        path = "./"; buffsize = 65536;
        treeroot = {{ name = "merge"; input = {{ name = "agg"; input = {{ name = "join";
                      build = {{ name = "r"; }}; probe = {{ name = "s"; }}; }}; }}; }};
        merge = {{ type = "merge"; threads = 3; }};
        agg = {{ {agg} }};
        join = {{ type = "gpuhashjoin"; hash = {{ fn = "modulo"; buckets = 4096; }};
                  threadgroups = ( [ 0, 1, 2 ] ); tuplesperbucket = 4; allocpolicy = "local";
                  buildjattr = 0; probejattr = 0; projection = ( "B$1", "B$2" );
                  gpu = {{ resident = true; probebatch = 50000L; }}; }};
        r = {{ type = "memscan"; table = "gb_r"; schema = ( "int", "int", "long" ); }};
        s = {{ type = "memscan"; table = "gb_s"; schema = ( "int", "int" ); }};
    """)
    try:
        P.register_table("gb_r", R)
        P.register_table("gb_s", S)
        out, meta = P.run(plan)
        assert meta["schema"] == "int,long"
        # the rows were grouped where the join wrote them: none reached the host
        assert meta["gpu_fused"] == P.FUSED_ROWS and meta["join_d2h_bytes"] == 0
        got = {}
        for g, v in zip(out["c0"].tolist(), out["c1"].tolist()):
            if global_:
                assert g not in got  # every group once
            got[g] = (got.get(g, 0) + v + 2**63) % 2**64 - 2**63
        assert got == want
    finally:
        P.clear_tables()
        P.clear_resident()


def test_budget_change_while_resident_copies_are_held():
    """The device context held by resident copies keeps the HBM budget it was
    created with: a plan asking for another budget fails with a clear error
    instead of running under the old one; once the copies are released the
    new budget applies (a 1-byte budget then fails its first allocation)."""
    rng = np.random.default_rng(5)
    rdt = np.dtype({"names": ["k", "v"], "formats": [np.int32, np.int32], "offsets": [0, 4], "itemsize": 8})
    R = np.empty(1000, rdt)
    R["k"], R["v"] = rng.permutation(1000), rng.integers(0, 100, 1000)
    S = np.empty(5000, rdt)
    S["k"], S["v"] = rng.integers(0, 1000, 5000), 0
    tight = _resident_agg_plan("aggregate_sum", 1).replace("probebatch = 30000L;", "probebatch = 30000L; hbmbudget = 1L;")
    assert "hbmbudget = 1L" in tight
    try:
        P.register_table("res_r", R)
        P.register_table("res_s", S)
        out, _ = P.run(_resident_agg_plan("aggregate_sum", 1))
        by_key = np.empty(1000, np.int64)
        by_key[R["k"]] = R["v"]
        assert int(out["c0"].astype(np.int64).sum()) == int(by_key[S["k"]].sum())
        with pytest.raises(P.QueryError, match="in use under an HBM budget"):
            P.run(tight)
        P.clear_resident()
        with pytest.raises(P.QueryError, match="budget"):
            P.run(tight)
    finally:
        P.clear_tables()
        P.clear_resident()


@pytest.mark.parametrize("case", ["dec-dec", "long-dec", "dec-int", "char-char", "char5-char9"])
def test_dec_and_char_join_keys_match_oracle(case):
    """Join keys beyond integers, compared as the reference's Comparator does
    (comparator.cpp:43-413, rawcompfns.cpp): dec keys as doubles (also
    against int / long keys; a NaN key matches nothing), char keys by strncmp
    over the shorter width.  Plans that hash the key bytes (`bytes`) or use
    one bucket (`alwayszero`, for keys whose equal values have different
    bytes) give the oracle's output multiset byte for byte."""
    rng = np.random.default_rng(abs(hash(case)) % 1000)
    nb, np_ = 2000, 6000
    bkt, pkt = case.split("-")
    width = {"char": 7, "char5": 5, "char9": 9}

    def col(kind):  # the oracle's char width counts the terminator: char(n) is ("char", n + 1)
        return ("char", width[kind] + 1) if kind.startswith("char") else (kind,)

    def keys(kind, v):
        if kind == "dec":
            out = v.astype(np.float64) * 0.5
            out[::97] = np.nan
            return out
        if kind.startswith("char"):
            return np.array([b"k%d" % x for x in v], dtype=f"S{width[kind] + 1}")
        return v.astype(np.int32 if kind == "int" else np.int64)

    bv = rng.integers(0, 700, nb) * (2 if pkt == "dec" and bkt != "dec" else 1)
    pv = rng.integers(0, 800, np_) * (2 if bkt == "dec" and pkt != "dec" else 1)
    bs = O.schema(col(bkt), ("long",))
    ps = O.schema(col(pkt), ("int",))
    brow = O.rows(bs, keys(bkt, bv), rng.integers(-2**60, 2**60, nb))
    prow = O.rows(ps, keys(pkt, pv), rng.integers(-2**30, 2**30, np_).astype(np.int32))
    same_bytes = bkt == pkt and not bkt.startswith("char5")
    fn, hfn = ("bytes", O.H_BYTES) if same_bytes else ("alwayszero", O.H_ALWAYSZERO)
    proj = ["P$1", "B$1", "B$0", "P$0"]
    d = O.join_desc(bs, ps, 0, 0, proj, hashfn=hfn, buckets=256, tuplesperbucket=4, threads=2)
    exp, _ = O.hashjoin(d, brow, prow, cap=len(prow) * 40)
    assert len(exp) > np_ // 4
    names = {"int": '"int"', "long": '"long"', "dec": '"dec"'}
    sch = lambda k, other: (f'"char({width[k]})"' if k.startswith("char") else names[k]) + ", " + other  # noqa: E731
    plan = textwrap.dedent(f"""
        path = "./"; buffsize = 4096;
        treeroot = {{ name = "merge"; input = {{ name = "join"; build = {{ name = "r"; }}; probe = {{ name = "s"; }}; }}; }};
        merge = {{ type = "merge"; threads = 2; }};
        join = {{ type = "gpuhashjoin"; hash = {{ fn = "{fn}"; buckets = 256; }}; threadgroups = ( [ 0, 1 ] );
                 tuplesperbucket = 4; allocpolicy = "local"; buildjattr = 0; probejattr = 0;
                 projection = ( "P$1", "B$1", "B$0", "P$0" ); gpu = {{ probebatch = 1000L; }}; }};
        r = {{ type = "memscan"; table = "dk_r"; schema = ( {sch(bkt, '"long"')} ); }};
        s = {{ type = "memscan"; table = "dk_s"; schema = ( {sch(pkt, '"int"')} ); }};
    """)
    try:
        P.register_table("dk_r", brow)
        P.register_table("dk_s", prow)
        out, meta = P.run(plan)
        rb = meta["row_bytes"]
        assert len(out) == len(exp)
        got = sorted(out.tobytes()[i * rb:(i + 1) * rb] for i in range(len(out)))
        want = sorted(exp.tobytes()[i * rb:(i + 1) * rb] for i in range(len(exp)))
        assert got == want
    finally:
        P.clear_tables()


def test_incomparable_join_keys_are_refused():
    """A date key against a long key, or a char key against an int key, has
    no Comparator (comparator.cpp:376-378, :132-135): the plan is refused."""
    for bt, pt in (('"date(%Y-%m-%d)"', '"long"'), ('"char(4)"', '"int"')):
        plan = textwrap.dedent(f"""
            path = "./"; buffsize = 4096;
            treeroot = {{ name = "join"; build = {{ name = "r"; }}; probe = {{ name = "s"; }}; }};
            join = {{ type = "gpuhashjoin"; hash = {{ fn = "alwayszero"; buckets = 16; }}; threadgroups = ( [ 0 ] );
                     tuplesperbucket = 4; allocpolicy = "local"; buildjattr = 0; probejattr = 0;
                     projection = ( "B$0" ); }};
            r = {{ type = "memscan"; table = "ic_r"; schema = ( {bt} ); }};
            s = {{ type = "memscan"; table = "ic_s"; schema = ( {pt} ); }};
        """)
        try:
            P.register_table("ic_r", np.zeros(1, np.dtype((np.void, 5 if "char" in bt else 8))))
            P.register_table("ic_s", np.zeros(1, np.dtype((np.void, 4 if "int" in pt else 8))))
            with pytest.raises(P.QueryError, match="comparison"):
                P.run(plan)
        finally:
            P.clear_tables()
