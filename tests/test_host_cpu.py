"""The plan engine that hosts the GPU operators, on the CPU: plan parsing,
type dispatch, the reference's configuration checks, the merge / scan Page
protocol, and error paths that must surface as errors rather than hangs.
No device work runs here."""
import os
import textwrap

import numpy as np
import pytest

from pythia_amd import plan as P


def write_text(tmp_path, name, rows):
    p = tmp_path / name
    p.write_text("".join("|".join(str(x) for x in r) + "\n" for r in rows))
    return name


def scan_merge_plan(tmp_path, fname, threads, buffsize=16, checker=True):
    inner = 'name = "merge"; input = { name = "check2"; input = { name = "scan"; }; };'
    tree = f'name = "check"; input = {{ {inner} }};' if checker else inner
    return textwrap.dedent(f"""
        path = "{tmp_path}/";
        buffsize = {buffsize};
        treeroot = {{ {tree} }};
        check = {{ type = "checker_callstate"; }};
        merge = {{ type = "merge"; threads = {threads}; }};
        check2 = {{ type = "checker_callstate"; }};
        scan = {{ type = "parallelscan"; filetype = "text"; files = ( "{fname}" ); schema = ( "long", "dec" ); }};
    """)


@pytest.mark.parametrize("threads", [1, 2, 4, 7])
def test_scan_merge_protocol_returns_every_row_once(tmp_path, threads):
    f = write_text(tmp_path, "t.tbl", [(i, i + 0.25) for i in range(1, 1001)])
    out, meta = P.run(scan_merge_plan(tmp_path, f, threads))
    assert meta["schema"] == "long,dec"
    assert sorted(out["c0"].tolist()) == list(range(1, 1001))
    assert np.all(out["c1"] == out["c0"] + 0.25)


@pytest.mark.parametrize("threads,buffsize", [(1, 1 << 16), (4, 1 << 20), (3, (1 << 20) + 48)])
def test_large_pages_reach_the_driver_whole(threads, buffsize):
    """Pages of 256 KB and more are copied into the driver's result buffer by
    several threads at once; pages below that by the driver alone.  Every
    row arrives once and unchanged, query after query (the helpers are
    reused), and a small query after a large one still returns only its own
    rows."""
    n = 700_001
    dt = np.dtype({"names": ["k", "v"], "formats": [np.int64, np.int64], "offsets": [0, 8], "itemsize": 16})
    T = np.empty(n, dt)
    T["k"] = np.arange(n)
    T["v"] = (np.arange(n, dtype=np.uint64) * np.uint64(0x9E3779B97F4A7C15)).view(np.int64)
    P.register_table("big_pages", T)
    try:
        plan = textwrap.dedent(f"""
            path = "./"; buffsize = {buffsize};
            treeroot = {{ name = "merge"; input = {{ name = "s"; }}; }};
            merge = {{ type = "merge"; threads = {threads}; }};
            s = {{ type = "memscan"; table = "big_pages"; schema = ( "long", "long" ); }};
        """)
        for _ in range(3):
            out, meta = P.run(plan)
            assert len(out) == n
            order = np.argsort(out["c0"], kind="stable")
            assert np.array_equal(out["c0"][order], T["k"])
            assert np.array_equal(out["c1"][order], T["v"])
        P.register_table("small_pages", T[:1000])
        out, _ = P.run(plan.replace("big_pages", "small_pages"))
        assert sorted(out["c0"].tolist()) == list(range(1000))
    finally:
        P.clear_tables()


@pytest.mark.parametrize("threads,buffsize", [(1, 20), (3, 20), (5, 4096)])
def test_threadidprepend(tmp_path, threads, buffsize):
    """threadidprepend (operators/threadidprepend.cpp:42-110): output =
    (int worker id, input columns); each worker's rows carry its own id."""
    f = write_text(tmp_path, "t.tbl", [(i, i ^ 0xABCDEF) for i in range(1, 2001)])
    tl = ", ".join(str(i) for i in range(threads))
    plan = textwrap.dedent(f"""
        path = "{tmp_path}/"; buffsize = {buffsize};
        treeroot = {{ name = "merge"; input = {{ name = "tid"; input = {{ name = "scan"; }}; }}; }};
        merge = {{ type = "merge"; threads = {threads}; }};
        tid = {{ type = "threadidprepend"; }};
        scan = {{ type = "parallelscan"; filetype = "text"; files = ( "{f}" ); mapping = ( [ {tl} ] );
                  schema = ( "long", "long" ); }};
    """)
    out, meta = P.run(plan)
    assert meta["schema"] == "int,long,long"
    assert sorted(out["c1"].tolist()) == list(range(1, 2001))
    assert np.all((out["c2"] ^ 0xABCDEF) == out["c1"])
    assert np.all((out["c0"] >= 0) & (out["c0"] < threads))


def test_config_syntax_features(tmp_path):
    f = write_text(tmp_path, "t.tbl", [(i, i) for i in range(5)])
    plan = f"""
        # comment
        path = "{tmp_path}" "/";   // adjacent string literals concatenate
        buffsize = 0x10;           /* hex */
        treeroot : {{ name = "merge"; input = {{ name = "scan"; }}; }}
        merge = {{ type = "merge"; threads = 2; }};
        scan = {{ type = "scan"; filetype = "text"; files = [ "{f}" ]; schema = ( "int", "long" );
                 mapping = ( ( 0, 1 ) ); extra = 5L; ratio = 1.5e0; flag = TRUE; }};
    """
    out, meta = P.run(plan)
    assert meta["schema"] == "int,long" and sorted(out["c0"].tolist()) == [0, 1, 2, 3, 4]


@pytest.mark.parametrize("bad,needle", [
    ("buffsize = 16; treeroot = { name = \"x\"; } x = { type = \"scan\" ", "parse error"),
    ("buffsize = 16; treeroot = { name = \"x\"; input = { name = \"y\"; }; }; x = { type = \"nosuchop\"; };",
     "neither a built-in"),
    ("buffsize = 16; x = { type = \"scan\"; };", "treeroot"),
    ("buffsize = 16; treeroot = { name = \"x\"; }; x = { type = \"merge\"; threads = 2; };", "`input'"),
    ("buffsize = 16; treeroot = { name = \"x\"; }; x = { type = \"hashjoin\"; };", "gpuhashjoin"),
])
def test_plan_errors(bad, needle):
    with pytest.raises(P.QueryError) as e:
        P.run(bad)
    assert needle in str(e.value)


def join_plan(tmp_path, fb, fp, threads, projection='"B$1", "P$1"', hashnode='fn = "modulo"; buckets = 16;',
              bschema='"long", "long"', pschema='"long", "dec"', extra="", groups=None, bmap="", pmap=""):
    groups = groups or f"[ {', '.join(str(i) for i in range(threads))} ]"
    return textwrap.dedent(f"""
        path = "{tmp_path}/";
        buffsize = 16;
        treeroot = {{ name = "merge"; input = {{ name = "join"; build = {{ name = "b"; }}; probe = {{ name = "p"; }}; }}; }};
        merge = {{ type = "merge"; threads = {threads}; }};
        join = {{ type = "gpuhashjoin"; hash = {{ {hashnode} }}; threadgroups = ( {groups} );
                 tuplesperbucket = 2; allocpolicy = "local"; buildjattr = 0; probejattr = 0;
                 projection = ( {projection} ); {extra} }};
        b = {{ type = "parallelscan"; filetype = "text"; files = ( "{fb}" ); schema = ( {bschema} ); {bmap} }};
        p = {{ type = "parallelscan"; filetype = "text"; files = ( "{fp}" ); schema = ( {pschema} ); {pmap} }};
    """)


@pytest.mark.parametrize("kw,needle", [
    (dict(projection='"X$1"'), "projection"),
    (dict(projection='"B1"'), "projection"),
    (dict(projection='"B$9"'), "out of range"),
    (dict(hashnode='fn = "modulo"; buckets = 0;'), "buckets cannot be zero"),
    (dict(hashnode='fn = "nosuchfn"; buckets = 16;'), "unknown hash"),
    (dict(hashnode='fn = "modulo"; buckets = 16; field = 0;'), "field"),
    (dict(bschema='"dec", "long"'), "illegal schema"),
    (dict(groups="[ 0, 1 ], [ 1 ]"), "two thread groups"),
])
def test_gpuhashjoin_init_validates_like_hashjoin(tmp_path, kw, needle):
    fb = write_text(tmp_path, "b.tbl", [(1, 1)])
    fp = write_text(tmp_path, "p.tbl", [(1, 1.5)])
    with pytest.raises(P.QueryError) as e:
        P.run(join_plan(tmp_path, fb, fp, 2, **kw))
    assert needle in str(e.value).lower() or needle in str(e.value)


def test_gpuaggregate_presorted_not_implemented(tmp_path):
    f = write_text(tmp_path, "t.tbl", [(1, 1)])
    plan = f"""
        path = "{tmp_path}/"; buffsize = 16; treeroot = {{ name = "agg"; input = {{ name = "s"; }}; }};
        agg = {{ type = "gpuaggregate_sum"; field = 0; sumfield = 1; presorted = true;
                hash = {{ fn = "modulo"; buckets = 4; field = 0; }}; }};
        s = {{ type = "scan"; filetype = "text"; files = ( "{f}" ); schema = ( "long", "long" ); }};
    """
    with pytest.raises(P.QueryError) as e:
        P.run(plan)
    assert "not yet implemented" in str(e.value)


@pytest.mark.timeout(120)
def test_device_failure_is_an_error_not_a_hang(tmp_path):
    """An HBM budget of 1 byte: creating the device table fails in one thread
    of a 4-thread group; the others must not stay at the group barrier."""
    fb = write_text(tmp_path, "b.tbl", [(i, i) for i in range(1, 50)])
    fp = write_text(tmp_path, "p.tbl", [(i, i + 0.1) for i in range(1, 50)])
    with pytest.raises(P.QueryError):
        P.run(join_plan(tmp_path, fb, fp, 4, extra="gpu = { hbmbudget = 1L; };"))


@pytest.mark.parametrize("case", ["mapped", "groups", "not_scan"])
def test_resident_join_needs_unmapped_scans_and_one_group(tmp_path, case):
    f = write_text(tmp_path, "t.tbl", [(i, i) for i in range(10)])
    groups = "( [ 0 ], [ 1 ] )" if case == "groups" else "( [ 0, 1 ] )"
    mapping = "mapping = ( [ 0, 1 ] );" if case == "mapped" else ""
    probe = 'name = "chk"; input = { name = "p"; };' if case == "not_scan" else 'name = "p";'
    plan = textwrap.dedent(f"""
        path = "{tmp_path}/"; buffsize = 64;
        treeroot = {{ name = "merge"; input = {{ name = "join"; build = {{ name = "b"; }}; probe = {{ {probe} }}; }}; }};
        merge = {{ type = "merge"; threads = 2; }};
        chk = {{ type = "checker_callstate"; }};
        join = {{ type = "gpuhashjoin"; hash = {{ fn = "modulo"; buckets = 16; }}; threadgroups = {groups};
                  tuplesperbucket = 2; allocpolicy = "local"; buildjattr = 0; probejattr = 0;
                  projection = ( "B$1" ); gpu = {{ resident = true; }}; }};
        b = {{ type = "parallelscan"; filetype = "text"; files = ( "{f}" ); schema = ( "long", "long" ); {mapping} }};
        p = {{ type = "parallelscan"; filetype = "text"; files = ( "{f}" ); schema = ( "long", "long" ); }};
    """)
    with pytest.raises(P.QueryError, match="gpu.resident"):
        P.run(plan)
