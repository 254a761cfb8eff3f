"""Group-by aggregation through the plan API where a workgroup first folds
its rows in LDS (few groups: the flags of TPC-H Q1's GROUP BY are the
reference's case) and where the keys outnumber the LDS table (rows folded
into the HBM table directly).  Reference behaviour: GenericAggregate::remember
(operators/genericaggregate.cpp:240-282) folding through AggregateSum /
AggregateCount (aggregatesum.cpp:53-107, aggregatecount.cpp:44-52); expected
groups from the oracle's restatement of that operator and from numpy."""
import textwrap

import numpy as np
import pytest

from oracle import oracle as O
from pythia_amd import plan as P

pytestmark = pytest.mark.gpu

ROW = np.dtype({"names": ["g", "h", "vi", "vl", "vd"], "formats": [np.int64, np.int32, np.int32, np.int64, np.float64],
                "offsets": [0, 8, 12, 16, 24], "itemsize": 32})
SCHEMA = '"long", "int", "int", "long", "dec"'


def plan(kind, fields, sumfield, threads=4):
    sf = f"sumfield = {sumfield};" if kind == "gpuaggregate_sum" else ""
    last = fields[-1]
    return textwrap.dedent(f"""
        path = "./"; buffsize = 1048576;
        treeroot = {{ name = "merge"; input = {{ name = "agg"; input = {{ name = "scan"; }}; }}; }};
        merge = {{ type = "merge"; threads = {threads}; }};
        agg = {{ type = "{kind}"; fields = ( {", ".join(map(str, fields))} ); {sf} global = true;
                threads = {threads}; hash = {{ fn = "bytes"; buckets = 4096; fieldrange = ( 0, {last} ); }}; }};
        scan = {{ type = "memscan"; table = "agg_t"; schema = ( {SCHEMA} ); }};
    """)


def table(n, g, h, seed):
    rng = np.random.default_rng(seed)
    t = np.zeros(n, ROW)
    t["g"], t["h"] = g, h
    t["vi"] = rng.integers(-2**31, 2**31, n)
    t["vl"] = rng.integers(-2**62, 2**62, n)
    t["vd"] = rng.integers(-2**20, 2**20, n).astype(np.float64)  # integer-valued: every order sums exactly
    return t


def oracle_groups(t, fields, sumfield, kind):
    s = O.schema(("long",), ("int",), ("int",), ("long",), ("dec",))
    d = O.agg_desc(s, fields, sumfield if kind == "gpuaggregate_sum" else -1, hashfn=O.H_ALWAYSZERO, buckets=1,
                   global_=True, threads=1)
    out = O.aggregate(d, np.frombuffer(t.tobytes(), dtype=s.dtype()))
    return {tuple(int(r[f"c{i}"]) for i in range(len(fields))): r[f"c{len(fields)}"] for r in out}


def run(t, kind, fields, sumfield):
    P.register_table("agg_t", t)
    try:
        return P.run(plan(kind, fields, sumfield))
    finally:
        P.clear_tables()


def as_dict(out, nkeys):
    return {tuple(int(r[f"c{i}"]) for i in range(nkeys)): r[f"c{nkeys}"] for r in out}


@pytest.mark.parametrize("sumfield,kind", [(2, "gpuaggregate_sum"), (3, "gpuaggregate_sum"),
                                          (4, "gpuaggregate_sum"), (None, "gpuaggregate_count")])
def test_few_groups_every_accumulator(sumfield, kind):
    """5 groups over 3M rows: every workgroup folds in LDS and adds 5
    partials to the HBM table; int sums wrap in 32 bits, long in 64, dec
    sums of integer values are exact, the count counts."""
    n = 3_000_000
    rng = np.random.default_rng(7)
    t = table(n, rng.integers(0, 5, n), 0, 8)
    out, meta = run(t, kind, (0,), sumfield)
    got = as_dict(out, 1)
    want = oracle_groups(t, (0,), sumfield, kind)
    assert set(got) == set(want) == {(g,) for g in range(5)}
    for k in want:
        if sumfield == 4:
            assert float(got[k]) == float(want[k])
        else:
            assert int(got[k]) == int(want[k])
    if sumfield == 2:  # and the wrap itself, from the column
        for (g,) in got:
            s = int(t["vi"][t["g"] == g].astype(np.int64).sum())
            assert int(got[(g,)]) == ((s + 2**31) % 2**32) - 2**31


def test_three_word_composite_key():
    """GROUP BY (long, int): 12 key bytes (two 64-bit words), 40 groups,
    with keys that differ only in the second word."""
    n = 1_500_000
    rng = np.random.default_rng(9)
    g = rng.integers(0, 8, n) * (1 << 40) - 3
    h = rng.integers(-2, 3, n).astype(np.int32)
    t = table(n, g, h, 10)
    out, _ = run(t, "gpuaggregate_sum", (0, 1), 3)
    got = as_dict(out, 2)
    want = oracle_groups(t, (0, 1), 3, "gpuaggregate_sum")
    assert len(want) == 40 and set(got) == set(want)
    assert all(int(got[k]) == int(want[k]) for k in want)


def test_more_keys_than_the_lds_table():
    """2^20 rows over 2^19 distinct keys in one append: each workgroup's
    1024 rows hold ~1000 keys, so many rows find no LDS slot and are folded
    into the HBM table directly; every group still appears once with its
    exact sum and count."""
    n = 1 << 20
    rng = np.random.default_rng(11)
    keys = rng.permutation(np.repeat(np.arange(n // 2, dtype=np.int64) * 7919 - 12345, 2))
    t = table(n, keys, 0, 12)
    out, _ = run(t, "gpuaggregate_sum", (0,), 3)
    got = as_dict(out, 1)
    order = np.argsort(t["g"], kind="stable")
    gs, vs = t["g"][order], t["vl"][order].view(np.uint64)
    starts = np.r_[0, np.nonzero(np.diff(gs))[0] + 1]
    sums = np.add.reduceat(vs, starts)
    assert len(got) == n // 2
    want = {(int(k),): int(np.int64(s.view(np.int64))) for k, s in zip(gs[starts], sums)}
    assert {k: int(v) for k, v in got.items()} == want
    cnt, _ = run(t, "gpuaggregate_count", (0,), None)
    assert set(int(v) for v in as_dict(cnt, 1).values()) == {2}
