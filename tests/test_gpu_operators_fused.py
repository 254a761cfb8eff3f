"""The drop-in operators on the partitioned kernels: an aggregate directly
above a resident `gpuhashjoin` runs with it as one device computation.

* no GROUP BY, count or sum of an integer column of the key's width: the
  2-pass radix join with its fused count + sum (the queryhashjoin plan,
  unit_tests/queryhashjoin.cpp), at the full C1 shape against the closed form
  (count exact, the int32 sum wrapped as the reference's CtInt accumulator);
* sum of a dec probe column grouped by an int build column: the partitioned
  group join (hashjoin -> aggregate_sum(field), genericaggregate.cpp:240-282),
  at the full C3 shape against exact integer cents (1e-12 relative);
* any other consumer keeps the row-format path (join rows folded in HBM).

Each test also checks which form ran (pythia_result.gpu_fused) and that no
join output crossed to the host."""
import os
import textwrap

import numpy as np
import pytest

import pythia_amd as pa
from oracle import oracle as O
from pythia_amd import plan as P

pytestmark = pytest.mark.gpu

THREADS = min(16, os.cpu_count() or 1)


def _download_columns(ctx, specs):
    """Generates device columns (name, n, dtype, kwargs) and returns them on the host."""
    out = []
    for name, n, dt, kw in specs:
        b = ctx.alloc(n * np.dtype(dt).itemsize)
        ctx.gen(name, n, b, **kw)
        ctx.sync()
        out.append(b.download(dt, n))
        b.free()
    return out


def _agg_join_plan(agg, threads, rschema, sschema, projection, agg_keys, global_=False):
    groups = ", ".join(str(i) for i in range(threads))
    extra = f" global = true; threads = {threads};" if global_ else ""
    return textwrap.dedent(f"""
        path = "./"; buffsize = 1048576;
        treeroot = {{ name = "merge"; input = {{ name = "agg"; input = {{ name = "join";
                      build = {{ name = "r"; }}; probe = {{ name = "s"; }}; }}; }}; }};
        merge = {{ type = "merge"; threads = {threads}; }};
        agg = {{ type = "gpu{agg}"; {agg_keys} {extra} }};
        join = {{ type = "gpuhashjoin"; hash = {{ fn = "modulo"; buckets = 65536; }};
                  threadgroups = ( [ {groups} ] ); tuplesperbucket = 16; allocpolicy = "local";
                  buildjattr = 0; probejattr = 0; projection = ( {projection} );
                  gpu = {{ resident = true; }}; }};
        r = {{ type = "memscan"; table = "fz_r"; schema = ( {rschema} ); }};
        s = {{ type = "memscan"; table = "fz_s"; schema = ( {sschema} ); }};
    """)


def test_c1_full_shape_through_the_operators(ctx):
    """The C1 shape (R 2^20 x S 2^24, int32 keys and payloads) through
    gpuhashjoin -> gpuaggregate_sum / gpuaggregate_count -> merge: the count
    and the wrapped int32 sum of B$1 (and, projected instead, of P$1) equal
    the closed form, on the first run and on a repeated one."""
    nR, nS = 1 << 20, 1 << 24
    rk, rv, sk, sv = _download_columns(ctx, [("C1_RKEY", nR, np.int32, {"n_build": nR}),
                                             ("C1_RPAY", nR, np.int32, {}),
                                             ("C1_SFK", nS, np.int32, {"n_build": nR}),
                                             ("C1_SPAY", nS, np.int32, {})])
    dt = np.dtype({"names": ["k", "v"], "formats": [np.int32, np.int32], "offsets": [0, 4], "itemsize": 8})
    R, S = np.empty(nR, dt), np.empty(nS, dt)
    R["k"], R["v"], S["k"], S["v"] = rk, rv, sk, sv
    ref = O.c1_closed(nR, nS, threads=THREADS)
    wrap = lambda x: int(np.array([x % 2**32], np.uint64).astype(np.uint32).view(np.int32)[0])  # noqa: E731
    try:
        P.register_table("fz_r", R)
        P.register_table("fz_s", S)
        for side, want in (("B$1", ref.sum_build), ("P$1", ref.sum_probe)):
            plan = _agg_join_plan("aggregate_sum", THREADS, '"int", "int"', '"int", "int"', f'"{side}"',
                                  "fields = (); sumfield = 0;")
            for _ in range(2):
                out, meta = P.run(plan)
                assert meta["gpu_fused"] == P.FUSED_COUNT_SUM
                assert meta["join_d2h_bytes"] == 0
                assert len(out) == 1  # one thread emits the total; no group elsewhere
                assert int(out["c0"][0]) == wrap(want)
        cnt, meta = P.run(_agg_join_plan("aggregate_count", THREADS, '"int", "int"', '"int", "int"', '"B$1"',
                                         "fields = ();"))
        assert meta["gpu_fused"] == P.FUSED_COUNT_SUM
        assert int(cnt["c0"].sum()) == ref.count == nS
    finally:
        P.clear_tables()
        P.clear_resident()


@pytest.mark.parametrize("global_", [False, True])
def test_count_sum_fusion_small_cases(global_):
    """Fused count / sum on small int64 relations against numpy: duplicate
    build keys (each match counted), misses, negative keys, the int64 minimum
    as a real key, long payloads wrapping mod 2^64; an empty probe side
    emits nothing."""
    rng = np.random.default_rng(7 if global_ else 8)
    dt = np.dtype({"names": ["k", "v"], "formats": [np.int64, np.int64], "offsets": [0, 8], "itemsize": 16})
    nR, nS = 70_000, 300_000
    R, S = np.empty(nR, dt), np.empty(nS, dt)
    R["k"] = rng.integers(-40_000, 40_000, nR)
    R["k"][:5] = np.iinfo(np.int64).min
    R["v"] = rng.integers(-2**62, 2**62, nR)
    S["k"] = rng.integers(-50_000, 50_000, nS)
    S["k"][:7] = np.iinfo(np.int64).min
    S["v"] = rng.integers(-2**62, 2**62, nS)
    order = np.argsort(R["k"], kind="stable")
    keys, first = np.unique(R["k"][order], return_index=True)
    cnt = np.diff(np.append(first, nR))
    bsum = np.add.reduceat(R["v"][order].view(np.uint64), first) if nR else np.zeros(0, np.uint64)
    pos = np.searchsorted(keys, S["k"])
    pos[pos >= len(keys)] = 0
    hit = keys[pos] == S["k"]
    m = cnt[pos] * hit
    with np.errstate(over="ignore"):
        want_b = int(np.sum(bsum[pos] * hit.astype(np.uint64), dtype=np.uint64))
        want_p = int(np.sum(S["v"].view(np.uint64) * m.astype(np.uint64), dtype=np.uint64))
    sw = lambda x: (x + 2**63) % 2**64 - 2**63  # noqa: E731
    try:
        P.register_table("fz_r", R)
        P.register_table("fz_s", S)
        for side, want in (("B$1", want_b), ("P$1", want_p)):
            out, meta = P.run(_agg_join_plan("aggregate_sum", 3, '"long", "long"', '"long", "long"', f'"{side}"',
                                             "fields = (); sumfield = 0;", global_))
            assert meta["gpu_fused"] == P.FUSED_COUNT_SUM
            assert len(out) == 1 and int(out["c0"][0]) == sw(want)
        out, _ = P.run(_agg_join_plan("aggregate_count", 3, '"long", "long"', '"long", "long"', '"B$1"',
                                      "fields = ();", global_))
        assert int(out["c0"].sum()) == int(m.sum())
        P.register_table("fz_s", S[:0])
        out, meta = P.run(_agg_join_plan("aggregate_sum", 3, '"long", "long"', '"long", "long"', '"B$1"',
                                         "fields = (); sumfield = 0;", global_))
        assert len(out) == 0
    finally:
        P.clear_tables()
        P.clear_resident()


def test_c3_full_shape_through_the_operators(ctx):
    """The C3 query through the operators: F (orderkey, partkey, quantity,
    price) join D (orderkey, custkey), sum(price) group by custkey, at the
    full 15M x 60M shape: every custkey group once, within 1e-12 relative of
    its exact integer-cent sum."""
    nD, nF = 15_000_000, 60_000_000
    dk, dc, fk, fpk, fq, fp = _download_columns(ctx, [
        ("C3_D_ORDERKEY", nD, np.int32, {"n_build": nD}), ("C3_D_CUSTKEY", nD, np.int32, {}),
        ("C3_F_ORDERKEY", nF, np.int32, {"n_build": nD}), ("C3_F_PARTKEY", nF, np.int32, {}),
        ("C3_F_QUANTITY", nF, np.int32, {}), ("C3_F_PRICE", nF, np.float64, {})])
    ddt = np.dtype({"names": ["k", "c"], "formats": [np.int32, np.int32], "offsets": [0, 4], "itemsize": 8})
    fdt = np.dtype({"names": ["k", "p", "q", "v"], "formats": [np.int32, np.int32, np.int32, np.float64],
                    "offsets": [0, 4, 8, 12], "itemsize": 20})
    D, F = np.empty(nD, ddt), np.empty(nF, fdt)
    D["k"], D["c"] = dk, dc
    F["k"], F["p"], F["q"], F["v"] = fk, fpk, fq, fp
    del dk, dc, fk, fpk, fq, fp
    _, cents = O.c3_closed(nD, nF, norders=nD, threads=THREADS)
    try:
        P.register_table("fz_r", D)
        P.register_table("fz_s", F)
        del D, F
        plan = _agg_join_plan("aggregate_sum", THREADS, '"int", "int"', '"int", "int", "int", "dec"',
                              '"B$1", "P$3"', "field = 0; sumfield = 1; hash = { fn = \"modulo\"; buckets = 2097152; "
                              "field = 0; };")
        for _ in range(2):
            out, meta = P.run(plan)
            assert meta["schema"] == "int,dec"
            assert meta["gpu_fused"] == P.FUSED_GROUP and meta["join_d2h_bytes"] == 0
            assert len(out) == np.count_nonzero(cents)
            assert len(np.unique(out["c0"])) == len(out)  # every group once
            exact = cents[out["c0"]] / 100.0
            assert np.all(np.abs(out["c1"] - exact) <= 1e-12 * np.abs(exact))
    finally:
        P.clear_tables()
        P.clear_resident()


def test_group_fusion_matches_oracle_on_edge_keys():
    """Fused group sums on small relations against the oracle's hashjoin ->
    aggregate restatement: duplicate dimension keys (the general group-join
    path), negative keys and groups, unmatched fact rows; Global mode emits
    each group once."""
    rng = np.random.default_rng(9)
    nD, nF = 3000, 40_000
    ddt = np.dtype({"names": ["k", "c"], "formats": [np.int32, np.int32], "offsets": [0, 4], "itemsize": 8})
    fdt = np.dtype({"names": ["k", "v"], "formats": [np.int32, np.float64], "offsets": [0, 4], "itemsize": 12})
    D, F = np.empty(nD, ddt), np.empty(nF, fdt)
    D["k"] = rng.integers(-2000, 2000, nD)  # duplicates
    D["c"] = rng.integers(-50, 50, nD)
    F["k"] = rng.integers(-2500, 2500, nF)
    F["v"] = rng.integers(1, 10**6, nF) / 100.0
    exp = {}
    by_key = {}
    for k, c in zip(D["k"].tolist(), D["c"].tolist()):
        by_key.setdefault(k, []).append(c)
    cents = {}
    for k, v in zip(F["k"].tolist(), np.rint(F["v"] * 100).astype(np.int64).tolist()):
        for c in by_key.get(k, ()):
            cents[c] = cents.get(c, 0) + v
    exp = {c: x / 100.0 for c, x in cents.items()}
    try:
        P.register_table("fz_r", D)
        P.register_table("fz_s", F)
        out, meta = P.run(_agg_join_plan("aggregate_sum", 4, '"int", "int"', '"int", "dec"', '"B$1", "P$1"',
                                         "field = 0; sumfield = 1; hash = { fn = \"modulo\"; buckets = 1024; "
                                         "field = 0; };", global_=True))
        assert meta["gpu_fused"] == P.FUSED_GROUP
        got = dict(zip(out["c0"].tolist(), out["c1"].tolist()))
        assert len(got) == len(out) == len(exp)
        for c, s in got.items():
            assert abs(s - exp[c]) <= 1e-12 * abs(exp[c])
    finally:
        P.clear_tables()
        P.clear_resident()


def test_other_consumers_keep_the_row_path():
    """A count grouped by a build column, or a sum of a build-side long under
    int keys, is not one of the fused forms: the join's rows are folded in
    HBM (gpu_fused = 1) and nothing crosses to the host."""
    rng = np.random.default_rng(10)
    rdt = np.dtype({"names": ["k", "g", "v"], "formats": [np.int32, np.int32, np.int64], "offsets": [0, 4, 8],
                    "itemsize": 16})
    sdt = np.dtype({"names": ["k", "p"], "formats": [np.int32, np.int32], "offsets": [0, 4], "itemsize": 8})
    R, S = np.empty(5000, rdt), np.empty(60_000, sdt)
    R["k"], R["g"], R["v"] = rng.permutation(5000), rng.integers(0, 40, 5000), rng.integers(-9, 9, 5000)
    S["k"], S["p"] = rng.integers(0, 6000, 60_000), 0
    hit = S["k"] < 5000
    pos = np.empty(5000, np.int64)
    pos[R["k"]] = np.arange(5000)
    m = pos[S["k"][hit]]
    try:
        P.register_table("fz_r", R)
        P.register_table("fz_s", S)
        out, meta = P.run(_agg_join_plan("aggregate_count", 2, '"int", "int", "long"', '"int", "int"',
                                         '"B$1", "B$2"', "field = 0; hash = { fn = \"modulo\"; buckets = 64; field = 0; };",
                                         global_=True))
        assert meta["gpu_fused"] == P.FUSED_ROWS and meta["join_d2h_bytes"] == 0
        assert dict(zip(out["c0"].tolist(), out["c1"].tolist())) == \
            {int(g): int(c) for g, c in enumerate(np.bincount(R["g"][m], minlength=40)) if c}
        out, meta = P.run(_agg_join_plan("aggregate_sum", 2, '"int", "int", "long"', '"int", "int"',
                                         '"B$2"', "fields = (); sumfield = 0;"))
        assert meta["gpu_fused"] == P.FUSED_ROWS and meta["join_d2h_bytes"] == 0
        assert int(out["c0"].sum()) == int(R["v"][m].sum())
    finally:
        P.clear_tables()
        P.clear_resident()
