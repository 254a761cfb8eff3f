"""GPU parity of the serving engine: resident tables, HBM-budget admission and
continuous batching must not change any query's result.

Expected values come from the oracle's closed forms (exact count and int64
sum of each query over its table) and from exact numpy joins for tables with
duplicate keys and the minimum integer as a real key.
"""
import os

import numpy as np
import pytest

import pythia_amd as pa
from oracle import oracle as O

pytestmark = pytest.mark.gpu
THREADS = min(16, os.cpu_count() or 1)
NT = 8


@pytest.fixture(scope="module")
def world(ctx):
    """NT tables (C2 family, sub = table) and 96 queries with uniform FKs."""
    sizes = pa.sizes(5, NT, 10, 17)
    tabs = []
    for t, n in enumerate(int(x) for x in sizes):
        k, p = ctx.alloc(8 * n), ctx.alloc(8 * n)
        ctx.gen("C2_RKEY", n, k, sub=t)
        ctx.gen("C2_RPAY", n, p, sub=t)
        tabs.append((k, p, n))
    qt, qm, _ = pa.c5_trace(96, 1000.0, ntables=NT, lo=8, hi=17)
    probes = []
    for q, (t, m) in enumerate(zip(qt.tolist(), qm.tolist())):
        b = ctx.alloc(8 * m)
        ctx.gen("C2_SFK_UNIFORM", m, b, sub=t, n_build=tabs[t][2], offset=q << 32)
        probes.append(b)
    ctx.sync()
    exp = [O.c5_closed(t, tabs[t][2], q << 32, m, threads=THREADS) for q, (t, m) in
           enumerate(zip(qt.tolist(), qm.tolist()))]
    yield tabs, qt, qm, probes, exp
    for b in probes:
        b.free()
    for k, p, _ in tabs:
        k.free()
        p.free()


def _server(ctx, tabs, **kw):
    s = pa.Server(ctx, **kw)
    ids = [s.add_table(k, p, n) for k, p, n in tabs]
    assert ids == list(range(len(tabs)))
    return s


@pytest.mark.parametrize("part_min", [1, 0, 0xFFFFFFFF], ids=["partitioned", "by-size", "random-probe"])
def test_served_results_alone_and_in_the_mix(ctx, world, part_min):
    """Every query gives its closed-form result when the whole stream arrives
    at once (continuous batching mixes them into shared launches) and when
    it runs alone, on the partitioned path, the random-probe path, or as the
    size rule splits them."""
    tabs, qt, qm, probes, exp = world
    s = _server(ctx, tabs, part_min_rows=part_min)
    mix = s.replay(qt, probes, qm, np.zeros(len(qt)))
    assert [(r.matches, r.sum_build) for r in mix] == [(e.count, e.sum_build) for e in exp]
    assert max(r.batch_size for r in mix) > 1  # queries shared launches
    assert all(r.t_arrival <= r.t_admit <= r.t_done for r in mix)
    alone = []
    for q in range(len(qt)):
        s.submit(int(qt[q]), probes[q], int(qm[q]), tag=q)
        s.drain()
        (r,) = s.poll()
        assert r.batch_size == 1 and r.tag == q
        alone.append((r.matches, r.sum_build))
    assert alone == [(r.matches, r.sum_build) for r in mix]
    st = s.stats()
    assert st["completed"] == st["submitted"] == 2 * len(qt) and st["reserved_bytes"] == 0
    if part_min == 1:
        assert st["partitioned"] == 2 * len(qt)
    elif part_min == 0xFFFFFFFF:
        assert st["partitioned"] == 0
    s.close()


def test_poisson_replay_latencies(ctx, world):
    tabs, qt, qm, probes, exp = world
    _, _, off = pa.c5_trace(len(qt), 2000.0, ntables=NT, lo=8, hi=17)
    s = _server(ctx, tabs, slots=2, part_min_rows=4096)
    res = s.replay(qt, probes, qm, off)
    assert [(r.matches, r.sum_build) for r in res] == [(e.count, e.sum_build) for e in exp]
    lat = np.array([r.t_done - r.t_arrival for r in res])
    assert (lat > 0).all() and np.median(lat) < 0.5
    s.close()


def test_plain_mode_gives_the_same_results(ctx, world):
    tabs, qt, qm, probes, exp = world
    s = _server(ctx, tabs, mode=pa.SERVE_PLAIN)
    s.prepare_plain(int(qm.max()))
    res = s.replay(qt[:32], probes[:32], qm[:32], np.zeros(32))
    assert [(r.matches, r.sum_build) for r in res] == [(e.count, e.sum_build) for e in exp[:32]]
    assert all(r.batch_size == 1 for r in res)
    s.close()


def test_hbm_budget_admission(ctx, world):
    """A budget that holds the tables plus about one probe batch admits the
    stream a few queries at a time; results do not change, and a probe that
    can never fit is refused."""
    tabs, qt, qm, probes, exp = world
    probe = pa.Server(ctx)
    for k, p, n in tabs:
        probe.add_table(k, p, n)
    resident = probe.stats()["resident_bytes"]
    probe.close()
    budget = resident + int(qm.max()) * 8 + 64
    s = _server(ctx, tabs, hbm_budget=budget)
    res = s.replay(qt, probes, qm, np.zeros(len(qt)))
    assert [(r.matches, r.sum_build) for r in res] == [(e.count, e.sum_build) for e in exp]
    big = int(qm.max())
    # admitted batches never held more probe bytes than the budget leaves
    by_batch = {}
    for r in res:
        by_batch.setdefault(r.batch, 0)
        by_batch[r.batch] += r.m * 8 + 64
    assert max(by_batch.values()) <= budget - resident
    with pytest.raises(pa.PgpuError):
        s.submit(0, probes[0], (budget - resident) // 8 + 1)
    assert big * 8 + 64 <= budget - resident
    s.close()


def test_duplicates_and_empty_marker_key(ctx):
    """Tables with repeated keys fold them into (count, sum); the minimum
    int64 (the table's empty marker) is still a joinable key."""
    rng = np.random.default_rng(5)
    lo = np.iinfo(np.int64).min
    bk = rng.integers(-2000, 2000, 30000).astype(np.int64)
    bk[:4] = lo
    bp = rng.integers(-2**40, 2**40, 30000).astype(np.int64)
    pk = rng.integers(-2500, 2500, 70000).astype(np.int64)
    pk[:6] = lo
    k, p, q = ctx.upload(bk), ctx.upload(bp), ctx.upload(pk)
    got_by_path = []
    for part_min in (1, 0xFFFFFFFF):
        s = pa.Server(ctx, part_min_rows=part_min)
        t = s.add_table(k, p, len(bk))
        s.submit(t, q, len(pk))
        s.submit(t, q, 1000)
        s.drain()
        got_by_path.append(sorted((r.m, r.matches, r.sum_build) for r in s.poll()))
        s.close()
    got = got_by_path[0]
    assert got_by_path[1] == got
    keys, inv = np.unique(bk, return_inverse=True)
    cnt = np.bincount(inv)
    sums = np.zeros(len(keys), np.uint64)
    np.add.at(sums, inv, bp.view(np.uint64))

    def exact(probe):
        idx = np.searchsorted(keys, probe)
        ok = idx < len(keys)
        ok[ok] = keys[idx[ok]] == probe[ok]
        return int(cnt[idx[ok]].sum()), int(sums[idx[ok]].sum(dtype=np.uint64))

    assert got == sorted([(len(pk),) + exact(pk), (1000,) + exact(pk[:1000])])
    # the chained-bucket restatement agrees on count and the int64 sum
    bs, ps = O.schema(("long",), ("long",)), O.schema(("long",))
    d = O.join_desc(bs, ps, 0, 0, ["B$1"], buckets=1 << 10, tuplesperbucket=4, threads=2)
    cs, _ = O.hashjoin_count_sum(d, O.rows(bs, bk, bp), O.rows(ps, pk), 0)
    assert (cs.count, cs.isum % 2**64) == exact(pk)
    for b in (k, p, q):
        b.free()


@pytest.fixture(scope="module")
def c5_world(ctx):
    """The C5 configuration's tables at full scale: 64 tables of
    round(2^U(16, 23)) int64 rows (about 1.1e8 rows), and a query pool from
    the C5 trace (probe batches of round(2^U(14, 25)) rows): its first 40
    queries plus the first 8 whose probe batch holds 2^24 rows or more."""
    ntab = 64
    sizes = pa.sizes(5, ntab, 16, 23)
    tabs = []
    for t, n in enumerate(int(x) for x in sizes):
        k, p = ctx.alloc(8 * n), ctx.alloc(8 * n)
        ctx.gen("C2_RKEY", n, k, sub=t)
        ctx.gen("C2_RPAY", n, p, sub=t)
        tabs.append((k, p, n))
    qt, qm, _ = pa.c5_trace(2000, 1.0, ntables=ntab, lo=14, hi=25)
    big = [int(q) for q in np.nonzero(qm >= (1 << 24))[0][:8]]
    sel = sorted(set(range(40)) | set(big))
    probes = []
    for q in sel:
        t, m = int(qt[q]), int(qm[q])
        b = ctx.alloc(8 * m)
        ctx.gen("C2_SFK_UNIFORM", m, b, sub=t, n_build=tabs[t][2], offset=q << 32)
        probes.append(b)
    ctx.sync()
    st, sm = qt[sel], qm[sel]
    exp = [O.c5_closed(int(t), tabs[int(t)][2], q << 32, int(m), threads=THREADS) for q, t, m in zip(sel, st, sm)]
    yield tabs, st, sm, probes, exp, len(big)
    for b in probes:
        b.free()
    for k, p, _ in tabs:
        k.free()
        p.free()


def test_c5_full_scale_poisson_2000_alone_and_in_the_mix(ctx, c5_world):
    """At the C5 configuration's scale (64 resident tables up to 2^23 rows,
    probe batches up to 2^25 rows, the ones that set the p99) every query's
    count and int64 sum equals its closed form in a Poisson replay at
    2000 queries/s and when it runs alone."""
    tabs, qt, qm, probes, exp, nbig = c5_world
    assert nbig == 8 and qm.max() >= 1 << 24
    _, _, off = pa.c5_trace(len(qt), 2000.0, ntables=len(tabs), lo=14, hi=25)
    s = _server(ctx, tabs)
    res = s.replay(qt, probes, qm, off)
    assert [(r.matches, r.sum_build) for r in res] == [(e.count, e.sum_build) for e in exp]
    alone = []
    for q in range(len(qt)):
        s.submit(int(qt[q]), probes[q], int(qm[q]), tag=q)
        s.drain()
        (r,) = s.poll()
        assert r.batch_size == 1
        alone.append((r.matches, r.sum_build))
    assert alone == [(e.count, e.sum_build) for e in exp]
    st = s.stats()
    assert st["partitioned"] >= 2 * nbig  # the large batches took the partitioned path
    s.close()


def test_warm_start_from_saved_segments(ctx, world):
    """Every table of a batched server saved as a memory segment (the
    HashTable::serialize route, util/hashtable.cpp:99-201) and attached by a
    fresh server with no build kernel and no columns: every query gives its
    closed form, alone and in the mix; tables with repeated keys and the
    empty marker keep their counts and side words; a segment that is not a
    table is refused."""
    tabs, qt, qm, probes, exp = world
    rng = np.random.default_rng(4)
    dk = rng.integers(-300, 300, 5000).astype(np.int64)
    dk[:4] = np.iinfo(np.int64).min
    dp = rng.integers(-2**40, 2**40, 5000).astype(np.int64)
    pk = rng.integers(-350, 350, 40000).astype(np.int64)
    pk[:3] = np.iinfo(np.int64).min
    dup = [ctx.upload(dk), ctx.upload(dp), ctx.upload(pk)]
    order = np.argsort(dk, kind="stable")
    ukeys, first = np.unique(dk[order], return_index=True)
    cnt = np.diff(np.append(first, len(dk)))
    ssum = np.add.reduceat(dp[order].view(np.uint64), first)
    pos = np.clip(np.searchsorted(ukeys, pk), 0, len(ukeys) - 1)
    hit = ukeys[pos] == pk
    want_dup = (int((cnt[pos] * hit).sum()), int(np.sum(ssum[pos] * hit.astype(np.uint64), dtype=np.uint64)))
    prefix = f"/dev/shm/pgpu_seg_{os.getpid()}"
    paths = [f"{prefix}_{t}" for t in range(len(tabs) + 1)]
    try:
        s = _server(ctx, tabs)
        s.add_table(dup[0], dup[1], len(dk))
        for t, path in enumerate(paths):
            s.save_table(t, path)
        s.close()
        w = pa.Server(ctx)
        ids = [w.load_table(path) for path in paths]
        assert ids == list(range(len(paths)))
        mix = w.replay(qt, probes, qm, np.zeros(len(qt)))
        assert [(r.matches, r.sum_build) for r in mix] == [(e.count, e.sum_build) for e in exp]
        w.submit(len(tabs), dup[2], len(pk), tag=7)
        w.drain()
        (r,) = w.poll()
        assert (r.matches, r.sum_build) == want_dup
        with open(f"{prefix}_bad", "wb") as f:
            f.write(b"not a table segment" * 500)
        with pytest.raises(pa.PgpuError):
            w.load_table(f"{prefix}_bad")
        w.close()
    finally:
        for path in paths + [f"{prefix}_bad"]:
            if os.path.exists(path):
                os.unlink(path)
        for b in dup:
            b.free()


def test_lone_queries_on_large_tables_take_the_radix_route(ctx):
    """Tables of more than 1024 x 3584 rows also keep their rows in the radix
    layout; a query that runs alone is partitioned in two radix passes and
    joined in LDS against it.  Every probe size, a table with repeated keys
    and the minimum integer as a real key, and the same queries batched
    together (the segmented path) give the exact result."""
    n = (1 << 22) + 12345
    k, p = ctx.alloc(8 * n), ctx.alloc(8 * n)
    ctx.gen("C2_RKEY", n, k, sub=0)
    ctx.gen("C2_RPAY", n, p, sub=0)
    rng = np.random.default_rng(21)
    nd = 1 << 22
    dk = (rng.integers(0, 1 << 20, nd) * 7919 - 5).astype(np.int64)
    dk[:5] = np.iinfo(np.int64).min
    dp = rng.integers(-2**40, 2**40, nd).astype(np.int64)
    m_dup = 300_000
    pk = np.concatenate([dk[rng.integers(0, nd, m_dup - 1000)], rng.integers(-2**40, 2**40, 997),
                         np.full(3, np.iinfo(np.int64).min)]).astype(np.int64)
    order = np.argsort(dk, kind="stable")
    ukeys, first = np.unique(dk[order], return_index=True)
    cnt = np.diff(np.append(first, nd))
    ssum = np.add.reduceat(dp[order].view(np.uint64), first)
    pos = np.clip(np.searchsorted(ukeys, pk), 0, len(ukeys) - 1)
    hit = ukeys[pos] == pk
    want_dup = (int((cnt[pos] * hit).sum()), int(np.sum(ssum[pos] * hit.astype(np.uint64), dtype=np.uint64)))
    bufs = [k, p, ctx.upload(dk), ctx.upload(dp), ctx.upload(pk)]
    sizes = [1, 999, (1 << 20) + 7, 1 << 22]
    probes = []
    for q, m in enumerate(sizes):
        b = ctx.alloc(8 * m)
        ctx.gen("C2_SFK_UNIFORM", m, b, sub=0, n_build=n, offset=q << 32)
        probes.append(b)
    ctx.sync()
    exp = [O.c5_closed(0, n, q << 32, m, threads=THREADS) for q, m in enumerate(sizes)]
    try:
        for part_min in (1, 0):
            s = pa.Server(ctx, part_min_rows=part_min)
            assert s.add_table(k, p, n) == 0 and s.add_table(bufs[2], bufs[3], nd) == 1
            for q, m in enumerate(sizes):  # alone
                s.submit(0, probes[q], m, tag=q)
                s.drain()
                (r,) = s.poll()
                assert (r.matches, r.sum_build) == (exp[q].count, exp[q].sum_build), (part_min, m)
            s.submit(1, bufs[4], m_dup, tag=9)
            s.drain()
            (r,) = s.poll()
            assert (r.matches, r.sum_build) == want_dup
            mix = s.replay([0] * len(sizes) + [1], probes + [bufs[4]], sizes + [m_dup], np.zeros(len(sizes) + 1))
            assert [(r.matches, r.sum_build) for r in mix] == [(e.count, e.sum_build) for e in exp] + [want_dup]
            s.close()
    finally:
        for b in bufs + probes:
            b.free()
