"""GPU parity of the batched join plans (many independent joins per launch).

Expected values come from the oracle: the C4 generators' closed forms
(exact int64 count / sums per join) and the chained-bucket restatement of
HashJoinOp on random inputs with duplicate keys and the minimum integer as a
real key.  Integer results must be bit-exact.
"""
import os

import numpy as np
import pytest

import pythia_amd as pa
from oracle import oracle as O

pytestmark = pytest.mark.gpu
THREADS = min(16, os.cpu_count() or 1)


def _c4_batch(ctx, first, sizes, probe_mult=8):
    """Generates joins first.. of the C4 family in HBM; returns (joins, buffers)."""
    joins, bufs = [], []
    for i, n in enumerate(int(x) for x in sizes):
        m = probe_mult * n
        rk, rp, sk, sp = ctx.alloc(4 * n), ctx.alloc(4 * n), ctx.alloc(4 * m), ctx.alloc(4 * m)
        ctx.gen("C4_RKEY", n, rk, sub=first + i, n_build=n)
        ctx.gen("C4_RPAY", n, rp, sub=first + i)
        ctx.gen("C4_SFK", m, sk, sub=first + i, n_build=n)
        ctx.gen("C4_SPAY", m, sp, sub=first + i)
        joins.append((rk, rp, n, sk, sp, m))
        bufs += [rk, rp, sk, sp]
    ctx.sync()
    return joins, bufs


def _free(bufs):
    for b in bufs:
        b.free()


def _check(res, ref):
    got = [(r.matches, r.sum_build, r.sum_probe) for r in res]
    exp = [(c.count, c.sum_build, c.sum_probe) for c in ref]
    assert got == exp


@pytest.mark.parametrize("variant,part_rows,graph", [(1, 0, True), (2, 0, True), (2, 1024, False), (1, 2048, True),
                                                     (3, 0, True), (3, 4096, False)])
def test_batch_c4_family_vs_closed_form(ctx, variant, part_rows, graph):
    sizes = pa.sizes(4, 40, 10, 17)
    joins, bufs = _c4_batch(ctx, 0, sizes)
    plan = pa.BatchPlan(ctx, joins, key_bytes=4, probe_payload=True, variant=variant, part_rows=part_rows,
                        graph=graph)
    res = plan.run()
    ref = O.c4_closed_many(0, sizes, 8 * sizes, threads=THREADS)
    _check(res, ref)
    assert max(r.radix_bits for r in res) >= 5  # the larger joins are partitioned
    assert all(r.lds_passes == 1 for r in res)  # uniform keys: every partition fits one LDS piece
    info = plan.info()
    assert info.graph == (1 if graph else 0) and info.items >= len(sizes)
    plan.close()
    _free(bufs)


def _exact(bk, bp, pk, pp):
    """Exact count and 64-bit wrapped sums of the equi-join (numpy)."""
    keys, inv = np.unique(bk, return_inverse=True)
    cnt = np.bincount(inv, minlength=len(keys)).astype(np.uint64)
    sums = np.zeros(len(keys), np.uint64)
    np.add.at(sums, inv, bp.astype(np.int64).view(np.uint64))
    idx = np.searchsorted(keys, pk)
    ok = idx < len(keys)
    ok[ok] = keys[idx[ok]] == pk[ok]
    c = cnt[idx[ok]]
    with np.errstate(over="ignore"):
        return (int(c.sum()), int(sums[idx[ok]].sum(dtype=np.uint64)),
                int((c * pp[ok].astype(np.int64).view(np.uint64)).sum(dtype=np.uint64)))


@pytest.mark.parametrize("kb", [4, 8])
def test_batch_random_keys_vs_oracle(ctx, kb):
    """Duplicates on both sides, the minimum integer as a real key, a hot key
    whose partition overflows one LDS piece, empty sides, and columns that
    are not 16-byte aligned (the scalar load path)."""
    rng = np.random.default_rng(kb)
    dt = np.int32 if kb == 4 else np.int64
    lo = np.iinfo(dt).min
    specs = []
    for t in range(7):
        nR = [3000, 20000, 0, 500, 60000, 1, 9000][t]
        nS = [40000, 0, 700, 9000, 90000, 5, 30000][t]
        bk = rng.integers(-4000, 4000, nR).astype(dt)
        pk = rng.integers(-4500, 4500, nS).astype(dt)
        if t == 4:
            bk[: 12000] = 77  # one hot key: its partition needs several LDS pieces
            pk[: 5000] = 77
        if nR > 10:
            bk[:5] = lo
        if nS > 10:
            pk[:9] = lo
        bp = rng.integers(np.iinfo(dt).min // 4, np.iinfo(dt).max // 4, nR).astype(dt)
        pp = rng.integers(np.iinfo(dt).min // 4, np.iinfo(dt).max // 4, nS).astype(dt)
        specs.append((bk, bp, pk, pp))
    joins, bufs, exp = [], [], []
    for t, (bk, bp, pk, pp) in enumerate(specs):
        off = kb if t == 6 else 0  # misaligned by one element
        arrs = []
        for a in (bk, bp, pk, pp):
            buf = ctx.alloc(a.nbytes + 64)
            host = np.zeros(len(a) + 64 // kb, dt)
            host[off // kb: off // kb + len(a)] = a
            buf.upload(host)
            arrs.append(buf.ptr + off)
            bufs.append(buf)
        joins.append((arrs[0], arrs[1], len(bk), arrs[2], arrs[3], len(pk)))
        exp.append(_exact(bk, bp, pk, pp))
        # the reference restatement: count and the column-type (wrapped) sums
        ct = "int" if kb == 4 else "long"
        bs, ps = O.schema((ct,), (ct,)), O.schema((ct,), (ct,))
        d = O.join_desc(bs, ps, 0, 0, ["B$1", "P$1"], buckets=1 << 10, tuplesperbucket=4, threads=2)
        if len(bk) and len(pk):
            csb, _ = O.hashjoin_count_sum(d, O.rows(bs, bk, bp), O.rows(ps, pk, pp), 0)
            csp, _ = O.hashjoin_count_sum(d, O.rows(bs, bk, bp), O.rows(ps, pk, pp), 1)
            mask = (1 << (8 * kb)) - 1
            assert csb.count == exp[-1][0]
            assert csb.isum & mask == exp[-1][1] & mask and csp.isum & mask == exp[-1][2] & mask
    # variant 3 (4-byte keys) hands the partitions with repeated build keys
    # (all of them here) and the hot key's oversized one to variant 2's kernel
    for variant in (1, 2, 3) if kb == 4 else (1, 2):
        plan = pa.BatchPlan(ctx, joins, key_bytes=kb, probe_payload=True, variant=variant)
        res = plan.run()
        assert [(r.matches, r.sum_build, r.sum_probe) for r in res] == exp
        assert res[4].lds_passes > 1
        plan.close()
    _free(bufs)


def test_batch_unique_keys_variant3_vs_exact(ctx):
    """Unique build keys (variant 3's 8-byte slots) with the minimum integer
    as a real key on both sides, probe keys that miss, a partition over
    6144 rows (handed to variant 2's kernel), an empty probe side and a join
    small enough for one partition."""
    rng = np.random.default_rng(33)
    lo = np.iinfo(np.int32).min
    specs = []
    for nR, nS in [(200000, 900000), (50000, 0), (3000, 40000), (7000, 7000)]:
        bk = rng.permutation(np.arange(-nR, nR, 2, dtype=np.int64))[:nR].astype(np.int32)
        if nR > 10:
            bk[3] = lo
        pk = rng.choice(bk, nS) if nS else np.zeros(0, np.int32)
        if nS > 100:
            pk[:50] = lo
            pk[50:90] += 1  # odd keys: no match
        bp = rng.integers(-2 ** 30, 2 ** 30, nR).astype(np.int32)
        pp = rng.integers(-2 ** 30, 2 ** 30, nS).astype(np.int32)
        specs.append((bk, bp, pk.astype(np.int32), pp))
    joins, bufs, exp = [], [], []
    for bk, bp, pk, pp in specs:
        arrs = []
        for a in (bk, bp, pk, pp):
            buf = ctx.alloc(max(a.nbytes, 16))
            if len(a):
                buf.upload(a)
            arrs.append(buf)
            bufs.append(buf)
        joins.append((arrs[0], arrs[1], len(bk), arrs[2], arrs[3], len(pk)))
        exp.append(_exact(bk, bp, pk, pp))
    for variant, part_rows in ((3, 0), (3, 4096), (2, 0)):
        plan = pa.BatchPlan(ctx, joins, key_bytes=4, probe_payload=True, variant=variant, part_rows=part_rows)
        res = plan.run()
        assert [(r.matches, r.sum_build, r.sum_probe) for r in res] == exp
        plan.close()
    # one partition of more than 6144 rows: 8000 build rows take 1 bit at 4096 rows per partition, and these
    # keys all hash to partition 0 (the top bit of key x 0x9E3779B97F4A7C15 is clear)
    cand = np.arange(0, 40000, dtype=np.int64)
    with np.errstate(over="ignore"):
        top = (cand.astype(np.uint64) * np.uint64(0x9E3779B97F4A7C15)) >> np.uint64(63)
    bk = rng.permutation(cand[top == 0][:8000]).astype(np.int32)
    assert len(bk) == 8000
    pk = rng.choice(bk, 30000).astype(np.int32)
    bp = rng.integers(-1000, 1000, 8000).astype(np.int32)
    pp = rng.integers(-1000, 1000, 30000).astype(np.int32)
    arrs = []
    for a in (bk, bp, pk, pp):
        buf = ctx.alloc(a.nbytes)
        buf.upload(a)
        arrs.append(buf)
        bufs.append(buf)
    plan = pa.BatchPlan(ctx, [(arrs[0], arrs[1], 8000, arrs[2], arrs[3], 30000)], key_bytes=4, probe_payload=True,
                        variant=3, part_rows=4096)
    res = plan.run()
    assert [(r.matches, r.sum_build, r.sum_probe) for r in res] == [_exact(bk, bp, pk, pp)]
    plan.close()
    _free(bufs)


def test_batch_graph_replays_match_eager(ctx):
    """The captured graph replayed several times, with other work on the
    stream between replays, gives the eager result every time."""
    sizes = pa.sizes(4, 16, 10, 16, first=100)
    joins, bufs = _c4_batch(ctx, 100, sizes)
    eager = pa.BatchPlan(ctx, joins, graph=False)
    want = [(r.matches, r.sum_build, r.sum_probe) for r in eager.run()]
    g = pa.BatchPlan(ctx, joins, graph=True)
    other = pa.JoinPlan(ctx, 4, int(sizes[0]), 8 * int(sizes[0]), probe_payload=True)
    for _ in range(4):
        got = [(r.matches, r.sum_build, r.sum_probe) for r in g.run()]
        assert got == want
        other.count_sum(joins[0][0], joins[0][1], joins[0][2], joins[0][3], joins[0][4], joins[0][5])
    ref = O.c4_closed_many(100, sizes, 8 * sizes, threads=THREADS)
    _check(eager.run(), ref)
    for p in (eager, g, other):
        p.close()
    _free(bufs)


def test_batch_autotune_picks_a_candidate_with_same_results(ctx):
    sizes = pa.sizes(4, 24, 10, 18, first=200)
    joins, bufs = _c4_batch(ctx, 200, sizes)
    best, table = pa.batch_autotune(ctx, joins, key_bytes=4, probe_payload=True, reps=1)
    assert len(table) == 6 and best in [(v, r) for v, r, _ in table]
    assert min(table, key=lambda e: e[2])[:2] == best
    plan = pa.BatchPlan(ctx, joins, variant=best[0], part_rows=best[1])
    _check(plan.run(), O.c4_closed_many(200, sizes, 8 * sizes, threads=THREADS))
    plan.close()
    _free(bufs)


def test_batch_rejects_bad_arguments(ctx):
    b = ctx.alloc(1024)
    j = [(b, b, 10, b, b, 10)]
    with pytest.raises(pa.PgpuError):
        pa.BatchPlan(ctx, j, key_bytes=2)
    with pytest.raises(pa.PgpuError):
        pa.BatchPlan(ctx, j, variant=1, part_rows=4096)  # larger than that variant's LDS piece
    with pytest.raises(pa.PgpuError):
        pa.BatchPlan(ctx, j, key_bytes=8, variant=3)  # the 8-byte slots hold 4-byte keys
    with pytest.raises(pa.PgpuError):
        pa.BatchPlan(ctx, [(None, b, 10, b, b, 10)])
    with pytest.raises(pa.PgpuError):
        pa.BatchPlan(ctx, [])
    b.free()


@pytest.mark.parametrize("variant,part_rows", [(1, 1024), (2, 2048)])
def test_batch_wide_fanout_vs_closed_form(ctx, variant, part_rows):
    """A join large enough for more than 2^10 partitions (the 12-bit cap at
    1024 rows per partition, 11 bits at 2048) sends the whole launch through
    the 64 KB-tile scatter; small joins ride along in the same launch."""
    sizes = np.array([(1 << 22) + 12345, 3000, 1 << 21, 700], dtype=np.uint64)
    joins, bufs = _c4_batch(ctx, 100, sizes)
    plan = pa.BatchPlan(ctx, joins, key_bytes=4, probe_payload=True, variant=variant, part_rows=part_rows)
    res = plan.run()
    ref = O.c4_closed_many(100, sizes, 8 * sizes, threads=THREADS)
    _check(res, ref)
    assert max(r.radix_bits for r in res) >= 11
    plan.close()
    _free(bufs)


def test_batch_full_c4_shape_vs_closed_form(ctx):
    """The full batch: 1024 joins, build sizes round(2^U(10, 22)), probe 8x,
    every join's count and sums against its closed form."""
    sizes = pa.sizes(4, 1024, 10, 22)
    assert 4.0e8 < sizes.sum() < 6.5e8
    joins, bufs = _c4_batch(ctx, 0, sizes)
    plan = pa.BatchPlan(ctx, joins, key_bytes=4, probe_payload=True)
    res = plan.run()
    ref = O.c4_closed_many(0, sizes, 8 * sizes, threads=THREADS)
    _check(res, ref)
    plan.close()
    _free(bufs)
