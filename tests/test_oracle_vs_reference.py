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
