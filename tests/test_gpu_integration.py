"""The reference engine itself with the GPU operators swapped in: the binary
oracle/_ref/refgpujoin links the reference's objects (compiled from its
sources), the binding integration/gpujoin.cpp and libpgpu.so, and runs each
plan once with HashJoinOp / AggregateSum / AggregateCount and once with
GpuHashJoinOp / GpuAggregateSum / GpuAggregateCount in their place -- what
changing `type = "hashjoin"` to `type = "gpuhashjoin"` does.  The results
must be identical: C1 count and wrapped int32 sums, the join output multiset
(mixed projection, int build key against long probe key, duplicate keys,
misses) and exact group sums in both aggregation modes."""
import json
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REFGPU = os.path.join(ROOT, "oracle", "_ref", "refgpujoin")


def run(*a):
    assert os.path.exists(REFGPU), "oracle/_ref/refgpujoin not built (oracle/ref.mk)"
    print(f"refgpujoin {' '.join(map(str, a))}", flush=True)
    out = subprocess.run([REFGPU, *map(str, a)], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True,
                         timeout=300, cwd="/tmp")
    assert out.returncode == 0, out.stderr
    return json.loads(out.stdout)


@pytest.mark.parametrize("agg", ["sum", "count", "probesum"])
@pytest.mark.parametrize("threads", [1, 4])
def test_queryhashjoin_plan_same_with_gpu_operators(agg, threads):
    cpu = run("c1", 14, 19, threads, "cpu", agg)
    gpu = run("c1", 14, 19, threads, "gpu", agg)
    assert gpu["value"] == cpu["value"]


@pytest.mark.parametrize("threads", [1, 3])
def test_join_output_multiset_same_with_gpu_operators(threads):
    cpu = run("rows", 30000, 200000, threads, "cpu")
    gpu = run("rows", 30000, 200000, threads, "gpu")
    assert cpu["rows"] > 100000
    assert (gpu["rows"], gpu["row_bytes"], gpu["sum"], gpu["xor"]) == \
        (cpu["rows"], cpu["row_bytes"], cpu["sum"], cpu["xor"])


@pytest.mark.parametrize("keys,nr,ns", [("dec", 30000, 200000), ("mixed", 3000, 20000), ("char", 30000, 200000),
                                        ("charw", 3000, 20000)])
def test_dec_and_char_keys_same_with_gpu_operators(keys, nr, ns):
    """Join keys the reference compares through Comparator (comparator.cpp:
    43-413): dec against dec (NaN keys match nothing), an int build key
    against a dec probe key (as doubles), char(7) against char(7) and
    char(5) against char(9) (strncmp over the shorter width)."""
    cpu = run("rows", nr, ns, 3, "cpu", keys)
    gpu = run("rows", nr, ns, 3, "gpu", keys)
    assert cpu["rows"] > ns // 2
    assert (gpu["rows"], gpu["row_bytes"], gpu["sum"], gpu["xor"]) == \
        (cpu["rows"], cpu["row_bytes"], cpu["sum"], cpu["xor"])


@pytest.mark.parametrize("global_", [0, 1])
def test_join_groupby_same_with_gpu_operators(global_):
    cpu = run("groups", 50000, 400000, 4, "cpu", global_)
    gpu = run("groups", 50000, 400000, 4, "gpu", global_)
    assert cpu["groups"] > 4000
    assert (gpu["groups"], gpu["sum"], gpu["xor"]) == (cpu["groups"], cpu["sum"], cpu["xor"])
