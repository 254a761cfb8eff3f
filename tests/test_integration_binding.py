"""The reference-side binding (integration/gpujoin.{h,cpp}): the GPU plan
operators as classes of the reference engine, compiled in its dialect
(gnu++98) against its own headers (operators/operators.h:79-186, schema.h,
hash.h, util/buffer.h) and the C ABI of include/pgpu.h.  The GPU runs of the
linked binary are in tests/test_gpu_integration.py."""
import json
import os
import subprocess

import pytest

from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REFSRC = "/root/reference"
REFGPU = os.path.join(ROOT, "oracle", "_ref", "refgpujoin")


@pytest.mark.skipif(not os.path.isdir(REFSRC), reason="reference sources absent")
def test_binding_compiles_against_the_reference_headers():
    cmd = ["g++", "-std=gnu++98", "-fsyntax-only", "-Wall", "-Werror", "-DSTATS_ALLOCATE",
           f"-I{REFSRC}", f"-I{REFSRC}/externals/libconfig-1.2", f"-I{REFSRC}/externals/bzip2-1.0.5",
           f"-I{ROOT}/include", os.path.join(ROOT, "integration", "gpujoin.cpp")]
    out = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert out.returncode == 0, out.stdout


@pytest.mark.skipif(not os.path.exists(REFGPU), reason="oracle/_ref/refgpujoin not built (oracle/ref.mk)")
def test_linked_engine_cpu_operators_match_the_oracle():
    """The driver's CPU plans (the reference's own HashJoinOp /
    AggregateSum / AggregateCount, the GPU classes linked in but unused) give
    the C1 closed form; thread-local and Global group sums agree."""
    def run(*a):
        out = subprocess.run([REFGPU, *map(str, a)], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True,
                             timeout=300, cwd="/tmp")
        assert out.returncode == 0, out.stderr
        return json.loads(out.stdout)

    ref = O.c1_closed(1 << 12, 1 << 17)
    wrap = lambda x: (x + 2**31) % 2**32 - 2**31  # noqa: E731
    assert run("c1", 12, 17, 4, "cpu", "count")["value"] == ref.count
    assert run("c1", 12, 17, 4, "cpu", "sum")["value"] == wrap(ref.sum_build)
    assert run("c1", 12, 17, 4, "cpu", "probesum")["value"] == wrap(ref.sum_probe)
    g0, g1 = run("groups", 20000, 100000, 4, "cpu", 0), run("groups", 20000, 100000, 4, "cpu", 1)
    assert (g0["groups"], g0["sum"], g0["xor"]) == (g1["groups"], g1["sum"], g1["xor"])
