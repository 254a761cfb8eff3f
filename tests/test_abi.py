"""The C-ABI library loads and exports every function include/*.h declares.
No compute calls here: this runs without a GPU."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "pythia_amd", "libpgpu.so")
HOSTLIB = os.path.join(ROOT, "pythia_amd", "libpythia_host.so")
HEADERS = [(os.path.join(ROOT, "include", "pgpu.h"), LIB, "pgpu_"),
           (os.path.join(ROOT, "include", "pythia_host.h"), HOSTLIB, "pythia_")]


def declared_functions(path, prefix):
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    pat = r"^\s*(?:const\s+)?[a-z_0-9]+\s*\**\s*(" + prefix + r"[a-z_0-9]+)\s*\("
    return sorted(set(re.findall(pat, text, flags=re.M)))


def nm_exports(path):
    import subprocess

    out = subprocess.run(["nm", "-D", "--defined-only", path], stdout=subprocess.PIPE, text=True, check=True).stdout
    return {l.split()[-1] for l in out.splitlines() if l.strip()}


@pytest.mark.parametrize("header,lib,prefix", HEADERS)
def test_library_exports_every_declared_symbol(header, lib, prefix):
    assert os.path.exists(lib), f"{lib} not built"
    C.CDLL(LIB, mode=C.RTLD_GLOBAL)
    C.CDLL(lib)  # loads (HIP runtime initialises lazily; no device needed)
    names = declared_functions(header, prefix)
    assert len(names) >= 5
    missing = [n for n in names if n not in nm_exports(lib)]
    assert not missing, f"declared but not exported: {missing}"


def test_exports_are_c_linkage():
    exports = nm_exports(LIB)
    assert "pgpu_join_count_sum" in exports  # unmangled
    assert not any(e.startswith("_Z") and "pgpu_join" in e and "plan" in e and e.endswith("create") for e in exports)


def test_error_path_without_device():
    """Argument errors are reported through the status code and
    pgpu_last_error() before any device work."""
    l = C.CDLL(LIB)
    l.pgpu_last_error.restype = C.c_char_p
    rc = l.pgpu_join_plan_create(None, None, C.c_uint64(0), C.c_uint64(0), None)
    assert rc == -1
    assert b"null" in l.pgpu_last_error()


def test_host_generator_helpers_match_oracle():
    """The library's host-side Zipf CDF is the same array the oracle builds,
    and log-uniform sizes stay in [2^lo, 2^hi] with a log-mean near the middle."""
    import numpy as np

    import pythia_amd as pa
    from oracle import oracle as O

    a, b = pa.zipf_cdf(5000, 1.0), O.zipf_cdf(5000, 1.0)
    assert np.array_equal(a, b)
    # closed form: CDF(1) = 1 / H_n for s = 1
    assert abs(a[0] - 1.0 / np.sum(1.0 / np.arange(1, 5001))) < 1e-15
    sz = pa.sizes(4, 4096, 10, 22)
    assert sz.min() >= 1 << 10 and sz.max() <= 1 << 22
    assert abs(np.log2(sz.astype(np.float64)).mean() - 16.0) < 0.2
    assert np.array_equal(sz[100:200], pa.sizes(4, 100, 10, 22, first=100))


def test_radix_plan_size_limits_without_device():
    """The radix plan keeps histogram columns and in-partition offsets in 32
    bits, so each side must hold fewer than 2^32 rows: larger plans are
    refused by the argument checks, which run before the context is used
    (the context pointer below is never dereferenced)."""
    from pythia_amd import _lib

    l = C.CDLL(LIB)
    l.pgpu_last_error.restype = C.c_char_p
    l.pgpu_join_plan_create.argtypes = [C.c_void_p, C.POINTER(_lib.JoinCfg), C.c_uint64, C.c_uint64,
                                        C.POINTER(C.c_void_p)]
    fake_ctx = C.create_string_buffer(64)
    out = C.c_void_p()
    for kb in (4, 8):
        cfg = _lib.JoinCfg(kb, 0, 0, _lib.ALGO_RADIX, 0)
        for nb, npr in [(1 << 20, 1 << 32), (1 << 32, 1 << 20), (1 << 32, 1 << 40)]:
            rc = l.pgpu_join_plan_create(C.addressof(fake_ctx), C.byref(cfg), nb, npr, C.byref(out))
            assert rc == -1, (kb, nb, npr)
            assert b"fewer than 2^32" in l.pgpu_last_error()
            assert not out.value
    bad = _lib.JoinCfg(8, 0, 0, _lib.ALGO_RADIX, 9)  # unknown LDS join shape
    assert l.pgpu_join_plan_create(C.addressof(fake_ctx), C.byref(bad), 16, 16, C.byref(out)) == -1
    assert b"variant" in l.pgpu_last_error()
