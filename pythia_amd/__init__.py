"""pythia_amd: MI355X-native hash join / hash aggregation for the Pythia
query engine's `gpuhashjoin` / `gpuaggregate_sum` operators.

The compute lives in libpgpu.so (hand-written gfx950 HIP kernels behind the C
ABI of include/pgpu.h).  This module is a thin Python handle over that ABI
used by tests and the benchmark; the C++ operator host in host/ uses the same
ABI directly.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import ALGO_AUTO, ALGO_GLOBAL, ALGO_RADIX, ALGO_RADIX1, GEN, PgpuError, call, lib  # noqa: F401

__all__ = ["Context", "DeviceBuffer", "JoinPlan", "GroupJoin", "Partition", "BatchPlan", "batch_autotune", "Server",
           "c5_trace", "device_count", "PgpuError", "GEN", "ALGO_AUTO", "ALGO_RADIX", "ALGO_GLOBAL", "ALGO_RADIX1",
           "SERVE_BATCHED", "SERVE_PLAIN"]

SERVE_BATCHED, SERVE_PLAIN = 0, 1
GJ_PARTITIONED, GJ_GENERAL = 1, 2


def device_count() -> int:
    n = C.c_int(0)
    rc = lib().pgpu_device_count(C.byref(n))
    return n.value if rc == 0 else 0


def zipf_cdf(n: int, s: float) -> np.ndarray:
    """Zipf(s) CDF over n ranks, as the generators' Zipf column expects it."""
    out = np.empty(n, np.float64)
    call("pgpu_zipf_cdf", n, s, out.ctypes.data_as(C.c_void_p))
    return out


def sizes(family: int, count: int, lo: float, hi: float, seed: int = 42, first: int = 0) -> np.ndarray:
    """round(2^U(lo, hi)) sizes of a family (4: batched joins, 5: resident
    tables, 6: probe batches)."""
    out = np.empty(count, np.uint64)
    call("pgpu_sizes", seed, family, first, count, lo, hi, out.ctypes.data_as(C.c_void_p))
    return out


class DeviceBuffer:
    """A budgeted HBM allocation owned by a Context."""

    def __init__(self, ctx: "Context", nbytes: int):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        call("pgpu_alloc", ctx.handle, self.nbytes, C.byref(p))
        self.ptr = p.value

    def free(self):
        if self.ptr:
            call("pgpu_free", self.ctx.handle, C.c_void_p(self.ptr))
            self.ptr = None

    def upload(self, arr: np.ndarray):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes
        call("pgpu_h2d", self.ctx.handle, C.c_void_p(self.ptr), arr.ctypes.data_as(C.c_void_p), arr.nbytes)

    def download(self, dtype, count: int) -> np.ndarray:
        out = np.empty(count, dtype=dtype)
        assert out.nbytes <= self.nbytes
        call("pgpu_d2h", self.ctx.handle, out.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr), out.nbytes)
        return out


class Context:
    def __init__(self, device: int = 0, hbm_budget: int = 0):
        h = C.c_void_p()
        call("pgpu_ctx_create", device, hbm_budget, C.byref(h))
        self.handle = h

    def close(self):
        if self.handle:
            lib().pgpu_ctx_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def alloc(self, nbytes: int) -> DeviceBuffer:
        return DeviceBuffer(self, nbytes)

    def upload(self, arr: np.ndarray) -> DeviceBuffer:
        b = DeviceBuffer(self, max(arr.nbytes, 1))
        b.upload(arr)
        return b

    def sync(self):
        call("pgpu_sync", self.handle)

    def host_alloc(self, nbytes: int) -> int:
        """Pinned host memory (pgpu_host_alloc); returns its address."""
        p = C.c_void_p()
        call("pgpu_host_alloc", self.handle, nbytes, C.byref(p))
        return p.value

    def host_free(self, addr: int):
        call("pgpu_host_free", self.handle, C.c_void_p(addr))

    def stream(self) -> int:
        s = C.c_void_p()
        call("pgpu_ctx_stream", self.handle, C.byref(s))
        return s.value

    def membw(self, nbytes: int = 1 << 30, iters: int = 10):
        """(read GB/s, copy GB/s) this device sustains for streaming kernels."""
        r, c = C.c_double(), C.c_double()
        call("pgpu_membw", self.handle, nbytes, iters, C.byref(r), C.byref(c))
        return r.value, c.value

    def hbm_usage(self):
        u, b, p = C.c_uint64(), C.c_uint64(), C.c_uint64()
        call("pgpu_hbm_usage", self.handle, C.byref(u), C.byref(b), C.byref(p))
        return u.value, b.value, p.value

    def gen(self, column: str, n: int, out: DeviceBuffer, seed: int = 42, sub: int = 0, n_build: int = 0,
            offset: int = 0, cdf: DeviceBuffer | None = None):
        """Fills n rows of a generated column into `out` (a DeviceBuffer or a
        raw device address)."""
        spec = _lib.GenSpec(seed, sub, n_build, offset, C.c_void_p(cdf.ptr) if cdf else None)
        dst = out.ptr if isinstance(out, DeviceBuffer) else int(out)
        call("pgpu_gen", self.handle, GEN[column], C.byref(spec), n, C.c_void_p(dst))


class JoinPlan:
    """Preallocated join (count + sum) over device-resident columns."""

    def __init__(self, ctx: Context, key_bytes: int, max_build: int, max_probe: int, radix_bits: int = 0,
                 probe_payload: bool = False, algo: int = ALGO_AUTO, join_variant: int = 0):
        self.ctx = ctx
        cfg = _lib.JoinCfg(key_bytes, radix_bits, 1 if probe_payload else 0, algo, join_variant)
        h = C.c_void_p()
        call("pgpu_join_plan_create", ctx.handle, C.byref(cfg), max_build, max_probe, C.byref(h))
        self.handle = h

    def close(self):
        if self.handle:
            lib().pgpu_join_plan_destroy(self.handle)
            self.handle = None

    @staticmethod
    def _p(b):
        if b is None:
            return C.c_void_p(None)
        return C.c_void_p(b.ptr if isinstance(b, DeviceBuffer) else int(b))

    def enqueue(self, rkey, rpay, n_build, skey, spay, n_probe):
        call("pgpu_join_count_sum_async", self.handle, self._p(rkey), self._p(rpay), n_build, self._p(skey),
             self._p(spay), n_probe)

    def result(self):
        r = _lib.JoinResult()
        call("pgpu_join_plan_result", self.handle, C.byref(r))
        return r

    def count_sum(self, rkey, rpay, n_build, skey, spay, n_probe):
        self.enqueue(rkey, rpay, n_build, skey, spay, n_probe)
        return self.result()

    def timing(self, on: bool):
        call("pgpu_join_plan_timing", self.handle, 1 if on else 0)

    def times(self):
        arr = (_lib.KernelTime * 64)()
        n = C.c_int()
        call("pgpu_join_plan_times", self.handle, arr, 64, C.byref(n))
        return {arr[i].name.decode(): (arr[i].launches, arr[i].total_ms) for i in range(min(n.value, 64))}


def _times(fn, handle):
    arr = (_lib.KernelTime * 64)()
    n = C.c_int()
    call(fn, handle, arr, 64, C.byref(n))
    return {arr[i].name.decode(): (arr[i].launches, arr[i].total_ms) for i in range(min(n.value, 64))}


GJ_ROW_SUM, GJ_ROW_COUNT = 1, 2


class GroupJoin:
    """SUM(F.val) GROUP BY D.grp over F JOIN D ON key (int32 keys / groups, FP64 sums)."""

    def __init__(self, ctx: Context, max_dim: int, max_fact: int, max_groups: int):
        self.ctx = ctx
        h = C.c_void_p()
        call("pgpu_groupjoin_create", ctx.handle, max_dim, max_fact, max_groups, C.byref(h))
        self.handle = h

    def close(self):
        if self.handle:
            lib().pgpu_groupjoin_destroy(self.handle)
            self.handle = None

    def run(self, dkey, dgrp, n_dim, fkey, fval, n_fact):
        ng, nj = C.c_uint64(), C.c_uint64()
        call("pgpu_groupjoin_run", self.handle, C.c_void_p(dkey.ptr), C.c_void_p(dgrp.ptr), n_dim,
             C.c_void_p(fkey.ptr), C.c_void_p(fval.ptr), n_fact, C.byref(ng), C.byref(nj))
        return ng.value, nj.value

    def fetch(self, n):
        k, s, c = np.empty(n, np.int32), np.empty(n, np.float64), np.empty(n, np.uint64)
        call("pgpu_groupjoin_fetch", self.handle, k.ctypes.data_as(C.c_void_p), s.ctypes.data_as(C.c_void_p),
             c.ctypes.data_as(C.c_void_p), n)
        return k, s, c

    def rows(self, acc: int, first: int, n: int, dst: int):
        """Result rows [first, first + n) as 12-byte (int32 key, FP64 sum or
        int64 count) rows into host memory at address dst."""
        call("pgpu_groupjoin_rows", self.handle, acc, first, n, C.c_void_p(dst))

    def path(self) -> int:
        """GJ_PARTITIONED or GJ_GENERAL: the algorithm of the last run."""
        v = C.c_int()
        call("pgpu_groupjoin_path", self.handle, C.byref(v))
        return v.value

    def timing(self, on: bool):
        call("pgpu_groupjoin_timing", self.handle, 1 if on else 0)

    def times(self):
        return _times("pgpu_groupjoin_times", self.handle)


COL_INT32, COL_INT64, COL_F64 = 1, 2, 3


class Partition:
    """Stable exactrange partitioning of packed rows (the gpupartition
    operator's device side): rows appended per source, partitioned on
    finish(), pulled per bucket."""

    def __init__(self, ctx: Context, row_bytes: int, key_offset: int, key_bytes: int, range_min: int,
                 range_max: int, buckets: int, sort_offset: int | None = None, sort_type: int = COL_INT64):
        self.ctx, self.row_bytes, self.buckets = ctx, row_bytes, buckets
        d = _lib.PartDesc(row_bytes=row_bytes, key_offset=key_offset, key_bytes=key_bytes, buckets=buckets,
                          range_min=range_min, range_max=range_max, sort=0 if sort_offset is None else 1,
                          sort_offset=sort_offset or 0, sort_type=sort_type)
        h = C.c_void_p()
        call("pgpu_part_create", ctx.handle, C.byref(d), C.byref(h))
        self.handle = h
        self.counts = None

    def close(self):
        if self.handle:
            lib().pgpu_part_destroy(self.handle)
            self.handle = None

    def append(self, source: int, rows: np.ndarray):
        rows = np.ascontiguousarray(rows)
        assert rows.dtype.itemsize == self.row_bytes
        call("pgpu_part_append", self.handle, source, rows.ctypes.data_as(C.c_void_p), len(rows))

    def finish(self) -> np.ndarray:
        counts = np.zeros(self.buckets, np.uint64)
        call("pgpu_part_finish", self.handle, counts.ctypes.data_as(C.c_void_p))
        self.counts = counts
        return counts

    def pull(self, bucket: int, dtype) -> np.ndarray:
        n = int(self.counts[bucket])
        out = np.zeros(n, dtype=dtype)
        call("pgpu_part_pull", self.handle, bucket, 0, n, out.ctypes.data_as(C.c_void_p))
        return out

    def reset(self):
        call("pgpu_part_reset", self.handle)
        self.counts = None


def _batch_array(joins):
    """joins: sequence of (rkey, rpay, n_build, skey, spay, n_probe) with
    DeviceBuffer / raw int device pointers / None."""
    def ptr(b):
        if b is None:
            return None
        return b.ptr if isinstance(b, DeviceBuffer) else int(b)

    arr = (_lib.BatchJoin * len(joins))()
    for i, (rk, rp, nb, sk, sp, ns) in enumerate(joins):
        arr[i] = _lib.BatchJoin(ptr(rk), ptr(rp), nb, ptr(sk), ptr(sp), ns)
    return arr


class BatchPlan:
    """Many independent joins bound into one plan (one hipGraph replay per run)."""

    def __init__(self, ctx: Context, joins, key_bytes: int = 4, probe_payload: bool = True, variant: int = 0,
                 part_rows: int = 0, graph: bool = True):
        self.ctx = ctx
        self.n = len(joins)
        self._arr = _batch_array(joins)
        cfg = _lib.BatchCfg(key_bytes, 1 if probe_payload else 0, variant, part_rows, 1 if graph else 0)
        h = C.c_void_p()
        call("pgpu_batch_plan_create", ctx.handle, C.byref(cfg), self._arr, self.n, C.byref(h))
        self.handle = h

    def close(self):
        if self.handle:
            lib().pgpu_batch_plan_destroy(self.handle)
            self.handle = None

    def enqueue(self):
        call("pgpu_batch_run_async", self.handle)

    def wait(self):
        out = (_lib.JoinResult * self.n)()
        call("pgpu_batch_wait", self.handle, out)
        return list(out)

    def run(self):
        self.enqueue()
        return self.wait()

    def info(self):
        i = _lib.BatchInfo()
        call("pgpu_batch_info", self.handle, C.byref(i))
        return i

    def timing(self, on: bool):
        call("pgpu_batch_timing", self.handle, 1 if on else 0)

    def times(self):
        return _times("pgpu_batch_times", self.handle)


def batch_autotune(ctx: Context, joins, key_bytes: int = 4, probe_payload: bool = True, reps: int = 2):
    """Times every (variant, part_rows) candidate on this batch; returns
    (best (variant, part_rows), [(variant, part_rows, ms), ...])."""
    arr = _batch_array(joins)
    base = _lib.BatchCfg(key_bytes, 1 if probe_payload else 0, 0, 0, 1)
    best = _lib.BatchCfg()
    ent = (_lib.TuneEntry * 16)()
    n = C.c_int()
    call("pgpu_batch_autotune", ctx.handle, C.byref(base), arr, len(joins), reps, C.byref(best), ent, 16,
         C.byref(n))
    return (best.variant, best.part_rows), [(ent[i].variant, ent[i].part_rows, ent[i].ms) for i in range(n.value)]


def _addr(b):
    if b is None:
        return None
    return b.ptr if isinstance(b, DeviceBuffer) else int(b)


def c5_trace(count: int, rate: float, ntables: int = 64, lo: float = 14, hi: float = 25, seed: int = 42):
    """(tables, probe sizes, arrival offsets in seconds) of the served query stream."""
    t = np.empty(count, np.uint32)
    m = np.empty(count, np.uint64)
    off = np.empty(count, np.float64)
    call("pgpu_c5_trace", seed, count, ntables, lo, hi, rate, t.ctypes.data_as(C.c_void_p),
         m.ctypes.data_as(C.c_void_p), off.ctypes.data_as(C.c_void_p))
    return t, m, off


class Server:
    """Join + sum serving engine (resident tables, HBM-budget admission,
    continuous batching; or plain per-query joins with mode=SERVE_PLAIN)."""

    def __init__(self, ctx: Context, mode: int = SERVE_BATCHED, hbm_budget: int = 0, slots: int = 0,
                 max_batch: int = 0, max_items: int = 0, part_min_rows: int = 0):
        self.ctx = ctx
        cfg = _lib.ServerCfg(hbm_budget, 8, mode, slots, max_batch, max_items, part_min_rows)
        h = C.c_void_p()
        call("pgpu_server_create", ctx.handle, C.byref(cfg), C.byref(h))
        self.handle = h

    def close(self):
        if self.handle:
            lib().pgpu_server_destroy(self.handle)
            self.handle = None

    def add_table(self, key, pay, n: int) -> int:
        t = C.c_uint32()
        call("pgpu_server_add_table", self.handle, C.c_void_p(_addr(key)), C.c_void_p(_addr(pay)), n, C.byref(t))
        return t.value

    def save_table(self, table: int, path: str):
        """Writes the table's built form to a memory segment (/dev/shm/<name>
        or a file)."""
        call("pgpu_server_save_table", self.handle, table, path.encode())

    def load_table(self, path: str) -> int:
        """Attaches a saved table without rebuilding it; returns its id."""
        t = C.c_uint32()
        call("pgpu_server_load_table", self.handle, path.encode(), C.byref(t))
        return t.value

    def prepare_plain(self, max_probe: int):
        call("pgpu_server_prepare_plain", self.handle, max_probe)

    def submit(self, table: int, probe, m: int, tag: int = 0, t_arrival: float = 0.0):
        call("pgpu_server_submit", self.handle, table, C.c_void_p(_addr(probe)), m, tag, t_arrival)

    def drain(self):
        call("pgpu_server_drain", self.handle)

    def poll(self, cap: int = 4096):
        out = (_lib.ServedResult * cap)()
        n = C.c_int()
        call("pgpu_server_poll", self.handle, out, cap, C.byref(n))
        return [out[i] for i in range(n.value)]

    def now(self) -> float:
        t = C.c_double()
        call("pgpu_server_now", self.handle, C.byref(t))
        return t.value

    def replay(self, tables, probes, ms, offsets):
        n = len(tables)
        t = np.ascontiguousarray(tables, dtype=np.uint32)
        p = np.ascontiguousarray([_addr(x) for x in probes], dtype=np.uint64)
        m = np.ascontiguousarray(ms, dtype=np.uint64)
        o = np.ascontiguousarray(offsets, dtype=np.float64)
        out = (_lib.ServedResult * max(n, 1))()
        call("pgpu_server_replay", self.handle, t.ctypes.data_as(C.c_void_p), p.ctypes.data_as(C.c_void_p),
             m.ctypes.data_as(C.c_void_p), o.ctypes.data_as(C.c_void_p), n, out)
        return [out[i] for i in range(n)]

    def stats(self):
        st = _lib.ServerStats()
        call("pgpu_server_stats_get", self.handle, C.byref(st))
        return {f: getattr(st, f) for f, _ in st._fields_}
