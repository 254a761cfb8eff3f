"""ctypes binding of libpgpu.so (the C ABI declared in include/pgpu.h).

There is no fallback: if the library is missing or cannot be loaded the
import of the bound functions raises, naming the file to build.
"""
from __future__ import annotations

import ctypes as C
import os

from .build import LIBPGPU

PGPU_OK = 0
STATUS = {
    -1: "PGPU_EINVAL", -2: "PGPU_ENOMEM", -3: "PGPU_EBUDGET", -4: "PGPU_EDEVICE",
    -5: "PGPU_EOVERFLOW", -6: "PGPU_ESTATE", -7: "PGPU_EUNSUPPORTED",
}

# enum pgpu_gen_column
GEN = {name: i for i, name in enumerate([
    "C1_RKEY", "C1_RPAY", "C1_SFK", "C1_SPAY", "C2_RKEY", "C2_RPAY", "C2_SFK_UNIFORM", "C2_SFK_ZIPF",
    "C3_D_ORDERKEY", "C3_D_CUSTKEY", "C3_F_ORDERKEY", "C3_F_PARTKEY", "C3_F_QUANTITY", "C3_F_PRICE",
    "C4_RKEY", "C4_RPAY", "C4_SFK", "C4_SPAY"], start=1)}

ALGO_AUTO, ALGO_RADIX, ALGO_GLOBAL, ALGO_RADIX1 = 0, 1, 2, 3


class GenSpec(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("sub", C.c_uint64), ("n_build", C.c_uint64),
                ("offset", C.c_uint64), ("d_cdf", C.c_void_p)]


class JoinCfg(C.Structure):
    _fields_ = [("key_bytes", C.c_uint32), ("radix_bits", C.c_uint32),
                ("probe_payload", C.c_uint32), ("algo", C.c_uint32), ("join_variant", C.c_uint32)]


class JoinResult(C.Structure):
    _fields_ = [("matches", C.c_uint64), ("sum_build", C.c_uint64), ("sum_probe", C.c_uint64),
                ("radix_bits", C.c_uint32), ("lds_passes", C.c_uint32)]


class BatchJoin(C.Structure):
    _fields_ = [("d_rkey", C.c_void_p), ("d_rpay", C.c_void_p), ("n_build", C.c_uint64),
                ("d_skey", C.c_void_p), ("d_spay", C.c_void_p), ("n_probe", C.c_uint64)]


class BatchCfg(C.Structure):
    _fields_ = [("key_bytes", C.c_uint32), ("probe_payload", C.c_uint32), ("variant", C.c_uint32),
                ("part_rows", C.c_uint32), ("graph", C.c_uint32)]


class BatchInfo(C.Structure):
    _fields_ = [("cfg", BatchCfg), ("items", C.c_uint32), ("grid", C.c_uint32), ("build_rows", C.c_uint64),
                ("probe_rows", C.c_uint64), ("chunks", C.c_uint64), ("graph", C.c_uint32)]


class TuneEntry(C.Structure):
    _fields_ = [("variant", C.c_uint32), ("part_rows", C.c_uint32), ("ms", C.c_double)]


class ServerCfg(C.Structure):
    _fields_ = [("hbm_budget", C.c_uint64), ("key_bytes", C.c_uint32), ("mode", C.c_uint32), ("slots", C.c_uint32),
                ("max_batch", C.c_uint32), ("max_items", C.c_uint32), ("part_min_rows", C.c_uint32)]


class ServedResult(C.Structure):
    _fields_ = [("tag", C.c_uint64), ("table", C.c_uint32), ("batch_size", C.c_uint32), ("m", C.c_uint64),
                ("matches", C.c_uint64), ("sum_build", C.c_uint64), ("t_arrival", C.c_double),
                ("t_admit", C.c_double), ("t_done", C.c_double), ("batch", C.c_uint64)]


class ServerStats(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("tables", "resident_bytes", "reserved_bytes", "budget", "submitted",
                                          "completed", "batches", "queued", "partitioned")]


class PartDesc(C.Structure):
    _fields_ = [("row_bytes", C.c_uint32), ("key_offset", C.c_uint32), ("key_bytes", C.c_uint32),
                ("buckets", C.c_uint32), ("range_min", C.c_int64), ("range_max", C.c_int64), ("sort", C.c_uint32),
                ("sort_offset", C.c_uint32), ("sort_type", C.c_uint32), ("pad", C.c_uint32)]


class KernelTime(C.Structure):
    _fields_ = [("name", C.c_char * 48), ("launches", C.c_uint64), ("total_ms", C.c_double)]


class PgpuError(RuntimeError):
    def __init__(self, fn: str, code: int, msg: str):
        super().__init__(f"{fn} -> {STATUS.get(code, code)}: {msg}")
        self.code = code


_lib = None

# name -> (restype, argtypes)
_PROTOS = {
    "pgpu_version": (C.c_char_p, []),
    "pgpu_last_error": (C.c_char_p, []),
    "pgpu_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "pgpu_ctx_create": (C.c_int, [C.c_int, C.c_uint64, C.POINTER(C.c_void_p)]),
    "pgpu_ctx_destroy": (C.c_int, [C.c_void_p]),
    "pgpu_ctx_stream": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p)]),
    "pgpu_alloc": (C.c_int, [C.c_void_p, C.c_uint64, C.POINTER(C.c_void_p)]),
    "pgpu_free": (C.c_int, [C.c_void_p, C.c_void_p]),
    "pgpu_hbm_usage": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                                 C.POINTER(C.c_uint64)]),
    "pgpu_h2d": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64]),
    "pgpu_d2h": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64]),
    "pgpu_sync": (C.c_int, [C.c_void_p]),
    "pgpu_membw": (C.c_int, [C.c_void_p, C.c_uint64, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "pgpu_gen": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(GenSpec), C.c_uint64, C.c_void_p]),
    "pgpu_zipf_cdf": (C.c_int, [C.c_uint64, C.c_double, C.c_void_p]),
    "pgpu_sizes": (C.c_int, [C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64, C.c_double, C.c_double,
                             C.c_void_p]),
    "pgpu_join_plan_create": (C.c_int, [C.c_void_p, C.POINTER(JoinCfg), C.c_uint64, C.c_uint64,
                                        C.POINTER(C.c_void_p)]),
    "pgpu_join_plan_destroy": (C.c_int, [C.c_void_p]),
    "pgpu_join_count_sum_async": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64,
                                            C.c_void_p, C.c_void_p, C.c_uint64]),
    "pgpu_join_plan_result": (C.c_int, [C.c_void_p, C.POINTER(JoinResult)]),
    "pgpu_join_count_sum": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p,
                                      C.c_void_p, C.c_uint64, C.POINTER(JoinResult)]),
    "pgpu_join_plan_timing": (C.c_int, [C.c_void_p, C.c_int]),
    "pgpu_join_plan_times": (C.c_int, [C.c_void_p, C.POINTER(KernelTime), C.c_int, C.POINTER(C.c_int)]),
    "pgpu_batch_plan_create": (C.c_int, [C.c_void_p, C.POINTER(BatchCfg), C.POINTER(BatchJoin), C.c_uint32,
                                         C.POINTER(C.c_void_p)]),
    "pgpu_batch_plan_destroy": (C.c_int, [C.c_void_p]),
    "pgpu_batch_run_async": (C.c_int, [C.c_void_p]),
    "pgpu_batch_wait": (C.c_int, [C.c_void_p, C.c_void_p]),
    "pgpu_batch_info": (C.c_int, [C.c_void_p, C.POINTER(BatchInfo)]),
    "pgpu_batch_timing": (C.c_int, [C.c_void_p, C.c_int]),
    "pgpu_batch_times": (C.c_int, [C.c_void_p, C.POINTER(KernelTime), C.c_int, C.POINTER(C.c_int)]),
    "pgpu_batch_autotune": (C.c_int, [C.c_void_p, C.POINTER(BatchCfg), C.POINTER(BatchJoin), C.c_uint32, C.c_int,
                                      C.POINTER(BatchCfg), C.POINTER(TuneEntry), C.c_int, C.POINTER(C.c_int)]),
    "pgpu_server_create": (C.c_int, [C.c_void_p, C.POINTER(ServerCfg), C.POINTER(C.c_void_p)]),
    "pgpu_server_destroy": (C.c_int, [C.c_void_p]),
    "pgpu_server_add_table": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64, C.POINTER(C.c_uint32)]),
    "pgpu_server_prepare_plain": (C.c_int, [C.c_void_p, C.c_uint64]),
    "pgpu_server_submit": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint64, C.c_uint64, C.c_double]),
    "pgpu_server_drain": (C.c_int, [C.c_void_p]),
    "pgpu_server_poll": (C.c_int, [C.c_void_p, C.POINTER(ServedResult), C.c_int, C.POINTER(C.c_int)]),
    "pgpu_server_now": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    "pgpu_server_replay": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64,
                                     C.POINTER(ServedResult)]),
    "pgpu_server_stats_get": (C.c_int, [C.c_void_p, C.POINTER(ServerStats)]),
    "pgpu_server_save_table": (C.c_int, [C.c_void_p, C.c_uint32, C.c_char_p]),
    "pgpu_server_load_table": (C.c_int, [C.c_void_p, C.c_char_p, C.POINTER(C.c_uint32)]),
    "pgpu_c5_trace": (C.c_int, [C.c_uint64, C.c_uint64, C.c_uint32, C.c_double, C.c_double, C.c_double,
                                C.c_void_p, C.c_void_p, C.c_void_p]),
    "pgpu_part_create": (C.c_int, [C.c_void_p, C.POINTER(PartDesc), C.POINTER(C.c_void_p)]),
    "pgpu_part_destroy": (C.c_int, [C.c_void_p]),
    "pgpu_part_append": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint64]),
    "pgpu_part_finish": (C.c_int, [C.c_void_p, C.c_void_p]),
    "pgpu_part_pull": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint64, C.c_uint64, C.c_void_p]),
    "pgpu_part_reset": (C.c_int, [C.c_void_p]),
    "pgpu_groupjoin_create": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint64, C.c_uint64, C.POINTER(C.c_void_p)]),
    "pgpu_groupjoin_destroy": (C.c_int, [C.c_void_p]),
    "pgpu_groupjoin_run": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p, C.c_void_p,
                                     C.c_uint64, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "pgpu_groupjoin_fetch": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64]),
    "pgpu_groupjoin_rows": (C.c_int, [C.c_void_p, C.c_int, C.c_uint64, C.c_uint64, C.c_void_p]),
    "pgpu_host_alloc": (C.c_int, [C.c_void_p, C.c_uint64, C.POINTER(C.c_void_p)]),
    "pgpu_host_free": (C.c_int, [C.c_void_p, C.c_void_p]),
    "pgpu_groupjoin_path": (C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    "pgpu_groupjoin_timing": (C.c_int, [C.c_void_p, C.c_int]),
    "pgpu_groupjoin_times": (C.c_int, [C.c_void_p, C.POINTER(KernelTime), C.c_int, C.POINTER(C.c_int)]),
}


def lib():
    """Loads libpgpu.so once; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPGPU):
            raise RuntimeError(f"{LIBPGPU} is missing: build it with __graft_entry__.build()")
        l = C.CDLL(LIBPGPU)
        for name, (res, args) in _PROTOS.items():
            f = getattr(l, name)
            f.restype = res
            f.argtypes = args
        _lib = l
    return _lib


def check(fn: str, rc: int) -> None:
    if rc != PGPU_OK:
        raise PgpuError(fn, rc, lib().pgpu_last_error().decode(errors="replace"))


def call(fn: str, *args) -> None:
    check(fn, getattr(lib(), fn)(*args))
