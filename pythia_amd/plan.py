"""ctypes binding of libpythia_host.so (include/pythia_host.h): run a plan in
libconfig syntax through the C++ engine that hosts the GPU operators."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .build import LIBHOST, LIBPGPU

_lib = None

NP = {"int": np.int32, "long": np.int64, "dec": np.float64, "date": np.int64}


class Result(C.Structure):
    _fields_ = [("rows", C.c_void_p), ("nrows", C.c_uint64), ("row_bytes", C.c_uint32),
                ("schema", C.c_char * 256), ("seconds", C.c_double), ("getnext_calls", C.c_uint64),
                ("gpu_fused", C.c_uint32), ("join_d2h_bytes", C.c_uint64)]

# pythia_result.gpu_fused bits
FUSED_ROWS, FUSED_COUNT_SUM, FUSED_GROUP = 1, 2, 4


class QueryError(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is None:
        for f in (LIBPGPU, LIBHOST):
            if not os.path.exists(f):
                raise RuntimeError(f"{f} is missing: build it with __graft_entry__.build()")
        C.CDLL(LIBPGPU, mode=C.RTLD_GLOBAL)
        l = C.CDLL(LIBHOST)
        l.pythia_last_error.restype = C.c_char_p
        l.pythia_register_table.argtypes = [C.c_char_p, C.c_void_p, C.c_uint64, C.c_uint32]
        l.pythia_run_query.argtypes = [C.c_char_p, C.POINTER(Result)]
        l.pythia_free_result.argtypes = [C.POINTER(Result)]
        _lib = l
    return _lib


def dtype_of(schema: str) -> np.dtype:
    """Packed-row numpy dtype of an engine schema string such as "int,long,dec"."""
    names, formats, offsets, off = [], [], [], 0
    for i, t in enumerate(s for s in schema.split(",") if s):
        if t.startswith("char("):
            w = int(t[5:-1]) + 1
            fmt = f"S{w}"
        else:
            w, fmt = (4 if t == "int" else 8), NP[t]
        names.append(f"c{i}")
        formats.append(fmt)
        offsets.append(off)
        off += w
    return np.dtype({"names": names, "formats": formats, "offsets": offsets, "itemsize": max(off, 1)})


def register_table(name: str, rows: np.ndarray):
    rows = np.ascontiguousarray(rows)
    rc = lib().pythia_register_table(name.encode(), rows.ctypes.data_as(C.c_void_p), len(rows), rows.dtype.itemsize)
    if rc != 0:
        raise QueryError(lib().pythia_last_error().decode())


def clear_tables():
    lib().pythia_clear_tables()


def clear_resident():
    """Frees the HBM copies kept by plans with gpu { resident = true; }."""
    lib().pythia_clear_resident()


def run(plan: str):
    """Runs a plan; returns (rows as a structured array, Result metadata)."""
    r = Result()
    rc = lib().pythia_run_query(plan.encode(), C.byref(r))
    if rc != 0:
        raise QueryError(lib().pythia_last_error().decode(errors="replace"))
    try:
        schema = r.schema.decode()
        dt = dtype_of(schema)
        n = r.nrows
        buf = (C.c_char * (n * r.row_bytes)).from_address(r.rows) if n else b""
        out = np.frombuffer(bytes(buf), dtype=dt, count=n).copy() if n else np.zeros(0, dtype=dt)
        meta = {"schema": schema, "seconds": r.seconds, "getnext_calls": r.getnext_calls, "row_bytes": r.row_bytes,
                "gpu_fused": r.gpu_fused, "join_d2h_bytes": r.join_d2h_bytes}
    finally:
        lib().pythia_free_result(C.byref(r))
    return out, meta
