"""TEST INFRASTRUCTURE ONLY: ctypes binding of oracle/liboracle.so.

The oracle is the CPU restatement of the reference's hash join / hash
aggregation (see oracle.h).  Only tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg load it, always as the checker or the timed
CPU baseline, never as the thing measured or shipped.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

INT, LONG, DEC, CHAR, DATE = 0, 1, 2, 3, 4
H_MODULO, H_PARAMMODULO, H_KNUTH, H_RANGE, H_EXACTRANGE, H_BYTES, H_ALWAYSZERO, H_WILLIS, H_TPCH, H_TPCHQ1 = range(10)

NP_OF = {INT: np.int32, LONG: np.int64, DEC: np.float64, DATE: np.int64}


class Schema(C.Structure):
    _fields_ = [("ncols", C.c_int), ("type", C.c_int * 32), ("size", C.c_int * 32),
                ("offset", C.c_int * 32), ("tuplesize", C.c_int)]

    def dtype(self) -> np.dtype:
        """numpy structured dtype of the packed row (no padding)."""
        names, formats, offsets = [], [], []
        for i in range(self.ncols):
            names.append(f"c{i}")
            t = self.type[i]
            formats.append(f"S{self.size[i]}" if t == CHAR else NP_OF[t])
            offsets.append(self.offset[i])
        return np.dtype({"names": names, "formats": formats, "offsets": offsets, "itemsize": self.tuplesize})


class Hasher(C.Structure):
    _fields_ = [("fn", C.c_int), ("k", C.c_int64), ("min", C.c_int64), ("max", C.c_int64),
                ("bucketrange", C.c_int64), ("skipbits", C.c_int), ("buckets", C.c_uint),
                ("offset", C.c_int), ("size", C.c_int)]


class JoinDesc(C.Structure):
    _fields_ = [("build", Schema), ("probe", Schema), ("buildjattr", C.c_int), ("probejattr", C.c_int),
                ("nproj", C.c_int), ("proj_side", C.c_int * 32), ("proj_col", C.c_int * 32),
                ("hashfn", C.c_int), ("buckets", C.c_uint), ("rmin", C.c_int64), ("rmax", C.c_int64),
                ("skipbits", C.c_int), ("tuplesperbucket", C.c_uint), ("threads", C.c_int)]


class JoinStats(C.Structure):
    _fields_ = [("matches", C.c_uint64), ("spills", C.c_uint64), ("build_s", C.c_double),
                ("probe_s", C.c_double)]


class CountSum(C.Structure):
    _fields_ = [("count", C.c_uint64), ("isum", C.c_int64), ("dsum", C.c_double)]


class AggDesc(C.Structure):
    _fields_ = [("in_", Schema), ("nfields", C.c_int), ("fields", C.c_int * 8), ("sumfield", C.c_int),
                ("hashfn", C.c_int), ("buckets", C.c_uint), ("rmin", C.c_int64), ("rmax", C.c_int64),
                ("global_", C.c_int), ("threads", C.c_int)]


class CS64(C.Structure):
    _fields_ = [("count", C.c_uint64), ("sum_build", C.c_uint64), ("sum_probe", C.c_uint64)]


_lib = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)
    return LIB


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        l = C.CDLL(LIB)
        P = C.c_void_p
        protos = {
            "orc_schema_add": (None, [C.POINTER(Schema), C.c_int, C.c_int]),
            "orc_getlogarithm": (C.c_uint, [C.c_uint]),
            "orc_hasher_init": (C.c_int, [C.POINTER(Hasher), C.c_int, C.c_uint, C.c_int64, C.c_int64, C.c_int,
                                          C.POINTER(Schema), C.c_int, C.c_int]),
            "orc_hash_value": (C.c_uint, [C.POINTER(Hasher), C.c_int64]),
            "orc_hash_tuple": (C.c_uint, [C.POINTER(Hasher), P]),
            "orc_exactrange_minimum": (C.c_int64, [C.POINTER(Hasher), C.c_uint]),
            "orc_generate": (C.c_int, [C.c_uint, C.c_int, C.c_uint, C.POINTER(C.c_uint), C.POINTER(C.c_int)]),
            "orc_ht_create": (P, [C.c_uint, C.c_uint, C.c_uint]),
            "orc_ht_destroy": (None, [P]),
            "orc_ht_allocate": (P, [P, C.c_uint]),
            "orc_ht_bucket_tuples": (C.c_uint, [P, C.c_uint, C.POINTER(P), C.c_uint]),
            "orc_ht_spills": (C.c_uint64, [P]),
            "orc_ht_serialize": (C.c_int, [P, C.c_char_p]),
            "orc_ht_deserialize": (C.c_int, [P, C.c_char_p]),
            "orc_join_outschema": (None, [C.POINTER(JoinDesc), C.POINTER(Schema)]),
            "orc_hashjoin": (C.c_int64, [C.POINTER(JoinDesc), P, C.c_uint64, P, C.c_uint64, P, C.c_uint64,
                                         C.POINTER(JoinStats)]),
            "orc_hashjoin_count_sum": (C.c_int, [C.POINTER(JoinDesc), P, C.c_uint64, P, C.c_uint64, C.c_int,
                                                 C.POINTER(CountSum), C.POINTER(JoinStats)]),
            "orc_agg_outschema": (None, [C.POINTER(AggDesc), C.POINTER(Schema)]),
            "orc_aggregate": (C.c_int64, [C.POINTER(AggDesc), P, C.c_uint64, P, C.c_uint64]),
            "orc_partition": (C.c_int, [C.POINTER(Schema), C.c_int, C.c_int64, C.c_int64, C.c_uint, P, C.c_uint64,
                                        P, P]),
            "orc_c1_closed": (None, [C.c_uint64, C.c_uint64, C.c_uint64, C.c_int, C.POINTER(CS64)]),
            "orc_c2_closed": (None, [C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64, C.c_double, C.c_int,
                                     C.POINTER(CS64)]),
            "orc_c3_closed": (C.c_uint64, [C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64, C.c_int,
                                           P]),
            "orc_c4_closed": (None, [C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64, C.POINTER(CS64)]),
            "orc_c4_closed_many": (None, [C.c_uint64, C.c_uint64, C.c_uint64, P, P, C.c_int, C.POINTER(CS64)]),
            "orc_gen_c1": (None, [C.c_uint64, C.c_uint64, C.c_uint64, P, P, P, P]),
            "orc_gen_c2": (None, [C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64, C.c_double, P, P, P]),
            "orc_gen_c3": (None, [C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64, P, P, P, P]),
            "orc_zipf_cdf": (None, [C.c_uint64, C.c_double, P]),
            "orc_c5_closed": (None, [C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64, C.c_int,
                                     C.POINTER(CS64)]),
            "orc_gen_c4": (None, [C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64, P, P, P, P]),
            "orc_baseline_join": (C.c_double, [C.c_int, P, P, C.c_uint64, P, C.c_uint64, C.c_uint, C.c_uint,
                                               C.c_int, C.POINTER(CS64)]),
        }
        for name, (res, args) in protos.items():
            f = getattr(l, name)
            f.restype = res
            f.argtypes = args
        _lib = l
    return _lib


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def schema(*cols) -> Schema:
    """schema(("int",), ("long",), ("dec",), ("char", 10)) -> packed Schema."""
    s = Schema()
    names = {"int": INT, "long": LONG, "dec": DEC, "date": DATE, "char": CHAR}
    for c in cols:
        t = names[c[0]]
        lib().orc_schema_add(C.byref(s), t, c[1] if len(c) > 1 else 0)
    return s


def rows(s: Schema, *columns) -> np.ndarray:
    """Packs columns into a packed-row array of schema s."""
    n = len(columns[0])
    r = np.zeros(n, dtype=s.dtype())
    for i, c in enumerate(columns):
        r[f"c{i}"] = c
    return r


def hasher(fn, buckets, s: Schema, field, rmin=0, rmax=0, skipbits=0, fieldmax=None) -> Hasher:
    h = Hasher()
    rc = lib().orc_hasher_init(C.byref(h), fn, buckets, rmin, rmax, skipbits, C.byref(s), field,
                               field if fieldmax is None else fieldmax)
    if rc != 0:
        raise ValueError(f"orc_hasher_init -> {rc}")
    return h


def join_desc(build: Schema, probe: Schema, bj: int, pj: int, projection, hashfn=H_MODULO, buckets=16,
              tuplesperbucket=2, threads=1, rmin=0, rmax=0) -> JoinDesc:
    d = JoinDesc()
    d.build, d.probe, d.buildjattr, d.probejattr = build, probe, bj, pj
    d.nproj = len(projection)
    for i, p in enumerate(projection):
        side, col = p[0], int(p[2:])
        assert p[1] == "$" and side in "BP"
        d.proj_side[i] = 0 if side == "B" else 1
        d.proj_col[i] = col
    d.hashfn, d.buckets, d.tuplesperbucket, d.threads = hashfn, buckets, tuplesperbucket, threads
    d.rmin, d.rmax = rmin, rmax
    return d


def join_outschema(d: JoinDesc) -> Schema:
    s = Schema()
    lib().orc_join_outschema(C.byref(d), C.byref(s))
    return s


def hashjoin(d: JoinDesc, build_rows: np.ndarray, probe_rows: np.ndarray, cap: int | None = None):
    out_s = join_outschema(d)
    if cap is None:
        cap = max(1, len(build_rows) * len(probe_rows)) if len(build_rows) * len(probe_rows) < 4_000_000 \
            else 4 * (len(probe_rows) + len(build_rows))
    out = np.zeros(cap, dtype=out_s.dtype())
    st = JoinStats()
    n = lib().orc_hashjoin(C.byref(d), _ptr(build_rows), len(build_rows), _ptr(probe_rows), len(probe_rows),
                           _ptr(out), cap, C.byref(st))
    if n < 0:
        raise RuntimeError(f"orc_hashjoin -> {n}")
    return out[:n], st


def hashjoin_count_sum(d: JoinDesc, build_rows, probe_rows, sumcol: int):
    cs, st = CountSum(), JoinStats()
    rc = lib().orc_hashjoin_count_sum(C.byref(d), _ptr(build_rows), len(build_rows), _ptr(probe_rows),
                                      len(probe_rows), sumcol, C.byref(cs), C.byref(st))
    if rc != 0:
        raise RuntimeError(f"orc_hashjoin_count_sum -> {rc}")
    return cs, st


def agg_desc(s: Schema, fields, sumfield, hashfn=H_MODULO, buckets=4, global_=False, threads=1) -> AggDesc:
    d = AggDesc()
    d.in_ = s
    d.nfields = len(fields)
    for i, f in enumerate(fields):
        d.fields[i] = f
    d.sumfield, d.hashfn, d.buckets, d.global_, d.threads = sumfield, hashfn, buckets, int(global_), threads
    return d


def aggregate(d: AggDesc, in_rows: np.ndarray, cap: int | None = None) -> np.ndarray:
    out_s = Schema()
    lib().orc_agg_outschema(C.byref(d), C.byref(out_s))
    cap = cap or max(1, len(in_rows) * max(1, d.threads))
    out = np.zeros(cap, dtype=out_s.dtype())
    n = lib().orc_aggregate(C.byref(d), _ptr(in_rows), len(in_rows), _ptr(out), cap)
    if n < 0:
        raise RuntimeError(f"orc_aggregate -> {n}")
    return out[:n]


def partition(s: Schema, attr: int, rmin: int, rmax: int, buckets: int, in_rows: np.ndarray):
    out = np.zeros_like(in_rows)
    counts = np.zeros(buckets, dtype=np.uint64)
    rc = lib().orc_partition(C.byref(s), attr, rmin, rmax, buckets, _ptr(in_rows), len(in_rows), _ptr(out),
                             _ptr(counts))
    if rc != 0:
        raise RuntimeError(f"orc_partition -> {rc}")
    return out, counts


def generate(buckets: int, skipbits: int, passes: int):
    b = (C.c_uint * passes)()
    s = (C.c_int * passes)()
    lib().orc_generate(buckets, skipbits, passes, b, s)
    return list(b), list(s)


# ---------------------------------------------------------------- generators / closed forms
def gen_c1(nR, nS, seed=42):
    rk, rp = np.empty(nR, np.int32), np.empty(nR, np.int32)
    sk, sp = np.empty(nS, np.int32), np.empty(nS, np.int32)
    lib().orc_gen_c1(seed, nR, nS, _ptr(rk), _ptr(rp), _ptr(sk), _ptr(sp))
    return rk, rp, sk, sp


def gen_c2(nR, nS, zipf_s=0.0, seed=42, sub=0):
    rk, rp, sk = np.empty(nR, np.int64), np.empty(nR, np.int64), np.empty(nS, np.int64)
    lib().orc_gen_c2(seed, sub, nR, nS, zipf_s, _ptr(rk), _ptr(rp), _ptr(sk))
    return rk, rp, sk


def gen_c3(nD, nF, norders=15_000_000, seed=42):
    dk, dc = np.empty(nD, np.int32), np.empty(nD, np.int32)
    fk, fp = np.empty(nF, np.int32), np.empty(nF, np.float64)
    lib().orc_gen_c3(seed, nD, nF, norders, _ptr(dk), _ptr(dc), _ptr(fk), _ptr(fp))
    return dk, dc, fk, fp


def gen_c4(join, n, m, seed=42):
    rk, rp, sk, sp = (np.empty(x, np.int32) for x in (n, n, m, m))
    lib().orc_gen_c4(seed, join, n, m, _ptr(rk), _ptr(rp), _ptr(sk), _ptr(sp))
    return rk, rp, sk, sp


def zipf_cdf(n, s):
    cdf = np.empty(n, np.float64)
    lib().orc_zipf_cdf(n, s, _ptr(cdf))
    return cdf


def c1_closed(nR, nS, seed=42, threads=8) -> CS64:
    r = CS64()
    lib().orc_c1_closed(seed, nR, nS, threads, C.byref(r))
    return r


def c2_closed(nR, nS, zipf_s=0.0, seed=42, sub=0, threads=8) -> CS64:
    r = CS64()
    lib().orc_c2_closed(seed, sub, nR, nS, zipf_s, threads, C.byref(r))
    return r


def c3_closed(nD, nF, norders=15_000_000, ncust=1_500_000, seed=42, threads=8):
    cents = np.zeros(ncust + 1, np.int64)
    joined = lib().orc_c3_closed(seed, nD, nF, norders, ncust, threads, _ptr(cents))
    return joined, cents


def c4_closed(join, n, nprobe, seed=42) -> CS64:
    r = CS64()
    lib().orc_c4_closed(seed, join, n, nprobe, C.byref(r))
    return r


def c4_closed_many(first, n, m, seed=42, threads=8):
    """Closed forms of joins first.. with build sizes n[i], probe sizes m[i]."""
    n = np.ascontiguousarray(n, dtype=np.uint64)
    m = np.ascontiguousarray(m, dtype=np.uint64)
    out = (CS64 * len(n))()
    lib().orc_c4_closed_many(seed, first, len(n), _ptr(n), _ptr(m), threads, out)
    return list(out)


def c5_closed(table, n, first, m, seed=42, threads=8) -> CS64:
    """Query over probe rows first.. of table `table`'s uniform FK column."""
    r = CS64()
    lib().orc_c5_closed(seed, table, n, first, m, threads, C.byref(r))
    return r


def baseline_join(kb, rk, rp, sk, buckets, tuplesperbucket, threads):
    r = CS64()
    secs = lib().orc_baseline_join(kb, _ptr(rk), _ptr(rp), len(rk), _ptr(sk), len(sk), buckets,
                                   tuplesperbucket, threads, C.byref(r))
    return secs, r
