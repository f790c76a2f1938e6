"""ORACLE -- test infrastructure only.

Python access to the CPU oracle (oracle/oracle.cpp, a restatement of the
reference's CPU engine handlers, core/engine.hpp:96-255,500-570) and to an
independent brute-force evaluator over the raw triples (oracle/brute.py).
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use
this package, and only as the checker (or the timed CPU baseline); the
product never imports it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "liboracle.so")

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        P = C.POINTER
        img = [C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint64, C.c_void_p, C.c_int]
        L.or_run_query.restype = C.c_int
        L.or_run_query.argtypes = img + [P(C.c_int64), C.c_int, C.c_int, P(P(C.c_int32)), P(C.c_uint64),
                                         P(C.c_int), P(C.c_uint64), P(C.c_double)]
        L.or_run_query_mt.restype = C.c_int
        L.or_run_query_mt.argtypes = img + [P(C.c_int64), C.c_int, C.c_int, C.c_int, P(P(C.c_int32)),
                                            P(C.c_uint64), P(C.c_int), P(C.c_uint64), P(C.c_double)]
        L.or_run_step.restype = C.c_int
        L.or_run_step.argtypes = img + [P(C.c_int64), C.c_int, C.c_int, P(C.c_int32), C.c_uint64, C.c_int,
                                        C.c_int, P(P(C.c_int32)), P(C.c_uint64), P(C.c_int), P(C.c_uint64)]
        L.or_run_many.restype = C.c_int
        L.or_run_many.argtypes = img + [P(C.c_int64), P(C.c_int64), C.c_int, C.c_int, P(C.c_double),
                                        P(C.c_uint64), P(C.c_uint64), P(C.c_uint64), P(C.c_double)]
        L.or_get_edges.restype = C.c_int
        L.or_get_edges.argtypes = img + [C.c_int64, C.c_int64, C.c_int64, P(C.c_uint64), P(C.c_uint64)]
        L.or_hash_u64.restype = C.c_uint64
        L.or_hash_u64.argtypes = [C.c_uint64]
        L.or_inverse_hash_u64.restype = C.c_uint64
        L.or_inverse_hash_u64.argtypes = [C.c_uint64]
        L.or_free.restype = None
        L.or_free.argtypes = [C.c_void_p]
        _lib = L
    return _lib


def _img(gs):
    s, ns, e, ne, m, nm = gs.oracle_args()
    return [C.cast(s, C.c_void_p), ns, C.cast(e, C.c_void_p), ne, m, nm]


def _take(ptr, rows, cols):
    if not ptr or rows == 0 or cols == 0:
        return np.zeros((rows, cols), dtype=np.int32)
    a = np.ctypeslib.as_array(ptr, shape=(rows * cols,)).reshape(rows, cols).copy()
    lib().or_free(C.cast(ptr, C.c_void_p))
    return a


class OracleError(RuntimeError):
    pass


def run_query(gs, cmd, stop_step=0, threads=1):
    """Runs a query (cmd_chains) on the CPU oracle: returns (table, expanded,
    seconds). threads > 1 runs each known-start step over row ranges in
    parallel (same result)."""
    L = lib()
    a = np.ascontiguousarray(cmd, dtype=np.int64)
    out = C.POINTER(C.c_int32)()
    rows, cols, ex, sec = C.c_uint64(), C.c_int(), C.c_uint64(), C.c_double()
    rc = L.or_run_query_mt(*_img(gs), a.ctypes.data_as(C.POINTER(C.c_int64)), len(a), stop_step, threads,
                           C.byref(out), C.byref(rows), C.byref(cols), C.byref(ex), C.byref(sec))
    if rc != 0:
        raise OracleError(f"oracle rejected the query ({rc})")
    return _take(out, rows.value, cols.value), ex.value, sec.value


def run_step(gs, cmd, step, table, threads=1):
    """One step of a query on a given input table: returns (table, expanded).
    threads > 1 runs a known-start step over row ranges in parallel (same
    result)."""
    L = lib()
    a = np.ascontiguousarray(cmd, dtype=np.int64)
    t = np.ascontiguousarray(table, dtype=np.int32)
    if t.ndim != 2:
        raise ValueError("table must be 2-D")
    out = C.POINTER(C.c_int32)()
    rows, cols, ex = C.c_uint64(), C.c_int(), C.c_uint64()
    rc = L.or_run_step(*_img(gs), a.ctypes.data_as(C.POINTER(C.c_int64)), len(a), step,
                       t.ctypes.data_as(C.POINTER(C.c_int32)), t.shape[0], t.shape[1], threads, C.byref(out),
                       C.byref(rows), C.byref(cols), C.byref(ex))
    if rc != 0:
        raise OracleError(f"oracle rejected the step ({rc})")
    return _take(out, rows.value, cols.value), ex.value


def run_many(gs, queries, threads=1, digests=False):
    """Runs queries to completion on `threads` threads: (latencies, rows,
    expanded, wall), plus each result table's table_digest when digests."""
    L = lib()
    off = np.zeros(len(queries) + 1, dtype=np.int64)
    for i, q in enumerate(queries):
        off[i + 1] = off[i] + len(q)
    flat = np.concatenate([np.asarray(q, dtype=np.int64) for q in queries]) if queries else np.zeros(0, np.int64)
    n = len(queries)
    lat = np.zeros(max(n, 1))
    rows = np.zeros(max(n, 1), dtype=np.uint64)
    ex = np.zeros(max(n, 1), dtype=np.uint64)
    dg = np.zeros(max(n, 1), dtype=np.uint64)
    wall = C.c_double()
    u64p = C.POINTER(C.c_uint64)
    rc = L.or_run_many(*_img(gs), flat.ctypes.data_as(C.POINTER(C.c_int64)),
                       off.ctypes.data_as(C.POINTER(C.c_int64)), n, threads,
                       lat.ctypes.data_as(C.POINTER(C.c_double)), rows.ctypes.data_as(u64p),
                       ex.ctypes.data_as(u64p), dg.ctypes.data_as(u64p) if digests else None, C.byref(wall))
    if rc != 0:
        raise OracleError(f"oracle rejected a query ({rc})")
    if digests:
        return lat[:n], rows[:n], ex[:n], wall.value, dg[:n]
    return lat[:n], rows[:n], ex[:n], wall.value


_GOLD = np.uint64(0x9E3779B97F4A7C15)


def _fmix64(x):
    x = x ^ (x >> np.uint64(33))
    x = x * np.uint64(0xff51afd7ed558ccd)
    x = x ^ (x >> np.uint64(33))
    x = x * np.uint64(0xc4ceb9fe1a85ec53)
    return x ^ (x >> np.uint64(33))


def table_digest(table, chunk=1 << 24):
    """Order-sensitive 64-bit digest of a row-major int32 table, the same
    function the oracle computes for its own tables (oracle.cpp
    table_digest): fmix64(rows * 256 + cols) + sum over cells i of
    fmix64((i + 1) * golden ^ cell_i), mod 2^64."""
    t = np.ascontiguousarray(table)
    rows, cols = (t.shape[0], t.shape[1]) if t.ndim == 2 else (0, 0)
    flat = t.reshape(-1).view(np.uint32)
    with np.errstate(over="ignore"):
        h = _fmix64(np.array([rows * 256 + cols], dtype=np.uint64))[0]
        for i0 in range(0, flat.size, chunk):
            c = flat[i0:i0 + chunk].astype(np.uint64)
            idx = np.arange(i0 + 1, i0 + 1 + c.size, dtype=np.uint64) * _GOLD
            h = h + np.sum(_fmix64(idx ^ c), dtype=np.uint64)
    return int(h)


def get_edges(gs, vid, d, pid):
    """Raw image lookup through the oracle's own bucket walk: edge list or None."""
    L = lib()
    off, size = C.c_uint64(), C.c_uint64()
    found = L.or_get_edges(*_img(gs), vid, d, pid, C.byref(off), C.byref(size))
    if not found:
        return None
    return gs.edges()[off.value:off.value + size.value].copy()


def hash_u64(x):
    return int(lib().or_hash_u64(x))


def inverse_hash_u64(x):
    return int(lib().or_inverse_hash_u64(x))
