"""Device context, result tables and the query engine (Python side).

Device / Table wrap libwkg_hip.so (include/wukong_gpu.h): one call per
RCache handler. GPUEngine wraps libwkg_engine.so (include/wukong_engine.h),
the C++ engine that keeps GPU_Engine::execute_one_step's dispatch
(core/gpu_engine.hpp:395-467).
"""
import ctypes as C

import numpy as np

from . import _native as N
from ._native import STEP_C2U, STEP_I2U, STEP_K2C, STEP_K2K, STEP_K2U


class DeviceError(RuntimeError):
    def __init__(self, rc, msg):
        super().__init__(f"status {rc}: {msg}")
        self.rc = rc


def _i64p(a):
    return a.ctypes.data_as(C.POINTER(C.c_int64))


class Table:
    """A device-resident row-major int32 result table."""

    def __init__(self, dev, handle):
        self.dev = dev
        self._h = C.c_void_p(handle)

    @property
    def shape(self):
        r, c = C.c_uint64(), C.c_int()
        N.hip_lib().wkg_table_shape(self._h, C.byref(r), C.byref(c))
        return int(r.value), int(c.value)

    def to_host(self):
        rows, cols = self.shape
        out = np.empty((rows, cols), dtype=np.int32)
        if rows and cols:
            self.dev._check(N.hip_lib().wkg_table_to_host(self.dev._h, self._h,
                                                          out.ctypes.data_as(C.POINTER(C.c_int32)), None))
        return out

    def free(self):
        if self._h and self.dev._h:
            N.hip_lib().wkg_table_free(self.dev._h, self._h)
        self._h = None

    __del__ = free


class Device:
    """A wkg_ctx: one GPU, one loaded gstore image, a table HBM budget."""

    def __init__(self, device=0, table_budget_bytes=0):
        L = N.hip_lib()
        h = C.c_void_p()
        rc = L.wkg_init(device, table_budget_bytes, C.byref(h))
        if rc != 0:
            raise DeviceError(rc, "wkg_init failed (no usable GPU?)")
        self._h = h
        self._image = None

    def _check(self, rc):
        if rc != 0:
            raise DeviceError(rc, N.hip_lib().wkg_last_error(self._h).decode())

    def close(self):
        if getattr(self, "_h", None):
            N.hip_lib().wkg_destroy(self._h)
            self._h = None

    __del__ = close

    def load(self, gstore, graph_budget_bytes=0, sorted_lists=True):
        """Uploads a GStore image (whole-image residency unless the budget is smaller)."""
        L = N.hip_lib()
        self._image = gstore  # the framed mode reads the host image later
        G = N.gstore_lib()
        self._check(L.wkg_load_graph(self._h, G.wks_slots(gstore._h), gstore.num_slots,
                                     G.wks_edges(gstore._h), gstore.num_edges,
                                     gstore.metas_ptr(), gstore.num_metas, graph_budget_bytes,
                                     1 if sorted_lists else 0))

    def build(self, triples, main_load=0.75):
        """Builds the gstore image on the device from ID triples (the image
        GStore.build makes, slot for slot) and keeps it resident. Returns the
        build time in seconds, upload included."""
        s, p, o = triples.arrays()
        s, p, o = (np.ascontiguousarray(x, dtype=np.uint32) for x in (s, p, o))
        t = C.c_double()
        u32p = C.POINTER(C.c_uint32)
        self._check(N.hip_lib().wkg_build_graph(self._h, s.ctypes.data_as(u32p), p.ctypes.data_as(u32p),
                                                 o.ctypes.data_as(u32p), len(s), float(main_load), C.byref(t)))
        return float(t.value)

    def image(self):
        """The resident image as host arrays: (slots (n, 2) u64, edges u32, metas (m, 8) u64)."""
        L = N.hip_lib()
        ns, ne, nm = C.c_uint64(), C.c_uint64(), C.c_int()
        self._check(L.wkg_graph_shape(self._h, C.byref(ns), C.byref(ne), C.byref(nm)))
        slots = np.empty((ns.value, 2), dtype=np.uint64)
        edges = np.empty(ne.value, dtype=np.uint32)
        metas = np.empty((nm.value, 8), dtype=np.uint64)
        self._check(L.wkg_graph_download(self._h, slots.ctypes.data_as(C.c_void_p), edges.ctypes.data_as(C.c_void_p),
                                         metas.ctypes.data_as(C.c_void_p)))
        return slots, edges, metas

    def profile(self, on=True):
        """Turns per-kernel HIP-event timing on (clearing totals) or off."""
        self._check(N.hip_lib().wkg_profile_enable(self._h, 1 if on else 0))

    def profile_read(self, kernel):
        """(launches, total_ms, algorithmic_bytes) for a kernel name."""
        n, ms, b = C.c_uint64(), C.c_double(), C.c_double()
        self._check(N.hip_lib().wkg_profile_read(self._h, kernel.encode(), C.byref(n), C.byref(ms), C.byref(b)))
        return int(n.value), float(ms.value), float(b.value)

    def copy_bandwidth(self, nbytes=1 << 30, iters=20):
        """GB/s (read + write) of a plain device copy kernel."""
        g = C.c_double()
        self._check(N.hip_lib().wkg_bandwidth_probe(self._h, nbytes, iters, C.byref(g)))
        return float(g.value)

    def set_table_limit(self, nbytes):
        """Largest step output table; bigger ones are produced in row chunks (0: none)."""
        self._check(N.hip_lib().wkg_set_table_limit(self._h, nbytes))

    def set_probe_tile_rows(self, rows):
        """Rows per probe tile of the per-step path: 256, 512, 1024, 2048, or 0 (by input size)."""
        self._check(N.hip_lib().wkg_set_probe_tile_rows(self._h, int(rows)))

    def pool_stats(self):
        """Table pool: live and cached bytes, hipMalloc calls, cache drops."""
        v = [C.c_uint64() for _ in range(4)]
        self._check(N.hip_lib().wkg_pool_stats(self._h, *[C.byref(x) for x in v]))
        return {k: int(x.value) for k, x in zip(("live_bytes", "cached_bytes", "mallocs", "trims"), v)}

    def set_launch_rows(self, rows):
        """Input rows per probe/compaction launch of the calling thread's steps (0: the default, 2^26)."""
        self._check(N.hip_lib().wkg_thread_launch_rows(self._h, int(rows)))

    def set_small_steps(self, on):
        """Known-start steps over at most 512 rows in one launch (on by default)."""
        self._check(N.hip_lib().wkg_set_small_steps(self._h, 1 if on else 0))

    def set_lineage(self, on):
        """Probe a star's next arm once per input row of the fan-out before it (default on)."""
        self._check(N.hip_lib().wkg_set_lineage(self._h, 1 if on else 0))

    def thread_context_begin(self):
        """The calling thread's per-step calls get a stream and read-back block of their own."""
        self._check(N.hip_lib().wkg_thread_context_begin(self._h))

    def thread_context_end(self):
        self._check(N.hip_lib().wkg_thread_context_end(self._h))

    def set_fanout_blocks(self, on, kept_cap=0):
        """Long shared lists: 64 parents per workgroup with a kept-row list of
        kept_cap rows (0: 2^22); off: one parent per workgroup."""
        self._check(N.hip_lib().wkg_set_fanout_blocks(self._h, 1 if on else 0, int(kept_cap)))

    def set_group_filter_rows(self, rows):
        """Smallest table a filter from a star arm's new column decides once per
        lineage group (default 2^22 rows; 0: any size)."""
        self._check(N.hip_lib().wkg_set_group_filter_rows(self._h, int(rows)))

    def random_access_rate(self, nbytes=4 << 30, iters=10):
        """Random 128-byte line reads per second (the hash probe's roofline)."""
        g = C.c_double()
        self._check(N.hip_lib().wkg_random_access_probe(self._h, nbytes, iters, C.byref(g)))
        return float(g.value)

    def stats(self):
        a, b, c, d = C.c_uint64(), C.c_uint64(), C.c_uint64(), C.c_uint64()
        self._check(N.hip_lib().wkg_graph_stats(self._h, C.byref(a), C.byref(b), C.byref(c)))
        self._check(N.hip_lib().wkg_graph_prefetches(self._h, C.byref(d)))
        return {"resident_bytes": a.value, "segment_loads": b.value, "segment_evictions": c.value,
                "prefetches": d.value}

    def ensure(self, pid):
        """Makes predicate pid's segment resident (framed graphs)."""
        self._check(N.hip_lib().wkg_ensure_predicate(self._h, int(pid)))

    def resident(self, pid):
        """True when predicate pid's whole segment is in frames (always, for a resident image)."""
        rc = N.hip_lib().wkg_predicate_resident(self._h, int(pid))
        if rc < 0:
            self._check(rc)
        return rc == 1

    def prefetch(self, pid, in_use=-1):
        """Starts loading pid's segment on the copy stream (framed graphs)."""
        self._check(N.hip_lib().wkg_prefetch_predicate(self._h, int(pid), int(in_use)))

    def preload(self, pids=None):
        """Loads the listed predicates' segments now (None: every predicate id from 1 up)."""
        if pids is None:
            self._check(N.hip_lib().wkg_preload_predicates(self._h, None, -1))
        else:
            a = (C.c_int32 * len(pids))(*[int(p) for p in pids])
            self._check(N.hip_lib().wkg_preload_predicates(self._h, a, len(pids)))

    def table(self, rows):
        """Host rows (2-D int array) -> device Table."""
        a = np.ascontiguousarray(rows, dtype=np.int32)
        if a.ndim != 2:
            raise ValueError("table must be 2-D")
        h = C.c_void_p()
        self._check(N.hip_lib().wkg_table_from_host(self._h, a.ctypes.data_as(C.POINTER(C.c_int32)),
                                                     a.shape[0], a.shape[1], None, C.byref(h)))
        return Table(self, h.value)

    def wrap_device(self, t, rows=None, cols=None):
        """A Table viewing the device cells of table t (or its first rows x
        cols cells) without owning them (wkg_table_wrap_device)."""
        L = N.hip_lib()
        r0, c0 = t.shape
        h = C.c_void_p()
        self._check(L.wkg_table_wrap_device(self._h, L.wkg_table_device_ptr(t._h), r0 if rows is None else rows,
                                            c0 if cols is None else cols, C.byref(h)))
        v = Table(self, h.value)
        v._parent = t  # the view's memory is t's
        return v

    def table_split(self, t, col, nparts):
        """Fork-join split of a device table by the owner (cell % nparts) of
        column `col`: (Table holding the sub-tables back to back, rows per
        sub-table, first row of each)."""
        h = C.c_void_p()
        rows = np.zeros(max(nparts, 1), dtype=np.uint64)
        head = np.zeros(max(nparts, 1), dtype=np.uint64)
        u64p = C.POINTER(C.c_uint64)
        self._check(N.hip_lib().wkg_table_split(self._h, t._h, int(col), int(nparts), None, C.byref(h),
                                                 rows.ctypes.data_as(u64p), head.ctypes.data_as(u64p)))
        return Table(self, h.value), rows[:nparts], head[:nparts]

    def table_concat(self, tables):
        """Sub-replies merged: the tables' rows appended in order."""
        arr = (C.c_void_p * max(len(tables), 1))(*[t._h.value for t in tables])
        h = C.c_void_p()
        self._check(N.hip_lib().wkg_table_concat(self._h, len(tables), arr, None, C.byref(h)))
        return Table(self, h.value)

    def table_slice(self, t, row0, nrows):
        """View of rows [row0, row0 + nrows) of t (free it before t)."""
        h = C.c_void_p()
        self._check(N.hip_lib().wkg_table_slice(self._h, t._h, int(row0), int(nrows), C.byref(h)))
        v = Table(self, h.value)
        v._parent = t  # the view's memory is t's
        return v

    def _step(self, fn, *args):
        h = C.c_void_p()
        ex = C.c_uint64()
        self._check(fn(self._h, *args, None, C.byref(h), C.byref(ex)))
        return Table(self, h.value), int(ex.value)

    # one call per RCache handler; each returns (Table, expanded)
    def i2u(self, idx, d):
        return self._step(N.hip_lib().wkg_step_i2u, idx, d)

    def c2u(self, vid, pid, d):
        return self._step(N.hip_lib().wkg_step_c2u, vid, pid, d)

    def k2u(self, t, start_col, pid, d):
        return self._step(N.hip_lib().wkg_step_k2u, t._h, start_col, pid, d)

    def k2u_noalloc(self, t, start_col, pid, d, out):
        """known_to_unknown into the caller's buffer `out` (a Table of at
        least the output's rows, cols + 1 columns) with the library's fixed
        scratch (wkg_step_k2u_noalloc): returns (rows written, expanded)."""
        rows, cols = t.shape
        cap, ocols = out.shape
        if ocols != cols + 1:
            raise ValueError("the output buffer needs cols + 1 columns")
        L = N.hip_lib()
        n, ex = C.c_uint64(), C.c_uint64()
        self._check(L.wkg_step_k2u_noalloc(self._h, L.wkg_table_device_ptr(t._h), rows, cols, start_col, pid, d,
                                           L.wkg_table_device_ptr(out._h), cap, None, C.byref(n), C.byref(ex)))
        return int(n.value), int(ex.value)

    # ---- small queries (wkg_query_small*) ----
    def set_small_query_server(self, workgroups):
        """Persistent small-query server workgroups (0: one launch per query)."""
        self._check(N.hip_lib().wkg_set_small_query_server(self._h, workgroups))

    def set_small_query_server_times(self, idle_s, life_s):
        """A server workgroup's idle exit and life (seconds)."""
        self._check(N.hip_lib().wkg_set_small_query_server_times(self._h, float(idle_s), float(life_s)))

    @staticmethod
    def _plan_array(plan):
        arr = (N.PlanStep * max(len(plan), 1))()
        for i, s in enumerate(plan):
            arr[i].kind, arr[i].pid, arr[i].dir = s["kind"], s["pid"], s["dir"]
            arr[i].start_col, arr[i].end_col = s.get("start_col", 0), s.get("end_col", 0)
            arr[i].vid = s.get("vid", 0)
        return arr

    _SQ_CELLS = 8192

    def _sq_result(self, rc, buf, rows, cols, ex):
        if rc == -4:  # WKG_ERANGE: declined, run it step by step
            return "declined", None, int(ex.value)
        self._check(rc)
        n = rows.value * cols.value
        return "ok", np.ctypeslib.as_array(buf)[:n].reshape(rows.value, cols.value).copy(), int(ex.value)

    def query_small(self, plan):
        """One small query (a plan of step dicts): ("ok", table, expanded) or
        ("declined", None, expanded so far)."""
        buf = (C.c_int32 * self._SQ_CELLS)()
        rows, cols, ex = C.c_uint64(), C.c_int(), C.c_uint64()
        rc = N.hip_lib().wkg_query_small(self._h, self._plan_array(plan), len(plan), None, buf, self._SQ_CELLS,
                                         C.byref(rows), C.byref(cols), C.byref(ex))
        return self._sq_result(rc, buf, rows, cols, ex)

    def query_small_submit(self, plan):
        """Submits a small query to the server; returns its ticket, or None
        when every slot is in flight."""
        t = C.c_uint64()
        rc = N.hip_lib().wkg_query_small_submit(self._h, self._plan_array(plan), len(plan), C.byref(t))
        if rc == -6:  # WKG_EAGAIN
            return None
        self._check(rc)
        return int(t.value)

    def query_small_collect(self, ticket, wait=True):
        """("ok", table, expanded), ("declined", None, expanded) or, with
        wait False, ("pending", None, 0) while the ticket runs."""
        buf = (C.c_int32 * self._SQ_CELLS)()
        rows, cols, ex = C.c_uint64(), C.c_int(), C.c_uint64()
        rc = N.hip_lib().wkg_query_small_collect(self._h, ticket, 1 if wait else 0, buf, self._SQ_CELLS,
                                                 C.byref(rows), C.byref(cols), C.byref(ex))
        if rc == -6:
            return "pending", None, 0
        return self._sq_result(rc, buf, rows, cols, ex)

    def k2u_filter(self, t, start_col, pid, d, filter_kind, filter_pid, filter_dir, end_col=-1, const=-1):
        """known_to_unknown then a filter on its new column (wkg_step_k2u_filter):
        (Table, fused, expanded of the k2u, expanded of the filter)."""
        out, fused, e1, e2 = C.c_void_p(), C.c_int(), C.c_uint64(), C.c_uint64()
        self._check(N.hip_lib().wkg_step_k2u_filter(self._h, t._h, start_col, pid, d, filter_kind, filter_pid,
                                                     filter_dir, end_col, const, None, C.byref(out),
                                                     C.byref(fused), C.byref(e1), C.byref(e2)))
        return Table(self, out.value), bool(fused.value), int(e1.value), int(e2.value)

    def k2k(self, t, start_col, pid, d, end_col):
        return self._step(N.hip_lib().wkg_step_k2k, t._h, start_col, pid, d, end_col)

    def k2c(self, t, start_col, pid, d, const):
        return self._step(N.hip_lib().wkg_step_k2c, t._h, start_col, pid, d, const)

    def step_batch(self, steps):
        """steps: list of dicts {kind, pid, dir, start_col, end_col, vid, input(Table)}."""
        n = len(steps)
        arr = (N.StepDesc * n)()
        for i, s in enumerate(steps):
            arr[i].kind = s["kind"]
            arr[i].pid = s["pid"]
            arr[i].dir = s["dir"]
            arr[i].start_col = s.get("start_col", 0)
            arr[i].end_col = s.get("end_col", 0)
            arr[i].vid = s.get("vid", 0)
            t = s.get("input")
            arr[i].input = t._h.value if t is not None else None
        outs = (C.c_void_p * n)()
        ex = (C.c_uint64 * n)()
        self._check(N.hip_lib().wkg_step_batch(self._h, n, arr, None, outs, ex))
        return [Table(self, outs[i]) for i in range(n)], [int(x) for x in ex]


def plan_of(cmd):
    """cmd_chain -> plan step dicts {kind, pid, dir, start_col, end_col, vid}:
    GPU_Engine::execute_one_step's dispatch (core/gpu_engine.hpp:395-467)
    with the table's column bookkeeping (a known variable -k is column k-1)."""
    steps, cols = [], 0
    for i in range(0, len(cmd), 4):
        s, p, d, e = (int(x) for x in cmd[i:i + 4])
        if i == 0 and 1 < s < (1 << 17):
            steps.append(dict(kind=STEP_I2U, pid=s, dir=d))
            cols = 1
        elif s >= 0:
            steps.append(dict(kind=STEP_C2U, pid=p, dir=d, vid=s))
            cols = 1
        elif e >= 0:
            steps.append(dict(kind=STEP_K2C, pid=p, dir=d, start_col=-s - 1, vid=e))
        elif e < 0 and -e <= cols:
            steps.append(dict(kind=STEP_K2K, pid=p, dir=d, start_col=-s - 1, end_col=-e - 1))
        else:
            steps.append(dict(kind=STEP_K2U, pid=p, dir=d, start_col=-s - 1))
            cols += 1
    return steps


def pack_queries(queries):
    """List of cmd_chains (sequences of int) -> (flat int64 array, int64 offsets).
    A (flat, offsets) pair of arrays passes through."""
    if isinstance(queries, tuple) and len(queries) == 2 and isinstance(queries[0], np.ndarray):
        return np.ascontiguousarray(queries[0], dtype=np.int64), np.ascontiguousarray(queries[1], dtype=np.int64)
    off = np.zeros(len(queries) + 1, dtype=np.int64)
    for i, q in enumerate(queries):
        off[i + 1] = off[i] + len(q)
    flat = np.zeros(int(off[-1]), dtype=np.int64)
    for i, q in enumerate(queries):
        flat[off[i]:off[i + 1]] = q
    return flat, off


class QueryResult:
    __slots__ = ("table", "status", "expanded", "latency_s", "rows")

    def __init__(self, table, status, expanded, latency_s, rows=None):
        self.table, self.status, self.expanded, self.latency_s = table, status, expanded, latency_s
        self.rows = table.shape[0] if rows is None else rows


class GPUEngine:
    """The C++ engine: plain, batched and served query execution."""

    def __init__(self, dev):
        L = N.engine_lib()
        h = C.c_void_p()
        rc = L.wke_create(dev._h, C.byref(h))
        if rc != 0:
            raise DeviceError(rc, "wke_create failed")
        self.dev = dev
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            N.engine_lib().wke_destroy(self._h)
            self._h = None

    __del__ = close

    def _collect(self, res, n):
        L = N.engine_lib()
        out = []
        for i in range(n):
            r = res[i]
            if r.table and r.rows and r.cols:
                t = np.ctypeslib.as_array(r.table, shape=(r.rows * r.cols,)).reshape(r.rows, r.cols).copy()
                L.wke_free(C.cast(r.table, C.c_void_p))
            else:
                t = np.zeros((r.rows, max(r.cols, 0)), dtype=np.int32)
            out.append(QueryResult(t, r.status, r.expanded, r.latency_s))
        return out

    def run_query(self, cmd):
        L = N.engine_lib()
        a = np.ascontiguousarray(cmd, dtype=np.int64)
        res = N.Result()
        rc = L.wke_run_query(self._h, _i64p(a), len(a), C.byref(res))
        if rc != 0:
            raise DeviceError(rc, L.wke_last_error(self._h).decode())
        return self._collect([res], 1)[0]

    def _many(self, fn, queries):
        flat, off = pack_queries(queries)
        n = len(queries)
        res = (N.Result * max(n, 1))()
        fn(self._h, _i64p(flat), _i64p(off), n, res)
        return self._collect(res, n)

    def run_plain(self, queries):
        return self._many(N.engine_lib().wke_run_plain, queries)

    def run_batched(self, queries):
        """Device-resident executor: whole queries on the device, one captured launch sequence."""
        return self._many(N.engine_lib().wke_run_batched, queries)

    def run_batched_packed(self, queries):
        """run_batched with all tables returned in one buffer. Returns (results,
        seconds spent in the native call); the tables are views of one array."""
        import time
        L = N.engine_lib()
        flat, off = pack_queries(queries)
        n = len(queries)
        res = (N.Result * max(n, 1))()
        coff = np.zeros(max(n, 1), dtype=np.uint64)
        cells = C.POINTER(C.c_int32)()
        ncells = C.c_uint64()
        t0 = time.perf_counter()
        L.wke_run_batched_packed(self._h, _i64p(flat), _i64p(off), n, res,
                                 coff.ctypes.data_as(C.POINTER(C.c_uint64)), C.byref(cells), C.byref(ncells))
        wall = time.perf_counter() - t0
        buf = (np.ctypeslib.as_array(cells, shape=(ncells.value,)).copy() if ncells.value
               else np.zeros(0, np.int32))
        if cells:
            L.wke_free(C.cast(cells, C.c_void_p))
        out = []
        for i in range(n):
            r = res[i]
            c0 = int(coff[i])
            t = buf[c0:c0 + r.rows * r.cols].reshape(r.rows, max(r.cols, 0))
            out.append(QueryResult(t, r.status, r.expanded, r.latency_s))
        return out, wall

    def run_rounds(self, queries):
        """Step-synchronous batching: one fused launch set per step round."""
        return self._many(N.engine_lib().wke_run_rounds, queries)

    def set_small_queries(self, on):
        """Try a whole small query in one launch before the step-by-step path (default on)."""
        N.engine_lib().wke_set_small_queries(self._h, 1 if on else 0)

    def splits(self):
        """Row-half splits made so far by the plain path's HBM-budgeted chunking."""
        n = C.c_uint64()
        N.engine_lib().wke_stats(self._h, C.byref(n))
        return int(n.value)

    def configure(self, max_queries=1024, max_patterns=8, arena_bytes=16 << 30):
        L = N.engine_lib()
        rc = L.wke_configure(self._h, max_queries, max_patterns, arena_bytes)
        if rc != 0:
            raise DeviceError(rc, "wke_configure failed")

    def set_heavy_mode(self, mode):
        """Heavy served queries: -1 auto, 0 a heavy executor lane, 1 the per-step worker thread."""
        L = N.engine_lib()
        rc = L.wke_set_heavy_mode(self._h, int(mode))
        if rc != 0:
            raise DeviceError(rc, "wke_set_heavy_mode failed")

    def set_heavy_workers(self, n):
        """Served heavy queries at once (worker threads, each with its own stream)."""
        rc = N.engine_lib().wke_set_heavy_workers(self._h, int(n))
        if rc != 0:
            raise DeviceError(rc, "wke_set_heavy_workers failed")

    def set_fused_filters(self, on):
        """known_to_unknown + a filter on its new column in one call (default on)."""
        rc = N.engine_lib().wke_set_fused_filters(self._h, 1 if on else 0)
        if rc != 0:
            raise DeviceError(rc, "wke_set_fused_filters failed")

    def set_pipeline(self, on):
        """Framed graphs: load a later pattern's predicate while a step runs (default on)."""
        rc = N.engine_lib().wke_set_pipeline(self._h, 1 if on else 0)
        if rc != 0:
            raise DeviceError(rc, "wke_set_pipeline failed")

    def set_serve_small(self, workgroups):
        """Small-query server workgroups for served light queries (0: light lanes only)."""
        rc = N.engine_lib().wke_set_serve_small(self._h, int(workgroups))
        if rc != 0:
            raise DeviceError(rc, "wke_set_serve_small failed")

    def serve_stats(self):
        """Counters of the last serve(): fallbacks, solo retries, worker queries, worker backlog,
        queries the small-query server answered and declined."""
        L = N.engine_lib()
        v = [C.c_uint64() for _ in range(6)]
        L.wke_serve_stats(self._h, *[C.byref(x) for x in v[:4]])
        L.wke_serve_small_stats(self._h, C.byref(v[4]), C.byref(v[5]))
        return {k: int(x.value) for k, x in zip(("plain_fallbacks", "solo_retries", "worker_queries",
                                                  "worker_backlog", "small_answered", "small_declined"), v)}

    def set_tuning(self, heavy=False, **shape):
        """Pins the executor's launch and tile shape (the keys of autotune's result)."""
        L = N.engine_lib()
        t = N.BatchTuning(**shape)
        rc = L.wke_set_tuning(self._h, C.byref(t), 1 if heavy else 0)
        if rc != 0:
            raise DeviceError(rc, L.wke_last_error(self._h).decode())

    def autotune(self, queries, reps=3, heavy=False):
        """Autotunes the executor's launch and tile shapes on a sample batch (for
        the heavy serving lane with heavy=True); returns (shape, ms)."""
        L = N.engine_lib()
        flat, off = pack_queries(queries)
        t = N.BatchTuning()
        ms = C.c_double()
        rc = L.wke_autotune(self._h, _i64p(flat), _i64p(off), len(queries), reps, 1 if heavy else 0,
                            C.byref(t), C.byref(ms))
        if rc != 0:
            raise DeviceError(rc, L.wke_last_error(self._h).decode())
        return t.as_dict(), float(ms.value)

    def serve(self, queries, arrival_s, heavy=None, table_budget_bytes=0, light_lanes=2):
        L = N.engine_lib()
        flat, off = pack_queries(queries)
        n = len(queries)
        arr = np.ascontiguousarray(arrival_s, dtype=np.float64)
        hv = np.ascontiguousarray(heavy if heavy is not None else np.zeros(n), dtype=np.uint8)
        res = (N.Result * max(n, 1))()
        wall = C.c_double()
        L.wke_serve(self._h, _i64p(flat), _i64p(off), n, arr.ctypes.data_as(C.POINTER(C.c_double)),
                    hv.ctypes.data_as(C.POINTER(C.c_uint8)), table_budget_bytes, light_lanes, res,
                    C.byref(wall))
        return self._collect(res, n), wall.value

    def serve_closed(self, queries, heavy, light_inflight, heavy_inflight, duration_s, table_budget_bytes=0,
                     light_lanes=2, keep=None, interval_s=0.0):
        """Closed-loop batch mode: (ClosedResults, issue_s, wall); issue_s[q] <
        0 for queries never issued. queries: a list of cmd_chains or a (flat,
        offsets) pair; keep: optional mask of the result tables to return (the
        others are blind: row counts only); interval_s: least time between
        two issue rounds."""
        L = N.engine_lib()
        flat, off = pack_queries(queries)
        n = len(off) - 1
        hv = np.ascontiguousarray(heavy, dtype=np.uint8)
        kp = None if keep is None else np.ascontiguousarray(keep, dtype=np.uint8)
        res = (N.Result * max(n, 1))()
        issue = np.zeros(max(n, 1), dtype=np.float64)
        wall = C.c_double()
        rc = L.wke_serve_closed(self._h, _i64p(flat), _i64p(off), n, hv.ctypes.data_as(C.POINTER(C.c_uint8)),
                                int(light_inflight), int(heavy_inflight), float(duration_s), float(interval_s),
                                table_budget_bytes,
                                light_lanes, None if kp is None else kp.ctypes.data_as(C.POINTER(C.c_uint8)), res,
                                issue.ctypes.data_as(C.POINTER(C.c_double)), C.byref(wall))
        if rc != 0:
            raise DeviceError(rc, L.wke_last_error(self._h).decode())
        return ClosedResults(res, n), issue[:n], wall.value


# Layout of wke_result (include/wukong_engine.h) as a numpy record.
_RESULT_DTYPE = np.dtype([("table", "<u8"), ("rows", "<u8"), ("cols", "<i4"), ("status", "<i4"),
                          ("expanded", "<u8"), ("latency_s", "<f8")])


class ClosedResults:
    """Results of a closed-loop run as per-query arrays (rows, cols, status,
    expanded, latency_s) and the returned tables by query index; indexing
    gives a QueryResult (table None for a blind query)."""

    def __init__(self, res, n):
        assert C.sizeof(N.Result) == _RESULT_DTYPE.itemsize
        rec = np.frombuffer((C.c_char * (n * _RESULT_DTYPE.itemsize)).from_address(C.addressof(res)),
                            dtype=_RESULT_DTYPE) if n else np.zeros(0, dtype=_RESULT_DTYPE)
        self.rows, self.cols = rec["rows"].copy(), rec["cols"].copy()
        self.status, self.expanded = rec["status"].copy(), rec["expanded"].copy()
        self.latency_s = rec["latency_s"].copy()
        self.tables = {}
        L = N.engine_lib()
        for q in np.nonzero(rec["table"])[0]:
            r = res[int(q)]
            t = np.ctypeslib.as_array(r.table, shape=(r.rows * r.cols,)).reshape(r.rows, r.cols).copy()
            L.wke_free(C.cast(r.table, C.c_void_p))
            self.tables[int(q)] = t

    def __len__(self):
        return len(self.rows)

    def __getitem__(self, q):
        t = self.tables.get(q)
        if t is None and self.rows[q] == 0:
            t = np.zeros((0, max(int(self.cols[q]), 0)), dtype=np.int32)
        return QueryResult(t, int(self.status[q]), int(self.expanded[q]), float(self.latency_s[q]),
                           int(self.rows[q]))

    def __iter__(self):
        return (self[q] for q in range(len(self)))


class Batch:
    """Direct handle on one device-resident executor (wkg_batch_*)."""

    def __init__(self, dev, max_queries, max_patterns, arena_bytes):
        self.dev = dev
        h = C.c_void_p()
        dev._check(N.hip_lib().wkg_batch_create(dev._h, max_queries, max_patterns, arena_bytes, None,
                                                 C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None) and self.dev._h:
            N.hip_lib().wkg_batch_destroy(self._h)
        self._h = None

    __del__ = close

    @staticmethod
    def _plans(plans):
        steps = [s for p in plans for s in p]
        arr = (N.PlanStep * max(len(steps), 1))()
        for i, s in enumerate(steps):
            arr[i].kind, arr[i].pid, arr[i].dir = s["kind"], s["pid"], s["dir"]
            arr[i].start_col, arr[i].end_col = s.get("start_col", 0), s.get("end_col", 0)
            arr[i].vid = s.get("vid", 0)
        off = (C.c_int32 * (len(plans) + 1))()
        for i, p in enumerate(plans):
            off[i + 1] = off[i] + len(p)
        return arr, off

    def set_tuning(self, **kw):
        t = N.BatchTuning(**kw)
        self.dev._check(N.hip_lib().wkg_batch_set_tuning(self._h, C.byref(t)))

    def reserve_results(self, nbytes):
        self.dev._check(N.hip_lib().wkg_batch_reserve_results(self._h, nbytes))

    def stats(self):
        """(arena high-water mark, result bytes) of the last batch."""
        p, r = C.c_uint64(), C.c_uint64()
        self.dev._check(N.hip_lib().wkg_batch_stats(self._h, C.byref(p), C.byref(r)))
        return int(p.value), int(r.value)

    def tuning(self):
        t = N.BatchTuning()
        self.dev._check(N.hip_lib().wkg_batch_get_tuning(self._h, C.byref(t)))
        return t.as_dict()

    def run(self, plans):
        """plans: list of lists of step dicts {kind, pid, dir, start_col, end_col, vid}.
        Returns a list of (status, table ndarray, expanded)."""
        arr, off = self._plans(plans)
        n = len(plans)
        res = (N.BatchResult * max(n, 1))()
        cells = C.POINTER(C.c_int32)()
        ncells = C.c_uint64()
        self.dev._check(N.hip_lib().wkg_batch_run(self._h, n, arr, off, res, C.byref(cells), C.byref(ncells)))
        # the cells live in the batch's pinned result buffer until the next launch
        flat = (np.ctypeslib.as_array(cells, shape=(ncells.value,)).copy() if ncells.value
                else np.zeros(0, np.int32))
        out = []
        for i in range(n):
            r = res[i]
            t = flat[r.offset:r.offset + r.rows * r.cols].reshape(r.rows, max(r.cols, 0)) if r.status == 0 \
                else None
            out.append((int(r.status), t, int(r.expanded)))
        return out
