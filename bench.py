#!/usr/bin/env python3
"""GPU graph-exploration benchmark on one MI355X.

The N=1 line is the largest single-GPU configuration: the 6-pattern cyclic
query (a triangle closed by known_to_known, then a known_to_const filter) on
the 2B-triple graph (200M entities, 512 predicates, Zipf 0.9, 2M-edge hubs;
an 89.8 GB image resident in HBM), on the plain GPU path with its ~1B-row
hub fan-out produced and consumed in HBM-budgeted 4 GB chunks. One "step" is
one such query end to end. `value` = expanded edges/s. Beside it: the same
query through the device-resident executor (batched with the 100k-member
sample cycle), both answers checked row for row against the CPU oracle, and
the oracle on one host thread as the reported CPU baseline.

The other configurations ride along as extra keys, each checked against the
oracle (the same keys `--config N` prints as a line of its own):
  config2_3  the heavy 6-pattern query on the 100M-triple graph (plain and
             executor), its 151M-row fan-out table, the 512-query batched mix
             (plain, step rounds, autotuned executor; every table checked),
             the served Poisson sweep, the closed-loop batch mode and the
             fork-join table split;
  config4    the served Poisson sweep and batch mode on the 1B-triple graph;
  config1    the 2-hop query from a constant on the 10M-triple graph.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--config all|1|2|4|5]
Progress goes to stderr; one JSON line goes to stdout.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "expanded edges/s + queries/s; p50/p99 latency, batched vs plain on 1 MI355X"
HBM_PEAK_GBS = 8000.0
ORACLE_THREADS = 16   # the GPU host's CPU share
C5_SAMPLE = 100_000   # members of the config-5 sample type (the bounded CPU baseline)
C5_LIMIT = 4 << 30    # config 5's output-table limit: the hub fan-out in 4 GB chunks
T_START = time.time()


def log(*a):
    print(f"[{time.time() - T_START:6.1f}s]", *a, file=sys.stderr, flush=True)


def pct(vals, q):
    """Percentile by the reference logger's index rule: sorted[int(q * n)]
    (core/logger.hpp:225-238)."""
    s = sorted(vals)
    if not s:
        return 0.0
    return s[min(int(q * len(s)), len(s) - 1)]


def _cpu_info():
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu_model": model, "nproc": os.cpu_count()}


def _same(xs, ys):
    import numpy as np
    return all(a.status == 0 and b.status == 0 and a.table.shape == b.table.shape and
               np.array_equal(a.table, b.table) and a.expanded == b.expanded for a, b in zip(xs, ys))


def _eq(res, want, wex):
    """A GPU result equals an oracle table, row for row, and its work count."""
    import numpy as np
    return bool(res.status == 0 and res.table.shape[0] == want.shape[0] and
                (want.size == 0 or np.array_equal(res.table.reshape(want.shape), want)) and res.expanded == wex)


def _oracle_digests(gs, queries, results):
    """Every result against the oracle's table for the same query: row count,
    work count and an order-sensitive digest of the cells. Returns (all
    equal, queries checked, oracle wall seconds)."""
    import oracle
    _, rows, ex, wall, dg = oracle.run_many(gs, queries, threads=ORACLE_THREADS, digests=True)
    ok = all(r.status == 0 and r.table.shape[0] == rows[i] and r.expanded == ex[i] and
             oracle.table_digest(r.table) == int(dg[i]) for i, r in enumerate(results))
    return bool(ok), len(queries), wall


def _kernel_table(dev, names, n):
    out = {}
    for k in names:
        v = dev.profile_read(k)
        if v[0]:
            out[k] = {"launches": v[0] / n, "ms": v[1] / n, "alg_GBps": (v[2] / (v[1] * 1e-3) / 1e9) if v[1] > 0 else 0.0}
    return out


STEP_KERNELS = ("probe", "probe_filter", "fanout_filter", "scan_tiles", "finalize", "compact", "partition", "expand",
                "probe_one", "copy_list", "small_step", "lineage")


def _expand_roofline(dev, n_queries, traffic_file, extra=None):
    """HBM roofline of k_expand from HIP events over the profiled queries:
    algorithmic bytes per launch / average launch time; traffic (PMC bytes
    per launch) from the committed rocprofv3 summary when there is one."""
    n_exp, ms_exp, b_exp = dev.profile_read("expand")
    achieved = b_exp / (ms_exp * 1e-3) / 1e9 if ms_exp > 0 else 0.0
    traffic, src = None, None
    tp = os.path.join(ROOT, "profiles", traffic_file)
    if os.path.exists(tp):
        with open(tp) as f:
            t = json.load(f)
        traffic, src = t["hbm_bytes_per_launch"], t["source"]
    r = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
         "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "kernel": "expand (k_expand)",
         "launches_per_query": n_exp / max(n_queries, 1), "alg_bytes_per_launch": b_exp / max(n_exp, 1),
         "avg_launch_ms": ms_exp / max(n_exp, 1), "traffic_source": src}
    if extra:
        r.update(extra)
    return r


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--config", default="all", choices=["all", "1", "2", "4", "5"],
                    help="all (default): config 5 as the line, configs 4, 2/3 and 1 as extra keys; "
                         "N: that configuration's line alone")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline and oracle parity legs")
    ap.add_argument("--time-budget", type=float, default=480.0,
                    help="seconds after which the default run leaves out its optional legs")
    ap.add_argument("--served-n", type=int, default=2000, help="queries per served rate (0: skip)")
    ap.add_argument("--served-rates", default="1000,2000,4000",
                    help="offered Poisson rates of the 100M-triple served mix, queries/s")
    ap.add_argument("--served-sweep-rates", default="4000,16000,32000,64000,128000,256000,1000000",
                    help="offered Poisson rates of the 1B-triple served sweep (to saturation), queries/s")
    ap.add_argument("--batch-s", type=float, default=2.0,
                    help="closed-loop batch mode: seconds per in-flight cap, a quarter of it warmup (0: skip)")
    ap.add_argument("--batch-p", default="20,200,1000",
                    help="closed-loop batch mode: light in-flight caps (heavy cap min(p, 50))")
    ap.add_argument("--batch-only", action="store_true",
                    help="run only the closed-loop batch mode on the 100M-triple graph")
    ap.add_argument("--exchange-only", action="store_true",
                    help="run only the fork-join table split (for profiler passes)")
    ap.add_argument("--served-heavy-only", action="store_true",
                    help="only the served mix's heavy queries, each alone on the plain path, with the filters "
                         "decided per lineage group and per row")
    ap.add_argument("--heavy-workers", type=int, default=0,
                    help="served heavy queries at once (worker threads; 0: the engine's default)")
    ap.add_argument("--group-filter-rows", type=int, default=1 << 22,
                    help="with --served-heavy-only: smallest table a filter decides per lineage group")
    ap.add_argument("--heavy-only", action="store_true",
                    help="run only the config-2 heavy query and the copy calibration (for profiler passes)")
    ap.add_argument("--probe-only", action="store_true",
                    help="run only the known_to_unknown probe over 64M random keys (for profiler passes)")
    ap.add_argument("--mix-only", action="store_true",
                    help="run only the 512-query mix through the executor, 5 batches (for profiler passes)")
    ap.add_argument("--cycle-only", action="store_true",
                    help="run only the config-5 cyclic query, timed and with HIP events (for profiler passes)")
    ap.add_argument("--framed-only", action="store_true",
                    help="run only the config-2/3 graph under a graph budget (framed residency)")
    ap.add_argument("--framed-budget", type=int, default=1 << 30,
                    help="graph budget of the framed-residency measurement, bytes (0: skip it)")
    ap.add_argument("--shape", default="",
                    help="with --mix-only: the executor shape as key=value pairs (probe_wg_per_cu=3,...) "
                         "instead of autotuning")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU's shard; --gpus {args.gpus} is not supported "
              "(this build runs on a single MI355X)", file=sys.stderr)
        sys.exit(2)
    if args.served_heavy_only:
        args.batch_only = True
    profiler_pass = (args.heavy_only or args.probe_only or args.mix_only or args.batch_only or args.exchange_only
                     or args.cycle_only or args.framed_only)
    if profiler_pass:
        args.no_cpu = True
    if args.exchange_only:
        from wukonggpu_amd import Device
        dev = Device(0)
        print(json.dumps({"table_exchange": table_exchange(dev, args.seed)}), flush=True)
        dev.close()
        return
    if args.cycle_only:
        print(json.dumps(run_config5(args, quick=True)), flush=True)
        return
    if profiler_pass:
        print(json.dumps(run_config23(args)), flush=True)
        return
    if args.config != "all":
        fn = {"1": run_config1, "2": run_config23, "4": run_config4, "5": run_config5}[args.config]
        print(json.dumps(fn(args)), flush=True)
        return

    line = run_config5(args)
    extra = {}
    for name, fn in (("config4", run_config4), ("config2_3", run_config23), ("config1", run_config1)):
        t = time.time()
        try:
            extra[name] = fn(args, nested=True)
        except Exception as e:  # a failing extra leg must not lose the headline line
            extra[name] = {"error": f"{type(e).__name__}: {e}"}
            log(f"{name} failed: {e}")
        extra[name]["leg_s"] = time.time() - t
    line.update(extra)
    line["bench_wall_s"] = time.time() - T_START
    print(json.dumps(line), flush=True)


def _line(value, unit, steps, warmup, ms, config, extra, higher=True, dtype="u32"):
    out = {"metric": METRIC, "value": value, "unit": unit, "n_gpus": 1, "steps": steps, "warmup": warmup,
           "ms_per_step": ms, "higher_is_better": higher, "scaling": "weak", "vs_baseline": None, "dtype": dtype,
           "data": "synthetic (seeded generator; graph resident in HBM)", "config": config}
    out.update(extra)
    return out


def time_left(args):
    return args.time_budget - (time.time() - T_START)


# ---- config 5: the cyclic query on the 2B-triple graph (the N=1 line) --------------------------------

def run_config5(args, quick=False, nested=False):
    """The cyclic 6-pattern query on the 2B-triple graph: plain path with a
    4 GB output-table limit (the ~1B-row hub fan-out produced and consumed in
    row chunks, depth-first), unchunked for comparison, the device-resident
    executor on a batch of the cycle from both start types, the CPU oracle
    on the full query (parallel, as the checker) and on the sample (one
    thread, as the baseline)."""
    import numpy as np
    import oracle
    from wukonggpu_amd import Device, GPUEngine, workloads
    t0 = time.time()
    w = workloads.config5(seed=5, sample_members=C5_SAMPLE)
    build_s = time.time() - t0
    log(f"config 5 graph: {w.info['triples']} triples, {w.info['keys']} keys, image "
        f"{w.info['image_bytes'] / 1e9:.2f} GB, generated and built in {build_s:.1f}s; hub {w.info['hub']} "
        f"degree {w.info['hub_degree']}")
    dev = Device(0, table_budget_bytes=96 << 30)
    t1 = time.time()
    dev.load(w.gstore)
    load_s = time.time() - t1
    eng = GPUEngine(dev)
    # a plain 1 GiB device copy (k_copy_calibrate): the reference point of
    # the HBM counters in a profiled run of this command
    copy_gbs = dev.copy_bandwidth(1 << 30, 3)
    full, sample = w.queries["full"], w.queries["sample"]
    dev.set_table_limit(C5_LIMIT)
    for _ in range(max(1, args.warmup)):
        r = eng.run_query(full)
    s0 = eng.splits()
    lat = []
    t_start = time.perf_counter()
    for _ in range(args.steps):
        q0 = time.perf_counter()
        r = eng.run_query(full)
        lat.append(time.perf_counter() - q0)
    elapsed = time.perf_counter() - t_start
    splits = (eng.splits() - s0) / max(args.steps, 1)
    value = r.expanded * args.steps / elapsed
    log(f"cycle, plain chunked: {r.table.shape[0]} rows, {r.expanded} expanded edges, "
        f"{elapsed / args.steps * 1e3:.2f} ms/query, {value / 1e9:.2f} G expanded edges/s, {splits:.0f} splits/query")
    n_prof = max(1, min(args.steps, 5))
    dev.profile(True)
    t_prof = time.perf_counter()
    for _ in range(n_prof):
        eng.run_query(full)
    prof_ms = (time.perf_counter() - t_prof) / n_prof * 1e3
    kernels = _kernel_table(dev, STEP_KERNELS, n_prof)
    fanout = {k: dev.profile_read("fanout_" + k)[2] / n_prof
              for k in ("filters", "parents", "lists", "list_entries", "work_items", "units", "row_tiles")}
    fanout["fused"] = dev.profile_read("fused_fanouts")[2] / n_prof
    nf, msf, bf = dev.profile_read("fanout_filter")
    dev.profile(False)
    log(f"fan-out filter per query: {fanout}")
    # the dominant kernel: the fan-out filter (k_fanout_blocks) when the
    # fan-out is fused with its filter (the longest launch of the query; the
    # four probe launches together take about as long), else the expand
    roof = None
    if msf > 0:
        ach = bf / (msf * 1e-3) / 1e9
        pairs = r.expanded  # (the fan-out's rows dominate the work count)
        tp = os.path.join(ROOT, "profiles", "fanout_cycle_2b.traffic.json")
        traffic, tsrc = None, None
        if os.path.exists(tp):
            with open(tp) as fh:
                tj = json.load(fh)
            traffic, tsrc = tj["hbm_bytes_per_launch"], tj["source"]
        roof = {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
                "traffic": traffic, "traffic_source": tsrc,
                "kernel": "fanout_filter (k_fanout_blocks, with its 1-word counter reset)",
                "launches_per_query": nf / n_prof, "avg_launch_ms": msf / nf, "alg_bytes_per_launch": bf / nf,
                "pair_tests_per_s": pairs / (msf / n_prof * 1e-3),
                "note": "the fused fan-out's filter tests each of the ~1.36G (parent row, list entry) pairs with a "
                        "register compare (64 rows per compare mask); its HBM bytes are each D entry once per "
                        "group of 64 parents and three counters per unit, so it is bound by each workgroup's "
                        "chain of dependent loads (list, parents, targets, entries), not HBM: pair_tests_per_s "
                        "against 39.3 T compare lanes/s (256 CUs x 4 SIMD16 x 2.4 GHz)"}
    else:
        roof = _expand_roofline(dev, n_prof, "expand_cycle_2b.traffic.json")
    log(f"{roof['kernel']}: {roof['launches_per_query']:.0f} launches/query, {roof['avg_launch_ms']:.3f} ms avg, "
        f"{roof['achieved']:.0f} GB/s algorithmic = {roof['frac']:.2f} of HBM peak"
        + (f", {roof['pair_tests_per_s'] / 1e12:.2f} T pair tests/s" if "pair_tests_per_s" in roof else ""))
    config = {"workload": "6-pattern cyclic query (triangle closed by known_to_known, known_to_const filter) on a "
                          "2B-triple graph: 200M entities, 512 predicates, Zipf 0.9, 2M-edge hubs, pc edges planted "
                          "between consecutive ph objects; 4 GB output-table limit (HBM-budgeted chunks); the plain "
                          "path fuses the hub fan-out with the filter on its new column (the fan-out is never "
                          "written; 'materialized' writes it)",
              **{k: v for k, v in w.info.items()}, "query": full, "result_rows": int(r.table.shape[0]),
              "expanded_edges_per_query": int(r.expanded), "row_half_splits_per_query": splits,
              "p50_ms": pct(lat, 0.5) * 1e3, "p99_ms": pct(lat, 0.99) * 1e3, "graph_build_s": build_s,
              "graph_load_s": load_s, "ms_per_query_with_kernel_events": prof_ms, "copy_kernel_gbs": copy_gbs,
              "kernels_per_query": kernels, "gpu_kernel_ms_per_query": sum(v["ms"] for v in kernels.values()),
              "fanout_filter_per_query": fanout}
    # the same query with the fan-out written (per-step known_to_unknown, then
    # the filter on the written table, in row chunks under the table limit):
    # the k_expand roofline of the materialising path
    eng.set_fused_filters(False)
    rm = eng.run_query(full)
    sm0 = eng.splits()
    tm = []
    for _ in range(3):
        q0 = time.perf_counter()
        rm = eng.run_query(full)
        tm.append(time.perf_counter() - q0)
    msplits = (eng.splits() - sm0) / 3
    dev.profile(True)
    for _ in range(2):
        eng.run_query(full)
    mkernels = _kernel_table(dev, STEP_KERNELS, 2)
    mroof = _expand_roofline(dev, 2, "expand_cycle_2b.traffic.json")
    dev.profile(False)
    mat_same = _same([r], [rm])
    materialized = {"ms_per_query": pct(tm, 0.5) * 1e3, "value": rm.expanded / pct(tm, 0.5),
                    "row_half_splits_per_query": msplits, "roofline": mroof, "kernels_per_query": mkernels,
                    "identical_to_fused": mat_same,
                    "what": "the fan-out written (k_expand, 4 GB table limit, row chunks) and then filtered"}
    log(f"cycle, fan-out written: {materialized['ms_per_query']:.2f} ms/query, "
        f"{materialized['value'] / 1e9:.2f} G expanded edges/s, k_expand {mroof['frac']:.2f} of HBM peak, "
        f"identical={mat_same}")
    if quick:
        out = _line(value, "expanded edges/s", args.steps, args.warmup, elapsed / args.steps * 1e3, config,
                    {"roofline": roof, "materialized": materialized})
        eng.close()
        dev.close()
        return out
    # the materialising path with whole tables (no output limit)
    dev.set_table_limit(0)
    rw = eng.run_query(full)
    tw = []
    for _ in range(3):
        q0 = time.perf_counter()
        rw = eng.run_query(full)
        tw.append(time.perf_counter() - q0)
    whole_same = _same([r], [rw])
    log(f"cycle, fan-out written whole: {pct(tw, 0.5) * 1e3:.2f} ms/query, identical={whole_same}")
    dev.set_table_limit(C5_LIMIT)
    eng.set_fused_filters(True)
    # the executor: the cycle from both start types in one device batch
    eng.configure(max_queries=16, max_patterns=8, arena_bytes=64 << 30)
    bx = eng.run_batched([full, sample])
    tb = []
    for _ in range(3):
        q0 = time.perf_counter()
        bx = eng.run_batched([full, sample])
        tb.append(time.perf_counter() - q0)
    rs = eng.run_query(sample)
    exec_same = _same([r, rs], bx)
    log(f"executor batch [full, sample]: {pct(tb, 0.5) * 1e3:.2f} ms, identical to plain={exec_same}")
    eng.close()
    dev.close()
    parity = {"executor_equals_plain": exec_same, "materialized_equals_fused": mat_same,
              "materialized_whole_equals_fused": whole_same}
    cpu = None
    if not args.no_cpu:
        want, wex, sec = oracle.run_query(w.gstore, full, threads=ORACLE_THREADS)
        parity["full_rows_equal_cpu"] = _eq(r, want, wex)
        parity["full_result_rows"] = int(want.shape[0])
        log(f"oracle, full cycle on {ORACLE_THREADS} threads: {want.shape[0]} rows in {sec:.1f} s, "
            f"equal={parity['full_rows_equal_cpu']}")
        del want
        sw, swex, ssec = oracle.run_query(w.gstore, sample)
        parity["sample_rows_equal_cpu"] = _eq(rs, sw, swex)
        parity["sample_result_rows"] = int(sw.shape[0])
        log(f"oracle, sample cycle on one thread: {sw.shape[0]} rows, {swex} expanded edges in {ssec:.1f} s, "
            f"equal={parity['sample_rows_equal_cpu']}")
        cpu = {"value": swex / ssec, "unit": "expanded edges/s", "cores": 1, "kind": "port",
               "sample": f"the same cycle from the {C5_SAMPLE}-member type t1 ({swex} expanded edges) once on one "
                         f"thread, {ssec:.1f} s", **_cpu_info()}
    w.gstore.close()
    return _line(value, "expanded edges/s", args.steps, args.warmup, elapsed / args.steps * 1e3, config,
                 {"roofline": roof, "cpu_baseline": cpu, "materialized": materialized,
                  "materialized_whole_tables_ms": pct(tw, 0.5) * 1e3,
                  "batched": {"queries": 2, "ms": pct(tb, 0.5) * 1e3, "identical_to_plain": exec_same,
                              "what": "the full cycle and the sample cycle in one device-resident executor batch"},
                  "parity": parity})


# ---- configs 2 and 3: the heavy query and the 512-query mix on the 100M-triple graph ------------------

def run_config23(args, nested=False):
    """The heavy 6-pattern query on the 100M-triple graph (plain path, timed
    uninstrumented, then with HIP events per launch; through the executor),
    its fan-out table against the oracle, the random-key probe roofline, the
    512-query batched mix three ways, the served sweep, the closed-loop
    batch mode and the fork-join split. The profiler passes (--heavy-only,
    --probe-only, --mix-only, --batch-only, --served-heavy-only) stop early."""
    import numpy as np
    import oracle
    from wukonggpu_amd import Device, GPUEngine, workloads
    t0 = time.time()
    w = workloads.config2_3(seed=args.seed)
    log(f"config 2/3 graph: {w.info['triples']} triples, {w.info['keys']} keys, image "
        f"{w.info['image_bytes'] / 1e9:.2f} GB built in {time.time() - t0:.1f}s")
    profiler_pass = args.heavy_only or args.probe_only or args.mix_only or args.batch_only or args.framed_only
    device_build = None
    if not profiler_pass and not nested:
        device_build = compare_device_build(w, Device)
    w.triples.close()
    w.triples = None
    dev = Device(0)
    dev.load(w.gstore)
    eng = GPUEngine(dev)
    if args.heavy_workers:
        eng.set_heavy_workers(args.heavy_workers)
    P = w.info["predicates"]
    type_ids = list(range(2 + P + 1, 2 + P + w.info["types"]))

    def done(obj):
        eng.close()
        dev.close()
        return obj

    if args.probe_only:
        return done({"probe_random_keys": random_key_probe(dev, w, args.seed, HBM_PEAK_GBS * 1e9 / 128)})
    if args.served_heavy_only:
        sq, sflags = workloads.served_mix(P, type_ids, w.info["t0"], args.served_n, args.seed + 11)
        hq = [q for q, f in zip(sq, sflags) if f]
        out = {"workload": f"the {len(hq)} heavy queries of the {len(sq)}-query served mix, each alone"}
        ref = None
        for name, rows in (("group_filters", args.group_filter_rows), ("per_row_filters", 1 << 62)):
            dev.set_group_filter_rows(rows)
            eng.run_plain(hq)  # warm-up
            res = eng.run_plain(hq)
            ms = [r.latency_s * 1e3 for r in res]
            if ref is None:
                ref = [r.table for r in res]
            same = all(np.array_equal(a.table, b) for a, b in zip(res, ref))
            out[name] = {"total_ms": sum(ms), "max_ms": max(ms), "p50_ms": sorted(ms)[len(ms) // 2],
                         "expanded": sum(int(r.expanded) for r in res), "identical": same}
            log(f"{name} (from {rows} rows): {len(hq)} heavy queries alone in {sum(ms):.1f} ms "
                f"(max {max(ms):.2f} ms), identical={same}")
        dev.set_group_filter_rows(1 << 22)
        return done({"served_heavy": out})
    if args.batch_only:
        eng.configure(max_queries=1024, max_patterns=8, arena_bytes=16 << 30)
        return done({"batch_mode": batch_mode(eng, w.gstore, P, type_ids, w.info["t0"], args)})
    if args.framed_only:
        heavy, mix = w.queries["heavy"], w.queries["batched"]
        ref_heavy = eng.run_query(heavy).table
        ref_mix = eng.run_plain(mix)
        return done({"framed": framed_mode(Device, GPUEngine, w.gstore, heavy, mix, ref_heavy, ref_mix,
                                           args.framed_budget or (1 << 30))})
    if args.mix_only:
        mix = w.queries["batched"]
        eng.configure(max_queries=1024, max_patterns=8, arena_bytes=32 << 30)
        if args.shape:
            tuning = {k: int(v) for k, v in (kv.split("=") for kv in args.shape.split(","))}
            eng.set_tuning(**tuning)
            log(f"executor shape {tuning}")
        else:
            tuning, tune_ms = eng.autotune(mix, reps=2)
            log(f"autotuned executor shape {tuning}: {tune_ms:.2f} ms per {len(mix)}-query batch")
        ms = []
        for _ in range(5):
            _, tb = eng.run_batched_packed(mix)
            ms.append(tb * 1e3)
        log(f"5 batches: {', '.join(f'{x:.3f}' for x in ms)} ms")
        return done({"mix_batch_ms": ms, "tuning": tuning})

    copy_gbs = dev.copy_bandwidth(1 << 30, 20)
    lines_ps = dev.random_access_rate(4 << 30, 10)
    log(f"copy kernel bandwidth: {copy_gbs:.0f} GB/s; random 128-byte line reads: {lines_ps / 1e9:.2f} G/s "
        f"({lines_ps * 128 / 1e9:.0f} GB/s)")
    heavy = w.queries["heavy"]
    for _ in range(args.warmup):
        eng.run_query(heavy)
    # the timed steps run uninstrumented; per-kernel HIP-event timing (an event
    # pair around every launch) runs over the same number of steps afterwards
    lat = []
    t_start = time.perf_counter()
    for _ in range(args.steps):
        q0 = time.perf_counter()
        r = eng.run_query(heavy)
        lat.append(time.perf_counter() - q0)
    elapsed = time.perf_counter() - t_start
    dev.profile(True)
    t_prof = time.perf_counter()
    for _ in range(args.steps):
        eng.run_query(heavy)
    prof_ms = (time.perf_counter() - t_prof) / args.steps * 1e3
    kernels = _kernel_table(dev, STEP_KERNELS, args.steps)
    probe_keys = dev.profile_read("probe_keys")[2]  # counts kept in the bytes field
    probe_rows = dev.profile_read("probe_rows")[2]
    ms_pr = dev.profile_read("probe")[1]
    alg_pr = dev.profile_read("probe")[2]
    roof = _expand_roofline(dev, args.steps, "expand_heavy_query.traffic.json", {"copy_kernel_gbs": copy_gbs})
    dev.profile(False)
    edges = r.expanded
    value = edges * args.steps / elapsed
    log(f"heavy query: {r.table.shape[0]} result rows, {edges} expanded edges, "
        f"{elapsed / args.steps * 1e3:.2f} ms/query, {value / 1e9:.2f} G expanded edges/s; k_expand "
        f"{roof['achieved']:.0f} GB/s = {roof['frac']:.2f} of peak")
    line_peak = HBM_PEAK_GBS * 1e9 / 128
    probe_roofline = {
        "bound": "random-access", "peak": line_peak, "unit": "128-byte bucket lines/s",
        "kernel": "probe (k_probe, known_to_unknown steps)", "measured_random_lines_4GiB": lines_ps,
        "heavy_query": {
            "keys_probed_per_query": probe_keys / args.steps, "rows_per_query": probe_rows / args.steps,
            "keys_per_s": probe_keys / (ms_pr * 1e-3) if ms_pr > 0 else 0.0,
            "rows_per_s": probe_rows / (ms_pr * 1e-3) if ms_pr > 0 else 0.0,
            "alg_GBps": alg_pr / (ms_pr * 1e-3) / 1e9 if ms_pr > 0 else 0.0,
            "note": "runs of equal keys share one bucket walk, so this probe is bound by reading its "
                    "input table (4 R c bytes), not by bucket lines; alg_GBps counts 4 R c + 128 per key "
                    "probed + 16 per row with edges",
        },
    }
    if args.heavy_only:
        return done(_line(value, "expanded edges/s", args.steps, args.warmup, elapsed / args.steps * 1e3,
                          {"workload": "config-2 heavy query"}, {"roofline": roof}))
    rk = random_key_probe(dev, w, args.seed, line_peak)
    rk["frac_of_random_lines_4GiB"] = rk["achieved"] / lines_ps
    probe_roofline.update({"achieved": rk["achieved"], "frac": rk["frac_of_peak"], "random_keys": rk})
    log(f"probe on random keys: {rk['keys_probed'] / 1e6:.1f} M keys in {rk['ms_per_launch']:.2f} ms = "
        f"{rk['achieved'] / 1e9:.2f} G keys/s = {rk['frac_of_peak']:.2f} of HBM peak / 128 B; random lines "
        f"over 4 GiB {lines_ps / 1e9:.1f} G/s, over the {rk['segment_bytes'] / 2**20:.0f} MiB segment "
        f"{rk['random_lines_same_working_set'] / 1e9:.1f} G/s ({rk['frac_of_random_lines_same_working_set']:.2f})")

    # ---- the heavy query through the device-resident executor ---------------------------
    eng.configure(max_queries=1024, max_patterns=8, arena_bytes=32 << 30)
    heavy_tuning, _ = eng.autotune([heavy], reps=2, heavy=True)  # shape of the heavy serving lane
    eng_h = GPUEngine(dev)
    eng_h.configure(max_queries=64, max_patterns=8, arena_bytes=32 << 30)
    eng_h.autotune([heavy], reps=2)
    hx = eng_h.run_batched([heavy])[0]
    hx_lat = []
    for _ in range(max(3, args.steps // 2)):
        q0 = time.perf_counter()
        hx = eng_h.run_batched([heavy])[0]
        hx_lat.append(time.perf_counter() - q0)
    eng_h.close()
    heavy_exec_same = _same([hx], [r])
    log(f"heavy query via the executor (shape {heavy_tuning}): {pct(hx_lat, 0.5) * 1e3:.2f} ms "
        f"(identical={heavy_exec_same})")

    # ---- batched mix: plain vs step rounds vs the autotuned executor ---------------------
    mix = w.queries["batched"]
    t_tune = time.perf_counter()
    tuning, tune_ms = eng.autotune(mix, reps=3)
    t_tune = time.perf_counter() - t_tune
    log(f"autotuned executor shape {tuning}: {tune_ms:.2f} ms per {len(mix)}-query batch ({t_tune:.1f} s)")
    eng.run_rounds(mix[:64])  # warm
    tp = time.perf_counter()
    plain = eng.run_plain(mix)
    t_plain = time.perf_counter() - tp
    tr_ = time.perf_counter()
    rounds = eng.run_rounds(mix)
    t_rounds = time.perf_counter() - tr_
    # the executor, timed around the native call (the Python conversion of the
    # 512 tables into numpy views happens after it)
    t_exec = []
    for _ in range(5):
        batched, tb = eng.run_batched_packed(mix)
        t_exec.append(tb)
    t_batched = sorted(t_exec)[len(t_exec) // 2]
    batched_1 = eng.run_batched(mix)  # the per-query-table form gives the same answers
    same_exec, same_rounds = _same(plain, batched) and _same(plain, batched_1), _same(plain, rounds)
    mix_edges = sum(x.expanded for x in batched)
    log(f"batched mix: plain {len(mix) / t_plain:.0f} q/s, step rounds {len(mix) / t_rounds:.0f} q/s, "
        f"executor {len(mix) / t_batched:.0f} q/s (x{t_plain / t_batched:.1f}); identical={same_exec and same_rounds}")

    # ---- served: open-loop Poisson arrivals, light + heavy lanes ----------------------------
    served, served_same = [], None
    if args.served_n > 0 and not nested:
        sq, sflags = workloads.served_mix(P, type_ids, w.info["t0"], args.served_n, args.seed + 11)
        alone = eng.run_plain(sq)
        served, served_same = served_sweep(eng, dev, sq, sflags, alone, args.served_rates, args.seed + 12,
                                           64 << 30)

    # ---- closed-loop batch mode: templates with random constants, in-flight caps -------------
    bm = None
    if args.batch_s > 0 and (not nested or time_left(args) > 120):
        bm = batch_mode(eng, w.gstore, P, type_ids, w.info["t0"], args)

    # ---- fork-join table exchange: the split a sharded deployment adds per step ---------------
    exchange = table_exchange(dev, args.seed)
    # ---- framed residency: the same graph under a graph budget below its image -----------------
    framed = None
    if args.framed_budget and (not nested or time_left(args) > 120):
        framed = framed_mode(Device, GPUEngine, w.gstore, heavy, mix, r.table, plain, args.framed_budget)
    # the 151M-row hub fan-out table (the heavy query's first four patterns), for the oracle below
    rp = eng.run_query(heavy[:16]) if not args.no_cpu else None
    eng.close()
    dev.close()

    # ---- the oracle: parity of every answer, the CPU baseline -----------------------------------
    cpu = None
    parity = {"heavy_executor_equals_plain": heavy_exec_same,
              "mix_plain_equals_executor": bool(same_exec), "mix_plain_equals_rounds": bool(same_rounds),
              "served_equals_alone": served_same, "batch_mode_equals_alone": bm["equals_alone"] if bm else None,
              "table_split_correct": all(x["correct"] for x in exchange)}
    if not args.no_cpu:
        want, wex, sec = oracle.run_query(w.gstore, heavy)
        parity["heavy_rows_equal_cpu"] = _eq(r, want, wex)
        parity["heavy_result_rows"] = int(want.shape[0])
        cpu = {"value": wex / sec, "unit": "expanded edges/s", "cores": 1, "kind": "port",
               "sample": f"the full heavy query once on one thread ({wex} expanded edges in {sec:.1f} s)",
               **_cpu_info()}
        log(f"cpu oracle: {wex / sec / 1e6:.1f} M expanded edges/s, heavy rows {want.shape[0]}, "
            f"parity={parity['heavy_rows_equal_cpu']}")
        ok, n, wall = _oracle_digests(w.gstore, mix, batched)
        parity["mix_tables_equal_cpu"] = ok
        parity["mix_tables_checked"] = n
        parity["mix_nonempty_answers"] = int(sum(1 for x in batched if x.table.shape[0]))
        log(f"512-query mix: every executor table against the oracle: {ok}")
        # the hub fan-out table, row for row
        pw, pex, psec = oracle.run_query(w.gstore, heavy[:16], threads=ORACLE_THREADS)
        parity["fanout_rows_equal_cpu"] = _eq(rp, pw, pex)
        parity["fanout_rows"] = int(pw.shape[0])
        log(f"fan-out table ({pw.shape[0]} rows x {pw.shape[1]}): equal to the oracle={parity['fanout_rows_equal_cpu']}")
        del pw, rp
        threads = ORACLE_THREADS
        _, _, _, wall_c = oracle.run_many(w.gstore, mix, threads=threads)
        cpu["mix_queries_per_s"] = len(mix) / wall_c
        cpu["mix_threads"] = threads
    w.gstore.close()
    return _line(value, "expanded edges/s", args.steps, args.warmup, elapsed / args.steps * 1e3, {
        "workload": "heavy 6-pattern query (index start, 2-hop chain, 2-arm star over a linked predicate pair, "
                    "closing known_to_known) on a 100M-triple graph: 10M entities, 64 predicates (linked in "
                    "pairs), Zipf s=1.0, 500k-member type",
        "triples": w.info["triples"], "keys": w.info["keys"], "edges": w.info["edges"],
        "query": heavy, "result_rows": int(r.table.shape[0]), "expanded_edges_per_query": int(edges),
        "p50_ms": pct(lat, 0.5) * 1e3, "p99_ms": pct(lat, 0.99) * 1e3,
        "gpu_kernel_ms_per_query": sum(v["ms"] for v in kernels.values()),
        "ms_per_query_with_kernel_events": prof_ms, "kernels_per_query": kernels}, {
        "roofline": roof, "probe_roofline": probe_roofline, "device_build": device_build, "cpu_baseline": cpu,
        "batched": {
            "workload": f"{len(mix)} selective 3-4 pattern queries from type indexes (sizes log-uniform 1..1e5)",
            "plain_queries_per_s": len(mix) / t_plain, "rounds_queries_per_s": len(mix) / t_rounds,
            "batched_queries_per_s": len(mix) / t_batched, "speedup": t_plain / t_batched,
            "executor": "device-resident: whole queries in one captured hipGraph, persistent tile kernels "
                        "(work queue or static deal, autotuned)",
            "tuning": tuning, "autotune_ms_per_batch": tune_ms, "autotune_s": t_tune,
            "plain_p50_ms": pct([x.latency_s for x in plain], 0.5) * 1e3,
            "plain_p99_ms": pct([x.latency_s for x in plain], 0.99) * 1e3,
            "batched_ms_per_batch": t_batched * 1e3, "expanded_edges": int(mix_edges)},
        "heavy_via_executor_ms": pct(hx_lat, 0.5) * 1e3, "heavy_tuning": heavy_tuning,
        "served": served, "batch_mode": bm, "table_exchange": exchange, "framed": framed, "parity": parity})


def served_sweep(eng, dev, sq, sflags, alone, rates, seed, budget, span_s=0.0):
    """Open-loop Poisson arrivals at each offered rate: achieved queries/s,
    p50/p99 overall and per class, each answer against the query run alone.
    With span_s, a rate serves the first max(2000, rate * span_s) queries of
    the stream (arrivals over about span_s seconds); else all of them."""
    import numpy as np
    hx = [int(a.expanded) for a, f in zip(alone, sflags) if f]
    hl = [a.latency_s for a, f in zip(alone, sflags) if f]
    log(f"served mix: {len(sq)} queries, {len(hx)} heavy expanding {min(hx, default=0)}..{max(hx, default=0)} "
        f"edges, {min(hl, default=0) * 1e3:.2f}..{max(hl, default=0) * 1e3:.2f} ms alone")
    out, all_same = [], True
    rng = np.random.default_rng(seed)
    for rate in [float(x) for x in rates.split(",") if x]:
        n = len(sq) if span_s <= 0 else min(len(sq), max(2000, int(rate * span_s)))
        qs, fl, al_ = sq[:n], sflags[:n], alone[:n]
        arr = np.cumsum(rng.exponential(1.0 / rate, size=n))
        pool0 = dev.pool_stats()
        res, wall = eng.serve(qs, arr, heavy=fl, table_budget_bytes=budget, light_lanes=2)
        pool1 = dev.pool_stats()
        ok = _same(al_, res)
        all_same = all_same and ok
        lat_l = [x.latency_s for x, f in zip(res, fl) if not f]
        lat_h = [x.latency_s for x, f in zip(res, fl) if f]
        rec = {"offered_qps": rate, "queries": n, "heavy": int(fl.sum()), "arrival_span_s": float(arr[-1]),
               "achieved_qps": n / wall, "p50_ms": pct([x.latency_s for x in res], 0.5) * 1e3,
               "p99_ms": pct([x.latency_s for x in res], 0.99) * 1e3,
               "light_p50_ms": pct(lat_l, 0.5) * 1e3, "light_p99_ms": pct(lat_l, 0.99) * 1e3,
               "heavy_p50_ms": pct(lat_h, 0.5) * 1e3, "heavy_p99_ms": pct(lat_h, 0.99) * 1e3,
               "identical_to_alone": bool(ok), "serve_stats": eng.serve_stats(),
               "pool": {"mallocs": pool1["mallocs"] - pool0["mallocs"], "trims": pool1["trims"] - pool0["trims"]}}
        out.append(rec)
        log(f"served @ {rate:.0f} q/s offered ({n} queries): {rec['achieved_qps']:.0f} q/s, p50 {rec['p50_ms']:.2f} "
            f"ms, p99 {rec['p99_ms']:.2f} ms (light p99 {rec['light_p99_ms']:.2f}, heavy p99 "
            f"{rec['heavy_p99_ms']:.2f}), identical={ok}")
    return out, all_same


def served_knee(served):
    """The highest offered rate the served loop kept up with (achieved at
    least 95% of offered), with its latencies, and the first rate past it."""
    kept = [r for r in served if r["achieved_qps"] >= 0.95 * r["offered_qps"]]
    past = [r for r in served if r["achieved_qps"] < 0.95 * r["offered_qps"]]
    k = max(kept, key=lambda r: r["offered_qps"]) if kept else None
    p = min(past, key=lambda r: r["offered_qps"]) if past else None
    pick = ("offered_qps", "achieved_qps", "p50_ms", "p99_ms", "light_p50_ms", "light_p99_ms", "heavy_p50_ms",
            "heavy_p99_ms")
    return {"last_kept_up": {x: k[x] for x in pick} if k else None,
            "first_saturated": {x: p[x] for x in pick} if p else None,
            "rule": "kept up: achieved >= 95% of offered"}


# ---- config 4: served on the 1B-triple graph -------------------------------------------------------

def run_config4(args, nested=False):
    """Served: open-loop Poisson arrivals over the 1B-triple graph, 95% light
    (2-3 patterns from small type indexes) and 5% heavy (the 6-pattern heavy
    shape from a 500k-member type, >= 10M-row tables), 64 GB table budget;
    every served answer against the query alone, all of them against the
    oracle's digests; the batch mode; the heavy queries' expand roofline."""
    import numpy as np
    import oracle
    from wukonggpu_amd import Device, GPUEngine, workloads
    t0 = time.time()
    w = workloads.config4(seed=4, keep_triples=not nested)
    log(f"config 4 graph: {w.info['triples']} triples, {w.info['keys']} keys, image "
        f"{w.info['image_bytes'] / 1e9:.2f} GB built in {time.time() - t0:.1f}s")
    device_build = None
    if not nested:
        device_build = compare_device_build(w, Device, host_gs=w.gstore, host_s=w.info["host_build_s"])
        w.triples.close()
        w.triples = None
    dev = Device(0)
    dev.load(w.gstore)
    eng = GPUEngine(dev)
    if args.heavy_workers:
        eng.set_heavy_workers(args.heavy_workers)
    P = w.info["predicates"]
    sq, flags = workloads.served_mix(P, w.info["light_types"], w.info["t0"], args.served_n * 10, args.seed + 41)
    heavy_idx = [i for i in range(len(sq)) if flags[i]]
    eng.configure(max_queries=1024, max_patterns=8, arena_bytes=16 << 30)
    light_idx = [i for i in range(len(sq)) if not flags[i]][:512]
    light = [sq[i] for i in light_idx]
    tuning, tune_ms = eng.autotune(light, reps=2)
    htuning, _ = eng.autotune([sq[i] for i in heavy_idx[:2]], reps=1, heavy=True)
    log(f"autotuned: light {tuning}, heavy {htuning}")
    tp = time.perf_counter()
    alone = eng.run_plain(sq)
    t_plain = time.perf_counter() - tp
    # the heavy queries' expand on the plain path, with HIP events
    hq = [sq[i] for i in heavy_idx]
    dev.profile(True)
    hres = eng.run_plain(hq)
    kernels = _kernel_table(dev, STEP_KERNELS, max(len(hq), 1))
    roof = _expand_roofline(dev, len(hq), "expand_served_1b.traffic.json",
                            {"workload": f"the {len(hq)} heavy queries of the served mix, each alone on the plain path"})
    dev.profile(False)
    hrows = [int(alone[i].expanded) for i in heavy_idx]
    log(f"{len(sq)} queries alone on the plain path: {len(sq) / t_plain:.0f} q/s; heavy queries expand "
        f"{min(hrows) if hrows else 0}..{max(hrows) if hrows else 0} edges; k_expand {roof['achieved']:.0f} GB/s "
        f"({roof['frac']:.2f} of peak)")
    tb = time.perf_counter()
    light_b = eng.run_batched(light)
    t_light = time.perf_counter() - tb
    # one untimed served run first: the lanes' arenas and captured sequences
    # are set up there, not in the first measured rate
    warm = np.cumsum(np.random.default_rng(args.seed + 40).exponential(1.0 / 4000, size=2000))
    eng.serve(sq[:2000], warm, heavy=flags[:2000], table_budget_bytes=64 << 30, light_lanes=2)
    served, all_same = served_sweep(eng, dev, sq, flags, alone, args.served_sweep_rates, args.seed + 42, 64 << 30,
                                    span_s=0.5)
    knee = served_knee(served)
    log(f"served knee: {knee}")
    # the same sweep with every light query on the light lanes' batches (no
    # small-query server): admission at batch boundaries only
    eng.set_serve_small(0)
    log("the same sweep without the small-query server:")
    served_lanes, same_lanes = served_sweep(eng, dev, sq, flags, alone, args.served_sweep_rates, args.seed + 42,
                                            64 << 30, span_s=0.5)
    eng.set_serve_small(16)
    knee_lanes = served_knee(served_lanes)
    log(f"served knee without the server: {knee_lanes}")
    all_same = all_same and same_lanes
    bm = None
    if args.batch_s > 0 and (not nested or time_left(args) > 200):
        bm = batch_mode(eng, w.gstore, P, w.info["light_types"], w.info["t0"], args)
    eng.close()
    dev.close()
    parity = {"served_equals_alone": all_same, "batch_mode_equals_alone": bm["equals_alone"] if bm else None,
              "light_batched_equals_plain": _same([alone[i] for i in light_idx], light_b),
              "heavy_profiled_equals_alone": _same(hres, [alone[i] for i in heavy_idx])}
    cpu = None
    if not args.no_cpu:
        ok, n, wall = _oracle_digests(w.gstore, sq, alone)
        parity["all_served_queries_equal_cpu"] = ok
        parity["queries_checked"] = n
        parity["heavy_nonempty_answers"] = int(sum(1 for i in heavy_idx if alone[i].table.shape[0]))
        log(f"oracle digests of all {n} served queries ({ORACLE_THREADS} threads, {wall:.1f} s): {ok}")
        samp = light[:256] + [sq[i] for i in heavy_idx[:1]]
        t = time.perf_counter()
        for q in samp:
            oracle.run_query(w.gstore, q)
        cpu_s = time.perf_counter() - t
        cpu = {"value": len(samp) / cpu_s, "unit": "queries/s", "cores": 1, "kind": "port",
               "sample": f"{len(samp)} of the served queries (256 light, 1 heavy) on one thread, {cpu_s:.1f} s",
               **_cpu_info()}
    best = max(served, key=lambda x: x["achieved_qps"]) if served else None
    w.gstore.close()
    return _line(best["achieved_qps"] if best else 0.0, "queries/s", len(sq), 0,
                 1e3 / best["achieved_qps"] if best else 0.0,
                 {"workload": "served open-loop Poisson mix (95% light 2-3 patterns, 5% heavy 6 patterns) on a "
                              "1B-triple graph: 100M entities, 256 predicates (linked in pairs), Zipf 1.0; 64 GB "
                              "table budget",
                  **{k: v for k, v in w.info.items() if k != "light_types"},
                  "queries_per_rate": len(sq), "heavy_queries": len(heavy_idx)},
                 {"roofline": roof, "kernels_per_heavy_query": kernels, "served": served, "served_knee": knee,
                  "served_lanes_only": served_lanes, "served_knee_lanes_only": knee_lanes,
                  "batch_mode": bm,
                  "plain_queries_per_s": len(sq) / t_plain, "device_build": device_build,
                  "light_batched_queries_per_s": len(light) / t_light, "tuning": tuning, "heavy_tuning": htuning,
                  "parity": parity, "cpu_baseline": cpu})


# ---- config 1: the 2-hop query from a constant on the 10M-triple graph -----------------------------------

def run_config1(args, nested=False):
    """Plain single query on the 10M-triple graph: the 2-hop chain from the
    constant of p1-degree ~1000, on the GPU (plain path and executor) and on
    the CPU oracle."""
    import oracle
    from wukonggpu_amd import Device, GPUEngine, workloads
    w = workloads.config1(seed=1)
    q = w.queries[0]
    dev = Device(0)
    dev.load(w.gstore)
    eng = GPUEngine(dev)
    steps = max(args.steps, 200)
    for _ in range(max(args.warmup, 10)):
        r = eng.run_query(q)
    lat = []
    for _ in range(steps):
        t = time.perf_counter()
        r = eng.run_query(q)
        lat.append(time.perf_counter() - t)
    el = sum(lat)
    dev.profile(True)
    for _ in range(20):
        eng.run_query(q)
    kernels = _kernel_table(dev, STEP_KERNELS + ("small_query",), 20)
    dev.profile(False)
    # the same query with one launch per query instead of the server
    dev.set_small_query_server(0)
    for _ in range(10):
        eng.run_query(q)
    lat_one = []
    for _ in range(steps):
        t = time.perf_counter()
        r1 = eng.run_query(q)
        lat_one.append(time.perf_counter() - t)
    dev.set_small_query_server(4)
    # the same query step by step (one launch sequence and read-back per step)
    eng.set_small_queries(False)
    for _ in range(10):
        rs = eng.run_query(q)
    lat_steps = []
    for _ in range(steps):
        t = time.perf_counter()
        rs = eng.run_query(q)
        lat_steps.append(time.perf_counter() - t)
    eng.set_small_queries(True)
    # the server's time on the query (claim to answer, device clock) against
    # the HBM roofline: algorithmic bytes of its two steps -- the constant's
    # bucket and list read and written (128 + 8 D), then known_to_unknown over
    # D rows (4 D read, 128 per key probed, the E gathered edges, E rows of 2
    # cells written, 8 D of offsets); the query is latency-bound
    D = int(w.info["const_degree"])
    E = int(r.table.shape[0])
    alg = 128 + 8 * D + 4 * D + 128 * D + 4 * E + 8 * E + 8 * D
    sq = kernels.get("small_query")
    roof = None
    if sq and sq["ms"] > 0:
        ach = alg / (sq["ms"] / sq["launches"] * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
                "traffic": None, "kernel": "small_query (k_sq_server: one query in one workgroup)",
                "launches_per_query": sq["launches"], "avg_launch_ms": sq["ms"] / sq["launches"],
                "alg_bytes_per_launch": alg,
                "note": "a 1.2k-edge query in one workgroup of the persistent server; time per query from the "
                        "device clock (claim to answer): a chain of dependent bucket and list reads, "
                        "latency-bound by design"}
    eng.configure(max_queries=64, max_patterns=4, arena_bytes=1 << 30)
    bx = eng.run_batched([q])[0]
    want, wex, sec = oracle.run_query(w.gstore, q)
    reps = max(1, int(2.0 / max(sec, 1e-4)))
    t = time.perf_counter()
    for _ in range(reps):
        oracle.run_query(w.gstore, q)
    cpu_s = (time.perf_counter() - t) / reps
    same = _eq(r, want, wex)
    log(f"config 1: {r.table.shape[0]} rows, {r.expanded} expanded edges, GPU p50 {pct(lat, 0.5) * 1e3:.3f} ms "
        f"(one launch {pct(lat_one, 0.5) * 1e3:.3f} ms, step by step {pct(lat_steps, 0.5) * 1e3:.3f} ms), "
        f"CPU {cpu_s * 1e3:.3f} ms/query, identical={same}")
    eng.close()
    dev.close()
    w.gstore.close()
    return _line(r.expanded * steps / el, "expanded edges/s", steps, max(args.warmup, 10), el / steps * 1e3,
                 {"workload": "plain 2-hop chain from one constant (p1 degree ~1k) on a 10M-triple graph: 1M "
                              "entities, 16 predicates, Zipf 1.1 capped at 10k",
                  **{k: v for k, v in w.info.items()}, "query": q, "result_rows": int(r.table.shape[0]),
                  "p50_ms": pct(lat, 0.5) * 1e3, "p99_ms": pct(lat, 0.99) * 1e3, "queries_per_s": steps / el,
                  "one_launch_p50_ms": pct(lat_one, 0.5) * 1e3, "step_by_step_p50_ms": pct(lat_steps, 0.5) * 1e3,
                  "path": "the whole query in one workgroup of the persistent small-query server "
                          "(wkg_query_small); one_launch_p50_ms: one launch per query; step_by_step_p50_ms: one "
                          "launch sequence and read-back per step",
                  "kernels_per_query": kernels},
                 {"roofline": roof,
                  "cpu_baseline": {"value": wex / cpu_s, "unit": "expanded edges/s", "cores": 1, "kind": "port",
                                   "sample": f"the full query, {reps} runs on one thread", "p50_ms": cpu_s * 1e3,
                                   **_cpu_info()},
                  "parity": {"rows_equal_cpu": same, "executor_equals_cpu": _eq(bx, want, wex),
                             "one_launch_equals_cpu": _eq(r1, want, wex),
                             "step_by_step_equals_cpu": _eq(rs, want, wex)}})


# Latency points of the reference logger's per-query-type CDF (core/logger.hpp:250-313).
CDF_RATES = [0.01] + [0.05 * i for i in range(1, 20)] + [0.95 + 0.01 * i for i in range(1, 6)]


def batch_mode(eng, gs, P, light_types, t0, args, budget=64 << 30):
    """Closed-loop batch mode (core/proxy.hpp:287-457) at each light in-flight
    cap p of --batch-p, the heavy cap min(p, 50) (global_heavy_flying_cnt of
    scripts/config), --batch-s seconds with the first quarter warmup. Per cap,
    as the reference's Logger reports it: light and heavy queries completed
    per second after the warmup, light and heavy median and 99th latency
    (sorted[n/2], sorted[int(0.99 n)], core/logger.hpp:196-238), and the
    25-point latency CDF per query type (core/logger.hpp:250-313). The first
    2048 light and 64 heavy issued queries are checked against the same
    query run alone on the per-step path."""
    import numpy as np
    from wukonggpu_amd import workloads
    dur = args.batch_s
    warm = dur / 4
    caps = [int(x) for x in args.batch_p.split(",") if x]
    n = 6_000_000
    t = time.perf_counter()
    flat, off, heavy, types = workloads.batch_stream(gs, P, light_types, t0, n, args.seed + 51)
    log(f"batch mode: {n} queries from 6 light templates and 4 heavy queries in {time.perf_counter() - t:.1f}s")
    light_pos, heavy_pos = np.nonzero(heavy == 0)[0], np.nonzero(heavy)[0]
    keep = np.zeros(n, dtype=bool)
    keep[light_pos[:2048]] = True
    keep[heavy_pos[:64]] = True
    kidx = np.nonzero(keep)[0]
    kq = [flat[off[i]:off[i + 1]].tolist() for i in kidx]
    alone = eng.run_plain(kq)
    out, all_same = [], True
    for p in caps:
        hcap = min(p, 50)
        res, issue, wall = eng.serve_closed((flat, off), heavy, p, hcap, dur, table_budget_bytes=budget,
                                            light_lanes=2, keep=keep)
        issued = issue >= 0
        lat = res.latency_s
        done = issue + lat
        st = res.status
        win = issued & (done >= warm) & (done <= dur)
        ok = bool((st[issued] == 0).all())
        for j, i in enumerate(kidx):
            if issued[i]:
                a, r = alone[j], res[int(i)]
                ok = ok and a.status == 0 and r.table.shape == a.table.shape and \
                    np.array_equal(r.table, a.table) and r.expanded == a.expanded
        all_same = all_same and ok
        rec = {"light_inflight": p, "heavy_inflight": hcap, "wall_s": wall,
               "issued": int(issued.sum()), "issued_heavy": int((issued & (heavy == 1)).sum()),
               "stream_exhausted": bool(issued[light_pos].all()),
               "light_qps": float((win & (heavy == 0)).sum() / (dur - warm)),
               "heavy_qps": float((win & (heavy == 1)).sum() / (dur - warm)), "equals_alone": ok}
        for cls, name in ((0, "light"), (1, "heavy")):
            v = np.sort(lat[issued & (heavy == cls)])
            rec[f"{name}_median_us"] = round(float(v[len(v) // 2] * 1e6), 1) if len(v) else None
            rec[f"{name}_p99_us"] = round(float(v[int(len(v) * 0.99)] * 1e6), 1) if len(v) else None
        cdf = {}
        for k in range(10):
            v = np.sort(lat[issued & (types == k)])
            if len(v):
                cdf[f"Q{k + 1}"] = [round(float(v[min(int(len(v) * r), len(v) - 1)] * 1e6), 1) for r in CDF_RATES]
        rec["cdf_us"] = cdf
        rec["cdf_rates"] = [round(r, 2) for r in CDF_RATES]
        out.append(rec)
        log(f"batch mode p={p} (heavy {hcap}): light {rec['light_qps']:.0f} q/s, heavy {rec['heavy_qps']:.0f} q/s; "
            f"light median {rec['light_median_us']} us p99 {rec['light_p99_us']} us, heavy median "
            f"{rec['heavy_median_us']} us p99 {rec['heavy_p99_us']} us; issued {rec['issued']} "
            f"({rec['issued_heavy']} heavy), exhausted={rec['stream_exhausted']}, identical={ok}")
        del res
    return {"duration_s": dur, "warmup_s": warm, "runs": out, "equals_alone": all_same,
            "templates": "6 light (constant start from a random type member: 2-hop, 3-arm star, 3-hop chain) and "
                         "4 heavy (index start, 6 patterns), loads of the reference's batch mix_config"}


def compare_device_build(w, Device, main_load=0.75, host_gs=None, host_s=None):
    """Builds the same graph on the device (wkg_build_graph) in a second
    context, times it against the host builder (all host threads; rebuilt
    here unless its image and time are given) and checks the two images are
    equal, slot for slot."""
    import numpy as np
    from wukonggpu_amd import GStore
    gs = host_gs
    if gs is None:
        t = time.perf_counter()
        gs = GStore.build(w.triples, main_load=main_load)
        host_s = time.perf_counter() - t
    d2 = Device(0)
    dev_s = d2.build(w.triples, main_load=main_load)
    slots, edges, metas = d2.image()
    same = bool(np.array_equal(metas, gs.metas()) and np.array_equal(edges, gs.edges()) and
                np.array_equal(slots, gs.slots().reshape(-1, 2)))
    d2.close()
    if host_gs is None:
        gs.close()
    n = len(w.triples)
    log(f"device build: {n} triples in {dev_s:.3f} s ({n / dev_s / 1e6:.0f} M triples/s, upload included); host "
        f"builder {host_s:.2f} s on {os.cpu_count()} threads; images identical={same}")
    return {"triples": n, "device_s": dev_s, "host_s": host_s, "host_threads": os.cpu_count(),
            "device_triples_per_s": n / dev_s, "image_equal_host": same}


def random_key_probe(dev, w, seed, line_peak):
    """One known_to_unknown step (predicate 2 OUT) over 64M rows of uniformly
    random vertex ids, timed with HIP events over 3 launches: keys probed per
    second against HBM peak / 128 B and against random 128-byte line reads
    over a buffer the size of the predicate's bucket segment."""
    import numpy as np
    rng_p = np.random.default_rng(seed + 7)
    n_rand = 1 << 26
    col = (1 << 17) + rng_p.integers(0, w.info["entities"], size=(n_rand, 1), dtype=np.int64)
    t_rand = dev.table(col.astype(np.int32))
    del col
    out_t, _ = dev.k2u(t_rand, 0, 2, 1)  # warm
    out_t.free()
    dev.profile(True)
    for _ in range(3):
        out_t, _ = dev.k2u(t_rand, 0, 2, 1)
        out_t.free()
    n_r, ms_r, _ = dev.profile_read("probe")
    keys_r = dev.profile_read("probe_keys")[2]
    dev.profile(False)
    t_rand.free()
    rand_rate = keys_r / (ms_r * 1e-3)
    m2 = w.gstore.metas()[2]  # partition_sz, main [start, end), indirect [start, end), ...
    seg_bytes = int(m2[0] + (m2[4] - m2[3])) * 128
    seg_lines = dev.random_access_rate(max(seg_bytes, 1 << 20), 10)
    return {"rows": n_rand, "keys_probed": keys_r / 3, "ms_per_launch": ms_r / max(n_r, 1),
            "achieved": rand_rate, "frac_of_peak": rand_rate / line_peak,
            "segment_bytes": seg_bytes, "random_lines_same_working_set": seg_lines,
            "frac_of_random_lines_same_working_set": rand_rate / seg_lines,
            "workload": "known_to_unknown over 64M uniformly random vertex ids, predicate 2 OUT"}


def framed_mode(Device, GPUEngine, gstore, heavy, mix, heavy_table, mix_ref, budget):
    """The config-2/3 graph under a graph budget below its image: predicate
    segments in frames loaded on demand, least recently used evicted
    (core/shard_manager.hpp), with the pipelined load of a later pattern's
    predicate (core/rcache.hpp:273-281) off and on. Per setting, on a fresh
    context: the heavy query cold (its segments loaded) and warm, then the
    512-query mix through the plain path; every answer is compared with the
    resident graph's."""
    import numpy as np
    out = {"budget_bytes": budget, "image_bytes": gstore.image_nbytes()}
    for on in (False, True):
        dev = Device(0)
        t0 = time.perf_counter()
        dev.load(gstore, graph_budget_bytes=budget)
        t_load = time.perf_counter() - t0
        eng = GPUEngine(dev)
        eng.set_pipeline(on)
        q0 = time.perf_counter()
        r1 = eng.run_query(heavy)
        cold = time.perf_counter() - q0
        st1 = dev.stats()
        q0 = time.perf_counter()
        r2 = eng.run_query(heavy)
        warm = time.perf_counter() - q0
        st2 = dev.stats()
        q0 = time.perf_counter()
        res = eng.run_plain(mix)
        t_mix = time.perf_counter() - q0
        st3 = dev.stats()
        same = bool(np.array_equal(r1.table, heavy_table) and np.array_equal(r2.table, heavy_table)
                    and _same(res, mix_ref))
        key = "pipeline" if on else "no_pipeline"
        out[key] = {
            "load_s": t_load, "resident_bytes": st1["resident_bytes"],
            "heavy_cold_ms": cold * 1e3, "heavy_cold_segment_loads": st1["segment_loads"],
            "heavy_warm_ms": warm * 1e3, "heavy_warm_segment_loads": st2["segment_loads"] - st1["segment_loads"],
            "mix_s": t_mix, "mix_q_per_s": len(mix) / t_mix,
            "mix_segment_loads": st3["segment_loads"] - st2["segment_loads"],
            "mix_evictions": st3["segment_evictions"] - st2["segment_evictions"],
            "prefetches": st3["prefetches"], "identical": same}
        log(f"framed ({budget / 2**30:.1f} GiB of {gstore.image_nbytes() / 2**30:.1f} GiB), pipeline {on}: "
            f"load {t_load:.1f} s, heavy cold {cold * 1e3:.1f} ms / warm {warm * 1e3:.2f} ms, mix "
            f"{len(mix) / t_mix:.0f} q/s ({out[key]['mix_segment_loads']} frame loads, "
            f"{out[key]['mix_evictions']} evictions, {st3['prefetches']} prefetches), identical={same}")
        eng.close()
        dev.close()
    return out


def table_exchange(dev, seed, rows=1 << 25, cols=4, parts=(8, 64, 1024)):
    """Fork-join split (wkg_table_split) of a rows x cols table of random
    vertex ids by the owner of column 0, for each part count (8: one 8-GPU
    node), HIP events over 5 launches after a warmup: split rows/s and the
    HBM roofline of the split (algorithmic bytes: the table read once and
    written once, 8 * rows * cols) over the count, scan and scatter kernels,
    and of the scatter kernel alone. Each split is checked by what defines
    it: sub-table p holds exactly the input rows whose column-0 cell is p mod
    the part count, in input order, at the offsets the sizes give; and the
    merge of its sub-tables (wkg_table_concat) is the split table."""
    import numpy as np
    rng = np.random.default_rng(seed + 61)
    t = ((1 << 17) + rng.integers(0, 100_000_000, size=(rows, cols), dtype=np.int64)).astype(np.int32)
    d = dev.table(t)
    alg = 8.0 * rows * cols
    out = []
    for n in parts:
        o, _, _ = dev.table_split(d, 0, n)  # warm
        o.free()
        dev.profile(True)
        tw = time.perf_counter()
        for _ in range(5):
            o, prow, phead = dev.table_split(d, 0, n)
            if _ < 4:
                o.free()
        wall = (time.perf_counter() - tw) / 5
        k = {name: dev.profile_read(name) for name in ("split_count", "split_scan", "split_scatter")}
        dev.profile(False)
        views = [dev.table_slice(o, int(h), int(c)) for h, c in zip(phead, prow)]
        merged = dev.table_concat(views).to_host()
        got = o.to_host()
        for v in views:
            v.free()
        o.free()
        owner = t[:, 0].view(np.uint32) % np.uint32(n)
        ok = bool(np.array_equal(merged, got) and int(prow.sum()) == rows and phead[0] == 0 and
                  (phead[1:] == np.cumsum(prow)[:-1]).all())
        for p in range(n) if ok else ():
            ok = ok and np.array_equal(got[int(phead[p]):int(phead[p] + prow[p])], t[owner == p])
        kms = sum(v[1] for v in k.values()) / 5
        sms = k["split_scatter"][1] / max(k["split_scatter"][0], 1)
        rec = {"parts": n, "rows": rows, "cols": cols, "wall_ms": wall * 1e3, "kernels_ms": kms,
               "rows_per_s": rows / (kms * 1e-3), "achieved_GBps": alg / (kms * 1e-3) / 1e9,
               "frac_of_peak": alg / (kms * 1e-3) / 1e9 / HBM_PEAK_GBS,
               "scatter_ms": sms, "scatter_GBps": alg / (sms * 1e-3) / 1e9,
               "per_kernel_ms": {name: v[1] / max(v[0], 1) for name, v in k.items()},
               "correct": ok}
        out.append(rec)
        log(f"table split {rows}x{cols} into {n}: {kms:.3f} ms in kernels ({rec['achieved_GBps']:.0f} GB/s, "
            f"{rec['frac_of_peak']:.2f} of peak; scatter {sms:.3f} ms = {rec['scatter_GBps']:.0f} GB/s), "
            f"{wall * 1e3:.3f} ms per call; correct={ok}")
    d.free()
    return out




if __name__ == "__main__":
    main()
