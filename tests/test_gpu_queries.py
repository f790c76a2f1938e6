"""Whole queries through the C++ engine on the GPU: plain, batched and served
results against the CPU oracle and the committed golden fixtures."""
import json
import os

import numpy as np
import pytest

import oracle
from tests.conftest import VB, random_queries
from wukonggpu_amd import Device, DeviceError, GPUEngine, GStore, Triples

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def eng_small(dev_small):
    e = GPUEngine(dev_small)
    yield e
    e.close()


def _check(gs, q, res):
    want, wex, _ = oracle.run_query(gs, q)
    assert res.status == 0, q
    assert res.table.shape[0] == want.shape[0], q
    if want.size:
        assert np.array_equal(res.table, want), q
    assert res.expanded == wex, q


def test_known_answer_tiny_on_gpu():
    with open(os.path.join(GOLDEN, "tiny_kat.json")) as f:
        kat = json.load(f)
    t = np.array(kat["triples"], dtype=np.uint32)
    gs = GStore.build(Triples.from_arrays(t[:, 0], t[:, 1], t[:, 2]))
    dev = Device(0)
    dev.load(gs)
    eng = GPUEngine(dev)
    for q in kat["queries"]:
        r = eng.run_query(q["cmd"])
        want = np.array(q["rows"], dtype=np.int32)
        assert r.table.shape[0] == len(q["rows"]), q["name"]
        if want.size:
            assert np.array_equal(r.table, want), q["name"]
        assert r.expanded == q["expanded"], q["name"]
    eng.close()
    dev.close()


def test_golden_random_fixture_on_gpu():
    z = np.load(os.path.join(GOLDEN, "random_small.npz"))
    gs = GStore.build(Triples.from_arrays(z["s"], z["p"], z["o"]))
    dev = Device(0)
    dev.load(gs)
    eng = GPUEngine(dev)
    qo, ro = z["q_off"], z["r_off"]
    queries = [z["queries"][qo[i]:qo[i + 1]] for i in range(len(qo) - 1)]
    for mode in ("plain", "batched"):
        res = eng.run_plain(queries) if mode == "plain" else eng.run_batched(queries)
        for i, r in enumerate(res):
            want = z["rows"][ro[i]:ro[i + 1]].reshape(-1, z["cols"][i])
            assert r.status == 0 and r.table.shape[0] == want.shape[0], (mode, i)
            if want.size:
                assert np.array_equal(r.table, want), (mode, i)
            assert r.expanded == z["expanded"][i], (mode, i)
    eng.close()
    dev.close()


def test_random_queries_plain(eng_small, graph_small):
    _, gs = graph_small
    rng = np.random.default_rng(12)
    for q in random_queries(rng, 300, 4, [6, 7], 120, max_steps=5):
        _check(gs, q, eng_small.run_query(q))


def test_random_queries_batched(eng_small, graph_small):
    _, gs = graph_small
    rng = np.random.default_rng(13)
    qs = random_queries(rng, 300, 4, [6, 7], 200, max_steps=5)
    for q, r in zip(qs, eng_small.run_batched(qs)):
        _check(gs, q, r)


def test_served_alone_equals_in_mix(eng_small, graph_small):
    _, gs = graph_small
    rng = np.random.default_rng(14)
    qs = random_queries(rng, 300, 4, [6, 7], 150, max_steps=5)
    arrivals = np.cumsum(rng.exponential(0.0005, size=len(qs)))
    res, wall = eng_small.serve(qs, arrivals)
    for q, r in zip(qs, res):
        alone = eng_small.run_query(q)
        assert r.status == 0 and np.array_equal(r.table, alone.table), q
        assert r.expanded == alone.expanded
        _check(gs, q, r)
    assert wall > 0


def test_engine_rejects_bad_plans(eng_small):
    from wukonggpu_amd import DeviceError
    for q in ([VB, 2, 1, VB + 1], [-1, 2, 1, -2], [2, 0, 1, -1, -1, 2, 1, -3], [2, 5, 1, -1]):
        with pytest.raises(DeviceError):
            eng_small.run_query(q)


def test_chunked_plain_path_matches_oracle(graph_hubs):
    """Outputs above a table limit are produced in row chunks, depth-first,
    and a table budget below a step's needs splits that step's input: every
    answer and work count is the unchunked one (the oracle's)."""
    _, gs = graph_hubs
    dev = Device(0, table_budget_bytes=4 << 20)
    dev.load(gs)
    eng = GPUEngine(dev)
    qs = [[2, 1, 0, -1, -1, 2, 1, -2],                       # 3000-member type index, fan-out
          [2, 1, 0, -1, -1, 2, 1, -2, -2, 3, 0, -3],          # through the hubs
          [2, 1, 0, -1, -1, 2, 1, -2, -2, 2, 0, -1],          # closing known_to_known
          [3, 0, 1, -1, -1, 2, 1, -2, -1, 3, 1, VB + 5],      # predicate index + known_to_const
          [2, 1, 0, -1, -1, 2, 1, -2, -2, 3, 0, -3, -3, 2, 1, -4]]
    rng = np.random.default_rng(31)
    qs += random_queries(rng, 6000, 3, [2, 3], 40, max_steps=4)
    dev.set_table_limit(64 << 10)
    print("plain path with a 64 KiB output table limit under a 4 MiB budget", flush=True)
    for q in qs:
        _check(gs, q, eng.run_query(q))
    assert eng.splits() > 0
    dev.set_table_limit(0)
    print("plain path under the 4 MiB budget alone", flush=True)
    for q in qs:
        _check(gs, q, eng.run_query(q))
    # the executor's arena is too small for most of them: they fall back to the chunked plain path
    eng.configure(max_queries=64, max_patterns=6, arena_bytes=1 << 20)
    dev.set_table_limit(64 << 10)
    print("executor with a 1 MiB arena", flush=True)
    for q, r in zip(qs, eng.run_batched(qs)):
        _check(gs, q, r)
    eng.close()
    dev.close()


def test_probe_tile_sizes_match_oracle(graph_hubs):
    """Every probe tile size of the per-step path (one to eight 64-row rounds
    per wave) gives the oracle's answers, on hub rows that span many tiles."""
    _, gs = graph_hubs
    dev = Device(0)
    dev.load(gs)
    eng = GPUEngine(dev)
    qs = [[2, 1, 0, -1, -1, 2, 1, -2],                       # 3000-member type index, fan-out
          [2, 1, 0, -1, -1, 2, 1, -2, -2, 3, 0, -3],          # through the hubs
          [2, 1, 0, -1, -1, 2, 1, -2, -2, 2, 0, -1],          # closing known_to_known
          [3, 0, 1, -1, -1, 2, 1, -2, -1, 3, 1, VB + 5],      # predicate index + known_to_const
          [2, 1, 0, -1, -1, 2, 1, -2, -2, 3, 0, -3, -3, 2, 1, -4]]
    rng = np.random.default_rng(77)
    qs += random_queries(rng, 6000, 3, [2, 3], 40, max_steps=5)
    for rows in (256, 512, 1024, 2048, 0):
        print(f"probe tile rows {rows}", flush=True)
        dev.set_probe_tile_rows(rows)
        for q in qs:
            _check(gs, q, eng.run_query(q))
    with pytest.raises(Exception):
        dev.set_probe_tile_rows(300)
    eng.close()
    dev.close()


def test_small_steps_match_oracle(graph_hubs):
    """Steps over at most 512 rows run in one launch: the answers equal the
    oracle's with it on and off, including small inputs whose output is
    above the in-launch expand's 2048 rows (hub rows of 5000 edges)."""
    _, gs = graph_hubs
    dev = Device(0)
    dev.load(gs)
    eng = GPUEngine(dev)
    qs = [[VB + 0, 2, 1, -1, -1, 3, 0, -2],                              # hub constant: 5000-row fan-out
          [VB + 0, 2, 1, -1, -1, 2, 0, -2, -2, 3, 1, -3],
          [3, 0, 1, -1, -1, 3, 1, VB + 5],                               # filter on a predicate index
          [VB + 1, 2, 1, -1, -1, 2, 0, -2, -2, 2, 1, -1],                # closing known_to_known
          [VB + 2, 3, 1, -1, -1, 2, 1, -2, -1, 3, 0, -3, -3, 2, 1, -4]]
    rng = np.random.default_rng(79)
    qs += random_queries(rng, 6000, 3, [2, 3], 60, max_steps=5)
    for on in (True, False):
        print(f"small steps {'on' if on else 'off'}", flush=True)
        dev.set_small_steps(on)
        for q in qs:
            _check(gs, q, eng.run_query(q))
    # one step of each kind on 1..512-row tables
    dev.set_small_steps(True)
    for rows in (1, 63, 257, 512, 513):
        t = (VB + rng.integers(0, 6000, size=(rows, 2))).astype(np.int32)
        t[0, 0] = VB  # a hub row
        for cmd, fn in (([-1, 2, 1, -3], lambda d: dev.k2u(d, 0, 2, 1)),
                        ([-1, 3, 0, -2], lambda d: dev.k2k(d, 0, 3, 0, 1)),
                        ([-2, 2, 1, VB + 7], lambda d: dev.k2c(d, 1, 2, 1, VB + 7))):
            want, wex = oracle.run_step(gs, [VB, 2, 1, -1] + cmd, 1, t)
            got, ex = fn(dev.table(t))
            assert np.array_equal(got.to_host(), want) and ex == wex, (rows, cmd)
    eng.close()
    dev.close()


def test_split_launches_match_oracle(graph_hubs):
    """Probe, compaction and expand passes split into launches of 4096 input
    rows (64 expand tiles) give the oracle's answers."""
    _, gs = graph_hubs
    dev = Device(0)
    dev.load(gs)
    eng = GPUEngine(dev)
    qs = [[2, 1, 0, -1, -1, 2, 1, -2],
          [2, 1, 0, -1, -1, 2, 1, -2, -2, 3, 0, -3],
          [2, 1, 0, -1, -1, 2, 1, -2, -2, 2, 0, -1],
          [2, 1, 0, -1, -1, 2, 1, -2, -2, 3, 0, -3, -3, 2, 1, -4, -3, 3, 1, -5]]
    rng = np.random.default_rng(81)
    qs += random_queries(rng, 6000, 3, [2, 3], 30, max_steps=5)
    try:
        for rows in (4096, 0):
            print(f"launch rows {rows}", flush=True)
            dev.set_launch_rows(rows)
            for q in qs:
                _check(gs, q, eng.run_query(q))
        with pytest.raises(Exception):
            dev.set_launch_rows(100)
    finally:
        dev.set_launch_rows(0)
    eng.close()
    dev.close()


def test_lineage_star_arms_match_oracle(graph_hubs):
    """A star's next arm after a fan-out probes once per input row of the
    fan-out (the output keeps which input row each row repeats): the answers
    equal the oracle's with it on and off, and it is taken."""
    _, gs = graph_hubs
    dev = Device(0)
    dev.load(gs)
    eng = GPUEngine(dev)
    qs = [[2, 1, 0, -1, -1, 2, 1, -2, -1, 3, 1, -3],                      # star from ?x
          [2, 1, 0, -1, -1, 2, 1, -2, -2, 3, 0, -3, -3, 2, 1, -4, -3, 3, 1, -5],  # heavy-query shape
          [2, 1, 0, -1, -1, 2, 1, -2, -1, 3, 1, -3, -1, 4, 1, -4],        # three arms
          [VB + 0, 2, 1, -1, -1, 3, 0, -2, -1, 3, 1, -3]]                  # hub constant start
    rng = np.random.default_rng(91)
    qs += random_queries(rng, 6000, 3, [2, 3], 80, max_steps=5)
    dev.profile(True)
    print("lineage on", flush=True)
    for q in qs:
        _check(gs, q, eng.run_query(q))
    n_lineage = dev.profile_read("lineage_steps")[2]
    dev.profile(False)
    assert n_lineage > 0
    dev.set_lineage(False)
    print("lineage off", flush=True)
    for q in qs:
        _check(gs, q, eng.run_query(q))
    eng.close()
    dev.close()


def test_lineage_group_filters_match_oracle(graph_hubs):
    """A filter from a star arm's new column (end column, if any, inherited
    from before the fan-out) decides each (input row, list entry) once and
    reads the decision back for every row holding it: the rows and the
    examined-edge count equal the oracle's, with it on at every size, at its
    default size and off, and it is taken."""
    _, gs = graph_hubs
    dev = Device(0)
    dev.load(gs)
    eng = GPUEngine(dev)
    qs = []
    for pa, pb, pc in [(4, 3, 2), (3, 4, 2), (4, 4, 3)]:
        head = [2, 1, 0, -1, -1, pa, 1, -2, -2, pb, 0, -3, -3, pb, 1, -4, -3, pc, 1, -5]
        qs += [head + [-5, pb, 0, -3],        # known_to_known back to the arm's start
               head + [-5, pa, 0, -2],        # end column from before the fan-out
               head + [-5, pc, 1, -1],
               head + [-5, pb, 0, -4],        # end column of the fan-out itself: per row
               head + [-5, pb, 0, VB + 7],    # known_to_const
               head + [-5, pc, 1, VB + 5999]]
    qs += [[2, 1, 0, -1, -1, 2, 1, -2, -1, 3, 1, -3, -3, 2, 0, -1],       # star from ?x, closed on ?x
           [2, 1, 0, -1, -1, 2, 1, -2, -1, 3, 1, -3, -3, 4, 0, -1, -3, 3, 0, -1]]  # then a second filter
    want = [oracle.run_query(gs, q)[:2] for q in qs]
    # (group filter rows, lineage, launch rows): 4096 input rows per launch
    # splits the group pass into launches that start past tile 0
    try:
        for rows, lineage, launch in [(0, True, 0), (0, True, 4096), (1 << 22, True, 0), (0, False, 0)]:
            dev.set_lineage(lineage)
            dev.set_group_filter_rows(rows)
            dev.set_launch_rows(launch)
            dev.profile(True)
            print("group filter rows", rows, "lineage", lineage, "launch rows", launch, flush=True)
            for q, (wt, wex) in zip(qs, want):
                res = eng.run_query(q)
                assert res.status == 0, q
                assert res.table.shape[0] == wt.shape[0], q
                assert np.array_equal(res.table.reshape(wt.shape), wt), q
                assert res.expanded == wex, q
            taken = dev.profile_read("lineage_filters")[2]
            dev.profile(False)
            if rows == 0 and lineage:
                assert taken > 0
            elif not lineage:
                assert taken == 0
    finally:
        dev.set_launch_rows(0)
        dev.set_group_filter_rows(1 << 22)
        dev.set_lineage(True)
    eng.close()
    dev.close()


@pytest.fixture(scope="module")
def graph_mid():
    from tests.conftest import small_graph
    # 11.5 MB image over 12 predicates and 3 types; its largest predicate
    # segment fits a 4 MiB graph budget, all of them together do not
    return small_graph(seed=12, n=20000, preds=12, triples=400000, s=1.0, cap=0, types=(5000, 300, 40), hubs=1,
                       hub_degree=20000)


def test_framed_residency_matches_oracle(graph_mid):
    """A graph budget below the image: per-predicate segments are loaded into
    power-of-two frames on demand and least recently used ones are evicted;
    every step kind, and the batched and served paths, still give the
    oracle's answers."""
    _, gs = graph_mid
    dev = Device(0)
    dev.load(gs, graph_budget_bytes=4 << 20)
    st = dev.stats()
    assert 0 < st["resident_bytes"] <= 4 << 20 < gs.image_nbytes()
    eng = GPUEngine(dev)
    rng = np.random.default_rng(41)
    qs = random_queries(rng, 20000, 12, [14, 15, 16], 90, max_steps=4)
    print("framed: plain path", flush=True)
    for q in qs:
        _check(gs, q, eng.run_query(q))
    st = dev.stats()
    assert st["segment_loads"] > 0 and st["segment_evictions"] > 0, st
    print("framed: step rounds", flush=True)
    for q, r in zip(qs, eng.run_rounds(qs)):
        _check(gs, q, r)
    print("framed: device-resident executor (batches whose predicates do not fit run on the plain path)",
          flush=True)
    eng.configure(max_queries=16, max_patterns=6, arena_bytes=64 << 20)
    for q, r in zip(qs, eng.run_batched(qs)):
        _check(gs, q, r)
    print("framed: served (heavy queries on a heavy executor lane, not the worker thread)", flush=True)
    heavy = np.array([len(q) >= 12 for q in qs], dtype=np.uint8)
    arrivals = np.cumsum(rng.exponential(0.0002, size=len(qs)))
    res, _ = eng.serve(qs, arrivals, heavy=heavy, table_budget_bytes=256 << 20, light_lanes=2)
    for q, r in zip(qs, res):
        _check(gs, q, r)
    eng.close()
    dev.close()


def test_framed_prefetch_and_preload(graph_mid):
    """Pipelined segment loads (a later pattern's predicate loads on the copy
    stream while a step runs, core/rcache.hpp:273-281) and start-up preloads
    (core/rcache.hpp:191-217) on a framed graph: the answers stay the
    oracle's with the pipeline on and off; a prefetched predicate is resident
    at once and the one in use is never its victim; a preloaded predicate is
    resident; on a resident image all three are no-ops."""
    _, gs = graph_mid
    dev = Device(0)
    dev.load(gs, graph_budget_bytes=4 << 20)
    eng = GPUEngine(dev)
    rng = np.random.default_rng(43)
    qs = random_queries(rng, 20000, 12, [14, 15, 16], 60, max_steps=4)
    for on in (True, False):
        print(f"framed: plain path, pipeline {on}", flush=True)
        eng.set_pipeline(on)
        before = dev.stats()["prefetches"]
        for q in qs:
            _check(gs, q, eng.run_query(q))
        made = dev.stats()["prefetches"] - before
        assert (made > 0) if on else (made == 0), (on, made)
    print("framed: explicit prefetches", flush=True)
    preds = list(range(2, 14))
    fetched = 0
    for in_use, p in zip(preds, preds[1:] + preds[:1]):
        dev.ensure(in_use)
        try:
            dev.prefetch(p, in_use=in_use)
        except DeviceError as e:  # the two segments do not fit the budget together
            assert e.rc == -2, e
            assert dev.resident(in_use) and not dev.resident(p), (p, in_use)
            continue
        fetched += 1
        assert dev.resident(p) and dev.resident(in_use), (p, in_use)
        dev.ensure(p)  # waits for the copies on the device
    assert fetched > 0
    for q in qs[:20]:
        _check(gs, q, eng.run_query(q))
    print("framed: preloads", flush=True)
    dev.preload()  # every predicate, later ones evicting earlier ones
    for p in preds:
        dev.preload([p])
        assert dev.resident(p)
    for q in qs[:20]:
        _check(gs, q, eng.run_query(q))
    eng.close()
    dev.close()
    dev = Device(0)
    dev.load(gs)
    dev.prefetch(3, in_use=2)
    dev.preload([2, 3])
    assert dev.resident(3) and dev.stats()["prefetches"] == 0
    dev.close()


def test_long_index_lists_resident(graph_mid):
    """Predicate-index lists longer than one probe-and-copy launch (8192
    entries) on a resident graph: each list and each query that starts from
    one, against the oracle."""
    _, gs = graph_mid
    dev = Device(0)
    dev.load(gs)
    assert dev.stats()["resident_bytes"] >= gs.image_nbytes()
    eng = GPUEngine(dev)
    longest = 0
    for pid in range(2, 14):
        for d in (0, 1):
            want, wex, _ = oracle.run_query(gs, [pid, 0, d, -1])
            longest = max(longest, want.shape[0])
            t, ex = dev.i2u(pid, d)
            assert np.array_equal(t.to_host(), want) and ex == wex, (pid, d)
            t.free()
            q = [pid, 0, d, -1, -1, 2 + (pid % 12), 1, -2]
            _check(gs, q, eng.run_query(q))
    assert longest > 8192
    eng.close()
    dev.close()


def test_fanout_filters_match_oracle(graph_hubs):
    """A known_to_known / known_to_const from a fan-out's new column (end
    column from before the fan-out, or a constant) is decided from the
    fan-out's lineage: each distinct list of its parents is probed once and
    every row tests its parent's target against its entry's list. Rows and
    work counts equal the oracle's, with it taken at every size, with split
    launches, in HBM-budgeted chunks (fan-outs of row slices) and off."""
    _, gs = graph_hubs
    dev = Device(0)
    dev.load(gs)
    eng = GPUEngine(dev)
    head = [7, 1, 0, -1, -1, 2, 1, -2, -2, 2, 0, -3, -3, 2, 1, -4]   # ... ?3 p2 ?4: the hub fan-out (675k rows)
    qs = [head + [-4, 3, 1, -2],            # triangle-style close on an inherited column
          head + [-4, 3, 0, -3],            # IN lists, end column the fan-out's parent column
          head + [-4, 2, 1, -1],
          head + [-4, 3, 1, VB + 17],       # known_to_const
          head + [-4, 2, 0, VB + 5999],
          head + [-4, 3, 1, -2, -3, 2, 1, VB + 3],   # then a per-row filter on the small result
          head + [-4, 2, 1, -4],            # end column the new column itself: per row
          [VB + 0, 2, 1, -1, -1, 2, 0, -2, -2, 3, 1, -3, -3, 3, 0, -1],  # constant start
          [6, 1, 0, -1, -1, 3, 1, -2, -2, 2, 1, -3, -3, 4, 1, -1]]
    # short lists shared by many parents (the hub repeated 5000 times with a
    # 2-entry list): row tiles instead of parent-aligned units
    hub = max(gs.get_edges(0, 2, 0).tolist(), key=lambda v: gs.degree(v, 2, 1))
    qs += [[hub, 2, 1, -1, -1, 2, 0, -2, -2, 3, 1, -3, -3, 3, 1, -1],
           [hub, 2, 1, -1, -1, 2, 0, -2, -2, 3, 1, -3, -3, 2, 0, -2]]
    want = [oracle.run_query(gs, q)[:2] for q in qs]
    try:
        for rows, lineage, launch, limit, fuse, blocks in [
                (0, True, 0, 0, True, 1), (0, True, 4096, 0, True, 1), (0, True, 0, 4 << 20, True, 1),
                (1 << 22, True, 0, 0, True, 1), (0, False, 0, 0, True, 1), (0, True, 0, 0, False, 1),
                (0, True, 0, 4 << 20, False, 1), (0, True, 0, 0, True, 0), (0, True, 4096, 0, True, 2)]:
            dev.set_lineage(lineage)
            dev.set_group_filter_rows(rows)
            dev.set_launch_rows(launch)
            dev.set_table_limit(limit)
            eng.set_fused_filters(fuse)
            # blocks: 1 the block filter, 0 one parent per workgroup, 2 the
            # block filter with a 3-row kept list (it overflows and reruns)
            dev.set_fanout_blocks(blocks != 0, 3 if blocks == 2 else 0)
            dev.profile(True)
            print("fan-out filters: group filter rows", rows, "lineage", lineage, "launch rows", launch,
                  "table limit", limit, "fused with the fan-out", fuse, "blocks", blocks, flush=True)
            for q, (wt, wex) in zip(qs, want):
                res = eng.run_query(q)
                assert res.status == 0, q
                assert res.table.shape[0] == wt.shape[0], q
                assert np.array_equal(res.table.reshape(wt.shape), wt), q
                assert res.expanded == wex, q
            taken = dev.profile_read("fanout_filters")[2]
            by_units, by_rows = dev.profile_read("fanout_units")[2], dev.profile_read("fanout_row_tiles")[2]
            fused = dev.profile_read("fused_fanouts")[2]
            overflow = dev.profile_read("fanout_kept_overflow")[2]
            dev.profile(False)
            assert (overflow > 0) if blocks == 2 else (overflow == 0), overflow
            if rows == 0 and lineage:
                assert taken > 0
                if limit == 0:
                    assert by_units > 0 and by_rows > 0, (by_units, by_rows)
                assert (fused > 0) if fuse else (fused == 0), fused
            elif not lineage:
                assert taken == 0 and fused == 0
    finally:
        dev.set_launch_rows(0)
        dev.set_group_filter_rows(1 << 22)
        dev.set_lineage(True)
        dev.set_table_limit(0)
        eng.set_fused_filters(True)
        dev.set_fanout_blocks(True)
    eng.close()
    dev.close()


def test_k2u_filter_call(graph_hubs):
    """wkg_step_k2u_filter against the two calls it stands for and the
    oracle: fused (the fan-out never written) for a large fan-out, not fused
    for a small input (the known_to_unknown output comes back and the filter
    runs on it), with known_to_known and known_to_const filters."""
    from wukonggpu_amd import STEP_K2C, STEP_K2K
    _, gs = graph_hubs
    dev = Device(0)
    dev.load(gs)
    dev.set_group_filter_rows(0)
    try:
        head = [7, 1, 0, -1, -1, 2, 1, -2, -2, 2, 0, -3]
        t1, _ = dev.i2u(7, 0)
        t2, _ = dev.k2u(t1, 0, 2, 1)
        t3, _ = dev.k2u(t2, 1, 2, 0)  # ?1 ?2 ?3: the fan-out's input (hub rows)
        _, wex_head, _ = oracle.run_query(gs, head)
        _, wex_fan, _ = oracle.run_query(gs, head + [-3, 2, 1, -4])
        cases = [(STEP_K2K, 3, 1, 1, -1, [-4, 3, 1, -2]),
                 (STEP_K2K, 3, 0, 2, -1, [-4, 3, 0, -3]),
                 (STEP_K2C, 3, 1, -1, VB + 17, [-4, 3, 1, VB + 17])]
        for kind, fp, fd, ec, cv, tail in cases:
            q = head + [-3, 2, 1, -4] + tail
            want, wex, _ = oracle.run_query(gs, q)
            out, fused, e1, e2 = dev.k2u_filter(t3, 2, 2, 1, kind, fp, fd, ec, cv)
            assert fused, (kind, fp, fd)
            assert np.array_equal(out.to_host().reshape(want.shape), want), q
            # the work counts of the two steps, as the oracle counts them
            assert e1 == wex_fan - wex_head and e1 + e2 == wex - wex_head, (e1, e2)
            out.free()
        # a small input: not fused, the known_to_unknown output comes back
        small, _ = dev.c2u(VB + 1, 2, 1)
        out, fused, e1, e2 = dev.k2u_filter(small, 0, 2, 1, STEP_K2C, 3, 1, -1, VB + 7)
        assert not fused and e2 == 0
        want, wex, _ = oracle.run_query(gs, [VB + 1, 2, 1, -1, -1, 2, 1, -2])
        assert np.array_equal(out.to_host().reshape(want.shape), want)
        _, wex1, _ = oracle.run_query(gs, [VB + 1, 2, 1, -1])
        assert e1 == wex - wex1
    finally:
        dev.set_group_filter_rows(1 << 22)
    dev.close()


def test_thread_contexts_run_queries_at_once(graph_hubs):
    """Four host threads, each with a stream and read-back block of its own
    (wkg_thread_context_begin), run queries on one context at the same time
    (fan-outs, lineage filters, small queries); every answer equals the
    oracle's. A thread cannot begin twice or end a context it has not
    begun."""
    import threading
    _, gs = graph_hubs
    dev = Device(0)
    dev.load(gs)
    eng = GPUEngine(dev)
    rng = np.random.default_rng(96)
    qs = random_queries(rng, 6000, 3, [5, 6, 7], 160, max_steps=5)
    head = [7, 1, 0, -1, -1, 2, 1, -2, -2, 2, 0, -3, -3, 2, 1, -4]
    qs += [head + [-4, 3, 1, -2], head + [-4, 3, 1, VB + 17]] * 4
    want = [oracle.run_query(gs, q)[:2] for q in qs]
    errors = []

    def work(part):
        try:
            dev.thread_context_begin()
            try:
                with pytest.raises(DeviceError):
                    dev.thread_context_begin()
                for i in part:
                    r = eng.run_query(qs[i])
                    wt, wex = want[i]
                    assert r.status == 0 and r.table.shape[0] == wt.shape[0], qs[i]
                    assert np.array_equal(r.table.reshape(wt.shape), wt) and r.expanded == wex, qs[i]
            finally:
                dev.thread_context_end()
        except BaseException as ex:  # reported below, on the test's thread
            errors.append(ex)

    ths = [threading.Thread(target=work, args=(range(k, len(qs), 4),)) for k in range(4)]
    for t in ths:
        t.start()
    for t in ths:
        t.join()
    assert not errors, errors[0]
    with pytest.raises(DeviceError):
        dev.thread_context_end()
    eng.close()
    dev.close()


def test_dataset_directory_and_persisted_image(tmp_path):
    """A graph read from a dataset directory in the reference's ID format
    (id_* files, str_index, str_normal), built, saved as an image file and
    loaded back, then loaded on the GPU: queries give the oracle's answers
    on the directly built image."""
    from tests.conftest import small_graph
    tr, gs = small_graph(seed=33, n=20000, preds=6, triples=300000, s=1.0, cap=0, types=(4000, 300, 30),
                         hubs=1, hub_degree=8000)
    d = tmp_path / "ds"
    tr.write_dataset(d, files=4)
    back, info = Triples.read_dataset(d)
    assert info["files"] == 4 and info["has_str_index"] and info["has_str_normal"]
    img = tmp_path / "graph.img"
    GStore.build(back).save(img)
    loaded = GStore.load(img)
    dev = Device(0)
    dev.load(loaded)
    eng = GPUEngine(dev)
    rng = np.random.default_rng(35)
    qs = random_queries(rng, 20000, 6, [8, 9, 10], 80, max_steps=5)
    print("queries on the file-loaded, persisted image", flush=True)
    for q in qs:
        _check(gs, q, eng.run_query(q))
    for q, r in zip(qs, eng.run_batched(qs)):
        _check(gs, q, r)
    eng.close()
    dev.close()


def test_small_queries_in_one_launch(graph_hubs):
    """Index- and constant-start queries whose tables stay within 4096 input
    rows a step and 8192 cells run all their steps in one launch
    (wkg_query_small, one workgroup with the tables in LDS); larger ones are
    declined and run step by step. The answers equal the oracle's either
    way, including filters over 3000 rows (three probe rounds a wave),
    empty tables and hub lists inside the limits."""
    _, gs = graph_hubs
    dev = Device(0)
    dev.load(gs)
    eng = GPUEngine(dev)
    hub = max(gs.get_edges(0, 2, 0).tolist(), key=lambda v: gs.degree(v, 2, 1))
    qs = [[5, 1, 0, -1, -1, 2, 1, -2],                        # 3000 members -> ~10k x 2: declined
          [5, 1, 0, -1, -1, 2, 1, VB + 7],                     # 3000 rows, known_to_const
          [5, 1, 0, -1, -1, 3, 0, -1],                         # 3000 rows, known_to_known on one column
          [7, 1, 0, -1, -1, 2, 1, -2, -2, 3, 0, -3],          # a chain
          [7, 1, 0, -1, -1, 2, 1, -2, -1, 3, 1, -3, -3, 2, 0, -1],  # a star, closed
          [VB + 1, 2, 1, -1, -1, 3, 1, VB + 7],                # constant start, known_to_const
          [VB + 99999, 2, 1, -1, -1, 3, 1, -2],                # a constant with no edges: empty
          [hub, 2, 1, -1, -1, 3, 1, -2],                       # a 5000-entry list: declined
          [2, 0, 1, -1, -1, 3, 1, -2]]                         # a predicate index
    rng = np.random.default_rng(93)
    qs += random_queries(rng, 6000, 3, [5, 6, 7], 80, max_steps=5)
    for on, server in ((True, 4), (True, 0), (False, 4)):
        eng.set_small_queries(on)
        dev.set_small_query_server(server)
        dev.profile(True)
        print("small queries", "on" if on else "off", "server workgroups", server, flush=True)
        for q in qs:
            _check(gs, q, eng.run_query(q))
        taken, declined = dev.profile_read("small_query_taken")[2], dev.profile_read("small_query_declined")[2]
        dev.profile(False)
        if on:
            assert taken > 0 and declined > 0, (taken, declined)
        else:
            assert taken == 0 and declined == 0
    eng.set_small_queries(True)
    dev.set_small_query_server(4)
    eng.close()
    dev.close()


def test_small_query_server_tickets(graph_hubs, graph_small):
    """The persistent server over many tickets in flight: up to all 64
    slots submitted at once and collected out of order, a full ring
    refusing one more, collection without waiting, a server that went idle
    and exited (the next ticket launches it again), one workgroup and
    eight, and a context destroyed while its server runs (it is stopped),
    then a new context on another graph. Every answer equals the oracle's;
    declined tickets are the ones the one-launch path declines."""
    import time
    from wukonggpu_amd import plan_of
    _, gs = graph_hubs
    dev = Device(0)
    dev.load(gs)
    rng = np.random.default_rng(94)
    qs = random_queries(rng, 6000, 3, [5, 6, 7], 80, max_steps=5)
    qs = [q for q in qs if q[0] > 1][:200]  # index starts

    def check(q, res):
        st, table, ex = res
        want, wex, _ = oracle.run_query(gs, q)
        if st == "declined":
            dev.set_small_query_server(0)
            st1 = dev.query_small(plan_of(q))[0]
            dev.set_small_query_server(wg)
            assert st1 == "declined", q
            return 0
        assert st == "ok", (q, st)
        assert table.shape == want.shape and np.array_equal(table, want), q
        assert ex == wex, (q, ex, wex)
        return 1

    for wg in (1, 8):
        dev.set_small_query_server(wg)
        print("server workgroups", wg, flush=True)
        ok = 0
        for start in range(0, len(qs), 64):
            chunk = qs[start:start + 64]
            tickets = [dev.query_small_submit(plan_of(q)) for q in chunk]
            assert all(t is not None for t in tickets)
            if len(chunk) == 64:
                assert dev.query_small_submit(plan_of(chunk[0])) is None  # the ring is full
            order = list(range(len(chunk)))
            rng.shuffle(order)
            for i in order:
                res = dev.query_small_collect(tickets[i], wait=bool(i % 2))
                while res[0] == "pending":
                    res = dev.query_small_collect(tickets[i], wait=False)
                ok += check(chunk[i], res)
        assert ok > len(qs) // 2, ok
        time.sleep(0.01)  # every workgroup idles out
        ok += check(qs[0], dev.query_small(plan_of(qs[0])))
    # a start list past the limits: declined by the server once, then at once
    from tests.conftest import small_graph
    _, gs_big = small_graph(seed=12, n=20000, preds=2, triples=40000, types=(10000,))
    dev2 = Device(0)
    dev2.load(gs_big)
    q_big = plan_of([4, 1, 0, -1])  # the 10000-member type's index
    dev2.profile(True)
    assert dev2.query_small(q_big)[0] == "declined"
    n1 = dev2.profile_read("small_query")[0]
    assert dev2.query_small(q_big)[0] == "declined"
    with pytest.raises(DeviceError) as ei:  # submitting it is declined at once as well
        dev2.query_small_submit(q_big)
    assert ei.value.rc == -4
    n2 = dev2.profile_read("small_query")[0]
    dev2.profile(False)
    assert n1 == 1 and n2 == 1, (n1, n2)  # no second run on the device
    eng2 = GPUEngine(dev2)
    want, wex, _ = oracle.run_query(gs_big, [4, 1, 0, -1])
    r = eng2.run_query([4, 1, 0, -1])
    assert np.array_equal(r.table, want) and r.expanded == wex
    eng2.close()
    dev2.close()
    # a context destroyed while its server runs, then a new one on another graph
    dev.query_small(plan_of(qs[1]))
    dev.close()
    _, gs = graph_small
    dev = Device(0)
    dev.load(gs)
    for q in random_queries(np.random.default_rng(95), 300, 4, [6, 7], 40, max_steps=4):
        check(q, dev.query_small(plan_of(q)))
    dev.close()
