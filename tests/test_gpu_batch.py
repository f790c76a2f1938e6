"""Device-resident batched executor (wkg_batch_*) and the served lanes.

Every query's table and work count must equal the plain per-step path's and
the CPU oracle's, whatever the launch/tile shape, the executor's capacity or
arena, and however often the captured launch sequence is replayed."""
import numpy as np
import pytest

import oracle
from tests.conftest import VB, random_queries
from wukonggpu_amd import (STEP_C2U, STEP_I2U, STEP_K2C, STEP_K2K, STEP_K2U, Batch, DeviceError, GPUEngine)

pytestmark = pytest.mark.gpu

SHAPES = [dict(probe_tile_rows=tr, expand_tile_items=ni, probe_wg_per_cu=pw, expand_wg_per_cu=xw, aux_wg_per_cu=aw,
               schedule=sc)
          for tr, ni, pw, xw, aw, sc in ((1024, 1024, 1, 1, 1, 0), (1024, 2048, 2, 4, 2, 0), (2048, 1024, 4, 2, 1, 0),
                                         (2048, 2048, 2, 2, 1, 0), (1024, 512, 2, 8, 1, 0), (2048, 512, 4, 8, 2, 0),
                                         (1024, 512, 3, 4, 1, 1), (2048, 1024, 2, 8, 2, 1))]


@pytest.fixture(scope="module")
def eng_small(dev_small):
    e = GPUEngine(dev_small)
    yield e
    e.close()


def _same(gs, q, r):
    want, wex, _ = oracle.run_query(gs, q)
    assert r.status == 0, (q, r.status)
    assert r.table.shape[0] == want.shape[0], q
    if want.size:
        assert np.array_equal(r.table, want), q
    assert r.expanded == wex, q


def test_executor_random_queries_match_oracle(eng_small, graph_small):
    _, gs = graph_small
    rng = np.random.default_rng(21)
    qs = random_queries(rng, 300, 4, [6, 7], 300, max_steps=6)
    eng_small.configure(max_queries=64, max_patterns=8, arena_bytes=1 << 30)  # several launches
    print("executor: 300 random queries, 64 per launch", flush=True)
    for q, r in zip(qs, eng_small.run_batched(qs)):
        _same(gs, q, r)


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: f"tr{s['probe_tile_rows']}-nv{s['expand_tile_items']}"
                                                        f"-{'static' if s['schedule'] else 'queue'}")
def test_executor_shapes_match_plain(dev_small, graph_small, shape):
    _, gs = graph_small
    rng = np.random.default_rng(22)
    qs = random_queries(rng, 300, 4, [6, 7], 120, max_steps=5)
    e = GPUEngine(dev_small)
    b = Batch(dev_small, 128, 6, 1 << 30)
    b.set_tuning(**shape)
    assert b.tuning() == shape
    plans = [_plan(q) for q in qs]
    print(f"executor shape {shape}", flush=True)
    for q, (st, tab, ex) in zip(qs, b.run(plans)):
        alone = e.run_query(q)
        assert st == 0, q
        assert tab.shape == alone.table.shape and np.array_equal(tab, alone.table), q
        assert ex == alone.expanded, q
    b.close()
    e.close()


def test_executor_hub_rows_and_big_index(graph_hubs):
    from wukonggpu_amd import Device
    _, gs = graph_hubs
    dev = Device(0)
    dev.load(gs)
    e = GPUEngine(dev)
    e.configure(max_queries=32, max_patterns=6, arena_bytes=1 << 30)
    qs = [[2, 1, 0, -1, -1, 2, 1, -2],                       # 3000-member type index, fan-out
          [2, 1, 0, -1, -1, 2, 1, -2, -2, 3, 0, -3],          # through the hubs
          [VB + 0, 2, 1, -1, -1, 3, 0, -2],                    # hub constant start
          [2, 1, 0, -1, -1, 2, 1, -2, -2, 2, 0, -1],          # closing known_to_known
          [3, 0, 1, -1, -1, 2, 1, -2, -1, 3, 1, VB + 5]]      # predicate index + known_to_const
    rng = np.random.default_rng(23)
    qs += random_queries(rng, 6000, 3, [2, 3], 60, max_steps=4)
    print("executor on the hub graph", flush=True)
    for q, r in zip(qs, e.run_batched(qs)):
        _same(gs, q, r)
    e.close()
    dev.close()


def test_executor_arena_exhaustion_falls_back(dev_small, graph_small):
    _, gs = graph_small
    rng = np.random.default_rng(24)
    qs = random_queries(rng, 300, 4, [6, 7], 100, max_steps=5)
    e = GPUEngine(dev_small)
    e.configure(max_queries=100, max_patterns=6, arena_bytes=1 << 20)  # 1 MiB: most queries do not fit
    b = Batch(dev_small, 100, 6, 1 << 20)
    recs = b.run([_plan(q) for q in qs])
    assert any(st == -2 for st, _, _ in recs), "expected some queries to exceed a 1 MiB arena"
    assert any(st == 0 for st, _, _ in recs)
    b.close()
    print("executor with a 1 MiB arena (fallback to the plain path)", flush=True)
    for q, r in zip(qs, e.run_batched(qs)):
        _same(gs, q, r)
    e.close()


def test_batch_replays_with_eager_work_between(dev_small, graph_small, eng_small):
    """The captured sequence replayed several times, with plain-path work and
    synchronisations between replays, and a smaller second batch."""
    _, gs = graph_small
    cmds = [[6, 1, 0, -1, -1, 2, 1, -2],
            [VB + 5, 3, 1, -1, -1, 2, 0, -2, -2, 2, 1, -1],
            [2, 0, 1, -1, -1, 3, 1, VB + 7],
            [7, 1, 0, -1, -1, 3, 1, -2, -2, 4, 0, -3, -1, 2, 1, -3]]
    plans = [_plan(c) for c in cmds]
    want = [eng_small.run_query(c) for c in cmds]
    b = Batch(dev_small, 8, 4, 64 << 20)
    for rep in range(4):
        n = len(cmds) if rep % 2 == 0 else 2
        print(f"replay {rep}: {n} queries", flush=True)
        got = b.run(plans[:n])
        for i in range(n):
            st, tab, ex = got[i]
            assert st == 0 and np.array_equal(tab, want[i].table) and ex == want[i].expanded, (rep, i)
        eng_small.run_query(cmds[rep % len(cmds)])  # eager work on another stream between replays
    b.close()


def test_result_buffer_overflow_then_growth(dev_small, eng_small):
    """Results beyond the host result buffer end in ENOMEM (in query order);
    the buffer grows for the next launch, which then answers every query."""
    rng = np.random.default_rng(27)
    qs = random_queries(rng, 300, 4, [6, 7], 80, max_steps=4)
    want = [eng_small.run_query(q) for q in qs]
    total = sum(w.table.nbytes for w in want)
    assert total > 16384
    b = Batch(dev_small, 80, 6, 256 << 20)
    b.reserve_results(8192)
    print("result buffer of 8 KiB", flush=True)
    got = b.run([_plan(q) for q in qs])
    st = [g[0] for g in got]
    assert -2 in st and 0 in st
    first_fail = st.index(-2)
    for i, (s_, tab, ex) in enumerate(got):
        if s_ == 0:
            assert np.array_equal(tab, want[i].table) and ex == want[i].expanded, i
    peak, rbytes = b.stats()
    assert peak > 0 and rbytes <= 8192
    got = b.run([_plan(q) for q in qs])  # grown
    for i, (s_, tab, ex) in enumerate(got):
        assert s_ == 0 and np.array_equal(tab, want[i].table) and ex == want[i].expanded, i
    assert first_fail >= 0
    b.close()


def test_batch_rejects_bad_plans(dev_small):
    b = Batch(dev_small, 4, 3, 16 << 20)
    bad = [
        [[dict(kind=STEP_K2U, pid=2, dir=1, start_col=0)]],                        # no start
        [[dict(kind=STEP_I2U, pid=6, dir=0), dict(kind=STEP_K2U, pid=2, dir=1, start_col=1)]],  # column 1 of 1
        [[dict(kind=STEP_I2U, pid=6, dir=0), dict(kind=STEP_I2U, pid=6, dir=0)]],  # index start at step 1
        [[dict(kind=STEP_I2U, pid=6, dir=0)] + [dict(kind=STEP_K2U, pid=2, dir=1, start_col=0)] * 3],  # 4 > 3
        [[dict(kind=STEP_C2U, pid=2, dir=1, vid=-5)]],                             # constant out of range
        [[dict(kind=STEP_I2U, pid=6, dir=2)]],                                     # direction
        [[dict(kind=STEP_I2U, pid=6, dir=0)]] * 5,                                 # 5 > 4 queries
    ]
    for plans in bad:
        with pytest.raises(DeviceError):
            b.run(plans)
    ok = b.run([[dict(kind=STEP_I2U, pid=6, dir=0)]])  # still usable
    assert ok[0][0] == 0
    b.close()


def test_served_lanes_alone_equals_in_mix(eng_small, graph_small):
    _, gs = graph_small
    rng = np.random.default_rng(25)
    qs = random_queries(rng, 300, 4, [6, 7], 200, max_steps=6)
    heavy = np.array([len(q) >= 20 for q in qs], dtype=np.uint8)
    arrivals = np.cumsum(rng.exponential(0.0002, size=len(qs)))
    eng_small.configure(max_queries=256, max_patterns=8, arena_bytes=1 << 30)
    # light queries through the small-query server (admitted one by one),
    # then through the light lanes only
    for wg in (16, 0):
        eng_small.set_serve_small(wg)
        print(f"served: 200 arrivals, small-query server {wg} workgroups, 2 light lanes + heavy lane", flush=True)
        res, wall = eng_small.serve(qs, arrivals, heavy=heavy, table_budget_bytes=2 << 30, light_lanes=2)
        assert wall > 0
        st = eng_small.serve_stats()
        assert (st["small_answered"] > 50) if wg else (st["small_answered"] == 0), st
        for q, r in zip(qs, res):
            _same(gs, q, r)  # the oracle's answer
            alone = eng_small.run_query(q)
            assert r.status == 0 and r.table.shape == alone.table.shape and np.array_equal(r.table, alone.table), q
            assert r.expanded == alone.expanded, q
            assert r.latency_s > 0
    eng_small.set_serve_small(16)
    # a budget too small for most batches still answers every query
    res2, _ = eng_small.serve(qs[:60], arrivals[:60] * 0, heavy=heavy[:60], table_budget_bytes=4 << 20,
                              light_lanes=2)
    for q, r in zip(qs[:60], res2):
        _same(gs, q, r)


def test_closed_loop_batch_mode(eng_small, graph_small):
    """The reference's batch driver semantics: light and heavy queries issued
    in order below per-class in-flight caps for a duration; every issued
    query completes with the answer it gets alone, none is issued after the
    duration, and the caps hold (a class never has more queries issued than
    completed plus its cap)."""
    _, gs = graph_small
    rng = np.random.default_rng(27)
    qs = random_queries(rng, 300, 4, [6, 7], 3000, max_steps=6)
    heavy = np.array([len(q) >= 20 for q in qs], dtype=np.uint8)
    eng_small.configure(max_queries=256, max_patterns=8, arena_bytes=1 << 30)
    print("closed loop: caps 16 light, 2 heavy, 0.25 s", flush=True)
    res, issue, wall = eng_small.serve_closed(qs, heavy, 16, 2, 0.25, table_budget_bytes=2 << 30, light_lanes=2)
    issued = issue >= 0
    assert issued.sum() > 0 and wall > 0
    assert (issue[issued] < 0.25).all()
    # the loop ends at the duration, or earlier once every query was issued
    assert issued.all() or wall >= 0.25
    # in order within a class: an issued query's class predecessors were issued
    for cls in (0, 1):
        f = issued[heavy == cls]
        assert not (f[1:] & ~f[:-1]).any()
    for q, r, f in zip(qs, res, issued):
        if not f:
            continue
        _same(gs, q, r)  # the oracle's answer
        alone = eng_small.run_query(q)
        assert r.status == 0 and r.table.shape == alone.table.shape and np.array_equal(r.table, alone.table), q
        assert r.expanded == alone.expanded and r.latency_s > 0
    # in-flight caps: at any issue time, earlier-issued queries of the class
    # that had not completed yet number fewer than the cap
    done_t = issue + np.array([r.latency_s for r in res])
    for cls, cap in ((0, 16), (1, 2)):
        idx = np.where(issued & (heavy == cls))[0]
        for i in idx:
            others = idx[(issue[idx] < issue[i]) & (done_t[idx] > issue[i])]
            assert len(others) < cap + 1, (cls, len(others))
    # a duration must be positive
    with pytest.raises(DeviceError):
        eng_small.serve_closed(qs[:50], heavy[:50], 16, 2, 0.0, table_budget_bytes=2 << 30, light_lanes=2)
    # a send interval: issue rounds at least 20 ms apart, so at most 6 in 0.1 s
    res_i, issue_i, _ = eng_small.serve_closed(qs, heavy, 16, 2, 0.1, table_budget_bytes=2 << 30, light_lanes=2,
                                               interval_s=0.02)
    rounds = np.unique(issue_i[issue_i >= 0])
    assert 1 <= len(rounds) <= 6 and (np.diff(rounds) >= 0.02 - 1e-9).all(), rounds
    assert (issue_i >= 0).sum() <= len(rounds) * 18  # each round fills at most the free slots of both caps


def test_closed_loop_template_stream(eng_small, graph_small):
    """The benchmark's batch-mode stream, packed, through the closed loop with
    only some result tables kept: kept tables equal the query run alone, the
    rest report the same row counts."""
    from wukonggpu_amd import workloads
    _, gs = graph_small
    preds = gs.num_metas - 4  # predicates 2..preds+1, then the two types
    t0 = 2 + preds
    flat, off, heavy, types = workloads.batch_stream(gs, preds, [t0 + 1], t0, 4000, seed=9)
    keep = np.zeros(len(heavy), dtype=bool)
    keep[::3] = True
    eng_small.configure(max_queries=256, max_patterns=8, arena_bytes=1 << 30)
    print("closed loop over the template stream: caps 64 light, 4 heavy, 0.1 s", flush=True)
    res, issue, _ = eng_small.serve_closed((flat, off), heavy, 64, 4, 0.1, table_budget_bytes=2 << 30,
                                           light_lanes=2, keep=keep)
    for i in np.nonzero(issue >= 0)[0]:
        want, wex, _ = oracle.run_query(gs, flat[off[i]:off[i + 1]])
        r = res[i]
        assert r.status == 0 and r.rows == want.shape[0] and r.expanded == wex, i
        if keep[i]:
            assert r.table.shape == want.shape and np.array_equal(r.table, want), i
        else:
            assert int(i) not in res.tables  # blind: row counts only


def test_autotune_keeps_answers(dev_small, graph_small):
    _, gs = graph_small
    rng = np.random.default_rng(26)
    qs = random_queries(rng, 300, 4, [6, 7], 64, max_steps=5)
    e = GPUEngine(dev_small)
    e.configure(max_queries=64, max_patterns=6, arena_bytes=256 << 20)
    print("autotune over 64 queries", flush=True)
    best, ms = e.autotune(qs, reps=1)
    assert best["probe_tile_rows"] in (1024, 2048) and best["expand_tile_items"] in (512, 1024, 2048)
    assert best["probe_wg_per_cu"] >= 1 and ms > 0
    for q, r in zip(qs, e.run_batched(qs)):
        _same(gs, q, r)
    e.close()


# ---- cmd_chains -> plan (GPU_Engine::execute_one_step's dispatch, for the direct Batch tests)

def _plan(cmd):
    steps, cols = [], 0
    for i in range(0, len(cmd), 4):
        s, p, d, e = (int(x) for x in cmd[i:i + 4])
        known = lambda v: v < 0 and -v <= cols  # noqa: E731
        if i == 0 and 1 < s < (1 << 17):
            steps.append(dict(kind=STEP_I2U, pid=s, dir=d))
            cols = 1
        elif s >= 0:
            steps.append(dict(kind=STEP_C2U, pid=p, dir=d, vid=s))
            cols = 1
        elif e >= 0:
            steps.append(dict(kind=STEP_K2C, pid=p, dir=d, start_col=-s - 1, vid=e))
        elif known(e):
            steps.append(dict(kind=STEP_K2K, pid=p, dir=d, start_col=-s - 1, end_col=-e - 1))
        else:
            steps.append(dict(kind=STEP_K2U, pid=p, dir=d, start_col=-s - 1))
            cols += 1
    return steps


def test_packed_results_match(eng_small, graph_small):
    """wke_run_batched_packed: one buffer for all tables, same answers."""
    _, gs = graph_small
    rng = np.random.default_rng(28)
    qs = random_queries(rng, 300, 4, [6, 7], 150, max_steps=5)
    eng_small.configure(max_queries=64, max_patterns=8, arena_bytes=1 << 30)
    packed, wall = eng_small.run_batched_packed(qs)
    assert wall > 0
    for q, r in zip(qs, packed):
        _same(gs, q, r)
    # with a 1 MiB arena most fall back to the plain path and are appended
    eng_small.configure(max_queries=64, max_patterns=8, arena_bytes=1 << 20)
    packed, _ = eng_small.run_batched_packed(qs)
    for q, r in zip(qs, packed):
        _same(gs, q, r)
