"""Every benchmark configuration's workload on the GPU against the CPU oracle.

  config 1  the full 10M-triple graph and its 2-hop query from a constant
  config 2  the full 100M-triple graph: the heavy 6-pattern query (non-empty
            answer), its 151M-row hub fan-out table row for row, and the
            500k-entry resident type index
  config 3  the same graph: all 512 mix queries through the device-resident
            executor and the plain path, each table against the oracle's
            (order-sensitive digest, row count, work count)
  config 4  the served mix at 1/10 scale (10M entities, 256 predicates,
            100M triples, the same generator rules): 400 Poisson arrivals
            with 5% heavy queries, each served answer against the oracle
  config 5  the cyclic query at 1/10 scale (20M entities, 512 predicates,
            200M triples, 200k-edge hubs, sibling plants): plain path in
            HBM-budgeted chunks and the executor, against the oracle
  plus tables above 2^31 cells, on the device (one step's output) and on
  the host (a chunked query's result).
"""
import numpy as np
import pytest

import oracle
from tests.conftest import VB
from wukonggpu_amd import Device, GPUEngine, GStore, Triples, workloads

pytestmark = pytest.mark.gpu
THREADS = 16  # oracle threads (the GPU host's CPU share)


def _same(res, want, wex, what):
    assert res.status == 0, what
    assert res.table.shape[0] == want.shape[0], (what, res.table.shape, want.shape)
    if want.size:
        assert np.array_equal(res.table.reshape(want.shape), want), what
    assert res.expanded == wex, what


def test_config1_full_graph():
    w = workloads.config1(seed=1)
    q = w.queries[0]
    dev = Device(0)
    dev.load(w.gstore)
    eng = GPUEngine(dev)
    want, wex, _ = oracle.run_query(w.gstore, q)
    assert want.shape[0] > 0
    print(f"config 1: {w.info['triples']} triples, query {q}: {want.shape[0]} rows", flush=True)
    _same(eng.run_query(q), want, wex, "plain")
    _same(eng.run_batched([q])[0], want, wex, "executor")
    eng.close()
    dev.close()


@pytest.fixture(scope="module")
def cfg23():
    w = workloads.config2_3(seed=2)
    print(f"config 2/3 graph: {w.info['triples']} triples, image {w.info['image_bytes'] / 1e9:.2f} GB", flush=True)
    w.triples.close()
    dev = Device(0)
    dev.load(w.gstore)
    eng = GPUEngine(dev)
    yield w, dev, eng
    eng.close()
    dev.close()


def test_config2_heavy_query_full(cfg23):
    w, dev, eng = cfg23
    gs, heavy = w.gstore, w.queries["heavy"]
    want, wex, _ = oracle.run_query(gs, heavy, threads=THREADS)
    assert want.shape[0] > 0, "the heavy query's answer should not be empty"
    print(f"heavy query: {want.shape[0]} rows, {wex} expanded edges", flush=True)
    _same(eng.run_query(heavy), want, wex, "heavy plain")
    dev.set_lineage(False)
    try:
        _same(eng.run_query(heavy), want, wex, "heavy plain, no lineage")
    finally:
        dev.set_lineage(True)
    _same(eng.run_batched([heavy])[0], want, wex, "heavy executor")
    # the hub fan-out (first four patterns) row for row: 151M rows x 4
    pre = heavy[:16]
    pw, pex, _ = oracle.run_query(gs, pre, threads=THREADS)
    assert pw.shape[0] > 100_000_000
    print(f"heavy prefix: {pw.shape[0]} rows", flush=True)
    _same(eng.run_query(pre), pw, pex, "fan-out table")
    del pw
    # the 500k-entry type index, resident (longer than one probe-and-copy launch)
    t0 = w.info["t0"]
    iw, iex, _ = oracle.run_query(gs, [t0, 1, 0, -1])
    assert iw.shape[0] == 500_000
    t, ex = dev.i2u(t0, 0)
    assert np.array_equal(t.to_host(), iw) and ex == iex
    t.free()


def test_config3_mix_full(cfg23):
    w, dev, eng = cfg23
    mix = w.queries["batched"]
    _, rows, ex, _, dg = oracle.run_many(w.gstore, mix, threads=THREADS, digests=True)
    assert (rows > 0).sum() > len(mix) // 4
    print(f"512-query mix: {int((rows > 0).sum())} non-empty answers, {int(rows.sum())} rows", flush=True)
    eng.configure(max_queries=1024, max_patterns=8, arena_bytes=16 << 30)
    for name, res in (("executor", eng.run_batched(mix)), ("plain", eng.run_plain(mix)),
                      ("packed", eng.run_batched_packed(mix)[0])):
        for i, r in enumerate(res):
            assert r.status == 0 and r.table.shape[0] == rows[i] and r.expanded == ex[i], (name, i)
            assert oracle.table_digest(r.table) == int(dg[i]), (name, i)


def test_config4_served_sample():
    w = workloads.config4(seed=4, N=10_000_000, P=256, T=100_000_000)
    P = w.info["predicates"]
    print(f"config 4 at 1/10 scale: {w.info['triples']} triples", flush=True)
    dev = Device(0)
    dev.load(w.gstore)
    eng = GPUEngine(dev)
    eng.configure(max_queries=1024, max_patterns=8, arena_bytes=4 << 30)
    sq, flags = workloads.served_mix(P, w.info["light_types"], w.info["t0"], 400, 45)
    _, rows, ex, _, dg = oracle.run_many(w.gstore, sq, threads=THREADS, digests=True)
    heavy_rows = rows[flags == 1]
    assert (heavy_rows > 0).sum() >= len(heavy_rows) // 3
    rng = np.random.default_rng(46)
    arr = np.cumsum(rng.exponential(1.0 / 2000, size=len(sq)))
    res, wall = eng.serve(sq, arr, heavy=flags, table_budget_bytes=16 << 30, light_lanes=2)
    print(f"served 400 queries in {wall:.3f} s", flush=True)
    for i, r in enumerate(res):
        assert r.status == 0 and r.table.shape[0] == rows[i] and r.expanded == ex[i], i
        assert oracle.table_digest(r.table) == int(dg[i]), i
    eng.close()
    dev.close()


def test_config5_cycle_sample():
    tr = workloads.config5_triples(seed=5, N=20_000_000, P=512, T=200_000_000, hub_degree=200_000,
                                   t0_members=200_000, sample_members=20_000)
    gs = GStore.build(tr)
    tr.close()
    qs, info = workloads.config5_queries(gs, 512)
    dev = Device(0)
    dev.load(gs)
    eng = GPUEngine(dev)
    for name in ("full", "sample"):
        q = qs[name]
        want, wex, _ = oracle.run_query(gs, q, threads=THREADS)
        fan = oracle.run_query(gs, q[:16], threads=THREADS)[0].shape[0]
        print(f"config 5 at 1/10 scale, {name}: {want.shape[0]} rows, {wex} expanded, fan-out {fan} rows",
              flush=True)
        if name == "full":
            assert want.shape[0] > 0 and fan > 5_000_000
        dev.set_table_limit(32 << 20)  # the fan-out in 32 MiB chunks
        s0 = eng.splits()
        _same(eng.run_query(q), want, wex, (name, "chunked"))
        assert name != "full" or eng.splits() > s0
        dev.set_table_limit(0)
        _same(eng.run_query(q), want, wex, (name, "whole"))
    eng.configure(max_queries=16, max_patterns=8, arena_bytes=8 << 30)
    for q, r in zip((qs["full"], qs["sample"]), eng.run_batched([qs["full"], qs["sample"]])):
        want, wex, _ = oracle.run_query(gs, q, threads=THREADS)
        _same(r, want, wex, "executor")
    eng.close()
    dev.close()


def test_step_output_above_2_31_cells(graph_hubs):
    """One known_to_unknown step whose output holds more than 2^31 cells (a
    hub's 5000-edge list joined to 220k rows: 1.1G rows x 2), then a filter
    over it, against the oracle."""
    _, gs = graph_hubs
    dev = Device(0)
    dev.load(gs)
    # the hub: predicate 2's subject with the longest list
    hub = max(gs.get_edges(0, 2, 0).tolist(), key=lambda v: gs.degree(v, 2, 1))
    hub_list = gs.get_edges(hub, 2, 1)
    assert len(hub_list) >= 5000
    R = 220_000
    d = len(hub_list)
    assert R * d * 2 > (1 << 31)
    t_in = dev.table(np.full((R, 1), hub, dtype=np.int32))
    out, ex = dev.k2u(t_in, 0, 2, 1)
    assert out.shape == (R * d, 2) and ex == R * d
    got = out.to_host().reshape(R, d, 2)
    assert (got[:, :, 0] == hub).all()
    assert (got[:, :, 1] == hub_list.astype(np.int32)[None, :]).all()
    del got
    # known_to_const over the 1.1G rows: keep rows whose ?o has a predicate-2 IN edge from the hub's last object
    const = int(hub_list[-1])
    filt, fex = dev.k2c(out, 1, 3, 1, const)
    cmd = [hub, 2, 1, -1, -1, 2, 1, -2, -2, 3, 1, const]
    base = np.full((R, 1), hub, dtype=np.int32)
    mid = np.stack([np.repeat(base[:, 0], d), np.tile(hub_list.astype(np.int32), R)], axis=1)
    want, wex = oracle.run_step(gs, cmd, 2, mid, threads=THREADS)
    del mid
    assert np.array_equal(filt.to_host(), want) and fex == wex
    for t in (filt, out, t_in):
        t.free()
    dev.close()


def test_chunked_query_result_above_2_31_cells():
    """A query whose answer holds 2.3G cells (576M rows x 4), produced in
    1 GiB output chunks and assembled on the host: equal to the oracle's."""
    K = 24_000
    h = VB
    objs = np.arange(VB + 1, VB + 1 + K, dtype=np.uint32)
    tgts = (VB + 1 + K + (objs % 100)).astype(np.uint32)
    s = np.concatenate([np.full(K, h, np.uint32), objs])
    p = np.concatenate([np.full(K, 2, np.uint32), np.full(K, 3, np.uint32)])
    o = np.concatenate([objs, tgts])
    gs = GStore.build(Triples.from_arrays(s, p, o))
    q = [h, 2, 1, -1, -1, 2, 0, -2, -2, 2, 1, -3, -1, 3, 1, -4]
    want, wex, _ = oracle.run_query(gs, q, threads=THREADS)
    assert want.shape == (K * K, 4) and want.size > (1 << 31)
    dev = Device(0)
    dev.load(gs)
    eng = GPUEngine(dev)
    dev.set_table_limit(1 << 30)
    r = eng.run_query(q)
    assert eng.splits() > 0
    _same(r, want, wex, "chunked")
    eng.close()
    dev.close()
