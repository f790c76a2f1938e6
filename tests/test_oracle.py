"""Pins the CPU oracle: known-answer tables worked out by hand, the hash against
the reference's own inverse, the brute-force evaluator over raw triples, and
the committed golden fixtures. None of these needs a GPU."""
import json
import os

import numpy as np
import pytest

import oracle
from oracle.brute import BruteGraph
from tests.conftest import VB, random_queries, small_graph
from wukonggpu_amd import GStore, Triples

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_hash_inverse_roundtrip():
    # inverse_hash_u64 (utils/mymath.hpp:73-107) undoes hash_u64 step by step;
    # a dropped term or a wrong shift in either breaks the round trip
    rng = np.random.default_rng(0)
    xs = [0, 1, 2, (1 << 64) - 1, 1 << 63, 0x123456789ABCDEF] + [int(x) for x in rng.integers(0, 1 << 63, 200)]
    for x in xs:
        assert oracle.inverse_hash_u64(oracle.hash_u64(x)) == x
    # and the hash is not the identity / not trivially linear
    assert oracle.hash_u64(1) != 1 and oracle.hash_u64(2) != 2 * oracle.hash_u64(1)


def test_known_answer_tiny():
    with open(os.path.join(GOLDEN, "tiny_kat.json")) as f:
        kat = json.load(f)
    t = np.array(kat["triples"], dtype=np.uint32)
    gs = GStore.build(Triples.from_arrays(t[:, 0], t[:, 1], t[:, 2]))
    bg = BruteGraph(t[:, 0], t[:, 1], t[:, 2])
    for q in kat["queries"]:
        want = np.array(q["rows"], dtype=np.int32)
        got, ex, _ = oracle.run_query(gs, q["cmd"])
        bt, bex = bg.run(q["cmd"])
        assert got.shape[0] == want.shape[0], q["name"]
        if want.size:
            assert np.array_equal(got, want), q["name"]
            assert np.array_equal(bt, want), q["name"]
        assert ex == q["expanded"] and bex == q["expanded"], q["name"]


def test_oracle_matches_brute_random():
    for seed in (3, 5):
        tr, gs = small_graph(seed=seed, n=250, preds=4, triples=2500, types=(30, 4, 1))
        s, p, o = tr.arrays()
        bg = BruteGraph(s, p, o)
        rng = np.random.default_rng(seed)
        for q in random_queries(rng, 250, 4, [6, 7, 8], 150):
            a, ea, _ = oracle.run_query(gs, q)
            b, eb = bg.run(q)
            assert a.shape == b.shape and np.array_equal(a, b) and ea == eb, q


def test_golden_random_fixture():
    z = np.load(os.path.join(GOLDEN, "random_small.npz"))
    gs = GStore.build(Triples.from_arrays(z["s"], z["p"], z["o"]))
    qo, ro = z["q_off"], z["r_off"]
    for i in range(len(qo) - 1):
        q = z["queries"][qo[i]:qo[i + 1]]
        want = z["rows"][ro[i]:ro[i + 1]].reshape(-1, z["cols"][i]) if z["cols"][i] else np.zeros((0, 0))
        got, ex, _ = oracle.run_query(gs, q)
        assert got.shape[0] == want.shape[0] and ex == z["expanded"][i], i
        if want.size:
            assert np.array_equal(got, want), i


def test_oracle_step_equals_query_prefix(graph_small):
    tr, gs = graph_small
    q = [2, 0, 1, -1, -1, 3, 1, -2, -2, 2, 0, -3, -1, 4, 1, -2]
    t1, _, _ = oracle.run_query(gs, q, stop_step=1)
    t2, _, _ = oracle.run_query(gs, q, stop_step=2)
    s2, _ = oracle.run_step(gs, q, 1, t1)
    assert np.array_equal(s2, t2)


def test_oracle_rejects_unsupported(graph_small):
    _, gs = graph_small
    with pytest.raises(oracle.OracleError):
        oracle.run_query(gs, [VB, 2, 1, VB + 1])  # const -> const
    with pytest.raises(oracle.OracleError):
        oracle.run_query(gs, [-1, 2, 1, -2])      # start from an unknown
    with pytest.raises(oracle.OracleError):
        oracle.run_query(gs, [VB, -3, 1, -1])     # predicate variable


def test_oracle_threads_agree(graph_small):
    _, gs = graph_small
    rng = np.random.default_rng(4)
    qs = random_queries(rng, 300, 4, [6, 7], 40)
    lat1, rows1, ex1, _ = oracle.run_many(gs, qs, threads=1)
    lat4, rows4, ex4, _ = oracle.run_many(gs, qs, threads=4)
    assert np.array_equal(rows1, rows4) and np.array_equal(ex1, ex4)


def test_oracle_threaded_steps_and_digest(graph_hubs):
    # known-start steps split over row ranges on 8 threads give the one-thread
    # table (each row's output depends on that row alone); the oracle's table
    # digest equals the numpy one over the same table, and moves when one
    # cell or the row order changes
    _, gs = graph_hubs
    tid = 2 + 3
    qs = [[tid, 1, 0, -1, -1, 3, 1, -2, -2, 3, 0, -3, -3, 2, 1, -4],
          [tid, 1, 0, -1, -1, 2, 1, -2, -2, 2, 0, -3, -3, 3, 1, -4, -4, 2, 0, -3],
          [tid, 1, 0, -1, -1, 2, 1, -2, -2, 2, 0, -3, -3, 2, 1, VB + 17]]
    big = 0
    for q in qs:
        a, ea, _ = oracle.run_query(gs, q)
        b, eb, _ = oracle.run_query(gs, q, threads=8)
        assert a.shape == b.shape and np.array_equal(a, b) and ea == eb, q
        big = max(big, oracle.run_query(gs, q, stop_step=3)[0].shape[0])
    assert big >= 1 << 16  # the threaded split ran
    _, rows, ex, _, dg = oracle.run_many(gs, qs, threads=2, digests=True)
    for q, d in zip(qs, dg):
        t, _, _ = oracle.run_query(gs, q)
        assert int(d) == oracle.table_digest(t)
    t, _, _ = oracle.run_query(gs, qs[0])
    assert t.shape[0] > 2
    h = oracle.table_digest(t)
    t2 = t.copy()
    t2[1, 0] ^= 1
    assert oracle.table_digest(t2) != h
    assert oracle.table_digest(t[::-1]) != h
    assert oracle.table_digest(t[:, :3]) != h
