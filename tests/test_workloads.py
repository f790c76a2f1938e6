"""Benchmark query streams (CPU): the closed-loop batch-mode stream draws
its templates by the reference's mix loads and instantiates each light
template with a member of the template's type."""
import numpy as np

import oracle
from tests.conftest import small_graph
from wukonggpu_amd import workloads
from wukonggpu_amd.workloads import BATCH_HEAVY_LOADS, BATCH_LIGHT_LOADS, IN


def test_batch_stream_mix_and_constants():
    preds = 6
    tr, gs = small_graph(seed=23, n=3000, preds=preds, triples=30000, s=1.0, cap=0, types=(500, 100, 200, 80, 3))
    t0 = 2 + preds
    light_types = list(range(t0 + 1, t0 + 5))
    n = 40000
    flat, off, heavy, types = workloads.batch_stream(gs, preds, light_types, t0, n, seed=5)
    assert len(off) == n + 1 and off[-1] == len(flat) and len(heavy) == n and len(types) == n
    loads = np.array(BATCH_LIGHT_LOADS + BATCH_HEAVY_LOADS, dtype=np.float64)
    frac = np.bincount(types, minlength=len(loads)) / n
    # binomial draws: within 5 standard deviations of the load shares
    want = loads / loads.sum()
    assert (np.abs(frac - want) <= 5 * np.sqrt(want * (1 - want) / n) + 1e-9).all(), (frac, want)
    assert np.array_equal(heavy, (types >= len(BATCH_LIGHT_LOADS)).astype(np.uint8))
    members = {t: set(int(x) for x in gs.get_edges(0, t, IN)) for t in light_types}
    for k in range(len(loads)):
        idx = np.nonzero(types == k)[0]
        tails = {tuple(flat[off[i] + 1:off[i + 1]]) for i in idx[:200]}
        assert len(tails) == 1  # one template per type
        if k < len(BATCH_LIGHT_LOADS):
            consts = {int(flat[off[i]]) for i in idx[:200]}
            # all constants of a template come from one type's members
            assert any(consts <= m for m in members.values()), k
        else:
            assert flat[off[idx[0]]] == t0 and flat[off[idx[0]] + 1] == 1  # index start
    # every template instance is a runnable query
    for k in range(len(loads)):
        i = int(np.nonzero(types == k)[0][0])
        table, _, _ = oracle.run_query(gs, flat[off[i]:off[i + 1]].tolist())
        assert table.ndim == 2
    tr.close()
