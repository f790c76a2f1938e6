"""The device gstore builder: the image it builds in HBM equals the host
builder's, slot for slot, and queries over it give the oracle's answers."""
import numpy as np
import pytest

import oracle
from tests.conftest import VB, random_queries, small_graph
from wukonggpu_amd import Device, GPUEngine, GStore, Triples

pytestmark = pytest.mark.gpu

GRAPHS = {
    "small": dict(),
    "hubs": dict(seed=11, n=6000, preds=3, triples=60000, s=1.0, cap=0, types=(3000, 1, 60), hubs=2,
                 hub_degree=5000, load=0.9),
    "mid": dict(seed=12, n=20000, preds=12, triples=400000, s=1.0, cap=0, types=(5000, 300, 40), hubs=1,
                hub_degree=20000),
}


def _host_image(gs):
    return gs.slots().reshape(-1, 2), gs.edges(), gs.metas()


@pytest.mark.parametrize("name", list(GRAPHS))
def test_device_build_equals_host_build(name):
    args = dict(GRAPHS[name])
    load = args.get("load", 0.75)
    tr, gs = small_graph(**args)
    dev = Device(0)
    secs = dev.build(tr, main_load=load)
    slots, edges, metas = dev.image()
    hs, he, hm = _host_image(gs)
    print(f"{name}: {len(tr)} triples, {hs.shape[0]} slots, {he.shape[0]} edges, device build {secs * 1e3:.1f} ms",
          flush=True)
    assert metas.shape == hm.shape and np.array_equal(metas, hm)
    assert edges.shape == he.shape and np.array_equal(edges, he)
    assert slots.shape == hs.shape and np.array_equal(slots, hs)
    dev.close()


def test_device_built_graph_answers_like_the_oracle():
    tr, gs = small_graph(**GRAPHS["hubs"])
    dev = Device(0)
    dev.build(tr, main_load=0.9)
    eng = GPUEngine(dev)
    qs = [[2, 1, 0, -1, -1, 2, 1, -2], [2, 1, 0, -1, -1, 2, 1, -2, -2, 3, 0, -3],
          [VB + 0, 2, 1, -1, -1, 3, 0, -2]]
    qs += random_queries(np.random.default_rng(5), 6000, 3, [2, 3], 40, max_steps=4)
    for q in qs:
        want, wex, _ = oracle.run_query(gs, q)
        r = eng.run_query(q)
        assert r.status == 0 and r.table.shape == want.shape and np.array_equal(r.table, want), q
        assert r.expanded == wex, q
    eng.close()
    dev.close()


def test_device_build_duplicates_and_bad_triples():
    # duplicate triples collapse, as on the host (core/dgraph.hpp:320-324)
    t = np.array([[VB + 0, 2, VB + 1], [VB + 0, 2, VB + 1], [VB + 1, 2, VB + 0], [VB + 0, 1, 3], [VB + 1, 1, 3],
                  [VB + 2, 1, 3], [VB + 2, 2, VB + 0]], dtype=np.uint32)
    tr = Triples.from_arrays(t[:, 0], t[:, 1], t[:, 2])
    gs = GStore.build(tr)
    dev = Device(0)
    dev.build(tr)
    slots, edges, metas = dev.image()
    hs, he, hm = _host_image(gs)
    assert np.array_equal(metas, hm) and np.array_equal(edges, he) and np.array_equal(slots, hs)
    dev.close()
    bad = Triples.from_arrays(np.array([VB, 5], dtype=np.uint32), np.array([2, 2], dtype=np.uint32),
                              np.array([VB + 1, VB], dtype=np.uint32))
    dev = Device(0)
    with pytest.raises(Exception, match="invalid triple 1"):
        dev.build(bad)
    dev.close()
