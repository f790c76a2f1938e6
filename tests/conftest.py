import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from wukonggpu_amd import GStore, Triples  # noqa: E402

VB = 1 << 17


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C ABI)")


def small_graph(seed=7, n=300, preds=4, triples=3000, s=1.1, cap=50, types=(40, 7), hubs=0, hub_degree=0,
                load=0.75):
    tr = Triples.generate(seed=seed, num_entities=n, num_preds=preds, num_triples=triples, zipf_s=s,
                          degree_cap=cap, type_sizes=list(types), num_hubs=hubs, hub_degree=hub_degree)
    gs = GStore.build(tr, main_load=load)
    return tr, gs


@pytest.fixture(scope="session")
def graph_small():
    return small_graph()


@pytest.fixture(scope="session")
def graph_hubs():
    # hub rows longer than one expand tile (2048 items) and a type index of 3000
    return small_graph(seed=11, n=6000, preds=3, triples=60000, s=1.0, cap=0, types=(3000, 1, 60), hubs=2,
                       hub_degree=5000, load=0.9)


def random_queries(rng, n_ent, preds, types, nq, max_steps=4):
    """Random index/const-start chains, stars and filters over a generated graph."""
    out = []
    for _ in range(nq):
        r = rng.random()
        if r < 0.4:
            p = int(rng.integers(2, 2 + preds))
            cmd = [p, 0, int(rng.integers(0, 2)), -1]
        elif r < 0.6 and types:
            cmd = [int(rng.choice(types)), 1, 0, -1]
        else:
            cmd = [VB + int(rng.integers(0, n_ent)), int(rng.integers(2, 2 + preds)), int(rng.integers(0, 2)), -1]
        ncol = 1
        for _ in range(int(rng.integers(1, max_steps))):
            x = rng.random()
            sc = -int(rng.integers(1, ncol + 1))
            pid = int(rng.integers(1, 2 + preds))
            d = int(rng.integers(0, 2))
            if x < 0.6:
                cmd += [sc, pid, d, -(ncol + 1)]
                ncol += 1
            elif x < 0.8 and ncol > 1:
                cmd += [sc, pid, d, -int(rng.integers(1, ncol + 1))]
            else:
                cmd += [sc, pid, d, VB + int(rng.integers(0, n_ent))]
        out.append(cmd)
    return out


@pytest.fixture(scope="session")
def dev_small(graph_small):
    from wukonggpu_amd import Device
    d = Device(0)
    d.load(graph_small[1])
    yield d
    d.close()
