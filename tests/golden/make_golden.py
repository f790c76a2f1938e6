"""Writes tests/golden/random_small.npz: a fixed triple set, random queries and
their expected result tables from the brute-force evaluator (oracle/brute.py),
which reads only the raw triples.

The reference ships no golden vectors for this path, so these fixtures pin
the engine to the triple-level semantics; they were produced by this script
and are committed as data. Run from the repo root:
    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle.brute import BruteGraph  # noqa: E402
from tests.conftest import random_queries  # noqa: E402
from wukonggpu_amd import Triples  # noqa: E402


def main():
    tr = Triples.generate(seed=2024, num_entities=400, num_preds=5, num_triples=4000, zipf_s=1.1,
                          degree_cap=60, type_sizes=[50, 9, 1])
    s, p, o = (np.array(a, dtype=np.uint32) for a in tr.arrays())
    bg = BruteGraph(s, p, o)
    rng = np.random.default_rng(99)
    types = [2 + 5, 2 + 5 + 1, 2 + 5 + 2]
    qs = random_queries(rng, 400, 5, types, 200)
    q_off = np.zeros(len(qs) + 1, dtype=np.int64)
    r_off = np.zeros(len(qs) + 1, dtype=np.int64)
    cols = np.zeros(len(qs), dtype=np.int32)
    expd = np.zeros(len(qs), dtype=np.int64)
    flat_q, flat_r = [], []
    for i, q in enumerate(qs):
        t, ex = bg.run(q)
        flat_q += q
        flat_r.append(t.reshape(-1))
        q_off[i + 1] = q_off[i] + len(q)
        r_off[i + 1] = r_off[i] + t.size
        cols[i] = t.shape[1]
        expd[i] = ex
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "random_small.npz")
    np.savez_compressed(out, s=s, p=p, o=o, queries=np.array(flat_q, dtype=np.int64), q_off=q_off,
                        rows=np.concatenate(flat_r).astype(np.int32), r_off=r_off, cols=cols, expanded=expd)
    print("wrote", out, len(s), "triples", len(qs), "queries")


if __name__ == "__main__":
    main()
