"""Seeded synthetic workloads: graphs and queries for the benchmark configs.

Graph model (all configs): vertex ids 2^17 + [0, N), predicates 2..P+1, types
after the predicates, rdf:type = 1 (datagen/generate_data.cpp:88-98). For each
predicate a seeded permutation ranks the subjects and rank r has out-degree
min(cap, floor(C r^-s)) with C solved so the predicate holds about T/P
triples; objects are uniform and distinct per subject.

Queries are cmd_chains (core/query.hpp:68): (start, predicate, direction,
end) per pattern, variables -1, -2, ... and dir IN = 0 / OUT = 1.
"""
import numpy as np

from .gstore import GStore, Triples, VID_BASE

IN, OUT = 0, 1


class Workload:
    def __init__(self, name, triples, gstore, queries, info):
        self.name, self.triples, self.gstore, self.queries, self.info = name, triples, gstore, queries, info


def config1(seed=1):
    """Plain single query: 1M entities, 16 predicates, ~10M triples, Zipf 1.1
    capped at 10k; the 2-hop chain `c p1 ?y . ?y p2 ?z` from the constant whose
    p1 out-degree is closest to 1000."""
    N, P = 1_000_000, 16
    tr = Triples.generate(seed=seed, num_entities=N, num_preds=P, num_triples=10_000_000, zipf_s=1.1,
                          degree_cap=10_000)
    gs = GStore.build(tr)
    rng = np.random.default_rng(seed)
    p1, p2 = (int(x) for x in 2 + rng.choice(P, size=2, replace=False))
    s, p, _ = tr.arrays()
    subj, deg = np.unique(s[p == p1], return_counts=True)
    c = int(subj[np.argmin(np.abs(deg.astype(np.int64) - 1000))])
    q = [c, p1, OUT, -1, -1, p2, OUT, -2]
    return Workload("config1", tr, gs, [q], {"entities": N, "predicates": P, "triples": len(tr), "p1": p1,
                                              "p2": p2, "const": c, "const_degree": int(deg.max() and
                                                                                        deg[subj == c][0])})


def _cfg23_graph(seed):
    """The 100M-triple graph shared by the heavy and batched configs: 10M
    entities, 64 predicates, Zipf 1.0 (uncapped), type t0 with 500k members and
    511 more types with log-uniform sizes in [1, 1e5]."""
    N, P = 10_000_000, 64
    rng = np.random.default_rng(seed)
    extra = np.exp(rng.uniform(0.0, np.log(1e5), 511)).astype(np.int64)
    sizes = [500_000] + [int(x) for x in extra]
    tr = Triples.generate(seed=seed, num_entities=N, num_preds=P, num_triples=100_000_000, zipf_s=1.0,
                          type_sizes=sizes)
    gs = GStore.build(tr)
    return tr, gs, N, P, sizes


def heavy_query(P, t0, pa=2, pb=3, pc=4):
    """Config-2 heavy query from type index t0 (2-hop chain, star of 2 outgoing
    predicates, closing known_to_known):
        ?x rdf:type t0 .       index_to_unknown                 (500k rows)
        ?x pa ?y .             known_to_unknown  OUT
        ?z pb ?y .             known_to_unknown  IN   (?z: pb hubs)
        ?z pb ?u .             known_to_unknown  OUT  (hub fan-out, the big table)
        ?z pc ?w .             known_to_unknown  OUT  (second star arm)
        ?z pb ?w               known_to_known, evaluated from ?w's IN list
    """
    return [t0, 1, IN, -1,
            -1, pa, OUT, -2,
            -2, pb, IN, -3,
            -3, pb, OUT, -4,
            -3, pc, OUT, -5,
            -5, pb, IN, -3]


def batched_queries(P, type_ids, n, seed):
    """Config-3 mix: n queries of 3-4 patterns (index start + 2-3 hops), chains
    or stars with random predicates and directions, each from a random type
    index (sizes log-uniform in [1, 1e5]). A hop never goes OUT over p from a
    variable that was reached IN over p (that pairs a size-biased hub set with
    its own lists and is not selective)."""
    rng = np.random.default_rng(seed)
    qs = []
    for _ in range(n):
        t = int(rng.choice(type_ids))
        q = [t, 1, IN, -1]
        via = {1: None}  # variable -> (pid, dir) it was reached by
        ncols = 1
        star = rng.random() < 0.5
        for _ in range(int(rng.integers(2, 4))):
            src = 1 if star else ncols
            while True:
                pid = int(rng.integers(2, 2 + P))
                d = int(rng.integers(0, 2))
                if not (via.get(src) == (pid, IN) and d == OUT):
                    break
            ncols += 1
            q += [-src, pid, d, -ncols]
            via[ncols] = (pid, d)
        qs.append(q)
    return qs


def config2_3(seed=2, n_batched=512):
    tr, gs, N, P, sizes = _cfg23_graph(seed)
    t0 = 2 + P
    heavy = heavy_query(P, t0)
    extra_types = list(range(2 + P + 1, 2 + P + len(sizes)))
    batch = batched_queries(P, extra_types, n_batched, seed + 1)
    info = {"entities": N, "predicates": P, "triples": len(tr), "types": len(sizes), "t0": t0,
            "t0_members": sizes[0], "keys": gs.num_keys, "edges": gs.num_edges,
            "image_bytes": gs.image_nbytes()}
    return Workload("config2_3", tr, gs, {"heavy": heavy, "batched": batch}, info)


def served_mix(P, type_ids, t0, n, seed, heavy_frac=0.05):
    """Config-4 style mix over the given graph: light 2-3 pattern queries from
    type indexes and a fraction of heavy 5-6 pattern ones (the heavy query with
    random predicates). Returns (queries, heavy flags)."""
    rng = np.random.default_rng(seed)
    light = batched_queries(P, type_ids, n, seed + 7)
    qs, flags = [], []
    for i in range(n):
        if rng.random() < heavy_frac:
            pa, pb, pc = (int(x) for x in 2 + rng.choice(P, size=3, replace=False))
            qs.append(heavy_query(P, t0, pa, pb, pc))
            flags.append(1)
        else:
            q = light[i]
            qs.append(q[:4 * int(rng.integers(2, 4))])
            flags.append(0)
    return qs, np.array(flags, dtype=np.uint8)


def config4(seed=4, n_light_types=1023):
    """Served graph: 100M entities, 256 predicates, ~1B triples (Zipf 1.0),
    type t0 with 500k members for the heavy queries and n_light_types types
    with log-uniform sizes in [1, 1e4] for the light ones. The triples are
    released once the image is built."""
    N, P = 100_000_000, 256
    rng = np.random.default_rng(seed)
    extra = np.exp(rng.uniform(0.0, np.log(1e4), n_light_types)).astype(np.int64)
    sizes = [500_000] + [int(x) for x in extra]
    tr = Triples.generate(seed=seed, num_entities=N, num_preds=P, num_triples=1_000_000_000, zipf_s=1.0,
                          type_sizes=sizes)
    n_tr = len(tr)
    gs = GStore.build(tr)
    tr.close()
    t0 = 2 + P
    info = {"entities": N, "predicates": P, "triples": n_tr, "types": len(sizes), "t0": t0,
            "t0_members": sizes[0], "light_types": list(range(t0 + 1, t0 + len(sizes))), "keys": gs.num_keys,
            "edges": gs.num_edges, "image_bytes": gs.image_nbytes()}
    return Workload("config4", None, gs, {}, info)


def cyclic_query(t, pa, ph, pc, pq, const):
    """Config-5 six-pattern cycle from type index t:
        ?a rdf:type t .   index_to_unknown
        ?a pa ?b .        known_to_unknown  OUT
        ?c ph ?b .        known_to_unknown  IN   (?c: ph's subjects, the hub among them)
        ?c ph ?d .        known_to_unknown  OUT  (hub fan-out: the ~1B-row table)
        ?d pc ?b .        known_to_known         (closes the triangle b-c-d)
        ?c pq const .     known_to_const
    """
    return [t, 1, IN, -1,
            -1, pa, OUT, -2,
            -2, ph, IN, -3,
            -3, ph, OUT, -4,
            -4, pc, OUT, -2,
            -3, pq, OUT, const]


def config5(seed=5, t0_members=2_000_000, sample_members=20_000):
    """Large graph: 200M entities, 512 predicates, ~2B triples, Zipf 0.9, and
    hubs: the rank-1 subject of each of the first 8 predicates has 2M
    edges. Types: t0 (t0_members) for the full query and t1
    (sample_members) for a parity sample small enough for the oracle's
    per-row work. The cyclic query goes through hub predicate ph = 2; the
    known_to_const constant is one of the hub's own ph objects."""
    N, P = 200_000_000, 512
    sizes = [t0_members, sample_members]
    tr = Triples.generate(seed=seed, num_entities=N, num_preds=P, num_triples=2_000_000_000, zipf_s=0.9,
                          type_sizes=sizes, num_hubs=8, hub_degree=2_000_000)
    n_tr = len(tr)
    gs = GStore.build(tr)
    tr.close()
    t0, t1 = 2 + P, 2 + P + 1
    ph, pa, pc, pq = 2, 100, 200, 2
    # the hub of ph: the subject with the longest ph OUT list; the constant is its first object
    hub, best = None, -1
    for s in gs.get_edges(0, ph, IN).tolist():
        d = gs.degree(s, ph, OUT)
        if d > best:
            hub, best = s, d
    const = int(gs.get_edges(hub, ph, OUT)[0])
    queries = {"full": cyclic_query(t0, pa, ph, pc, pq, const), "sample": cyclic_query(t1, pa, ph, pc, pq, const)}
    info = {"entities": N, "predicates": P, "triples": n_tr, "t0": t0, "t1": t1, "t0_members": t0_members,
            "t1_members": sample_members, "ph": ph, "pa": pa, "pc": pc, "pq": pq, "hub": hub, "hub_degree": best,
            "const": const, "keys": gs.num_keys, "edges": gs.num_edges, "image_bytes": gs.image_nbytes()}
    return Workload("config5", None, gs, queries, info)
