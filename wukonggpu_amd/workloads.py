"""Seeded synthetic workloads: graphs and queries for the benchmark configs.

Graph model (all configs): vertex ids 2^17 + [0, N), predicates 2..P+1, types
after the predicates, rdf:type = 1 (datagen/generate_data.cpp:88-98). For each
predicate a seeded permutation ranks the subjects and rank r has out-degree
min(cap, floor(C r^-s)) with C solved so the predicate holds about T/P
triples; objects are uniform and distinct per subject.

Queries are cmd_chains (core/query.hpp:68): (start, predicate, direction,
end) per pattern, variables -1, -2, ... and dir IN = 0 / OUT = 1.
"""
import time

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


LINK_FRAC = 0.25


def _cfg23_graph(seed):
    """The 100M-triple graph shared by the heavy and batched configs: 10M
    entities, 64 predicates, Zipf 1.0 (uncapped), type t0 with 500k members and
    511 more types with log-uniform sizes in [1, 1e5]. Predicates are linked
    in pairs (2k+3 to 2k+2, a quarter of a list from the same subject's
    source list), so a star whose arms follow a linked pair can close."""
    N, P = 10_000_000, 64
    rng = np.random.default_rng(seed)
    extra = np.exp(rng.uniform(0.0, np.log(1e5), 511)).astype(np.int64)
    sizes = [500_000] + [int(x) for x in extra]
    tr = Triples.generate(seed=seed, num_entities=N, num_preds=P, num_triples=100_000_000, zipf_s=1.0,
                          type_sizes=sizes, num_linked=P // 2, link_frac=LINK_FRAC)
    gs = GStore.build(tr)
    return tr, gs, N, P, sizes


def heavy_query(P, t0, pa=2, pb=4, pc=5):
    """Config-2 heavy query from type index t0 (2-hop chain, star of 2 outgoing
    predicates, closing known_to_known):
        ?x rdf:type t0 .       index_to_unknown                 (500k rows)
        ?x pa ?y .             known_to_unknown  OUT
        ?z pb ?y .             known_to_unknown  IN   (?z: pb hubs)
        ?z pb ?u .             known_to_unknown  OUT  (hub fan-out, the big table)
        ?z pc ?w .             known_to_unknown  OUT  (second star arm)
        ?z pb ?w               known_to_known, evaluated from ?w's IN list
    On a graph whose pc is linked to pb (part of a subject's pc list is taken
    from its pb list) the closing edge exists for about that part of the rows.
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


def linked_triple(rng, P):
    """(pa, pb, pc) for a heavy query: pb a source predicate of a linked pair
    (even offset from 2), pc its linked partner, pa any other predicate."""
    pb = 2 + 2 * int(rng.integers(0, P // 2))
    pa = int(rng.choice([p for p in range(2, 2 + P) if p not in (pb, pb + 1)]))
    return pa, pb, pb + 1


def served_mix(P, type_ids, t0, n, seed, heavy_frac=0.05):
    """Config-4 style mix over the given graph: light 2-3 pattern queries from
    type indexes and a fraction of heavy 5-6 pattern ones (the heavy query with
    random predicates). Returns (queries, heavy flags)."""
    rng = np.random.default_rng(seed)
    light = batched_queries(P, type_ids, n, seed + 7)
    qs, flags = [], []
    for i in range(n):
        if rng.random() < heavy_frac:
            pa, pb, pc = linked_triple(rng, P)
            qs.append(heavy_query(P, t0, pa, pb, pc))
            flags.append(1)
        else:
            q = light[i]
            qs.append(q[:4 * int(rng.integers(2, 4))])
            flags.append(0)
    return qs, np.array(flags, dtype=np.uint8)


# The reference's batch mix (scripts/batch/mix_config): six light templates
# (q1-q6) and four heavy queries (lubm_q1, q2, q3, q7) with these loads.
BATCH_LIGHT_LOADS = (2800, 2800, 290, 900, 2800, 200)
BATCH_HEAVY_LOADS = (5, 100, 100, 5)


def _best_hop(gs, rng, frm, P, tries=16):
    """(pid, dir) among `tries` random candidates whose lists from the
    vertices `frm` are most often non-empty (ties: larger total)."""
    best, key = None, (-1.0, -1)
    for _ in range(tries):
        pid, d = int(rng.integers(2, 2 + P)), int(rng.integers(0, 2))
        degs = np.array([gs.degree(int(v), pid, d) for v in frm])
        k = (float((degs > 0).mean()), int(degs.sum()))
        if k > key:
            best, key = (pid, d), k
    return best


def batch_stream(gs, P, light_types, t0, n, seed):
    """Closed-loop batch-mode stream modelled on the reference's batch mix
    (scripts/batch/mix_config), drawn by load as core/proxy.hpp:356-377 does
    (mymath::get_distribution). Each light template starts from a random
    constant of its own type (request_template::instantiate over the type's
    members, the candidates of core/proxy.hpp:75-115):
        q1, q2, q3, q5:  %c pa ?X . ?X pb ?Y                (const_to_unknown, known_to_unknown)
        q4:              %c pa ?X . ?X pb ?Y1 . ?X pc ?Y2 . ?X pd ?Y3   (a star of three arms)
        q6:              %c pa ?X . ?X pb ?Y . ?Y pc ?Z     (a chain)
    with predicates and directions picked on the graph so that the hops are
    mostly non-empty. The templates' rdf:type checks are left out: on this
    graph about 1% of the entities are typed, so they would empty nearly every
    answer. The heavy queries start from the t0 type index (start_from_index,
    which the reference counts as heavy): the heavy 6-pattern shape with four
    predicate triples. Returns (flat int64 cmd_chains, int64 offsets, heavy
    flags uint8, template index per query: 0-5 light, 6-9 heavy)."""
    rng = np.random.default_rng(seed)
    cand = [t for t in light_types if gs.degree(0, t, IN) >= 64] or \
        [t for t in light_types if gs.degree(0, t, IN) > 0]
    tmpls, members = [], []
    for k in range(len(BATCH_LIGHT_LOADS)):
        t = int(rng.choice(cand))
        m = gs.get_edges(0, t, IN).astype(np.int64)
        samp = rng.choice(m, size=min(32, len(m)), replace=False)
        pa, da = _best_hop(gs, rng, samp, P)
        xs = np.concatenate([gs.get_edges(int(c), pa, da) for c in samp] + [np.zeros(0, np.uint32)])
        xs = xs[:64] if len(xs) else samp[:1]
        q = [0, pa, da, -1]
        if k == 3:  # star: three arms from ?X
            for j in range(3):
                pb, db = _best_hop(gs, rng, xs, P)
                q += [-1, pb, db, -(2 + j)]
        else:
            pb, db = _best_hop(gs, rng, xs, P)
            q += [-1, pb, db, -2]
            if k == 5:  # chain: one more hop from ?Y
                ys = np.concatenate([gs.get_edges(int(x), pb, db) for x in xs] + [np.zeros(0, np.uint32)])
                pc, dc = _best_hop(gs, rng, ys[:64] if len(ys) else xs[:1], P)
                q += [-2, pc, dc, -3]
        tmpls.append(np.array(q, dtype=np.int64))
        members.append(m)
    trip = [(2, 4, 5)] + [linked_triple(rng, P) for _ in range(3)]
    for pa, pb, pc in trip:
        tmpls.append(np.array(heavy_query(P, t0, pa, pb, pc), dtype=np.int64))
    loads = np.array(BATCH_LIGHT_LOADS + BATCH_HEAVY_LOADS, dtype=np.float64)
    types = rng.choice(len(loads), size=n, p=loads / loads.sum())
    lens = np.array([len(x) for x in tmpls], dtype=np.int64)
    off = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(lens[types], out=off[1:])
    flat = np.empty(int(off[-1]), dtype=np.int64)
    for k in range(len(tmpls)):
        idx = np.nonzero(types == k)[0]
        if not idx.size:
            continue
        flat[off[idx][:, None] + np.arange(lens[k])] = tmpls[k]
        if k < len(BATCH_LIGHT_LOADS):
            flat[off[idx]] = rng.choice(members[k], size=idx.size)
    heavy = (types >= len(BATCH_LIGHT_LOADS)).astype(np.uint8)
    return flat, off, heavy, types.astype(np.int32)


def config4(seed=4, n_light_types=1023, keep_triples=False, N=100_000_000, P=256, T=1_000_000_000,
            t0_members=500_000, light_max=1e4):
    """Served graph: 100M entities, 256 predicates, ~1B triples (Zipf 1.0,
    predicates linked in pairs), type t0 with 500k members for the heavy
    queries and n_light_types types with log-uniform sizes in [1, 1e4] for
    the light ones. Smaller N, P, T give the same shape at a stated scale.
    The triples are released once the image is built unless kept."""
    rng = np.random.default_rng(seed)
    extra = np.exp(rng.uniform(0.0, np.log(light_max), n_light_types)).astype(np.int64)
    sizes = [t0_members] + [int(x) for x in extra]
    tr = Triples.generate(seed=seed, num_entities=N, num_preds=P, num_triples=T, zipf_s=1.0,
                          type_sizes=sizes, num_linked=P // 2, link_frac=LINK_FRAC)
    n_tr = len(tr)
    t_build = time.perf_counter()
    gs = GStore.build(tr)
    t_build = time.perf_counter() - t_build
    if not keep_triples:
        tr.close()
        tr = None
    t0 = 2 + P
    info = {"entities": N, "predicates": P, "triples": n_tr, "types": len(sizes), "t0": t0,
            "t0_members": sizes[0], "light_types": list(range(t0 + 1, t0 + len(sizes))), "keys": gs.num_keys,
            "edges": gs.num_edges, "image_bytes": gs.image_nbytes()}
    info["host_build_s"] = t_build
    return Workload("config4", tr, gs, {}, info)


def cyclic_query(t, pa, ph, pc, pq, const):
    """Config-5 six-pattern cycle from type index t:
        ?a rdf:type t .   index_to_unknown
        ?a pa ?b .        known_to_unknown  OUT
        ?c ph ?b .        known_to_unknown  IN   (?c: ph's subjects, the hub among them)
        ?c ph ?d .        known_to_unknown  OUT  (hub fan-out: the ~1B-row table)
        ?d pc ?b .        known_to_known         (closes the triangle b-c-d)
        ?c pq const .     known_to_const
    On a graph where ph plants pc edges between consecutive objects of each
    ph list (sibling plants), the triangle closes for about half of the
    (b, c) pairs, on the d that is b's planted partner.
    """
    return [t, 1, IN, -1,
            -1, pa, OUT, -2,
            -2, ph, IN, -3,
            -3, ph, OUT, -4,
            -4, pc, OUT, -2,
            -3, pq, OUT, const]


CONFIG5_PREDS = {"ph": 2, "pa": 100, "pc": 200, "pq": 2}


def config5_triples(seed=5, N=200_000_000, P=512, T=2_000_000_000, hub_degree=2_000_000, t0_members=2_000_000,
                    sample_members=20_000):
    """The config-5 triples: Zipf 0.9, the rank-1 subject of each of the
    first 8 predicates a hub of hub_degree edges, types t0 and t1, and pc
    edges planted between consecutive objects of every ph list."""
    c = CONFIG5_PREDS
    return Triples.generate(seed=seed, num_entities=N, num_preds=P, num_triples=T, zipf_s=0.9,
                            type_sizes=[t0_members, sample_members], num_hubs=8, hub_degree=hub_degree,
                            siblings=[(c["ph"], c["pc"])])


def config5_queries(gs, P):
    """The cycle from t0 (full) and from t1 (sample), through hub predicate
    ph; the known_to_const constant is one of the ph hub's own objects."""
    c = CONFIG5_PREDS
    t0, t1 = 2 + P, 2 + P + 1
    ph = c["ph"]
    hub, best = None, -1
    for s in gs.get_edges(0, ph, IN).tolist():
        d = gs.degree(s, ph, OUT)
        if d > best:
            hub, best = s, d
    const = int(gs.get_edges(hub, ph, OUT)[0])
    queries = {"full": cyclic_query(t0, c["pa"], ph, c["pc"], c["pq"], const),
               "sample": cyclic_query(t1, c["pa"], ph, c["pc"], c["pq"], const)}
    return queries, {"t0": t0, "t1": t1, **c, "hub": hub, "hub_degree": best, "const": const}


def config5(seed=5, t0_members=2_000_000, sample_members=20_000):
    """Large graph: 200M entities, 512 predicates, ~2B triples, Zipf 0.9, and
    hubs: the rank-1 subject of each of the first 8 predicates has 2M
    edges. Types: t0 (t0_members) for the full query and t1
    (sample_members) for a sample. The cyclic query goes through hub
    predicate ph = 2; the known_to_const constant is one of the hub's own
    ph objects."""
    N, P = 200_000_000, 512
    tr = config5_triples(seed, N, P, t0_members=t0_members, sample_members=sample_members)
    n_tr = len(tr)
    gs = GStore.build(tr)
    tr.close()
    queries, qi = config5_queries(gs, P)
    info = {"entities": N, "predicates": P, "triples": n_tr, "t0_members": t0_members,
            "t1_members": sample_members, **qi, "keys": gs.num_keys, "edges": gs.num_edges,
            "image_bytes": gs.image_nbytes()}
    return Workload("config5", None, gs, queries, info)
