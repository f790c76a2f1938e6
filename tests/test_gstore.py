"""Gstore builder and generator invariants (CPU).

The image is checked through the oracle's own bucket walk (its hash and
lookup share no code with the builder) and against the raw triples."""
import os

import numpy as np
import pytest

import oracle
from oracle.brute import BruteGraph
from tests.conftest import VB, small_graph
from wukonggpu_amd import GStore, Triples


def _lists_equal(gs, bg, vids, pids):
    for v in vids:
        for pid in pids:
            for d in (0, 1):
                want = bg.edges(v, pid, d)
                got = oracle.get_edges(gs, v, d, pid)
                got = [] if got is None else list(got)
                assert got == want, (v, pid, d)
                assert list(gs.get_edges(v, pid, d)) == want, (v, pid, d)


def test_every_key_found(graph_small):
    tr, gs = graph_small
    s, p, o = tr.arrays()
    bg = BruteGraph(s, p, o)
    _lists_equal(gs, bg, list(range(VB, VB + 300)) + [0], list(range(0, gs.num_metas)))


@pytest.mark.parametrize("load", [0.3, 1.0])
def test_heavy_chaining(load):
    # load 1.0 forces long chains through the indirect region
    tr, gs = small_graph(seed=19, n=200, preds=3, triples=2000, load=load)
    s, p, o = tr.arrays()
    _lists_equal(gs, BruteGraph(s, p, o), list(range(VB, VB + 200)) + [0], list(range(0, gs.num_metas)))
    if load == 1.0:
        assert gs.num_buckets_ext > 0


def test_absent_keys_miss(graph_small):
    _, gs = graph_small
    for v, pid, d in [(VB + 10 ** 6, 2, 1), (VB, 9999, 0), (VB + 5, 0, 1), (12345, 3, 0)]:
        assert oracle.get_edges(gs, v, d, pid) is None or len(oracle.get_edges(gs, v, d, pid)) == 0
        assert len(gs.get_edges(v, pid, d)) == 0


def test_metas_segments(graph_small):
    _, gs = graph_small
    m = gs.metas()
    used = m[m[:, 0] > 0]
    # main partitions are disjoint, consecutive and sized partition_sz
    assert np.all(used[:, 2] - used[:, 1] == used[:, 0])
    order = np.argsort(used[:, 1])
    assert np.all(used[order][1:, 1] >= used[order][:-1, 2])
    # edge segments are disjoint and cover the edge array
    segs = sorted((int(a), int(b)) for a, b in zip(m[:, 6], m[:, 7]) if b > a)
    assert segs[0][0] == 0 and segs[-1][1] == gs.num_edges
    assert all(segs[i][1] == segs[i + 1][0] for i in range(len(segs) - 1))
    # indirect ranges are disjoint and inside the indirect region
    ind = sorted((int(a), int(b)) for a, b in zip(m[:, 3], m[:, 4]) if b > a)
    assert all(a >= gs.num_buckets for a, _ in ind)
    assert all(ind[i][1] <= ind[i + 1][0] for i in range(len(ind) - 1))


def test_lists_sorted_and_within_segment(graph_small):
    _, gs = graph_small
    sl, e, m = gs.slots(), gs.edges(), gs.metas()
    keys, ptrs = sl[:, 0], sl[:, 1]
    data = (np.arange(len(keys)) % 8 != 7) & (keys != 0)
    for k, ptr in zip(keys[data], ptrs[data]):
        size, off = int(ptr) & ((1 << 28) - 1), int(ptr) >> 28
        pid = (int(k) >> 1) & ((1 << 17) - 1)
        assert m[pid, 6] <= off and off + size <= m[pid, 7]
        lst = e[off:off + size]
        assert np.all(lst[1:] > lst[:-1])


def test_chains_terminate_and_link_once(graph_small):
    _, gs = graph_small
    sl = gs.slots()
    links = sl[7::8, 0]
    tgt = (links[links != 0] >> 18).astype(np.int64)
    assert np.all(tgt >= gs.num_buckets) and np.all(tgt < gs.num_buckets + gs.num_buckets_ext)
    assert len(np.unique(tgt)) == len(tgt)


def test_key_packing(graph_small):
    # raw u64 of ikey_t(v, p, d) is v<<18 | p<<1 | d (bit fields dir:1, pid:17, vid:46)
    _, gs = graph_small
    sl = gs.slots()
    k = sl[:, 0][(np.arange(len(sl)) % 8 != 7) & (sl[:, 0] != 0)].astype(np.uint64)
    vid, pid, d = k >> np.uint64(18), (k >> np.uint64(1)) & np.uint64((1 << 17) - 1), k & np.uint64(1)
    assert np.all((vid == 0) | (vid >= VB))
    assert np.all((pid >= 1) & (pid < gs.num_metas)) and np.all(d <= 1)


def test_generator_properties():
    tr = Triples.generate(seed=1, num_entities=5000, num_preds=4, num_triples=40000, zipf_s=1.1, degree_cap=100,
                          type_sizes=[123, 1])
    s, p, o = (np.asarray(a) for a in tr.arrays())
    norm = p != 1
    # no duplicate triples, ids in range, the cap holds
    trip = np.stack([s, p, o], 1)
    assert len(np.unique(trip, axis=0)) == len(trip)
    assert s.min() >= VB and s.max() < VB + 5000 and o[norm].min() >= VB and o[norm].max() < VB + 5000
    for pid in range(2, 6):
        _, cnt = np.unique(s[p == pid], return_counts=True)
        assert cnt.max() <= 100
        assert abs(int((p == pid).sum()) - 10000) < 500  # about T/P per predicate
    # type sizes exact
    assert (p == 1).sum() == 124 and ((p == 1) & (o == 6)).sum() == 123
    # deterministic
    tr2 = Triples.generate(seed=1, num_entities=5000, num_preds=4, num_triples=40000, zipf_s=1.1, degree_cap=100,
                           type_sizes=[123, 1])
    assert all(np.array_equal(a, b) for a, b in zip(tr.arrays(), tr2.arrays()))


def test_generator_hubs():
    tr = Triples.generate(seed=3, num_entities=20000, num_preds=3, num_triples=30000, zipf_s=0.9, num_hubs=2,
                          hub_degree=8000)
    s, p, _ = (np.asarray(a) for a in tr.arrays())
    for pid in (2, 3):
        _, cnt = np.unique(s[p == pid], return_counts=True)
        assert cnt.max() >= 8000


def test_save_load_and_file_io(tmp_path, graph_small):
    tr, gs = graph_small
    path = os.path.join(tmp_path, "img.bin")
    gs.save(path)
    g2 = GStore.load(path)
    assert np.array_equal(g2.slots(), gs.slots()) and np.array_equal(g2.edges(), gs.edges())
    assert np.array_equal(g2.metas(), gs.metas())
    tp = os.path.join(tmp_path, "id_triples")
    tr.write(tp)
    t2 = Triples.read(tp)
    assert all(np.array_equal(a, b) for a, b in zip(tr.arrays(), t2.arrays()))


def test_builder_rejects_bad_triples():
    with pytest.raises(RuntimeError):
        GStore.build(Triples.from_arrays([VB], [1], [VB + 1]))  # rdf:type object must be a type id
    with pytest.raises(RuntimeError):
        GStore.build(Triples.from_arrays([5], [2], [VB]))       # subject must be a normal vertex
    with pytest.raises(RuntimeError):
        GStore.build(Triples.from_arrays([VB], [0], [VB]))      # predicate 0 is reserved


def test_generator_planted_links_and_siblings():
    # linked pairs (2k+3 takes a share of its lists from 2k+2's list of the
    # same subject) and sibling plants ((o_2i, dst, o_2i+1) for consecutive
    # objects of every source list) leave every other list and every degree
    # as the plain generator makes them
    import math
    kw = dict(seed=3, num_entities=20000, num_preds=6, num_triples=200000, zipf_s=1.0, type_sizes=[100])
    a = Triples.generate(**kw)
    b = Triples.generate(**kw, num_linked=2, link_frac=0.25, siblings=[(2, 7), (4, 3)])
    sa, pa, oa = (x.copy() for x in a.arrays())
    sb, pb, ob = (x.copy() for x in b.arrays())
    n = len(sa)

    def lists(s, p, o, pid):
        m = p == pid
        out = {}
        for x, y in zip(s[m].tolist(), o[m].tolist()):
            out.setdefault(x, []).append(y)
        return out

    for pid in range(2, 8):
        la, lb = lists(sa, pa, oa, pid), lists(sb[:n], pb[:n], ob[:n], pid)
        assert {k: len(v) for k, v in la.items()} == {k: len(v) for k, v in lb.items()}, pid
        if pid in (2, 4, 6, 7):  # sources and the unlinked pair
            assert la == lb, pid
    for src, dst in ((2, 3), (4, 5)):
        ls, ld = lists(sb[:n], pb[:n], ob[:n], src), lists(sb[:n], pb[:n], ob[:n], dst)
        both = 0
        for s, objs in ld.items():
            assert len(set(objs)) == len(objs)
            if s in ls:
                both += 1
                need = min(math.ceil(0.25 * len(objs)), len(ls[s]))
                assert len(set(objs) & set(ls[s])) >= need, (dst, s)
        assert both > 100
    planted = {}
    for x, p, y in zip(sb[n:].tolist(), pb[n:].tolist(), ob[n:].tolist()):
        planted.setdefault(p, []).append((x, y))
    for src, dst in ((2, 7), (4, 3)):
        want = []
        for s, objs in sorted(lists(sb[:n], pb[:n], ob[:n], src).items()):
            objs = sorted(objs)
            want += [(objs[i], objs[i + 1]) for i in range(0, len(objs) - 1, 2)]
        assert sorted(planted[dst]) == sorted(want), dst
    a.close()
    b.close()


def test_dataset_directory_roundtrip(tmp_path):
    # the reference's ID-format dataset (id_* files, str_index, str_normal):
    # written and read back it gives the same triples, in file order, and the
    # same image; a predicate missing from str_index is rejected
    tr, gs = small_graph(seed=31, n=400, preds=5, triples=4000, types=(40, 7))
    d = tmp_path / "ds"
    tr.write_dataset(d, files=3)
    names = sorted(p.name for p in d.iterdir())
    assert names == ["id_0000.nt", "id_0001.nt", "id_0002.nt", "str_index", "str_normal"]
    back, info = Triples.read_dataset(d)
    a, b = tr.arrays(), back.arrays()
    assert all(np.array_equal(x, y) for x, y in zip(a, b))
    s, p, o = a
    assert info["files"] == 3 and info["triples"] == len(s)
    ids = set(p.tolist()) | set(o[p == 1].tolist())
    assert info["has_str_index"] and info["index_ids"] == len(ids | {1}) and info["max_index_id"] == max(ids)
    normal = set(s.tolist()) | set(o[p != 1].tolist())
    assert info["has_str_normal"] and info["normal_ids"] == len(normal) and info["max_normal_id"] == max(normal)
    g2 = GStore.build(back)
    assert np.array_equal(g2.slots(), gs.slots()) and np.array_equal(g2.edges(), gs.edges())
    assert np.array_equal(g2.metas(), gs.metas())
    # str_index without predicate 3: the dataset is inconsistent
    lines = (d / "str_index").read_text().splitlines()
    (d / "str_index").write_text("\n".join(x for x in lines if not x.endswith("\t3")) + "\n")
    with pytest.raises(RuntimeError):
        Triples.read_dataset(d)
    # a directory with no id_* files
    (tmp_path / "empty").mkdir()
    with pytest.raises(RuntimeError):
        Triples.read_dataset(tmp_path / "empty")
