"""GPU parity of single pattern steps against the CPU oracle (bit-exact,
row for row: both keep input-row order and edge-list order)."""
import numpy as np
import pytest

import oracle
from tests.conftest import VB
from wukonggpu_amd import Device, STEP_I2U, STEP_C2U, STEP_K2U, STEP_K2K, STEP_K2C

pytestmark = pytest.mark.gpu


def _eq(a, b, what):
    assert a.shape == b.shape, (what, a.shape, b.shape)
    assert np.array_equal(a, b), what


def _oracle_step(gs, cmd, table):
    return oracle.run_step(gs, cmd, 1, table) if table is not None else oracle.run_query(gs, cmd)[:2]


def _input_table(rng, nrows, ncols, n_ent, miss_frac=0.1):
    t = (VB + rng.integers(0, n_ent, size=(nrows, ncols))).astype(np.int32)
    miss = rng.random((nrows, ncols)) < miss_frac
    t[miss] = VB + n_ent + 1000  # vertices with no edges
    return t


@pytest.fixture(scope="module")
def dev_hubs(graph_hubs):
    d = Device(0)
    d.load(graph_hubs[1])
    yield d
    d.close()


def test_i2u_all_indexes(dev_small, graph_small):
    _, gs = graph_small
    for idx in range(2, gs.num_metas):
        for d in (0, 1):
            t, ex = dev_small.i2u(idx, d)
            want, wex, _ = oracle.run_query(gs, [idx, 0, d, -1])
            _eq(t.to_host(), want, ("i2u", idx, d))
            assert ex == wex


def test_c2u(dev_small, graph_small):
    _, gs = graph_small
    for v in list(range(VB, VB + 300, 7)) + [VB + 99999]:
        for pid in (1, 2, 3, 4, 5):
            for d in (0, 1):
                t, ex = dev_small.c2u(v, pid, d)
                want, wex, _ = oracle.run_query(gs, [v, pid, d, -1])
                _eq(t.to_host(), want, ("c2u", v, pid, d))
                assert ex == wex


@pytest.mark.parametrize("ncols,nrows", [(1, 1), (1, 257), (3, 5000), (7, 20000), (40, 900)])
def test_k2u_random_tables(dev_small, graph_small, ncols, nrows):
    _, gs = graph_small
    rng = np.random.default_rng(ncols * 1000 + nrows)
    tab = _input_table(rng, nrows, ncols, 300)
    tab[rng.random((nrows, ncols)) < 0.02] = -5  # negative cells never match a key
    dt = dev_small.table(tab)
    for sc in sorted({0, ncols - 1}):
        for pid in (1, 2, 3, 4):
            for d in (0, 1):
                cmd = [VB, 2, 1, -1] + [-(sc + 1), pid, d, -(ncols + 1)]
                want, wex = oracle.run_step(gs, cmd, 1, tab)
                t, ex = dev_small.k2u(dt, sc, pid, d)
                _eq(t.to_host(), want, ("k2u", ncols, nrows, sc, pid, d))
                assert ex == wex


@pytest.mark.parametrize("kind", ["k2k", "k2c"])
def test_filters_random_tables(dev_small, graph_small, kind):
    _, gs = graph_small
    rng = np.random.default_rng(5)
    for nrows, ncols in [(1, 2), (300, 2), (4000, 3), (70000, 4)]:
        tab = _input_table(rng, nrows, ncols, 300, miss_frac=0.05)
        dt = dev_small.table(tab)
        for pid in (1, 2, 3):
            for d in (0, 1):
                if kind == "k2k":
                    cmd = [VB, 2, 1, -1] + [-1, pid, d, -ncols]
                    t, ex = dev_small.k2k(dt, 0, pid, d, ncols - 1)
                else:
                    const = int(tab[0, -1])
                    cmd = [VB, 2, 1, -1] + [-1, pid, d, const]
                    t, ex = dev_small.k2c(dt, 0, pid, d, const)
                want, wex = oracle.run_step(gs, cmd, 1, tab)
                _eq(t.to_host(), want, (kind, nrows, pid, d))
                assert ex == wex, (kind, nrows, pid, d, ex, wex)


def test_filter_hits_rows_built_from_edges(dev_small, graph_small):
    # rows (x, y) with y drawn from x's own list must all survive known_to_known
    _, gs = graph_small
    t1, _ = dev_small.i2u(2, 0)
    t2, _ = dev_small.k2u(t1, 0, 2, 1)
    t3, ex = dev_small.k2k(t2, 0, 2, 1, 1)
    a, b = t2.to_host(), t3.to_host()
    _eq(b, a, "k2k keeps every row")


def test_k2c_const_out_of_range(dev_small, graph_small):
    _, gs = graph_small
    tab = _input_table(np.random.default_rng(1), 500, 2, 300)
    dt = dev_small.table(tab)
    # cmd_chains cannot carry a negative constant (it would be a variable), so
    # both constants are checked against the oracle's answer for one that no
    # u32 edge can equal: no row survives and every edge is examined
    want, wex = oracle.run_step(gs, [VB, 2, 1, -1, -1, 2, 1, 1 << 33], 1, tab)
    for const in (-7, 1 << 33):
        t, ex = dev_small.k2c(dt, 0, 2, 1, const)
        _eq(t.to_host(), want, const)
        assert t.shape[0] == 0 and ex == wex


def test_empty_inputs(dev_small):
    e = dev_small.table(np.zeros((0, 3), dtype=np.int32))
    for t, ex in (dev_small.k2u(e, 1, 2, 1), dev_small.k2k(e, 0, 2, 1, 2), dev_small.k2c(e, 0, 2, 1, VB)):
        assert t.shape[0] == 0 and ex == 0
    assert dev_small.k2u(e, 1, 2, 1)[0].shape == (0, 4)


def test_hub_rows_span_tiles(dev_hubs, graph_hubs):
    # hub subjects have 5000 edges: their outputs cross several expand tiles
    _, gs = graph_hubs
    rng = np.random.default_rng(2)
    for pid in (2, 3):
        t_sub, _ = dev_hubs.i2u(pid, 0)
        subs = t_sub.to_host()
        tab = np.concatenate([subs, _input_table(rng, 3000, 1, 6000)])[rng.permutation(len(subs) + 3000)]
        tab = np.repeat(tab, 2, axis=1)
        dt = dev_hubs.table(tab)
        for d in (0, 1):
            cmd = [VB, 2, 1, -1, -2, pid, d, -3]
            want, wex = oracle.run_step(gs, cmd, 1, tab)
            t, ex = dev_hubs.k2u(dt, 1, pid, d)
            _eq(t.to_host(), want, ("hub k2u", pid, d))
            assert ex == wex
            # known_to_known against a column drawn from the expanded lists (both sorted-search paths)
            out = want
            cmd2 = [VB, 2, 1, -1, -2, 3, 1, -3, -1, pid, d, -3]
            w2, wex2 = oracle.run_step(gs, cmd2, 2, out)
            t2, ex2 = dev_hubs.k2k(t, 0, pid, d, 2)
            _eq(t2.to_host(), w2, ("hub k2k", pid, d))
            assert ex2 == wex2


def test_runs_of_equal_keys(dev_small, graph_small):
    # long runs of one start value (probed once per wave), runs crossing wave
    # and tile boundaries, misses and negative cells inside runs
    _, gs = graph_small
    rng = np.random.default_rng(21)
    vals = VB + rng.integers(0, 300, size=400)
    lens = rng.integers(1, 300, size=400)
    col = np.repeat(vals, lens).astype(np.int32)
    col[rng.random(len(col)) < 0.01] = -1
    col[rng.random(len(col)) < 0.01] = VB + 5000
    tab = np.stack([col, np.roll(col, 7)], axis=1)
    dt = dev_small.table(tab)
    for pid, d in ((2, 1), (3, 0), (1, 1)):
        want, wex = oracle.run_step(gs, [VB, 2, 1, -1, -2, pid, d, -3], 1, tab)
        t, ex = dev_small.k2u(dt, 1, pid, d)
        _eq(t.to_host(), want, ("runs k2u", pid, d))
        assert ex == wex
        want, wex = oracle.run_step(gs, [VB, 2, 1, -1, -1, pid, d, -2], 1, tab)
        t, ex = dev_small.k2k(dt, 0, pid, d, 1)
        _eq(t.to_host(), want, ("runs k2k", pid, d))
        assert ex == wex


def test_many_tiles_two_level_scan(dev_small, graph_small):
    # more than 2048 probe tiles of 2048 rows: the tile scan runs in two levels
    _, gs = graph_small
    rng = np.random.default_rng(22)
    n = 2049 * 2048 + 77
    tab = (VB + rng.integers(0, 300, size=(n, 1))).astype(np.int32)
    dt = dev_small.table(tab)
    want, wex = oracle.run_step(gs, [VB, 2, 1, -1, -1, 2, 1, -2], 1, tab)
    t, ex = dev_small.k2u(dt, 0, 2, 1)
    _eq(t.to_host(), want, "two-level k2u")
    assert ex == wex
    const = int(tab[5, 0])
    want, wex = oracle.run_step(gs, [VB, 2, 1, -1, -1, 3, 0, const], 1, tab)
    t, ex = dev_small.k2c(dt, 0, 3, 0, const)
    _eq(t.to_host(), want, "two-level k2c")
    assert ex == wex


def test_type_index_large(dev_hubs, graph_hubs):
    _, gs = graph_hubs
    tid = 2 + 3  # first type id after 3 predicates
    t, ex = dev_hubs.i2u(tid, 0)
    want, wex, _ = oracle.run_query(gs, [tid, 1, 0, -1])
    _eq(t.to_host(), want, "type index")
    assert want.shape[0] == 3000 and ex == wex


def test_step_batch_matches_plain(dev_small, graph_small):
    _, gs = graph_small
    rng = np.random.default_rng(8)
    tabs = [dev_small.table(_input_table(rng, n, c, 300)) for n, c in [(10, 1), (0, 2), (3000, 3), (1, 4), (700, 2)]]
    steps = [
        dict(kind=STEP_I2U, pid=3, dir=0),
        dict(kind=STEP_C2U, pid=2, dir=1, vid=VB + 3),
        dict(kind=STEP_K2U, pid=2, dir=1, start_col=0, input=tabs[0]),
        dict(kind=STEP_K2U, pid=3, dir=0, start_col=1, input=tabs[1]),
        dict(kind=STEP_K2U, pid=1, dir=1, start_col=2, input=tabs[2]),
        dict(kind=STEP_K2K, pid=2, dir=1, start_col=0, end_col=2, input=tabs[2]),
        dict(kind=STEP_K2C, pid=2, dir=0, start_col=3, vid=VB + 17, input=tabs[3]),
        dict(kind=STEP_K2U, pid=4, dir=0, start_col=1, input=tabs[4]),
        dict(kind=STEP_I2U, pid=6, dir=0),
    ]
    outs, exs = dev_small.step_batch(steps)
    for s, o, ex in zip(steps, outs, exs):
        k = s["kind"]
        if k == STEP_I2U:
            t, e = dev_small.i2u(s["pid"], s["dir"])
        elif k == STEP_C2U:
            t, e = dev_small.c2u(s["vid"], s["pid"], s["dir"])
        elif k == STEP_K2U:
            t, e = dev_small.k2u(s["input"], s["start_col"], s["pid"], s["dir"])
        elif k == STEP_K2K:
            t, e = dev_small.k2k(s["input"], s["start_col"], s["pid"], s["dir"], s["end_col"])
        else:
            t, e = dev_small.k2c(s["input"], s["start_col"], s["pid"], s["dir"], s["vid"])
        _eq(o.to_host(), t.to_host(), s)
        assert ex == e, s


def test_bad_arguments_rejected(dev_small):
    from wukonggpu_amd import DeviceError
    t = dev_small.table(np.full((4, 2), VB, dtype=np.int32))
    with pytest.raises(DeviceError):
        dev_small.k2u(t, 5, 2, 1)        # start column out of range
    with pytest.raises(DeviceError):
        dev_small.k2k(t, 0, 2, 1, 9)     # end column out of range
    with pytest.raises(DeviceError):
        dev_small.k2u(t, 0, 99999, 1)    # predicate outside the image
    with pytest.raises(DeviceError):
        dev_small.k2u(t, 0, 2, 3)        # direction
    with pytest.raises(DeviceError):
        dev_small.i2u(VB, 0)             # not an index id


def test_steps_on_wrapped_device_cells(dev_hubs, graph_hubs):
    """A table over device cells the caller owns (wkg_table_wrap_device: the
    reference's history buffers) runs every known-start step like a table
    of the library's own, and its fan-outs keep no lineage on it."""
    _, gs = graph_hubs
    rng = np.random.default_rng(71)
    tab = _input_table(rng, 30000, 3, 6000)
    owner = dev_hubs.table(tab)
    for rows in (30000, 1000):  # the whole buffer, then a prefix of it
        v = dev_hubs.wrap_device(owner, rows=rows)
        sub = tab[:rows]
        for cmd, fn in (([-1, 2, 1, -4], lambda t: dev_hubs.k2u(t, 0, 2, 1)),
                        ([-2, 3, 0, -1], lambda t: dev_hubs.k2k(t, 1, 3, 0, 0)),
                        ([-3, 2, 1, VB + 11], lambda t: dev_hubs.k2c(t, 2, 2, 1, VB + 11))):
            want, wex = oracle.run_step(gs, [VB, 2, 1, -1] + cmd, 1, sub)
            got, ex = fn(v)
            _eq(got.to_host(), want, ("wrapped", rows, cmd))
            assert ex == wex
            # a second step over the fan-out of a wrapped table (no lineage kept on it)
            if cmd[3] == -4:
                want2, wex2 = oracle.run_step(gs, [VB, 2, 1, -1, -1, 2, 1, -4] + [-1, 3, 1, -5], 2, want)
                got2, ex2 = dev_hubs.k2u(got, 0, 3, 1)
                _eq(got2.to_host(), want2, ("wrapped fan-out, star arm", rows))
                assert ex2 == wex2
                got2.free()
            got.free()
        v.free()
    owner.free()
