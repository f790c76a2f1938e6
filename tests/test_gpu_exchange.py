"""GPU parity of the fork-join table exchange (wkg_table_split,
wkg_table_concat) against the oracle's restatement: bit-exact, row for row
(both keep input order inside every sub-table)."""
import numpy as np
import pytest

from oracle.exchange import merge_replies, split_table
from wukonggpu_amd import DeviceError

pytestmark = pytest.mark.gpu


def _table(rng, rows, cols, lo=1 << 17, span=1 << 20):
    return rng.integers(lo, lo + span, size=(rows, cols), dtype=np.int64).astype(np.int32)


@pytest.mark.parametrize("rows,cols,col,nparts", [
    (0, 3, 1, 8), (1, 1, 0, 1), (63, 2, 0, 2), (64, 3, 2, 3), (65, 1, 0, 8), (1023, 4, 1, 8), (1024, 5, 4, 7),
    (1025, 2, 1, 64), (5000, 17, 9, 65), (20000, 3, 0, 1000), (70000, 6, 5, 1024), (300000, 4, 2, 8),
    (2000, 8, 6, 8), (9000, 8, 1, 100), (4000, 1, 0, 300), (3000, 12, 11, 5),
    (100000, 4, 3, 64), (50000, 2, 0, 33), (9000, 1, 0, 64), (4097, 4, 0, 64)])
def test_split_matches_oracle(dev_small, rows, cols, col, nparts):
    rng = np.random.default_rng(rows + 31 * nparts)
    t = _table(rng, rows, cols)
    print(f"split {rows}x{cols} on column {col} into {nparts}", flush=True)
    d = dev_small.table(t)
    out, prow, phead = dev_small.table_split(d, col, nparts)
    want, wrows, wheads = split_table(t, col, nparts)
    got = out.to_host()
    assert got.shape == want.shape and np.array_equal(got, want)
    assert np.array_equal(prow, wrows) and np.array_equal(phead, wheads)
    # each sub-table as a view, merged back in part order = the split table
    views = [dev_small.table_slice(out, int(h), int(n)) for h, n in zip(phead, prow)]
    merged = dev_small.table_concat(views).to_host()
    assert np.array_equal(merged, merge_replies([v.to_host() for v in views], cols))
    assert np.array_equal(merged, want)
    for v in views:
        v.free()


def test_split_skewed_and_negative_cells(dev_small):
    # all rows in one part, and cells above 2^31 (read as unsigned)
    rng = np.random.default_rng(3)
    t = _table(rng, 4096, 3)
    t[:, 1] = 8 * rng.integers(0, 1000, size=4096)  # every owner is part 0 of 8
    t[::7, 2] = -1 - rng.integers(0, 1000, size=len(t[::7]))
    for col, n in ((1, 8), (2, 5), (2, 1000)):
        out, prow, phead = dev_small.table_split(dev_small.table(t), col, n)
        want, wrows, wheads = split_table(t, col, n)
        assert np.array_equal(out.to_host(), want) and np.array_equal(prow, wrows) and np.array_equal(phead, wheads)


def test_split_many_tiles_property(dev_small):
    # size-independent checks on a larger table: ownership, stability and the
    # row multiset through an index column
    rng = np.random.default_rng(4)
    rows = 3_000_000
    t = np.empty((rows, 3), dtype=np.int32)
    t[:, 0] = _table(rng, rows, 1)[:, 0]
    t[:, 1] = np.arange(rows, dtype=np.int32)
    t[:, 2] = _table(rng, rows, 1)[:, 0]
    print(f"split {rows}x3 into 8", flush=True)
    out, prow, phead = dev_small.table_split(dev_small.table(t), 0, 8)
    got = out.to_host()
    for p in range(8):
        seg = got[int(phead[p]):int(phead[p] + prow[p])]
        assert ((seg[:, 0].view(np.uint32) % 8) == p).all()
        assert (np.diff(seg[:, 1]) > 0).all()
    assert np.array_equal(np.sort(got[:, 1]), t[:, 1])
    assert np.array_equal(got[np.argsort(got[:, 1])], t)


def test_concat_edges(dev_small):
    rng = np.random.default_rng(5)
    a, b = _table(rng, 10, 2), _table(rng, 3, 2)
    empty = dev_small.table(np.zeros((0, 5), dtype=np.int32))  # an empty reply: any column count
    got = dev_small.table_concat([dev_small.table(a), empty, dev_small.table(b)]).to_host()
    assert np.array_equal(got, np.vstack([a, b]))
    assert dev_small.table_concat([]).shape == (0, 0)
    with pytest.raises(DeviceError):
        dev_small.table_concat([dev_small.table(a), dev_small.table(_table(rng, 2, 3))])


def test_split_bad_arguments(dev_small):
    d = dev_small.table(np.ones((4, 2), dtype=np.int32))
    for col, n in ((2, 4), (-1, 4), (0, 0), (0, 1025)):
        with pytest.raises(DeviceError):
            dev_small.table_split(d, col, n)
