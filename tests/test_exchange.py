"""Fork-join table exchange oracle (CPU): known answer, brute force and the
invariants a wrong column, modulus, order or offset would break."""
import numpy as np
import pytest

from oracle.exchange import merge_replies, split_table


def test_known_answer():
    t = np.array([[5, 1], [8, 2], [3, 3], [6, 4], [9, 5], [2, 6]], dtype=np.int32)
    out, rows, heads = split_table(t, 0, 3)
    # owners 5%3=2, 8%3=2, 3%3=0, 6%3=0, 9%3=0, 2%3=2
    assert out.tolist() == [[3, 3], [6, 4], [9, 5], [5, 1], [8, 2], [2, 6]]
    assert rows.tolist() == [3, 0, 3] and heads.tolist() == [0, 3, 3]
    out1, rows1, _ = split_table(t, 1, 2)  # by the second column
    assert out1.tolist() == [[8, 2], [6, 4], [2, 6], [5, 1], [3, 3], [9, 5]] and rows1.tolist() == [3, 3]


def _brute(t, col, n):
    parts = [[] for _ in range(n)]
    for row in t.tolist():
        parts[(row[col] & 0xffffffff) % n].append(row)
    return parts


@pytest.mark.parametrize("rows,cols,col,n", [(0, 2, 1, 4), (1, 1, 0, 1), (97, 3, 2, 5), (500, 4, 0, 64),
                                             (300, 2, 1, 1000)])
def test_brute_force_and_invariants(rows, cols, col, n):
    rng = np.random.default_rng(rows * 7 + n)
    t = rng.integers(1 << 17, (1 << 17) + 5000, size=(rows, cols), dtype=np.int64).astype(np.int32)
    out, sizes, heads = split_table(t, col, n)
    parts = _brute(t, col, n)
    assert [len(p) for p in parts] == sizes.tolist()
    for p in range(n):
        got = out[int(heads[p]):int(heads[p] + sizes[p])]
        assert got.tolist() == parts[p]
        assert ((got[:, col].view(np.uint32) % n) == p).all()  # ownership
    assert int(sizes.sum()) == rows and (heads[1:] == np.cumsum(sizes)[:-1]).all()
    # the same multiset of rows
    assert sorted(map(tuple, out.tolist())) == sorted(map(tuple, t.tolist()))
    # split then merge in part order, then undo by each row's original index
    idx = np.arange(rows, dtype=np.int32)[:, None]
    out_i, _, _ = split_table(np.hstack([t, idx]), col, n)
    assert sorted(out_i[:, -1].tolist()) == list(range(rows))  # a permutation of the rows
    merged = merge_replies([out_i[int(h):int(h + s)] for h, s in zip(heads, sizes)], cols + 1)
    assert np.array_equal(merged[np.argsort(merged[:, -1])][:, :cols], t)
    # stable: original indices ascend inside every part
    for p in range(n):
        seg = out_i[int(heads[p]):int(heads[p] + sizes[p]), -1]
        assert (np.diff(seg) > 0).all()


def test_one_part_is_identity_and_negative_cells_unsigned():
    t = np.array([[-1, 4], [7, -5], [0, 0]], dtype=np.int32)
    out, sizes, heads = split_table(t, 0, 1)
    assert np.array_equal(out, t) and sizes.tolist() == [3] and heads.tolist() == [0]
    out, sizes, _ = split_table(t, 0, 10)
    # -1 is 0xffffffff = 4294967295 -> part 5; 7 -> 7; 0 -> 0
    assert out[:, 0].tolist() == [0, -1, 7] and sizes[5] == 1 and sizes[7] == 1


def test_merge_skips_empty_replies():
    a = np.array([[1, 2]], dtype=np.int32)
    assert merge_replies([np.zeros((0, 2), np.int32), a, [], a], 2).tolist() == [[1, 2], [1, 2]]
    assert merge_replies([], 3).shape == (0, 3)
