"""ORACLE -- test infrastructure only.

CPU restatement of the fork-join table exchange on one server's side:

* split_table: RCache::generate_sub_query (core/rcache.hpp:964-1043) with
  its kernels d_hash_dispatched_server_id, d_history_dispatch and
  d_update_result_table_sub (core/gpu_hash.cu:698-873): row i goes to
  sub-table table[i][col] % nparts (the owner rule hash_mod,
  utils/mymath.hpp:55-59), sub-tables back to back at the exclusive prefix
  sums of their sizes. The reference slots rows with an atomicAdd, so its
  order inside a sub-table is unspecified; this restatement keeps the input
  order (the stable partition), the order the device split promises.
  Cells are read as unsigned 32-bit vertex ids.
* merge_replies: RMap::put_reply (core/reply_map.hpp:43-80): sub-reply
  tables appended in arrival order, row counts summed.

Pinned in tests/test_exchange.py by a hand-written known answer, by
brute force on small tables and by invariants (ownership, row multiset,
stability, sizes, nparts = 1 identity, split-then-merge round trip).
"""
import numpy as np


def split_table(table, col, nparts):
    """(rows x cols int32 table, column, parts) -> (split table, rows per part, first row per part)."""
    t = np.ascontiguousarray(table, dtype=np.int32)
    if t.ndim != 2 or not 0 <= col < t.shape[1]:
        raise ValueError("bad table or column")
    if not 1 <= nparts:
        raise ValueError("nparts must be >= 1")
    owner = t[:, col].view(np.uint32).astype(np.uint64) % np.uint64(nparts)
    order = np.argsort(owner, kind="stable")
    sizes = np.bincount(owner.astype(np.int64), minlength=nparts).astype(np.uint64)
    heads = np.zeros(nparts, dtype=np.uint64)
    np.cumsum(sizes[:-1], out=heads[1:])
    return t[order], sizes, heads


def merge_replies(tables, cols):
    """Sub-reply tables appended in order; empty replies carry no cells."""
    parts = [np.asarray(x, dtype=np.int32).reshape(-1, cols) for x in tables if np.asarray(x).size]
    if not parts:
        return np.zeros((0, cols), dtype=np.int32)
    return np.concatenate(parts, axis=0)
