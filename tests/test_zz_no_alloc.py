"""wkg_step_k2u_noalloc: known_to_unknown into the caller's output buffer with
the library's fixed device scratch, no allocation. Its output must equal
wkg_step_k2u's row for row (and the oracle's), on two calls in a row, the
second on a smaller input, and the table pool must see no allocation."""
import numpy as np
import pytest

import oracle
from tests.conftest import VB
from wukonggpu_amd import Device, DeviceError

pytestmark = pytest.mark.gpu

NOALLOC_SCRATCH = 64 << 20  # WKG_NOALLOC_SCRATCH_BYTES


def test_noalloc_k2u_matches_k2u(graph_hubs):
    _, gs = graph_hubs
    dev = Device(0)
    dev.load(gs)
    hub = max(gs.get_edges(0, 2, 0).tolist(), key=lambda v: gs.degree(v, 2, 1))
    rng = np.random.default_rng(5)
    out_buf = None
    for R in (60_000, 2_500):  # the second call is smaller, into the same buffer
        t = (VB + rng.integers(0, 6000, size=(R, 2))).astype(np.int32)
        t[::997, 0] = hub  # hub rows: 5000-edge lists spanning expand tiles
        print(f"no-alloc k2u on {R} rows", flush=True)
        din = dev.table(t)
        ref, rex = dev.k2u(din, 0, 2, 1)
        want = ref.to_host()
        ref.free()
        if out_buf is None:
            out_buf = dev.table(np.zeros((want.shape[0] + 1000, 3), dtype=np.int32))
        before = dev.pool_stats()
        rows, ex = dev.k2u_noalloc(din, 0, 2, 1, out_buf)
        after = dev.pool_stats()
        assert after["mallocs"] == before["mallocs"] and after["live_bytes"] == before["live_bytes"]
        assert rows == want.shape[0] and ex == rex
        v = dev.table_slice(out_buf, 0, rows)
        got = v.to_host()
        v.free()
        assert np.array_equal(got, want)
        ow, owex = oracle.run_step(gs, [VB, 2, 1, -1, -1, 0, 0, -2, -1, 2, 1, -3], 2, t)
        assert np.array_equal(got, ow) and ex == owex
        din.free()
    # an output buffer too small: WKG_ERANGE, and the row count it needs
    din = dev.table(np.full((3, 1), hub, dtype=np.int32))
    small = dev.table(np.zeros((10, 2), dtype=np.int32))
    with pytest.raises(DeviceError) as e:
        dev.k2u_noalloc(din, 0, 2, 1, small)
    assert e.value.rc == -4
    # an input whose step scratch is above the fixed array: WKG_ERANGE before any launch
    big_rows = NOALLOC_SCRATCH // 16 + 1
    big = dev.table(np.full((big_rows, 1), VB, dtype=np.int32))
    with pytest.raises(DeviceError) as e:
        dev.k2u_noalloc(big, 0, 2, 1, small)
    assert e.value.rc == -4
    for x in (big, small, din, out_buf):
        x.free()
    dev.close()
