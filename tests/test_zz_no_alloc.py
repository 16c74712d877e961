"""The allocation-free compaction (wf_compact_u32_noalloc) against the workspace compaction
(wf_compact_u32), through the C ABI.

Two calls in a row share the library's static partial sums on the device, the second with a
smaller input, so a partial left over from the first call would show up in the second result.
Sizes above the static array's limit are refused with no side effects.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CANARY = 0xDEADBEEF


def compact_with_workspace(wf, vals, flags):
    from hipmem import DevArray
    n = vals.size
    ws = C.c_void_p()
    wf._check(wf.lib().wf_scan_workspace_create(n, C.byref(ws)))
    d_v, d_f = DevArray(vals), DevArray(flags)
    d_o = DevArray(np.full(n + 1, CANARY, np.uint32))
    d_c = DevArray(np.zeros(1, np.uint32))
    wf._check(wf.lib().wf_compact_u32(d_v.ptr, d_f.ptr, d_o.ptr, d_c.ptr, None, n, ws, None, 0))
    out, cnt = d_o.get(), int(d_c.get()[0])
    wf.lib().wf_scan_workspace_destroy(ws)
    return out, cnt


def compact_noalloc(wf, vals, flags, tile=0):
    from hipmem import DevArray
    n = vals.size
    d_v, d_f = DevArray(vals), DevArray(flags)
    d_o = DevArray(np.full(n + 1, CANARY, np.uint32))
    d_c = DevArray(np.full(1, CANARY, np.uint32))
    wf._check(wf.lib().wf_compact_u32_noalloc(d_v.ptr, d_f.ptr, d_o.ptr, d_c.ptr, n, None, tile))
    return d_o.get(), int(d_c.get()[0])


def inputs(n, seed, density):
    rs = np.random.default_rng(seed)
    flags = ((rs.random(n) < density) * rs.integers(1, 9, size=n)).astype(np.uint32)
    vals = rs.integers(0, 2**32 - 1, size=n, dtype=np.uint32)
    return vals, flags


@pytest.mark.parametrize("sizes", [(1_000_003, 65_537), (921_600, 777), (4096, 1)])
def test_noalloc_compaction_equals_workspace_compaction_twice_in_a_row(wf, sizes):
    for k, n in enumerate(sizes):
        print(f"[noalloc] call {k} n={n}", flush=True)
        # The first call is dense, the second sparse: leftover partial sums would shift positions.
        vals, flags = inputs(n, 100 + n, 0.9 if k == 0 else 0.2)
        want_out, want_cnt = compact_with_workspace(wf, vals, flags)
        got_out, got_cnt = compact_noalloc(wf, vals, flags)
        assert got_cnt == want_cnt == int((flags != 0).sum())
        assert np.array_equal(got_out[:got_cnt], want_out[:want_cnt])
        assert np.array_equal(got_out[:got_cnt], vals[flags != 0])
        assert got_out[got_cnt] == CANARY and got_out[n] == CANARY


@pytest.mark.parametrize("tile", [512, 1024, 2048, 4096])
def test_noalloc_compaction_every_tile(wf, tile):
    vals, flags = inputs(300_001, tile, 0.5)
    want_out, want_cnt = compact_with_workspace(wf, vals, flags)
    got_out, got_cnt = compact_noalloc(wf, vals, flags, tile)
    assert got_cnt == want_cnt
    assert np.array_equal(got_out[:got_cnt], want_out[:want_cnt])


def test_noalloc_compaction_empty_and_limits(wf):
    from hipmem import DevArray
    vals, flags = inputs(16, 1, 0.5)
    out, cnt = compact_noalloc(wf, vals[:0], flags[:0])
    assert cnt == 0
    d = DevArray(vals)
    d_c = DevArray(np.full(1, CANARY, np.uint32))
    # Above the static array's limit: refused, nothing written.
    rc = wf.lib().wf_compact_u32_noalloc(d.ptr, d.ptr, d.ptr, d_c.ptr, (1 << 26) + 1, None, 0)
    assert rc == wf.WF_ERR_INVALID
    assert "2^26" in wf.lib().wf_last_error().decode()
    assert int(d_c.get()[0]) == CANARY and np.array_equal(d.get(), vals)
    assert wf.lib().wf_compact_u32_noalloc(d.ptr, d.ptr, d.ptr, d_c.ptr, 16, None, 3000) == wf.WF_ERR_INVALID
    assert wf.lib().wf_compact_u32_noalloc(d.ptr, d.ptr, d.ptr, d_c.ptr, -1, None, 0) == wf.WF_ERR_INVALID
