"""The allocation-free flattening entry point (lyg_flatten_segments_device_noalloc, no lyon
counterpart) against lyg_flatten_segments_device, through the C ABI: the same point count,
offsets and point bytes on two calls in a row, the second with fewer segments (its fixed
scratch keeps the first call's contents, so anything it fails to clear shows up here), and
a clear capacity error beyond its maximum segment count."""
import ctypes as C

import numpy as np
import pytest

from lyon_amd import _lib as B
from lyon_amd import workloads
from lyon_amd.batch import DeviceBuffer, _Ctx
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu

LYG_ERR_CAPACITY = 16


def _call(fn, dk, dc, n, tol, out, doff):
    cnt = C.c_uint32(0)
    st = fn(_Ctx.get(), dk.ptr, dc.ptr, n, tol, out.ptr if out is not None else None,
            (out.nbytes // 8) if out is not None else 0, doff.ptr, C.byref(cnt))
    return st, cnt.value


def _flatten_both(batch, tol):
    kinds, ctrl = workloads.segments_of(batch)
    n = len(kinds)
    print(f"noalloc: {n} segments, tolerance {tol}", flush=True)
    L = B.lib()
    dk = DeviceBuffer.from_numpy(kinds)
    dc = DeviceBuffer.from_numpy(ctrl.reshape(-1))
    res = []
    for fn in (L.lyg_flatten_segments_device, L.lyg_flatten_segments_device_noalloc):
        doff = DeviceBuffer(4 * (n + 1))
        st, total = _call(fn, dk, dc, n, tol, None, doff)
        assert st == LYG_ERR_CAPACITY and total > 0
        out = DeviceBuffer(8 * total)
        st, total2 = _call(fn, dk, dc, n, tol, out, doff)
        assert st == 0 and total2 == total
        res.append((out.download(np.float32, 2 * total), doff.download(np.uint32, n + 1)))
    (pa, oa), (pb, ob) = res
    assert np.array_equal(oa, ob), "offsets differ"
    assert np.array_equal(pa.view(np.uint32), pb.view(np.uint32)), "points differ"
    return kinds, ctrl, pb.reshape(-1, 2), ob


def test_noalloc_flatten_matches_allocating_entry_point():
    # first call: more than 32768 segments (the scan runs over tiles of the scratch)
    big = workloads.stroke_batch(600, 8, 256, 2048.0, seed=11)
    kinds, ctrl, pts, offs = _flatten_both(big, 0.05)
    assert len(kinds) > 32768
    # second call, fewer segments and another tolerance: the single-block scan
    small = workloads.stroke_batch(40, 8, 64, 1024.0, seed=12)
    kinds2, ctrl2, pts2, offs2 = _flatten_both(small, 0.1)
    assert len(kinds2) < len(kinds)
    # and the points are lyon's (oracle) on a sample of segments
    sel = np.arange(0, len(kinds2), 5)
    rp, ro = O.flatten_segments(kinds2[sel], ctrl2.reshape(-1, 4, 2)[sel], 0.1)
    for j, i in enumerate(sel):
        assert np.array_equal(pts2[offs2[i]:offs2[i + 1]].view(np.uint32),
                              rp[ro[j]:ro[j + 1]].view(np.uint32)), f"segment {i}"


def test_noalloc_flatten_rejects_more_segments_than_its_scratch():
    L = B.lib()
    m = L.lyg_flatten_noalloc_max_segments()
    assert m == 1 << 20
    dk = DeviceBuffer(16)
    dc = DeviceBuffer(16)
    doff = DeviceBuffer(16)
    st, total = _call(L.lyg_flatten_segments_device_noalloc, dk, dc, m + 1, 0.1, None, doff)
    assert st == LYG_ERR_CAPACITY and total == 0
    assert b"scratch" in L.lyg_ctx_last_error(_Ctx.get())
