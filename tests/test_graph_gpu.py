"""HIP graph replay of the batched stroke, and the autotuner: results must not change."""
import numpy as np
import pytest

from lyon_amd import workloads
from lyon_amd.batch import (DeviceBatch, FillBatcher, StrokeBatcher, StrokeGraph, autotune,
                            synchronize)

pytestmark = pytest.mark.gpu


def snapshot(res):
    return (res.vertex_offsets.copy(), res.index_offsets.copy(), res.all_vertices().copy(),
            res.all_indices().copy())


def same(a, b):
    return (np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and
            np.array_equal(a[2].view(np.uint32), b[2].view(np.uint32)) and
            np.array_equal(a[3], b[3]))


def test_graph_replays_match_eager():
    batch = workloads.stroke_batch(600, 8, 256, 2048.0, seed=4)
    opts = workloads.stroke_options(600, tolerance=0.05, seed=4)
    db = DeviceBatch(batch, opts)
    eager = StrokeBatcher()
    ref = snapshot(eager.run(db, vertex_offset=3))
    g = StrokeGraph(db, int(ref[2].shape[0]) + 100, int(ref[3].shape[0]) + 100, vertex_offset=3)
    other = DeviceBatch(workloads.stroke_batch(200, 8, 64, 2048.0, seed=5),
                        workloads.stroke_options(200, seed=5))
    for k in range(4):
        g.launch()
        r = g.result()
        print(f"replay {k}: {r.n_vertices} vertices", flush=True)
        assert same(snapshot(r), ref), f"replay {k} differs from the eager batch"
        eager.run(other)  # eager work between replays
        synchronize()


def test_graph_reports_overflow():
    batch = workloads.stroke_batch(50, 8, 64, 2048.0, seed=6)
    db = DeviceBatch(batch, workloads.stroke_options(50, seed=6))
    g = StrokeGraph(db, 64, 64)
    g.launch()
    with pytest.raises(Exception):
        g.result()


def test_autotune_keeps_results():
    batch = workloads.stroke_batch(800, 8, 128, 2048.0, seed=7)
    db = DeviceBatch(batch, workloads.stroke_options(800, seed=7))
    b = StrokeBatcher()
    ref = snapshot(b.run(db))
    shape = autotune(b, db)
    print("stroke shape", shape, flush=True)
    assert shape["blocks_per_cu"] in (2, 4, 6, 8, 12, 16)
    assert shape["split_windows"] in (0, 16, 32, 64, 128)
    assert same(snapshot(b.run(db)), ref)

    fbatch = workloads.fill_batch(64, seed=7)
    fdb = DeviceBatch(fbatch, workloads.fill_options(64, seed=7))
    fb = FillBatcher()
    fref = snapshot(fb.run(fdb, check=False))
    fshape = autotune(fb, fdb)
    print("fill shape", fshape, flush=True)
    assert fshape["waves_per_block"] in (1, 2, 4)
    assert same(snapshot(fb.run(fdb, check=False)), fref)


def test_autotune_band_size_keeps_the_region():
    """The banded fill's band size changes its triangles, never the region: after the
    autotuner picks one, every path covers the same area (relative 1e-6 of the exact
    area, or the f32 bound for small shapes far from the origin)."""
    import numpy as np
    from lyon_amd.batch import set_band_size
    from oracle import pyoracle as O
    fbatch = workloads.fill_batch(48, seed=11)
    fopts = workloads.fill_options(48, seed=11)
    fdb = DeviceBatch(fbatch, fopts)
    fb = FillBatcher(banded=True)
    try:
        shape = autotune(fb, fdb)
        print("banded shape", shape, flush=True)
        assert shape["edges_per_band"] in (16, 32, 64, 128, 256)
        res = fb.run(fdb, check=False)
        assert np.all(res.status == 0)
        for k in range(48):
            v, idx = res.path(k)
            idx = idx - int(res.vertex_offsets[k])
            t = np.asarray(idx, np.int64).reshape(-1, 3)
            vv = np.asarray(v, np.float64)
            a, b, c = vv[t[:, 0]], vv[t[:, 1]], vv[t[:, 2]]
            area = float(np.abs((b - a)[:, 0] * (c - a)[:, 1] - (b - a)[:, 1] * (c - a)[:, 0]).sum() / 2)
            p = fbatch.path(k)
            exact, _ = O.fill_area(p, 0.05, int(fopts[k]["fill_rule"]))
            pts = np.asarray(p.points, np.float64)
            perim = float(np.sum(np.hypot(*np.diff(pts, axis=0).T)))
            tol = max(1e-6 * exact, perim * 0.5 * float(np.spacing(np.float32(pts.max()))))
            assert abs(area - exact) <= tol, (k, area, exact, tol)
    finally:
        set_band_size(256, 16384)
