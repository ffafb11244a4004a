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


def fill_paths(res):
    """Per path: status, vertex bits and indices relative to the path's first vertex."""
    st = res.status.copy()
    vo, io = res.vertex_offsets, res.index_offsets
    v = res.all_vertices().view(np.uint32)
    i = res.all_indices()
    out = []
    for p in range(res.n):
        out.append((int(st[p]), v[vo[p]:vo[p + 1]].copy(),
                    (i[io[p]:io[p + 1]].astype(np.int64) - int(vo[p])).copy()))
    return out


def test_fill_graph_replays_match_eager():
    """The captured first sweep round of the batched fill gives the eager call's bytes
    (offsets, vertices, indices, statuses) on every replay, with eager work between."""
    from lyon_amd.batch import FillGraph
    n = 384
    db = DeviceBatch(workloads.fill_batch(n, seed=21), workloads.fill_options(n, seed=21))
    eager = FillBatcher()
    res = eager.run(db, vertex_offset=5, check=False)
    ref = snapshot(res)
    ref_st = res.status.copy()
    g = FillGraph(db, int(ref[2].shape[0]) + 64, int(ref[3].shape[0]) + 64, vertex_offset=5)
    other = DeviceBatch(workloads.fill_batch(64, seed=22), workloads.fill_options(64, seed=22))
    for k in range(3):
        g.launch()
        r = g.result(check=False)
        print(f"fill replay {k}: {r.n_vertices} vertices", flush=True)
        assert same(snapshot(r), ref), f"fill replay {k} differs from the eager batch"
        assert np.array_equal(r.status, ref_st)
        eager.run(other, check=False)  # eager work between replays
        synchronize()


def test_fill_graph_follows_data_changed_in_place():
    """A replay reads the batch's buffers as they are: with every point halved after the
    capture (the same verbs; flattening needs fewer edges, so the planned room holds), a
    replay equals the eager fill of the new points."""
    from lyon_amd.batch import FillGraph
    n = 192
    batch = workloads.fill_batch(n, seed=23)
    db = DeviceBatch(batch, workloads.fill_options(n, seed=23))
    eager = FillBatcher()
    first = eager.run(db, check=False)
    cap_v, cap_i = first.n_vertices + 64, first.n_indices + 64
    g = FillGraph(db, cap_v, cap_i)
    db.points.upload(np.ascontiguousarray(batch.points.reshape(-1) * np.float32(0.5)))
    synchronize()
    ref = eager.run(db, check=False)
    ref_snap, ref_st = snapshot(ref), ref.status.copy()
    g.launch()
    r = g.result(check=False)
    assert same(snapshot(r), ref_snap)
    assert np.array_equal(r.status, ref_st)


def test_fill_graph_reports_paths_that_outgrow_their_room():
    """With the LDS capacities cut to an eighth (many paths outgrow them; the eager fill
    sweeps those again in the global form) a replay, which has no second round, leaves
    them empty with LYG_ERR_NO_SPACE; every other path equals the eager call's path, and
    result() raises NO_SPACE when asked to check."""
    from lyon_amd.batch import FillGraph, TessellationError, set_fill_lds_resident
    n = 256
    db = DeviceBatch(workloads.fill_batch(n, seed=24), workloads.fill_options(n, seed=24))
    try:
        set_fill_lds_resident(8)
        eager = FillBatcher()
        res = eager.run(db, check=False)
        ref = fill_paths(res)
        g = FillGraph(db, res.n_vertices, res.n_indices)
        g.launch()
        r = g.result(check=False)
        got = fill_paths(r)
        grown = [p for p in range(n) if got[p][0] == 20]
        print(f"{len(grown)} of {n} paths outgrew their room", flush=True)
        assert grown, "no path outgrew the cut LDS room"
        for p in range(n):
            if p in grown:
                assert got[p][1].size == 0 and got[p][2].size == 0
            else:
                assert got[p][0] == ref[p][0]
                assert np.array_equal(got[p][1], ref[p][1]), f"path {p}: vertices differ"
                assert np.array_equal(got[p][2], ref[p][2]), f"path {p}: indices differ"
        g.launch()
        with pytest.raises(TessellationError):
            g.result()
    finally:
        set_fill_lds_resident(True)


def test_fill_graph_reports_overflow():
    from lyon_amd.batch import FillGraph
    db = DeviceBatch(workloads.fill_batch(40, seed=25), workloads.fill_options(40, seed=25))
    g = FillGraph(db, 64, 64)
    g.launch()
    r = g.result(check=False)
    assert r.n_vertices > 64
    with pytest.raises(Exception):
        g.launch()
        g.result()
