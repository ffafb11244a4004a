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
    assert same(snapshot(b.run(db)), ref)

    fbatch = workloads.fill_batch(64, seed=7)
    fdb = DeviceBatch(fbatch, workloads.fill_options(64, seed=7))
    fb = FillBatcher()
    fref = snapshot(fb.run(fdb, check=False))
    fshape = autotune(fb, fdb)
    print("fill shape", fshape, flush=True)
    assert fshape["waves_per_block"] in (1, 2, 4)
    assert same(snapshot(fb.run(fdb, check=False)), fref)
