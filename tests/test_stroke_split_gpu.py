"""Intra-path parallelism of the batched stroke (lyg_ctx_set_stroke_split): long paths cut
into parts walked by different waves give the same bytes as the whole-path walk, also
where a part's prologue cannot rebuild the walk's state (points closer than the keep
threshold across a part boundary: the path is then walked again whole), with vertex
offsets, round joins and caps, closed and open paths."""
import numpy as np
import pytest

from lyon_amd import workloads
from lyon_amd.path import PathBatch, PathBuilder

pytestmark = pytest.mark.gpu


LAST = {}


def run(db, split, vertex_offset=0):
    from lyon_amd.batch import StrokeBatcher, set_stroke_split, stroke_split_stats
    set_stroke_split(split)
    try:
        r = StrokeBatcher().run(db, vertex_offset=vertex_offset)
        LAST.update(stroke_split_stats())
    finally:
        set_stroke_split(64)
    return (r.vertex_offsets.copy(), r.index_offsets.copy(),
            r.all_vertices().view(np.uint32).copy(), r.all_indices().copy())


def same(a, b, what):
    for x, y, name in zip(a, b, ("vertex offsets", "index offsets", "vertices", "indices")):
        assert x.shape == y.shape and np.array_equal(x, y), (what, name)


def long_paths(n, seed, closed):
    return workloads.stroke_batch(n, 256, 2048, 4096.0, seed=seed, closed=closed)


@pytest.mark.parametrize("closed", [False, True])
def test_split_matches_whole_walk(closed):
    from lyon_amd.batch import DeviceBatch
    batch = long_paths(24, 3 + closed, closed)
    opts = workloads.stroke_options(24, tolerance=0.05, seed=4)
    db = DeviceBatch(batch, opts)
    ref = run(db, 0, vertex_offset=5)
    for w in (1, 3, 16):
        print(f"split into parts of {w} windows, closed={closed}", flush=True)
        same(ref, run(db, w, vertex_offset=5), f"split {w}")
        assert LAST["items"] > len(batch), "no path was split"


def test_split_with_short_paths_mixed_in():
    """The 10,000-path batched stroke workload's first 2,000 paths (8-256 segments), most
    of them too short to split, with a few long ones."""
    from lyon_amd.batch import DeviceBatch
    short = workloads.stroke_batch(2000, 8, 256, 2048.0, seed=0)
    longs = long_paths(8, 11, False)
    batch = PathBatch.from_paths([short.path(i) for i in range(1000)] +
                                 [longs.path(i) for i in range(8)] +
                                 [short.path(i) for i in range(1000, 2000)])
    opts = workloads.stroke_options(len(batch), tolerance=0.05, seed=2)
    db = DeviceBatch(batch, opts)
    ref = run(db, 0)
    for w in (2, 8):
        print(f"mixed batch, parts of {w} windows", flush=True)
        same(ref, run(db, w), f"split {w}")


def near_duplicate_paths(n=6, steps=3000, seed=7):
    rng = np.random.default_rng(seed)
    paths = []
    for k in range(n):
        b = PathBuilder()
        x, y = 100.0, 100.0
        b.begin((x, y))
        for i in range(steps):
            if i % 340 < 300:
                # runs of 300 tiny steps, longer than a part's prologue: the keep filter
                # drops them against the run's first point, which a prologue starting
                # inside the run cannot know
                x += float(rng.uniform(-0.004, 0.004))
                y += float(rng.uniform(-0.004, 0.004))
            else:
                x += float(rng.uniform(-30, 30))
                y += float(rng.uniform(-30, 30))
            b.line_to((x, y))
        b.end(k % 2 == 1)
        paths.append(b.build())
    return paths


def test_split_where_prologues_cannot_rebuild_the_state():
    """Runs of points closer than the keep threshold (tolerance^2 / 2) across many part
    boundaries: a part's prologue then starts inside a run of dropped points, so some
    checks fail and those paths are walked again whole; the bytes stay the same."""
    from lyon_amd.batch import DeviceBatch
    batch = PathBatch.from_paths(near_duplicate_paths())
    opts = workloads.stroke_options(len(batch), tolerance=0.1, seed=1)
    db = DeviceBatch(batch, opts)
    ref = run(db, 0)
    again = 0
    for w in (1, 2, 5):
        print(f"near-duplicate points, parts of {w} windows", flush=True)
        same(ref, run(db, w), f"split {w}")
        print(f"  {LAST}", flush=True)
        again += LAST["walked_again"]
    assert again > 0, "no path was walked again: the check was not exercised"


def test_graph_replays_of_the_split_walk():
    """The split walk captured in a HIP graph (lyg_stroke_batch_graph with splitting on),
    replayed three times with eager work and synchronisations between: every replay gives
    the eager whole-path bytes, on long paths and on paths whose prologues fail the check
    (walked again by the graph's second-walk kernel, whose path count is on the device)."""
    from lyon_amd.batch import DeviceBatch, StrokeBatcher, StrokeGraph, set_stroke_split
    longs = long_paths(12, 21, False)
    batch = PathBatch.from_paths([longs.path(i) for i in range(12)] + near_duplicate_paths(4))
    opts = workloads.stroke_options(len(batch), tolerance=0.1, seed=3)
    db = DeviceBatch(batch, opts)
    ref = run(db, 0, vertex_offset=11)
    set_stroke_split(2)
    try:
        g = StrokeGraph(db, int(ref[0][-1]) + 64, int(ref[1][-1]) + 192, vertex_offset=11)
        for rep in range(3):
            g.launch()
            r = g.result()
            got = (r.vertex_offsets.copy(), r.index_offsets.copy(),
                   r.all_vertices().view(np.uint32).copy(), r.all_indices().copy())
            same(ref, got, f"graph replay {rep}")
            # eager work between replays
            StrokeBatcher().run(db)
    finally:
        set_stroke_split(64)
