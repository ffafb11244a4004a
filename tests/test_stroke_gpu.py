"""GPU parity of the stroke tessellator (plain and batched) against the oracle.

Bar: index topology bit-exact, positions within 1e-5 of the path's bbox diagonal.
"""
import itertools

import numpy as np
import pytest

from helpers import assert_same_geometry, bbox_diag, empty_caps_path, square_paths
from lyon_amd import (BuffersBuilder, LineCap, LineJoin, PathBuilder, StrokeOptions,
                      StrokeTessellator, VertexBuffers)
from lyon_amd import workloads
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu


def gpu_stroke(path, opts, vertex_offset=0):
    buf = VertexBuffers()
    StrokeTessellator().tessellate_path(path, opts,
                                        BuffersBuilder(buf).with_vertex_offset(vertex_offset))
    return buf.vertices, buf.indices


def ref_stroke(path, opts, vertex_offset=0):
    st, v, i = O.stroke(path, opts.line_width, opts.miter_limit, opts.tolerance,
                        int(opts.start_cap), int(opts.end_cap), int(opts.line_join),
                        vertex_offset)
    assert st == 0
    return v, i


def check(path, opts, what, vertex_offset=0):
    g = gpu_stroke(path, opts, vertex_offset)
    r = ref_stroke(path, opts, vertex_offset)
    print(f"{what}: {len(r[0])} vertices, {len(r[1]) // 3} triangles", flush=True)
    # Scale of the 1e-5 position bound: the path's bbox diagonal, or the line width for a
    # path of zero extent (its stroke is a cap of that size).
    return assert_same_geometry(g, r, max(bbox_diag(path.points), opts.line_width), what=what)


ALL_OPTS = [StrokeOptions(line_width=w, line_join=j, start_cap=c, end_cap=c2, tolerance=t,
                          miter_limit=m)
            for (w, j, c, c2, t, m) in [
                (1.0, LineJoin.Miter, LineCap.Butt, LineCap.Butt, 0.1, 4.0),
                (2.0, LineJoin.Bevel, LineCap.Square, LineCap.Square, 0.1, 4.0),
                (0.3, LineJoin.MiterClip, LineCap.Round, LineCap.Round, 0.1, 1.0),
                (1.0, LineJoin.Round, LineCap.Round, LineCap.Round, 0.001, 4.0),
                (4.0, LineJoin.Miter, LineCap.Butt, LineCap.Square, 0.05, 4.0),
                (8.0, LineJoin.Round, LineCap.Square, LineCap.Round, 0.05, 2.0),
            ]]


def test_square_kats():
    """stroke.rs:2852-2910: Miter+Butt -> 12 triangles, Bevel+Square -> 16."""
    p = square_paths()
    v, i = gpu_stroke(p, StrokeOptions(line_join=LineJoin.Miter, start_cap=LineCap.Butt,
                                       end_cap=LineCap.Butt))
    assert len(i) // 3 == 12
    v, i = gpu_stroke(p, StrokeOptions(line_join=LineJoin.Bevel, start_cap=LineCap.Square,
                                       end_cap=LineCap.Square))
    assert len(i) // 3 == 16
    for k, o in enumerate(ALL_OPTS):
        check(p, o, f"square opts#{k}")


def test_empty_caps_kats():
    """stroke.rs:2921-2974: Butt -> 0 triangles, Square -> 8."""
    p = empty_caps_path()
    assert len(gpu_stroke(p, StrokeOptions().with_line_cap(LineCap.Butt))[1]) == 0
    assert len(gpu_stroke(p, StrokeOptions().with_line_cap(LineCap.Square))[1]) // 3 == 8
    for cap in LineCap:
        check(p, StrokeOptions().with_line_cap(cap), f"empty caps {cap.name}")


def test_line_width_kat():
    """stroke.rs:3126-3154: vertices of a width-2 segment are the square's corners."""
    b = PathBuilder()
    b.begin((0, 1)); b.line_to((2, 1)); b.end(False)
    v, _ = gpu_stroke(b.build(), StrokeOptions(line_width=2.0))
    got = {tuple(np.round(x, 5)) for x in v.tolist()}
    assert got <= {(0.0, 0.0), (0.0, 2.0), (2.0, 0.0), (2.0, 2.0)}


def test_miter_clip_length_kat():
    """stroke.rs:3380-3416: max y == 300 + 30 * 3 * 0.5 exactly."""
    b = PathBuilder()
    b.begin((-70, -200)); b.line_to((0, 300)); b.line_to((70, -200)); b.end(False)
    v, _ = gpu_stroke(b.build(), StrokeOptions(line_width=30.0, line_join=LineJoin.MiterClip,
                                               miter_limit=3.0))
    assert float(v[:, 1].max()) == 345.0


def test_issue_819_close_merge():
    """stroke.rs:3289-3318: last point within merge range of the first two."""
    b = PathBuilder()
    b.begin((650.539978027344, 173.559997558594))
    b.line_to((650.609985351562, 173.529998779297))
    b.line_to((650.729980468750, 173.630004882812))
    b.line_to((650.559997558594, 173.570007324219))
    b.end(True)
    check(b.build(), StrokeOptions(tolerance=0.1), "issue_819")


def test_plain_workload_path():
    """The plain workload: one closed 64-segment path, width 4, Miter limit 4, Butt."""
    p = workloads.plain_path(0)
    frac = check(p, StrokeOptions(line_width=4.0, miter_limit=4.0, line_join=LineJoin.Miter,
                                  start_cap=LineCap.Butt, end_cap=LineCap.Butt), "plain")
    print("bit-exact vertex fraction", frac)


def test_random_paths_all_options():
    batch = workloads.stroke_batch(24, 1, 40, 600.0, seed=7)
    closed = workloads.stroke_batch(24, 1, 40, 600.0, seed=8, closed=True)
    for bi, b in enumerate((batch, closed)):
        for i in range(len(b)):
            o = ALL_OPTS[(i + bi) % len(ALL_OPTS)]
            check(b.path(i), o, f"random[{bi}][{i}]", vertex_offset=i * 13)


def test_degenerate_inputs():
    cases = []
    b = PathBuilder(); b.begin((5, 5)); b.line_to((5, 5)); b.line_to((5.01, 5)); b.end(False)
    cases.append(b.build())
    b = PathBuilder(); b.begin((0, 0)); b.cubic_bezier_to((0, 0), (0, 0), (0, 0)); b.end(True)
    cases.append(b.build())
    b = PathBuilder(); b.begin((0, 0)); b.quadratic_bezier_to((100, 0), (0, 0.001)); b.end(False)
    cases.append(b.build())
    b = PathBuilder(); b.begin((0, 0)); b.line_to((10, 0)); b.line_to((0, 0.0001)); b.end(False)
    cases.append(b.build())
    b = PathBuilder(); b.begin((0, 0)); b.line_to((10, 0)); b.line_to(( 0, 0)); b.end(True)
    cases.append(b.build())
    # cusp cubic and a sharp quadratic
    b = PathBuilder(); b.begin((0, 0)); b.cubic_bezier_to((100, 100), (0, 100), (100, 0))
    b.quadratic_bezier_to((-500, 3), (10, 1)); b.end(False)
    cases.append(b.build())
    for k, p in enumerate(cases):
        for j, o in enumerate(ALL_OPTS):
            check(p, o, f"degenerate[{k}] opts#{j}")


def test_batch_matches_sequential_plain_calls():
    """The batched form equals tessellate_path per path into one shared VertexBuffers."""
    from lyon_amd.batch import DeviceBatch, StrokeBatcher
    batch = workloads.stroke_batch(300, 8, 256, 2048.0, seed=0)
    opts = workloads.stroke_options(300, tolerance=0.05)
    res = StrokeBatcher().run(DeviceBatch(batch, opts))
    vo = res.vertex_offsets
    for i in range(len(batch)):
        o = opts[i]
        so = StrokeOptions(line_width=float(o["line_width"]), miter_limit=float(o["miter_limit"]),
                           tolerance=float(o["tolerance"]), start_cap=int(o["start_cap"]),
                           end_cap=int(o["end_cap"]), line_join=int(o["line_join"]))
        p = batch.path(i)
        r = ref_stroke(p, so, vertex_offset=int(vo[i]))
        g = res.path(i)
        assert_same_geometry(g, r, max(bbox_diag(p.points), so.line_width),
                             what=f"batch path {i}")


def test_batch_sizing_modes_agree():
    """The batched stroke's three sizing modes give the same bytes: one emit pass into
    per-path staging (default), the same with staging room too small for most paths (they
    are emitted again at their final place), and separate count and emit passes."""
    from lyon_amd.batch import DeviceBatch, StrokeBatcher, set_stroke_staging
    batch = workloads.stroke_batch(400, 4, 128, 1024.0, seed=5)
    opts = workloads.stroke_options(400, tolerance=0.05, seed=5)
    db = DeviceBatch(batch, opts)

    def run(enabled, room):
        set_stroke_staging(enabled, room)
        r = StrokeBatcher().run(db, vertex_offset=7)
        return (r.vertex_offsets, r.index_offsets, r.all_vertices().view(np.uint32),
                r.all_indices())

    try:
        staged = run(True, 1.0)
        squeezed = run(True, 0.05)
        two_pass = run(False, 1.0)
    finally:
        set_stroke_staging(True, 1.0)
    for name, other in (("room 0.05", squeezed), ("count + emit", two_pass)):
        for a, b, what in zip(staged, other, ("vertex offsets", "index offsets", "vertices",
                                               "indices")):
            assert np.array_equal(a, b), f"{name}: {what} differ from the staged run"
    # and the squeezed run (every path emitted twice) against the oracle on a sample
    vo = squeezed[0]
    for i in range(0, len(batch), 37):
        o = opts[i]
        so = StrokeOptions(line_width=float(o["line_width"]), miter_limit=float(o["miter_limit"]),
                           tolerance=float(o["tolerance"]), start_cap=int(o["start_cap"]),
                           end_cap=int(o["end_cap"]), line_join=int(o["line_join"]))
        p = batch.path(i)
        r = ref_stroke(p, so, vertex_offset=7 + int(vo[i]))
        v = squeezed[2].view(np.float32).reshape(-1, 2)[vo[i]:vo[i + 1]]
        io = squeezed[1]
        idx = squeezed[3][io[i]:io[i + 1]]
        assert_same_geometry((v, idx), r, max(bbox_diag(p.points), so.line_width),
                             what=f"squeezed batch path {i}")


def test_large_batch_offsets():
    """A batch of more than 32768 paths (its scans run across the whole GPU): offsets are the
    running sums of the per-path outputs and sampled paths equal the oracle's."""
    from lyon_amd.batch import DeviceBatch, StrokeBatcher
    n = 40000
    batch = workloads.stroke_batch(n, 1, 6, 512.0, seed=9)
    opts = workloads.stroke_options(n, tolerance=0.1, seed=9)
    res = StrokeBatcher().run(DeviceBatch(batch, opts))
    vo = res.vertex_offsets
    io = res.index_offsets
    assert vo[0] == 0 and io[0] == 0
    assert vo[-1] == res.n_vertices and io[-1] == res.n_indices
    assert np.all(np.diff(vo.astype(np.int64)) >= 0) and np.all(np.diff(io.astype(np.int64)) % 3 == 0)
    for i in list(range(0, n, 1733)) + [n - 1]:
        o = opts[i]
        so = StrokeOptions(line_width=float(o["line_width"]), miter_limit=float(o["miter_limit"]),
                           tolerance=float(o["tolerance"]), start_cap=int(o["start_cap"]),
                           end_cap=int(o["end_cap"]), line_join=int(o["line_join"]))
        p = batch.path(i)
        r = ref_stroke(p, so, vertex_offset=int(vo[i]))
        assert len(r[0]) == vo[i + 1] - vo[i] and len(r[1]) == io[i + 1] - io[i], f"path {i}"
        assert_same_geometry(res.path(i), r, max(bbox_diag(p.points), so.line_width),
                             what=f"large batch path {i}")


def test_threads_use_their_own_contexts():
    """Python threads calling the tessellators at once each get their own lyg context
    (a context is not thread-safe): results equal the single-threaded ones."""
    import threading
    from lyon_amd import FillOptions, FillTessellator
    batch = workloads.stroke_batch(48, 8, 96, 1024.0, seed=21)
    opts = StrokeOptions(line_width=3.0, line_join=LineJoin.Round, start_cap=LineCap.Round,
                         end_cap=LineCap.Square, tolerance=0.05)
    fopts = FillOptions(tolerance=0.05)

    def run(indices):
        out = {}
        for i in indices:
            p = batch.path(i)
            s = gpu_stroke(p, opts)
            fb = VertexBuffers()
            FillTessellator().tessellate_path(p, fopts, BuffersBuilder(fb))
            out[i] = (s[0].copy(), s[1].copy(), fb.vertices.copy(), fb.indices.copy())
        return out

    alone = run(range(len(batch)))
    results = [None, None]
    errors = []

    def worker(k):
        try:
            results[k] = run(range(k, len(batch), 2))
        except Exception as e:  # surfaced below
            errors.append(e)

    threads = [threading.Thread(target=worker, args=(k,)) for k in (0, 1)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert not errors, errors
    for k in (0, 1):
        for i, got in results[k].items():
            for a, b in zip(got, alone[i]):
                assert np.array_equal(a, b), f"path {i} differs when run in thread {k}"
