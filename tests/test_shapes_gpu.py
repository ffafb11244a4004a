"""lyon's basic shapes (fill.rs:644-690, stroke.rs:255-307, basic_shapes.rs,
builder.rs): direct fill constructions through the C ABI, and the path-built shapes
through the GPU tessellators."""
import math

import numpy as np
import pytest

from helpers import assert_same_geometry, bbox_diag
from lyon_amd import (BuffersBuilder, FillOptions, FillTessellator, StrokeOptions,
                      StrokeTessellator, VertexBuffers)
from lyon_amd.path import PathBuilder
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu


def fill_shape(method, *args, offset=0):
    vb = VertexBuffers()
    getattr(FillTessellator(), method)(*args, BuffersBuilder(vb).with_vertex_offset(offset))
    return vb.vertices, vb.indices


def test_rectangle_kat():
    # basic_shapes.rs:209-222
    v, i = fill_shape("tessellate_rectangle", (0.0, 1.0, 2.0, 4.0), FillOptions(), offset=5)
    assert v.tolist() == [[0.0, 1.0], [0.0, 4.0], [2.0, 4.0], [2.0, 1.0]]
    assert i.tolist() == [5, 6, 7, 5, 7, 8]


@pytest.mark.parametrize("radius,tol", [(100.0, 0.1), (3.0, 0.01), (-40.0, 0.5), (1e-3, 0.1)])
def test_circle_matches_restatement(radius, tol):
    v, i = fill_shape("tessellate_circle", (1.0, 2.0), radius, FillOptions(tolerance=tol))
    rv, ri = O.fill_circle((1.0, 2.0), radius, tol)
    assert np.array_equal(i, ri), "topology differs"
    r = abs(radius)
    assert np.max(np.abs(v - rv)) <= 1e-5 * 2 * r + 1e-6
    d = np.hypot(v[:, 0] - 1.0, v[:, 1] - 2.0)
    assert np.all(np.abs(d - r) <= 1e-5 * r + 1e-6)
    # the triangles tile the inscribed polygon: total area within the chord error of pi r^2
    t = v[i.reshape(-1, 3)].astype(np.float64)
    e1, e2 = t[:, 1] - t[:, 0], t[:, 2] - t[:, 0]
    area = 0.5 * np.abs(e1[:, 0] * e2[:, 1] - e1[:, 1] * e2[:, 0]).sum()
    # topology from the geometry alone: after the two base triangles (0,3,1), (1,3,2), the
    # k-th triangle adds vertex k+2 on the bisector of its two other corners
    tri = i.reshape(-1, 3).astype(np.int64)
    assert tri[:2].tolist() == [[0, 3, 1], [1, 3, 2]]
    assert np.array_equal(tri[2:, 1], np.arange(4, len(v)))
    ang = np.arctan2(v[:, 1].astype(np.float64) - 2.0, v[:, 0].astype(np.float64) - 1.0)
    a, b, m = ang[tri[2:, 2]], ang[tri[2:, 0]], ang[tri[2:, 1]]
    bis = a + 0.5 * np.angle(np.exp(1j * (b - a)))
    assert np.all(np.abs(np.angle(np.exp(1j * (m - bis)))) <= 1e-4)
    n = len(v)
    assert abs(area - 0.5 * n * r * r * math.sin(2 * math.pi / n)) <= 1e-4 * r * r


def test_circle_zero_radius_is_empty():
    v, i = fill_shape("tessellate_circle", (1.0, 2.0), 0.0, FillOptions())
    assert len(v) == 0 and len(i) == 0


def test_stroke_circle_path_and_result():
    # builder.rs:1508-1548: four cubics from (-r, 0) with d = r * 0.55191505
    b = PathBuilder()
    b.add_circle((10.0, 20.0), 5.0, True)
    p = b.build()
    f = np.float32
    d = f(5.0) * f(0.55191505)
    assert p.verbs.tolist() == [3, 2, 2, 2, 2, 4]
    assert p.points[0].tolist() == [5.0, 20.0]
    assert p.points[1].tolist() == [float(f(10.0) + f(-5.0)), float(f(20.0) + (-d))]
    vb = VertexBuffers()
    so = StrokeOptions(line_width=2.0, tolerance=0.05)
    StrokeTessellator().tessellate_circle((10.0, 20.0), 5.0, so, BuffersBuilder(vb))
    st, rv, ri = O.stroke(p, 2.0, so.miter_limit, 0.05, int(so.start_cap), int(so.end_cap),
                          int(so.line_join))
    assert st == 0
    assert_same_geometry((vb.vertices, vb.indices), (rv, ri), bbox_diag(p.points), what="circle")


def test_fill_polygon_and_ellipse_coverage():
    pts = [(0.0, 0.0), (50.0, 5.0), (20.0, 40.0), (60.0, 60.0), (-10.0, 30.0)]
    v, i = fill_shape("tessellate_polygon", pts, True, FillOptions(tolerance=0.05))
    b = PathBuilder()
    b.add_polygon(pts, True)
    path = b.build()
    st, rv, ri = O.fill(path, 0.05, 0)
    c = O.coverage_compare(rv, ri, v, i, n=1024)
    assert c["mismatch"] == 0
    v, i = fill_shape("tessellate_ellipse", (100.0, 50.0), (80.0, 30.0), 0.4, True,
                      FillOptions(tolerance=0.05))
    b = PathBuilder()
    b.add_ellipse((100.0, 50.0), (80.0, 30.0), 0.4, True)
    ep = b.build()
    w = O.winding_compare(ep, 0.05, 0, v, i, n=1024)
    assert w["mismatch"] == 0


def test_tessellate_over_events_matches_tessellate_path():
    """FillTessellator / StrokeTessellator::tessellate (fill.rs:578-598, stroke.rs:122-138)
    over a path's events give what tessellate_path gives."""
    from lyon_amd.workloads import stroke_batch
    sb = stroke_batch(6, 4, 40, 512.0, seed=3, closed=True)
    for i in range(len(sb)):
        p = sb.path(i)
        for tess, opts in ((FillTessellator(), FillOptions(tolerance=0.05)),
                           (StrokeTessellator(), StrokeOptions(line_width=3.0, tolerance=0.05))):
            a, b = VertexBuffers(), VertexBuffers()
            tess.tessellate_path(p, opts, BuffersBuilder(a))
            tess.tessellate(p.iter(), opts, BuffersBuilder(b))
            assert np.array_equal(a.vertices, b.vertices) and np.array_equal(a.indices, b.indices)
