"""Host-side path logic: PathEvent iteration (path.rs:817-909) and paths rebuilt from
events, as the tessellators' `tessellate(events)` entry points consume them."""
import numpy as np

from lyon_amd.path import Path, PathBuilder, PathEvent


def _sample():
    b = PathBuilder()
    b.begin((0.0, 0.0))
    b.line_to((1.0, 0.0))
    b.quadratic_bezier_to((2.0, 0.0), (2.0, 1.0))
    b.end(True)
    b.begin((5.0, 5.0))
    b.cubic_bezier_to((6.0, 5.0), (6.0, 6.0), (5.0, 6.0))
    b.end(False)
    return b.build()


def test_events_follow_path_iter():
    ev = list(_sample().iter())
    assert ev == [
        PathEvent.begin((0.0, 0.0)),
        PathEvent.line((0.0, 0.0), (1.0, 0.0)),
        PathEvent.quadratic((1.0, 0.0), (2.0, 0.0), (2.0, 1.0)),
        PathEvent.end((2.0, 1.0), (0.0, 0.0), True),
        PathEvent.begin((5.0, 5.0)),
        PathEvent.cubic((5.0, 5.0), (6.0, 5.0), (6.0, 6.0), (5.0, 6.0)),
        PathEvent.end((5.0, 6.0), (5.0, 5.0), False),
    ]


def test_from_events_round_trip():
    from lyon_amd.workloads import stroke_batch
    sb = stroke_batch(20, 4, 40, 512.0, seed=1, closed=True)
    paths = [sb.path(i) for i in range(len(sb))] + [_sample(), PathBuilder().build()]
    for p in paths:
        q = Path.from_events(p.iter())
        assert np.array_equal(q.verbs, p.verbs)
        assert np.array_equal(q.points.view(np.uint32), p.points.view(np.uint32))


def test_path_buffer():
    """PathBuffer (path_buffer.rs:14-104): paths appended through its builder get
    consecutive indices, read back unchanged, and pack into the batched layout."""
    from lyon_amd.path import PathBuffer
    buf = PathBuffer()
    assert buf.is_empty() and len(buf.as_batch()) == 0
    b = buf.builder()
    b.add_polygon([(0.0, 0.0), (1.0, 0.0), (1.0, 1.0)], True)
    assert b.build() == 0
    assert buf.push(_sample()) == 1
    assert list(buf.indices()) == [0, 1]
    batch = buf.as_batch()
    assert len(batch) == 2
    for i, p in enumerate(buf.iter()):
        q = batch.path(i)
        assert np.array_equal(q.verbs, p.verbs) and np.array_equal(q.points, p.points)
    assert buf.get(1).verbs.tolist() == _sample().verbs.tolist()
    again = PathBuffer.from_batch(batch)
    assert len(again) == 2 and np.array_equal(again.get(0).points, buf.get(0).points)
    buf.clear()
    assert len(buf) == 0
