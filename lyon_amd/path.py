"""lyon_path's flat path storage, mirrored in numpy.

A ``Path`` is the pair ``(verbs: u8[], points: f32[n, 2])`` exactly as lyon stores it
(crates/path/src/path.rs:29-36 for the verb order, :75-79 for the struct, :516-611
for what each builder call pushes).  ``PathBuilder`` reproduces ``BuilderImpl``:
``begin`` pushes the point, ``line_to`` one point, ``quadratic_bezier_to`` two,
``cubic_bezier_to`` three, and ``end(close=True)`` pushes a copy of the first point
with a ``CLOSE`` verb (path.rs:537-542).

``PathBatch`` packs many paths into one verb array and one point array with offset
tables; it is the wire format of the batched and served forms.
"""
from __future__ import annotations

import numpy as np

VERB_LINE = 0
VERB_QUADRATIC = 1
VERB_CUBIC = 2
VERB_BEGIN = 3
VERB_CLOSE = 4
VERB_END = 5


class Path:
    """An immutable lyon path: ``verbs`` (uint8) and ``points`` (float32, shape (n, 2))."""

    __slots__ = ("verbs", "points")

    def __init__(self, verbs, points):
        self.verbs = np.ascontiguousarray(verbs, dtype=np.uint8)
        self.points = np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 2)

    @staticmethod
    def builder() -> "PathBuilder":
        return PathBuilder()

    def __len__(self):
        return int(self.verbs.shape[0])

    def num_segments(self) -> int:
        return int(np.count_nonzero(self.verbs <= VERB_CUBIC))


class PathBuilder:
    """Mirror of lyon_path's ``BuilderImpl`` (path.rs:472-611)."""

    def __init__(self):
        self._verbs: list[int] = []
        self._points: list[tuple[float, float]] = []
        self._first = (0.0, 0.0)
        self._in_sub_path = False

    def begin(self, at):
        assert not self._in_sub_path, "begin() inside a sub-path"
        self._in_sub_path = True
        self._first = (float(at[0]), float(at[1]))
        self._points.append(self._first)
        self._verbs.append(VERB_BEGIN)
        return len(self._points) - 1

    def line_to(self, to):
        assert self._in_sub_path, "line_to() outside a sub-path"
        self._points.append((float(to[0]), float(to[1])))
        self._verbs.append(VERB_LINE)
        return len(self._points) - 1

    def quadratic_bezier_to(self, ctrl, to):
        assert self._in_sub_path
        self._points.append((float(ctrl[0]), float(ctrl[1])))
        self._points.append((float(to[0]), float(to[1])))
        self._verbs.append(VERB_QUADRATIC)
        return len(self._points) - 1

    def cubic_bezier_to(self, ctrl1, ctrl2, to):
        assert self._in_sub_path
        self._points.append((float(ctrl1[0]), float(ctrl1[1])))
        self._points.append((float(ctrl2[0]), float(ctrl2[1])))
        self._points.append((float(to[0]), float(to[1])))
        self._verbs.append(VERB_CUBIC)
        return len(self._points) - 1

    def end(self, close: bool = False):
        assert self._in_sub_path, "end() outside a sub-path"
        self._in_sub_path = False
        if close:
            self._points.append(self._first)
        self._verbs.append(VERB_CLOSE if close else VERB_END)

    def close(self):
        self.end(True)

    # ---- shapes (crates/path/src/builder.rs) ------------------------------------------
    def add_polygon(self, points, closed: bool = True):
        """BuilderImpl::add_polygon (builder.rs:560-573)."""
        if len(points) == 0:
            return
        self.begin(points[0])
        for p in points[1:]:
            self.line_to(p)
        self.end(closed)

    def add_rectangle(self, rect, positive: bool = True):
        """add_rectangle (builder.rs:649-680); rect = (min x, min y, max x, max y)."""
        x0, y0, x1, y1 = (np.float32(v) for v in rect)
        if positive:
            pts = [(x0, y0), (x1, y0), (x1, y1), (x0, y1)]
        else:
            pts = [(x0, y0), (x0, y1), (x1, y1), (x1, y0)]
        self.add_polygon(pts, True)

    def add_circle(self, center, radius, positive: bool = True):
        """add_circle (builder.rs:1508-1548): four cubics with the 0.55191505 factor, in f32."""
        f = np.float32
        cx, cy = f(center[0]), f(center[1])
        r = f(abs(f(radius)))
        d_ = f(1.0) if positive else f(-1.0)
        d = r * f(0.55191505)

        def pt(dx, dy):
            return (cx + f(dx), cy + f(dy))

        self.begin(pt(-r, f(0.0)))
        self.cubic_bezier_to(pt(-r, -d * d_), pt(-d, -r * d_), pt(f(0.0), -r * d_))
        self.cubic_bezier_to(pt(d, -r * d_), pt(r, -d * d_), pt(r, f(0.0)))
        self.cubic_bezier_to(pt(r, d * d_), pt(d, r * d_), pt(f(0.0), r * d_))
        self.cubic_bezier_to(pt(-d, r * d_), pt(-r, d * d_), pt(-r, f(0.0)))
        self.close()

    def add_ellipse(self, center, radii, x_rotation: float, positive: bool = True):
        """add_ellipse (builder.rs:606-636): the arc's quadratic approximation
        (Arc::for_each_quadratic_bezier, evaluated by the GPU arc kernel); the sub-path begins
        at the first quadratic's start, which samples the same angle 0 as arc.sample(0)."""
        from .tessellation import arc_to_quadratics
        sweep = np.float32(2.0) * np.float32(np.pi) * (np.float32(1.0) if positive else np.float32(-1.0))
        arc = np.array([[center[0], center[1], radii[0], radii[1], 0.0, sweep, x_rotation]],
                       np.float32)
        q, _ = arc_to_quadratics(arc)
        self.begin(q[0, 0])
        for k in range(len(q)):
            self.quadratic_bezier_to(q[k, 1], q[k, 2])
        self.end(True)

    def build(self) -> Path:
        assert not self._in_sub_path, "build() with an open sub-path"
        pts = np.asarray(self._points, dtype=np.float32).reshape(-1, 2)
        return Path(np.asarray(self._verbs, dtype=np.uint8), pts)


def polygon(points, closed: bool = True) -> Path:
    """A path made of one sub-path through ``points`` (lyon's ``Polygon`` events)."""
    b = PathBuilder()
    b.begin(points[0])
    for p in points[1:]:
        b.line_to(p)
    b.end(closed)
    return b.build()


class PathBatch:
    """Many paths packed end to end: the batched form's input layout.

    ``verbs``/``points`` hold every path back to back; ``verb_offsets`` and
    ``point_offsets`` have ``n + 1`` entries so that path ``i`` is
    ``verbs[verb_offsets[i]:verb_offsets[i+1]]``.
    """

    __slots__ = ("verbs", "points", "verb_offsets", "point_offsets")

    def __init__(self, verbs, points, verb_offsets, point_offsets):
        self.verbs = np.ascontiguousarray(verbs, dtype=np.uint8)
        self.points = np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 2)
        self.verb_offsets = np.ascontiguousarray(verb_offsets, dtype=np.uint32)
        self.point_offsets = np.ascontiguousarray(point_offsets, dtype=np.uint32)

    @staticmethod
    def from_paths(paths) -> "PathBatch":
        vo = [0]
        po = [0]
        for p in paths:
            vo.append(vo[-1] + len(p.verbs))
            po.append(po[-1] + len(p.points))
        verbs = np.concatenate([p.verbs for p in paths]) if paths else np.zeros(0, np.uint8)
        points = (np.concatenate([p.points for p in paths]) if paths
                  else np.zeros((0, 2), np.float32))
        return PathBatch(verbs, points, vo, po)

    def __len__(self):
        return int(self.verb_offsets.shape[0] - 1)

    def path(self, i: int) -> Path:
        a, b = int(self.verb_offsets[i]), int(self.verb_offsets[i + 1])
        c, d = int(self.point_offsets[i]), int(self.point_offsets[i + 1])
        return Path(self.verbs[a:b], self.points[c:d])
