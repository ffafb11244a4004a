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


class PathEvent:
    """lyon_path's PathEvent (crates/path/src/events.rs): ``kind`` is one of "begin",
    "line", "quadratic", "cubic", "end"; Begin carries ``at``; Line / Quadratic / Cubic
    carry ``from_``, ``ctrl1`` / ``ctrl2`` where they have them, and ``to``; End carries
    ``last``, ``first`` and ``close``."""

    __slots__ = ("kind", "at", "from_", "ctrl1", "ctrl2", "to", "last", "first", "close")

    def __init__(self, kind, at=None, from_=None, ctrl1=None, ctrl2=None, to=None, last=None,
                 first=None, close=False):
        self.kind, self.at, self.from_, self.ctrl1, self.ctrl2 = kind, at, from_, ctrl1, ctrl2
        self.to, self.last, self.first, self.close = to, last, first, close

    @staticmethod
    def begin(at):
        return PathEvent("begin", at=at)

    @staticmethod
    def line(from_, to):
        return PathEvent("line", from_=from_, to=to)

    @staticmethod
    def quadratic(from_, ctrl, to):
        return PathEvent("quadratic", from_=from_, ctrl1=ctrl, to=to)

    @staticmethod
    def cubic(from_, ctrl1, ctrl2, to):
        return PathEvent("cubic", from_=from_, ctrl1=ctrl1, ctrl2=ctrl2, to=to)

    @staticmethod
    def end(last, first, close):
        return PathEvent("end", last=last, first=first, close=bool(close))

    def _key(self):
        t = lambda p: None if p is None else (float(p[0]), float(p[1]))
        return (self.kind, t(self.at), t(self.from_), t(self.ctrl1), t(self.ctrl2), t(self.to),
                t(self.last), t(self.first), self.close)

    def __eq__(self, other):
        return isinstance(other, PathEvent) and self._key() == other._key()

    def __repr__(self):
        return "PathEvent" + repr(self._key())


class Path:
    """An immutable lyon path: ``verbs`` (uint8) and ``points`` (float32, shape (n, 2))."""

    __slots__ = ("verbs", "points", "attributes")

    def __init__(self, verbs, points, attributes=None):
        self.verbs = np.ascontiguousarray(verbs, dtype=np.uint8)
        self.points = np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 2)
        # per point, the custom attributes of the endpoint stored there (lyon's
        # builder_with_attributes; control points' rows are NaN): float32 (n, k) or None
        self.attributes = None
        if attributes is not None:
            a = np.ascontiguousarray(attributes, dtype=np.float32)
            self.attributes = a.reshape(self.points.shape[0], -1)

    @staticmethod
    def builder() -> "PathBuilder":
        return PathBuilder()

    def __len__(self):
        return int(self.verbs.shape[0])

    def num_segments(self) -> int:
        return int(np.count_nonzero(self.verbs <= VERB_CUBIC))

    def iter(self):
        """Path::iter (path.rs:817-909): the path's events.  Close reads the stored copy of
        the first point; End moves the current point back to the first."""
        pts = [tuple(p) for p in self.points.tolist()]
        k = 0

        def nxt():
            nonlocal k
            p = pts[k] if k < len(pts) else (float("nan"), float("nan"))
            k += 1
            return p

        current = first = (0.0, 0.0)
        for v in self.verbs.tolist():
            if v == VERB_BEGIN:
                current = first = nxt()
                yield PathEvent.begin(current)
            elif v == VERB_LINE:
                f, current = current, nxt()
                yield PathEvent.line(f, current)
            elif v == VERB_QUADRATIC:
                f, c = current, nxt()
                current = nxt()
                yield PathEvent.quadratic(f, c, current)
            elif v == VERB_CUBIC:
                f, c1, c2 = current, nxt(), nxt()
                current = nxt()
                yield PathEvent.cubic(f, c1, c2, current)
            elif v == VERB_CLOSE:
                nxt()
                yield PathEvent.end(current, first, True)
            else:
                last, current = current, first
                yield PathEvent.end(last, first, False)

    @staticmethod
    def from_events(events) -> "Path":
        """A path built from PathEvents the way the tessellators consume them (only the
        events' end and control points are read: Begin's ``at``, the segments' ``ctrl`` and
        ``to``, End's ``close``)."""
        b = PathBuilder()
        for e in events:
            if e.kind == "begin":
                b.begin(e.at)
            elif e.kind == "line":
                b.line_to(e.to)
            elif e.kind == "quadratic":
                b.quadratic_bezier_to(e.ctrl1, e.to)
            elif e.kind == "cubic":
                b.cubic_bezier_to(e.ctrl1, e.ctrl2, e.to)
            elif e.kind == "end":
                b.end(e.close)
            else:
                raise ValueError(f"unknown path event {e.kind!r}")
        return b.build()


class PathBuilder:
    """Mirror of lyon_path's ``BuilderImpl`` (path.rs:472-611)."""

    def __init__(self, num_attributes: int = 0):
        """num_attributes > 0: Path::builder_with_attributes (path.rs:115-118); every
        endpoint then takes that many custom attributes."""
        self._verbs: list[int] = []
        self._points: list[tuple[float, float]] = []
        self._first = (0.0, 0.0)
        self._in_sub_path = False
        self._n_attr = int(num_attributes)
        self._attrs: list[list[float]] = []  # one row per point
        self._first_attrs: list[float] = []

    def _row(self, attributes, control=0):
        if self._n_attr == 0:
            return []
        for _ in range(control):
            self._attrs.append([float("nan")] * self._n_attr)
        row = [float(a) for a in (attributes if attributes is not None else [])]
        assert len(row) == self._n_attr, "one value per custom attribute"
        self._attrs.append(row)
        return row

    def begin(self, at, attributes=None):
        assert not self._in_sub_path, "begin() inside a sub-path"
        self._in_sub_path = True
        self._first = (float(at[0]), float(at[1]))
        self._points.append(self._first)
        self._first_attrs = self._row(attributes)
        self._verbs.append(VERB_BEGIN)
        return len(self._points) - 1

    def line_to(self, to, attributes=None):
        assert self._in_sub_path, "line_to() outside a sub-path"
        self._points.append((float(to[0]), float(to[1])))
        self._row(attributes)
        self._verbs.append(VERB_LINE)
        return len(self._points) - 1

    def quadratic_bezier_to(self, ctrl, to, attributes=None):
        assert self._in_sub_path
        self._points.append((float(ctrl[0]), float(ctrl[1])))
        self._points.append((float(to[0]), float(to[1])))
        self._row(attributes, control=1)
        self._verbs.append(VERB_QUADRATIC)
        return len(self._points) - 1

    def cubic_bezier_to(self, ctrl1, ctrl2, to, attributes=None):
        assert self._in_sub_path
        self._points.append((float(ctrl1[0]), float(ctrl1[1])))
        self._points.append((float(ctrl2[0]), float(ctrl2[1])))
        self._points.append((float(to[0]), float(to[1])))
        self._row(attributes, control=2)
        self._verbs.append(VERB_CUBIC)
        return len(self._points) - 1

    def end(self, close: bool = False):
        assert self._in_sub_path, "end() outside a sub-path"
        self._in_sub_path = False
        if close:
            self._points.append(self._first)
            if self._n_attr:
                self._attrs.append(list(self._first_attrs))
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
        attrs = None
        if self._n_attr:
            attrs = np.asarray(self._attrs, dtype=np.float32).reshape(-1, self._n_attr)
        return Path(np.asarray(self._verbs, dtype=np.uint8), pts, attrs)


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


class PathBuffer:
    """lyon_path's ``PathBuffer`` (path_buffer.rs:14-104): many paths stored contiguously.
    Paths are appended with :meth:`builder` (whose ``build()`` returns the new path's
    index, path_buffer.rs:230-245) or :meth:`push`; :meth:`as_batch` gives the packed
    layout the batched tessellators and the SVG parser use (verbs and points end to end,
    n + 1 offsets)."""

    def __init__(self):
        self._paths: list[Path] = []
        self._batch = None

    def __len__(self):
        return len(self._paths)

    def is_empty(self) -> bool:
        return not self._paths

    def get(self, index: int) -> Path:
        return self._paths[index]

    def indices(self):
        return range(len(self._paths))

    def iter(self):
        return iter(self._paths)

    def clear(self):
        self._paths.clear()
        self._batch = None

    def push(self, path: Path) -> int:
        self._paths.append(path)
        self._batch = None
        return len(self._paths) - 1

    def builder(self) -> "PathBufferBuilder":
        return PathBufferBuilder(self)

    def as_batch(self) -> PathBatch:
        if self._batch is None:
            self._batch = PathBatch.from_paths(self._paths)
        return self._batch

    @staticmethod
    def from_batch(batch: PathBatch) -> "PathBuffer":
        b = PathBuffer()
        for i in range(len(batch)):
            b.push(batch.path(i))
        return b


class PathBufferBuilder(PathBuilder):
    """PathBuffer::builder (path_buffer.rs:195-260): a path builder whose ``build()``
    appends the path to its buffer and returns the path's index."""

    def __init__(self, buffer: PathBuffer):
        super().__init__()
        self._buffer = buffer

    def build(self) -> int:  # type: ignore[override]
        return self._buffer.push(PathBuilder.build(self))
