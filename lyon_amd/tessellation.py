"""Python mirror of lyon_tessellation's drop-in API, backed by liblyg.so on the GPU.

Names, defaults and error behaviour follow crates/tessellation/src/lib.rs (options,
enums), geometry_builder.rs (VertexBuffers, BuffersBuilder) and error.rs
(TessellationError).  ``StrokeTessellator.tessellate_path`` and
``FillTessellator.tessellate_path`` append to a ``VertexBuffers`` through a
``BuffersBuilder`` exactly as the Rust calls do; on error they raise
``TessellationError`` and leave the buffers untouched (abort_geometry).
"""
from __future__ import annotations

import ctypes as C
import threading
from enum import IntEnum

import numpy as np

from . import _lib as B
from .path import Path


class LineCap(IntEnum):  # crates/path/src/lib.rs:170-188
    Butt = 0
    Square = 1
    Round = 2


class LineJoin(IntEnum):  # crates/path/src/lib.rs:190-206
    Miter = 0
    MiterClip = 1
    Round = 2
    Bevel = 3


class FillRule(IntEnum):  # crates/path/src/lib.rs:262-280
    EvenOdd = 0
    NonZero = 1


class Orientation(IntEnum):  # crates/tessellation/src/lib.rs:302-308
    Horizontal = 0
    Vertical = 1


class TessellationError(Exception):
    """TessellationError (error.rs); ``code`` is the lyg_status value."""

    def __init__(self, code: int, detail: str = ""):
        self.code = int(code)
        msg = B.lib().lyg_status_string(self.code).decode()
        super().__init__(msg + (f" ({detail})" if detail else ""))


class StrokeOptions:
    """StrokeOptions (lib.rs:311-433)."""

    MINIMUM_MITER_LIMIT = 1.0
    DEFAULT_MITER_LIMIT = 4.0
    DEFAULT_LINE_CAP = LineCap.Butt
    DEFAULT_LINE_JOIN = LineJoin.Miter
    DEFAULT_LINE_WIDTH = 1.0
    DEFAULT_TOLERANCE = 0.1

    def __init__(self, start_cap=LineCap.Butt, end_cap=LineCap.Butt, line_join=LineJoin.Miter,
                 line_width=1.0, miter_limit=4.0, tolerance=0.1, variable_line_width=None):
        self.start_cap = LineCap(start_cap)
        self.end_cap = LineCap(end_cap)
        self.line_join = LineJoin(line_join)
        self.line_width = float(line_width)
        self.miter_limit = float(miter_limit)
        self.tolerance = float(tolerance)
        self.variable_line_width = variable_line_width

    @classmethod
    def default(cls):
        return cls()

    @classmethod
    def tolerance_(cls, tolerance):
        return cls(tolerance=tolerance)

    def _copy(self, **kw):
        d = dict(start_cap=self.start_cap, end_cap=self.end_cap, line_join=self.line_join,
                 line_width=self.line_width, miter_limit=self.miter_limit,
                 tolerance=self.tolerance, variable_line_width=self.variable_line_width)
        d.update(kw)
        return StrokeOptions(**d)

    def with_tolerance(self, t):
        return self._copy(tolerance=t)

    def with_line_cap(self, cap):
        return self._copy(start_cap=cap, end_cap=cap)

    def with_start_cap(self, cap):
        return self._copy(start_cap=cap)

    def with_end_cap(self, cap):
        return self._copy(end_cap=cap)

    def with_line_join(self, join):
        return self._copy(line_join=join)

    def with_line_width(self, w):
        return self._copy(line_width=w)

    def with_miter_limit(self, limit):
        # lib.rs:421-426 asserts the SVG minimum.
        assert limit >= self.MINIMUM_MITER_LIMIT
        return self._copy(miter_limit=limit)

    def to_c(self) -> B.StrokeOptionsC:
        return B.StrokeOptionsC(self.line_width, self.miter_limit, self.tolerance,
                                int(self.start_cap), int(self.end_cap), int(self.line_join), 0)


class FillOptions:
    """FillOptions (lib.rs:445-531)."""

    DEFAULT_TOLERANCE = 0.1
    DEFAULT_FILL_RULE = FillRule.EvenOdd
    DEFAULT_SWEEP_ORIENTATION = Orientation.Vertical

    def __init__(self, tolerance=0.1, fill_rule=FillRule.EvenOdd,
                 sweep_orientation=Orientation.Vertical, handle_intersections=True):
        self.tolerance = float(tolerance)
        self.fill_rule = FillRule(fill_rule)
        self.sweep_orientation = Orientation(sweep_orientation)
        self.handle_intersections = bool(handle_intersections)

    @classmethod
    def default(cls):
        return cls()

    @classmethod
    def even_odd(cls):
        return cls()

    @classmethod
    def non_zero(cls):
        return cls(fill_rule=FillRule.NonZero)

    @classmethod
    def tolerance_(cls, t):
        return cls(tolerance=t)

    def with_tolerance(self, t):
        return FillOptions(t, self.fill_rule, self.sweep_orientation, self.handle_intersections)

    def with_fill_rule(self, r):
        return FillOptions(self.tolerance, r, self.sweep_orientation, self.handle_intersections)

    def with_sweep_orientation(self, o):
        return FillOptions(self.tolerance, self.fill_rule, o, self.handle_intersections)

    def with_intersections(self, b):
        return FillOptions(self.tolerance, self.fill_rule, self.sweep_orientation, b)

    def to_c(self) -> B.FillOptionsC:
        return B.FillOptionsC(self.tolerance, int(self.fill_rule), int(self.sweep_orientation),
                              1 if self.handle_intersections else 0, 0)


class VertexBuffers:
    """VertexBuffers<[f32; 2], u32> (geometry_builder.rs:262-286)."""

    def __init__(self):
        self.vertices = np.zeros((0, 2), np.float32)
        self.indices = np.zeros(0, np.uint32)

    def clear(self):
        self.vertices = self.vertices[:0]
        self.indices = self.indices[:0]


class BuffersBuilder:
    """BuffersBuilder<[f32;2], u32, Positions> (geometry_builder.rs:300-351): ids are
    ``len(vertices) + vertex_offset`` at the time each vertex is added."""

    def __init__(self, buffers: VertexBuffers, ctor=None):
        if ctor is not None:
            raise NotImplementedError("only the Positions vertex constructor is supported")
        self.buffers = buffers
        self.vertex_offset = 0

    def with_vertex_offset(self, offset: int) -> "BuffersBuilder":
        self.vertex_offset = int(offset)
        return self


def simple_builder(buffers: VertexBuffers) -> BuffersBuilder:
    return BuffersBuilder(buffers)


class _CtxHandle:
    """Owns one lyg context; destroyed with the thread that made it."""

    def __init__(self, handle):
        self.handle = handle

    def __del__(self):
        try:
            B.lib().lyg_ctx_destroy(self.handle)
        except Exception:  # interpreter shutdown: the library may already be gone
            pass


class _Ctx:
    """One lyg context per Python thread (created on first use, device 0).

    A lyg context is not thread-safe (its scratch buffers, pinned staging and stream are
    shared by every call made through it), and ctypes releases the GIL during a call, so
    each thread gets its own context, kept in thread-local storage."""

    _local = threading.local()
    _lock = threading.Lock()

    @classmethod
    def get(cls):
        h = getattr(cls._local, "ctx", None)
        if h is None:
            with cls._lock:
                handle = B.lib().lyg_ctx_create(0)
            if not handle:
                raise RuntimeError("lyg_ctx_create failed (no usable HIP device?)")
            h = _CtxHandle(handle)
            cls._local.ctx = h
        return h.handle


def _path_view(path: Path) -> B.PathViewC:
    v = path.verbs
    p = path.points
    return B.PathViewC(v.ctypes.data_as(B.U8P), len(v), p.ctypes.data_as(B.F32P), len(p))


def _run(call, path: Path, copts, builder: BuffersBuilder):
    bufs = builder.buffers
    first = len(bufs.vertices) + builder.vertex_offset
    pv = _path_view(path)
    # Two-phase protocol: start with a generous guess, retry at the exact size.
    nseg = max(1, len(path.verbs))
    vcap, icap = 64 * nseg + 64, 192 * nseg + 192
    while True:
        v = np.empty((vcap, 2), np.float32)
        i = np.empty(icap, np.uint32)
        out = B.OutputC(v.ctypes.data_as(B.F32P), i.ctypes.data_as(B.U32P), vcap, icap,
                        first, 0, 0)
        st = call(_Ctx.get(), C.byref(pv), C.byref(copts), C.byref(out))
        if st == 16:  # LYG_ERR_CAPACITY
            vcap, icap = max(out.n_vertices, 1), max(out.n_indices, 1)
            continue
        if st != 0:
            raise TessellationError(st, B.lib().lyg_ctx_last_error(_Ctx.get()).decode())
        break
    bufs.vertices = np.concatenate([bufs.vertices, v[:out.n_vertices]])
    bufs.indices = np.concatenate([bufs.indices, i[:out.n_indices]])


def _shape_path(fn) -> Path:
    from .path import PathBuilder
    b = PathBuilder()
    fn(b)
    return b.build()


def _run_direct(call, builder: BuffersBuilder):
    """A direct construction (no path) through lyg_output, two-phase like _run."""
    bufs = builder.buffers
    first = len(bufs.vertices) + builder.vertex_offset
    vcap, icap = 256, 768
    while True:
        v = np.empty((vcap, 2), np.float32)
        i = np.empty(icap, np.uint32)
        out = B.OutputC(v.ctypes.data_as(B.F32P), i.ctypes.data_as(B.U32P), vcap, icap,
                        first, 0, 0)
        st = call(out)
        if st == 16:  # LYG_ERR_CAPACITY
            vcap, icap = max(out.n_vertices, 1), max(out.n_indices, 1)
            continue
        if st != 0:
            raise TessellationError(st)
        break
    bufs.vertices = np.concatenate([bufs.vertices, v[:out.n_vertices]])
    bufs.indices = np.concatenate([bufs.indices, i[:out.n_indices]])


class StrokeTessellator:
    """StrokeTessellator (stroke.rs:108-308) — fixed line width only."""

    def tessellate_path(self, path: Path, options: StrokeOptions, builder: BuffersBuilder):
        if options.variable_line_width is not None:
            # tessellate_with_ids_vw (stroke.rs:653-767): the width factor is the path's
            # custom attribute `variable_line_width` of each endpoint
            if path.attributes is None:
                raise ValueError("variable_line_width needs a path with custom attributes")
            attrs = np.ascontiguousarray(path.attributes, np.float32)
            idx = int(options.variable_line_width)
            if not 0 <= idx < attrs.shape[1]:
                raise ValueError("variable_line_width is not an attribute of the path")

            def call(ctx, pv, copts, out):
                return B.lib().lyg_stroke_path_vw(ctx, pv, attrs.ctypes.data_as(B.F32P),
                                                  attrs.shape[1], idx, copts, out, None)
            _run(call, path, options.to_c(), builder)
            return
        _run(B.lib().lyg_stroke_path, path, options.to_c(), builder)

    tessellate = tessellate_path

    # stroke.rs:255-307: the shapes are built as paths and stroked
    def tessellate(self, events, options: StrokeOptions, builder: BuffersBuilder):
        """StrokeTessellator::tessellate (stroke.rs:122-138) over PathEvents."""
        self.tessellate_path(Path.from_events(events), options, builder)

    def tessellate_polygon(self, points, closed, options: StrokeOptions, builder: BuffersBuilder):
        self.tessellate_path(_shape_path(lambda b: b.add_polygon(points, closed)), options, builder)

    def tessellate_rectangle(self, rect, options: StrokeOptions, builder: BuffersBuilder):
        self.tessellate_path(_shape_path(lambda b: b.add_rectangle(rect, True)), options, builder)

    def tessellate_circle(self, center, radius, options: StrokeOptions, builder: BuffersBuilder):
        self.tessellate_path(_shape_path(lambda b: b.add_circle(center, radius, True)), options,
                             builder)

    def tessellate_ellipse(self, center, radii, x_rotation, positive, options: StrokeOptions,
                           builder: BuffersBuilder):
        self.tessellate_path(
            _shape_path(lambda b: b.add_ellipse(center, radii, x_rotation, positive)), options,
            builder)


class FillTessellator:
    """FillTessellator (fill.rs:523-821).

    ``method="sweep"`` (default) is lyon's sweep, one path per wave: lyon's vertices and
    triangles.  ``method="banded"`` spreads one large path over the whole GPU
    (lyg_fill_path_banded): the same covered region, different triangles."""

    def __init__(self, method: str = "sweep"):
        if method not in ("sweep", "banded"):
            raise ValueError("method is 'sweep' or 'banded'")
        self.method = method

    def tessellate_path(self, path: Path, options: FillOptions, builder: BuffersBuilder):
        if not (options.tolerance > 0.0):
            raise TessellationError(4)
        fn = B.lib().lyg_fill_path_banded if self.method == "banded" else B.lib().lyg_fill_path
        _run(fn, path, options.to_c(), builder)

    tessellate = tessellate_path

    def tessellate(self, events, options: FillOptions, builder: BuffersBuilder):
        """FillTessellator::tessellate (fill.rs:578-598) over PathEvents."""
        self.tessellate_path(Path.from_events(events), options, builder)

    def tessellate_polygon(self, points, closed, options: FillOptions, builder: BuffersBuilder):
        """fill.rs:644-651: the polygon's path events."""
        self.tessellate_path(_shape_path(lambda b: b.add_polygon(points, closed)), options, builder)

    def tessellate_rectangle(self, rect, options: FillOptions, builder: BuffersBuilder):
        """fill.rs:653-661 -> basic_shapes::fill_rectangle (lyg_fill_rectangle)."""
        r = np.asarray(rect, np.float32)
        _run_direct(lambda out: B.lib().lyg_fill_rectangle(
            _Ctx.get(), r.ctypes.data_as(B.F32P), C.byref(out)), builder)

    def tessellate_circle(self, center, radius, options: FillOptions, builder: BuffersBuilder):
        """fill.rs:663-672 -> basic_shapes::fill_circle (lyg_fill_circle)."""
        o = options.to_c()
        _run_direct(lambda out: B.lib().lyg_fill_circle(
            _Ctx.get(), float(center[0]), float(center[1]), float(radius), C.byref(o),
            C.byref(out)), builder)

    def tessellate_ellipse(self, center, radii, x_rotation, positive, options: FillOptions,
                           builder: BuffersBuilder):
        """fill.rs:674-690: the ellipse path, filled with intersections off."""
        self.tessellate_path(
            _shape_path(lambda b: b.add_ellipse(center, radii, x_rotation, positive)),
            options.with_intersections(False), builder)


# ---- lyon_geom flattening (array form) --------------------------------------------------
def flatten_segments(kinds, ctrl, tolerance: float):
    """for_each_flattened_with_t over n segments on the GPU.

    ``kinds``: int array (0 line, 1 quadratic, 2 cubic); ``ctrl``: (n, 4, 2) float32 with
    from/ctrl1/ctrl2/to (a line uses slots 0 and 1).  Returns (points (m, 2), offsets (n+1)).
    """
    k = np.ascontiguousarray(kinds, np.int32)
    c = np.ascontiguousarray(ctrl, np.float32).reshape(-1, 8)
    n = len(k)
    offs = np.zeros(n + 1, np.uint32)
    npts = C.c_uint32(0)
    L = B.lib()
    st = L.lyg_flatten_segments(_Ctx.get(), k.ctypes.data_as(B.I32P), c.ctypes.data_as(B.F32P),
                                n, tolerance, None, 0, offs.ctypes.data_as(B.U32P),
                                C.byref(npts))
    if st not in (0, 16):
        raise TessellationError(st)
    pts = np.empty((max(npts.value, 1), 2), np.float32)
    st = L.lyg_flatten_segments(_Ctx.get(), k.ctypes.data_as(B.I32P), c.ctypes.data_as(B.F32P),
                                n, tolerance, pts.ctypes.data_as(B.F32P), npts.value,
                                offs.ctypes.data_as(B.U32P), C.byref(npts))
    if st != 0:
        raise TessellationError(st)
    return pts[:npts.value], offs


def _arcs(fn, arcs, tol, per):
    a = np.ascontiguousarray(arcs, np.float32).reshape(-1, 7)
    n = len(a)
    offs = np.zeros(n + 1, np.uint32)
    cnt = C.c_uint32(0)
    args = [_Ctx.get(), a.ctypes.data_as(B.F32P), n] + ([tol] if tol is not None else [])
    st = fn(*args, None, 0, offs.ctypes.data_as(B.U32P), C.byref(cnt))
    if st not in (0, 16):
        raise TessellationError(st)
    out = np.empty((max(cnt.value, 1), per), np.float32)
    st = fn(*args, out.ctypes.data_as(B.F32P), cnt.value, offs.ctypes.data_as(B.U32P),
            C.byref(cnt))
    if st != 0:
        raise TessellationError(st)
    return out[:cnt.value], offs


def flatten_arcs(arcs, tolerance: float):
    """Arc::for_each_flattened_with_t for arcs given as rows
    (cx, cy, rx, ry, start, sweep, x_rotation).  Returns (points (m, 2), offsets)."""
    return _arcs(B.lib().lyg_flatten_arcs, arcs, tolerance, 2)


def arc_to_quadratics(arcs):
    """arc_to_quadratic_beziers_with_t: returns (quads (m, 3, 2), offsets)."""
    q, o = _arcs(B.lib().lyg_arc_to_quadratics, arcs, None, 6)
    return q.reshape(-1, 3, 2), o
