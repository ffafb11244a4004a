"""Batched form: many paths per launch, device-resident inputs and outputs.

Device memory is allocated through liblyg.so's own entry points (``lyg_device_alloc`` and
friends), so the batched form needs nothing but the library.  The result of a batch equals
calling ``StrokeTessellator.tessellate_path`` / ``FillTessellator.tessellate_path`` for
every path in order into one shared ``VertexBuffers`` (geometry_builder.rs:300-351): path
i's vertices start at ``vertex_offsets[i]`` and its indices are absolute ids.  A fill path
that fails contributes nothing and its status is reported per path (fill.rs:788-793).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as B
from .path import PathBatch
from .tessellation import TessellationError, _Ctx

LYG_ERR_CAPACITY = 16


class DeviceBuffer:
    """A device allocation owned by the library's context (freed with the object)."""

    def __init__(self, nbytes: int):
        self.nbytes = int(max(nbytes, 16))
        p = C.c_void_p(0)
        st = B.lib().lyg_device_alloc(_Ctx.get(), self.nbytes, C.byref(p))
        if st != 0:
            raise TessellationError(st, B.lib().lyg_ctx_last_error(_Ctx.get()).decode())
        self.ptr = p.value

    @classmethod
    def from_numpy(cls, a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        b.upload(a)
        return b

    def upload(self, a: np.ndarray, offset: int = 0):
        a = np.ascontiguousarray(a)
        if offset + a.nbytes > self.nbytes:
            raise ValueError("upload beyond the buffer")
        st = B.lib().lyg_copy_to_device(_Ctx.get(), self.ptr + offset, a.ctypes.data, a.nbytes)
        if st != 0:
            raise TessellationError(st, B.lib().lyg_ctx_last_error(_Ctx.get()).decode())

    def download(self, dtype, count: int, offset: int = 0) -> np.ndarray:
        out = np.empty(count, dtype)
        if offset + out.nbytes > self.nbytes:
            raise ValueError("download beyond the buffer")
        st = B.lib().lyg_copy_to_host(_Ctx.get(), out.ctypes.data, self.ptr + offset, out.nbytes)
        if st != 0:
            raise TessellationError(st, B.lib().lyg_ctx_last_error(_Ctx.get()).decode())
        return out

    def __del__(self):
        try:
            if getattr(self, "ptr", None):
                B.lib().lyg_device_free(_Ctx.get(), self.ptr)
                self.ptr = None
        except Exception:
            pass


def synchronize():
    st = B.lib().lyg_ctx_synchronize(_Ctx.get())
    if st != 0:
        raise TessellationError(st, B.lib().lyg_ctx_last_error(_Ctx.get()).decode())


class DeviceBatch:
    """A PathBatch copied to device memory once (inputs stay resident in HBM), with
    per-path options: a structured array of lyg_stroke_options or lyg_fill_options
    records (or one record shared by every path)."""

    def __init__(self, batch: PathBatch, opts=None):
        self.n = len(batch)
        self.batch = batch
        self.verbs = DeviceBuffer.from_numpy(batch.verbs)
        self.points = DeviceBuffer.from_numpy(batch.points.reshape(-1))
        self.verb_offsets = DeviceBuffer.from_numpy(batch.verb_offsets)
        self.point_offsets = DeviceBuffer.from_numpy(batch.point_offsets)
        self.opts = None
        self.n_opts = 0
        if opts is not None:
            raw = np.ascontiguousarray(opts)
            self.opts = DeviceBuffer.from_numpy(raw.view(np.uint8).reshape(-1))
            self.n_opts = len(raw)

    def c_in(self) -> B.BatchInC:
        return B.BatchInC(self.verbs.ptr, self.points.ptr, self.verb_offsets.ptr,
                          self.point_offsets.ptr, self.n)


class BatchResult:
    """Device-resident output of one batch call."""

    def __init__(self, vertices, indices, voffs, ioffs, status, n_paths, n_vertices, n_indices):
        self.vertices = vertices      # DeviceBuffer, (capacity, 2) f32
        self.indices = indices        # DeviceBuffer, (capacity,) u32
        self.voffs = voffs            # DeviceBuffer, (n+1,) u32
        self.ioffs = ioffs
        self.status_buf = status      # DeviceBuffer, (n,) i32
        self.n = n_paths
        self.n_vertices = n_vertices
        self.n_indices = n_indices

    @property
    def vertex_offsets(self) -> np.ndarray:
        return self.voffs.download(np.uint32, self.n + 1)

    @property
    def index_offsets(self) -> np.ndarray:
        return self.ioffs.download(np.uint32, self.n + 1)

    @property
    def status(self) -> np.ndarray:
        return self.status_buf.download(np.int32, self.n)

    def all_vertices(self) -> np.ndarray:
        return self.vertices.download(np.float32, 2 * self.n_vertices).reshape(-1, 2)

    def all_indices(self) -> np.ndarray:
        return self.indices.download(np.uint32, self.n_indices)

    def path(self, i: int):
        """(vertices, indices) of path i as numpy arrays (indices absolute, u32)."""
        vo = self.vertex_offsets
        io = self.index_offsets
        v = self.vertices.download(np.float32, 2 * int(vo[i + 1] - vo[i]), 8 * int(vo[i]))
        idx = self.indices.download(np.uint32, int(io[i + 1] - io[i]), 4 * int(io[i]))
        return v.reshape(-1, 2), idx


class _Batcher:
    """Reusable output arena for repeated batch calls (no allocation per call once the
    arena is large enough)."""

    _entry = None

    def __init__(self):
        self.vertices = None
        self.indices = None
        self.voffs = None
        self.ioffs = None
        self.status = None
        self.n_cap = 0

    def _ensure(self, n_paths, nv, ni):
        if self.voffs is None or self.n_cap < n_paths:
            self.voffs = DeviceBuffer(4 * (n_paths + 1))
            self.ioffs = DeviceBuffer(4 * (n_paths + 1))
            self.status = DeviceBuffer(4 * max(n_paths, 1))
            self.n_cap = n_paths
        if self.vertices is None or self.vertices.nbytes < 8 * nv:
            self.vertices = DeviceBuffer(8 * max(nv, 1))
        if self.indices is None or self.indices.nbytes < 4 * ni:
            self.indices = DeviceBuffer(4 * max(ni, 1))

    def run(self, db: DeviceBatch, vertex_offset: int = 0, check=True) -> BatchResult:
        if db.opts is None:
            raise ValueError("DeviceBatch needs options")
        self._ensure(db.n, 0, 0)
        cin = db.c_in()
        fn = getattr(B.lib(), self._entry)
        for _ in range(3):
            out = B.BatchOutC(self.vertices.ptr, self.indices.ptr, self.vertices.nbytes // 8,
                              self.indices.nbytes // 4, self.voffs.ptr, self.ioffs.ptr,
                              vertex_offset, 0, 0, self.status.ptr)
            st = fn(_Ctx.get(), C.byref(cin), db.opts.ptr, db.n_opts, C.byref(out))
            if st == LYG_ERR_CAPACITY:
                self._ensure(db.n, out.n_vertices, out.n_indices)
                continue
            if st != 0 and (check or st in (17, 18, 19, 20)):
                raise TessellationError(st, B.lib().lyg_ctx_last_error(_Ctx.get()).decode())
            return BatchResult(self.vertices, self.indices, self.voffs, self.ioffs, self.status,
                               db.n, out.n_vertices, out.n_indices)
        raise RuntimeError("batch did not converge")


class StrokeBatcher(_Batcher):
    _entry = "lyg_stroke_batch"

    def count(self, db: DeviceBatch):
        nv = DeviceBuffer(4 * db.n)
        ni = DeviceBuffer(4 * db.n)
        tv = C.c_uint64(0)
        ti = C.c_uint64(0)
        cin = db.c_in()
        st = B.lib().lyg_stroke_batch_count(_Ctx.get(), C.byref(cin), db.opts.ptr, db.n_opts,
                                            nv.ptr, ni.ptr, C.byref(tv), C.byref(ti))
        if st != 0:
            raise TessellationError(st)
        return tv.value, ti.value


class StrokeVWBatcher(_Batcher):
    """Batched variable-width stroke (lyg_stroke_batch_vw): ``attributes`` holds, per point
    of the batch, the custom attributes of the endpoint stored there (n_points x k float32);
    column ``width_attribute`` is the width factor."""

    _entry = "lyg_stroke_batch_vw"

    def __init__(self, attributes, width_attribute: int = 0, vertex_attributes=False):
        """vertex_attributes: also write every vertex's interpolated attributes into
        ``self.vertex_attributes`` (a DeviceBuffer of n_vertices x k float32)."""
        super().__init__()
        a = np.ascontiguousarray(attributes, np.float32)
        self.n_attributes = 1 if a.ndim == 1 else int(a.shape[1])
        self.width_attribute = int(width_attribute)
        self.attributes = DeviceBuffer.from_numpy(a.reshape(-1))
        self.want_vertex_attributes = bool(vertex_attributes)
        self.vertex_attributes = None

    def _ensure(self, n_paths, nv, ni):
        super()._ensure(n_paths, nv, ni)
        if self.want_vertex_attributes:
            need = 4 * self.n_attributes * max(nv, 1, self.vertices.nbytes // 8)
            if self.vertex_attributes is None or self.vertex_attributes.nbytes < need:
                self.vertex_attributes = DeviceBuffer(need)

    def run(self, db: DeviceBatch, vertex_offset: int = 0, check=True) -> BatchResult:
        self._ensure(db.n, 0, 0)
        cin = db.c_in()
        for _ in range(3):
            out = B.BatchOutC(self.vertices.ptr, self.indices.ptr, self.vertices.nbytes // 8,
                              self.indices.nbytes // 4, self.voffs.ptr, self.ioffs.ptr,
                              vertex_offset, 0, 0, self.status.ptr)
            st = B.lib().lyg_stroke_batch_vw(_Ctx.get(), C.byref(cin), self.attributes.ptr,
                                             self.n_attributes, self.width_attribute,
                                             db.opts.ptr, db.n_opts, C.byref(out),
                                             self.vertex_attributes.ptr
                                             if self.vertex_attributes is not None else None)
            if st == LYG_ERR_CAPACITY:
                self._ensure(db.n, out.n_vertices, out.n_indices)
                continue
            if st != 0 and (check or st in (17, 18, 19)):
                raise TessellationError(st, B.lib().lyg_ctx_last_error(_Ctx.get()).decode())
            return BatchResult(self.vertices, self.indices, self.voffs, self.ioffs, self.status,
                               db.n, out.n_vertices, out.n_indices)
        raise RuntimeError("batch did not converge")


class FillBatcher(_Batcher):
    """Batched fill; with check=False a path's failure is reported in ``status`` only.
    ``banded=True``: every path through the banded fill (lyg_fill_batch_banded)."""

    _entry = "lyg_fill_batch"

    def __init__(self, banded: bool = False):
        super().__init__()
        self.banded = banded
        if banded:
            self._entry = "lyg_fill_batch_banded"


def flattening_errors(db: DeviceBatch, tolerance, stroke_rule: bool) -> np.ndarray:
    """lyg_flattening_errors: per path, the largest distance (f64) between its flattened
    curves' chords and the true curves, under the stroker's or the fill's flattening."""
    tol = np.atleast_1d(np.asarray(tolerance, np.float32))
    dt = DeviceBuffer.from_numpy(tol)
    out = DeviceBuffer(4 * max(db.n, 1))
    cin = db.c_in()
    st = B.lib().lyg_flattening_errors(_Ctx.get(), C.byref(cin), dt.ptr, len(tol),
                                       1 if stroke_rule else 0, out.ptr)
    if st != 0:
        raise TessellationError(st, B.lib().lyg_ctx_last_error(_Ctx.get()).decode())
    return out.download(np.float32, db.n)


def set_band_size(edges_per_band: int = 256, max_bands: int = 16384):
    """Band size of the banded fill (lyg_ctx_set_band_size)."""
    st = B.lib().lyg_ctx_set_band_size(_Ctx.get(), int(edges_per_band), int(max_bands))
    if st != 0:
        raise TessellationError(st)


class StrokeGraph:
    """One batched stroke captured into a HIP graph (lyg_stroke_batch_graph) on fixed
    device buffers: ``launch()`` replays count -> scan -> emit with no host sync inside;
    ``result()`` waits and returns a BatchResult (raises on an output overflow)."""

    _entry = "lyg_stroke_batch_graph"
    _status_out = False

    def __init__(self, db: DeviceBatch, vertex_capacity: int, index_capacity: int,
                 vertex_offset: int = 0):
        self.db = db
        self.vertices = DeviceBuffer(8 * max(vertex_capacity, 1))
        self.indices = DeviceBuffer(4 * max(index_capacity, 1))
        self.voffs = DeviceBuffer(4 * (db.n + 1))
        self.ioffs = DeviceBuffer(4 * (db.n + 1))
        self.status = DeviceBuffer(4 * max(db.n, 1))
        out = B.BatchOutC(self.vertices.ptr, self.indices.ptr, vertex_capacity, index_capacity,
                          self.voffs.ptr, self.ioffs.ptr, vertex_offset, 0, 0,
                          self.status.ptr if self._status_out else None)
        cin = db.c_in()
        h = C.c_void_p(0)
        st = getattr(B.lib(), self._entry)(_Ctx.get(), C.byref(cin), db.opts.ptr, db.n_opts,
                                           C.byref(out), C.byref(h))
        if st != 0:
            raise TessellationError(st, B.lib().lyg_ctx_last_error(_Ctx.get()).decode())
        self._h = h.value

    def launch(self):
        st = B.lib().lyg_graph_launch(self._h)
        if st != 0:
            raise TessellationError(st)

    def result(self, check=True) -> BatchResult:
        """With check=False a status that belongs to paths (a path's error, an output
        overflow, a path that outgrew its room) is left in ``status``, not raised."""
        nv = C.c_uint64(0)
        ni = C.c_uint64(0)
        st = B.lib().lyg_graph_result(self._h, C.byref(nv), C.byref(ni))
        if st != 0 and (check or st in (17, 18, 19)):
            raise TessellationError(st)
        return BatchResult(self.vertices, self.indices, self.voffs, self.ioffs, self.status,
                           self.db.n, nv.value, ni.value)

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                B.lib().lyg_graph_destroy(self._h)
                self._h = None
        except Exception:
            pass


class FillGraph(StrokeGraph):
    """The batched fill's first sweep round captured into HIP graphs
    (lyg_fill_batch_graph), planned from db at capture time: ``launch()`` replays the
    sweep launches, offset scans and copy with no host sync inside and writes each path's
    status; ``result()`` waits and returns a BatchResult (raises as lyg_fill_batch does,
    and with LYG_ERR_NO_SPACE when a path outgrew its planned room)."""

    _entry = "lyg_fill_batch_graph"
    _status_out = True


def autotune(batcher: _Batcher, db: DeviceBatch) -> dict:
    """Times db under each candidate launch shape (lyg_autotune_stroke / _fill / _banded)
    and keeps the fastest in the context; returns the chosen shape (stroke: blocks per CU
    and windows per part of split paths; fill: waves per block and LDS bytes per wave;
    banded fill: edges per band)."""
    if getattr(batcher, "banded", False):
        # the smallest band size makes the most triangles: size the arena for it
        B.lib().lyg_ctx_set_band_size(_Ctx.get(), 16, 16384)
    res = batcher.run(db, check=False)  # sizes the output arena
    out = B.BatchOutC(batcher.vertices.ptr, batcher.indices.ptr, batcher.vertices.nbytes // 8,
                      batcher.indices.nbytes // 4, batcher.voffs.ptr, batcher.ioffs.ptr, 0, 0, 0,
                      batcher.status.ptr)
    cin = db.c_in()
    t = B.TuningC()
    if isinstance(batcher, StrokeBatcher):
        fn = B.lib().lyg_autotune_stroke
    elif getattr(batcher, "banded", False):
        fn = B.lib().lyg_autotune_banded
    else:
        fn = B.lib().lyg_autotune_fill
    st = fn(_Ctx.get(), C.byref(cin), db.opts.ptr, db.n_opts, C.byref(out), C.byref(t))
    if st != 0:
        raise TessellationError(st, B.lib().lyg_ctx_last_error(_Ctx.get()).decode())
    del res
    return {k: getattr(t, k) for k, _ in t._fields_}


def set_tuning(stroke_blocks_per_cu=None, fill_waves_per_block=None, fill_lds_bytes=None):
    """Sets launch shapes explicitly (lyg_ctx_set_tuning)."""
    s = f = None
    if stroke_blocks_per_cu is not None:
        s = B.TuningC(stroke_blocks_per_cu, 4, 0, 0.0)
    if fill_waves_per_block is not None:
        f = B.TuningC(0, fill_waves_per_block, fill_lds_bytes or 0, 0.0)
    st = B.lib().lyg_ctx_set_tuning(_Ctx.get(), C.byref(s) if s else None, C.byref(f) if f else None)
    if st != 0:
        raise TessellationError(st)


def set_stroke_split(windows_per_part: int = 64):
    """Intra-path parallelism of the staged batched stroke (lyg_ctx_set_stroke_split): a
    long one-sub-path path is cut into parts of this many 64-step windows walked by
    different waves (default 64; 0: one wave per path).  The output is the same either
    way."""
    st = B.lib().lyg_ctx_set_stroke_split(_Ctx.get(), int(windows_per_part))
    if st != 0:
        raise TessellationError(st)


def set_fill_lds_resident(enabled=True):
    """Placement of the batched fill sweep's arrays (lyg_ctx_set_fill_lds_resident): by
    default the first round keeps every hot array of a path that fits 150 KiB in LDS;
    False: arrays start in LDS and move to global memory as they grow; an int > 1 divides
    the LDS capacities by it (paths outgrow them and are swept again).  Same output."""
    v = int(enabled) if not isinstance(enabled, bool) else (1 if enabled else 0)
    st = B.lib().lyg_ctx_set_fill_lds_resident(_Ctx.get(), v)
    if st != 0:
        raise TessellationError(st)


def stroke_split_stats() -> dict:
    """Of the last split batched stroke: items walked and paths walked again whole."""
    items = C.c_uint64(0)
    again = C.c_uint32(0)
    B.lib().lyg_ctx_stroke_split_stats(_Ctx.get(), C.byref(items), C.byref(again))
    return {"items": items.value, "walked_again": again.value}


def set_stroke_staging(enabled: bool = True, room_scale: float = 1.0):
    """How batched strokes size their output (lyg_ctx_set_stroke_staging): one emit pass
    into per-path staging room then a compaction copy (default), or count + emit passes.
    The output is the same either way."""
    st = B.lib().lyg_ctx_set_stroke_staging(_Ctx.get(), 1 if enabled else 0, float(room_scale))
    if st != 0:
        raise TessellationError(st)


def flatten_segments_device(kinds: DeviceBuffer, ctrl: DeviceBuffer, n: int, tolerance: float,
                            out: DeviceBuffer, offsets: DeviceBuffer) -> int:
    """lyg_flatten_segments_device: flattens n segments held in device memory into `out`
    (f32 x, y pairs) with per-segment offsets (n + 1, u32); returns the point count.
    Raises TessellationError(16) with the count in .needed when `out` is too small."""
    cnt = C.c_uint32(0)
    st = B.lib().lyg_flatten_segments_device(_Ctx.get(), kinds.ptr, ctrl.ptr, n, tolerance,
                                             out.ptr if out is not None else None,
                                             (out.nbytes // 8) if out is not None else 0,
                                             offsets.ptr, C.byref(cnt))
    if st != 0:
        e = TessellationError(st, B.lib().lyg_ctx_last_error(_Ctx.get()).decode())
        e.needed = cnt.value
        raise e
    return cnt.value
