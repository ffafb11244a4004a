"""ORACLE — test infrastructure only.

ctypes face of ``liboracle.so`` (the CPU restatement of lyon's tessellators).  Only
tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg import this
module, and only as the checker; the product (``lyon_amd``) never does.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

F32P = C.POINTER(C.c_float)
U32P = C.POINTER(C.c_uint32)
U8P = C.POINTER(C.c_uint8)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run `make -C oracle`)")
        L = C.CDLL(path)
        L.orc_result_free.argtypes = [C.c_void_p]
        L.orc_result_status.argtypes = [C.c_void_p]
        L.orc_result_status.restype = C.c_int
        for n in ("orc_result_n_vertices", "orc_result_n_indices", "orc_result_n_offsets"):
            getattr(L, n).argtypes = [C.c_void_p]
            getattr(L, n).restype = C.c_uint64
        L.orc_result_copy.argtypes = [C.c_void_p, F32P, U32P, U32P]
        L.orc_stroke_path.argtypes = [U8P, C.c_uint32, F32P, C.c_uint32, C.c_float, C.c_float,
                                      C.c_float, C.c_int, C.c_int, C.c_int, C.c_uint32]
        L.orc_stroke_path.restype = C.c_void_p
        L.orc_stroke_path_vw.argtypes = [U8P, C.c_uint32, F32P, C.c_uint32, F32P, C.c_float,
                                         C.c_float, C.c_float, C.c_int, C.c_int, C.c_int,
                                         C.c_uint32, F32P, C.c_uint32]
        L.orc_stroke_path_vw.restype = C.c_void_p
        L.orc_result_n_attrs.argtypes = [C.c_void_p]
        L.orc_result_n_attrs.restype = C.c_size_t
        L.orc_result_copy_attrs.argtypes = [C.c_void_p, F32P]
        L.orc_fill_path.argtypes = [U8P, C.c_uint32, F32P, C.c_uint32, C.c_float, C.c_int,
                                    C.c_int, C.c_int, C.c_uint32]
        L.orc_fill_path.restype = C.c_void_p
        L.orc_flatten_segments.argtypes = [C.POINTER(C.c_int), F32P, C.c_uint32, C.c_float]
        L.orc_flatten_segments.restype = C.c_void_p
        L.orc_flatten_arcs.argtypes = [F32P, C.c_uint32, C.c_float]
        L.orc_flatten_arcs.restype = C.c_void_p
        L.orc_arc_to_quads.argtypes = [F32P, C.c_uint32]
        L.orc_arc_to_quads.restype = C.c_void_p
        L.orc_batch_triangles.argtypes = [U8P, F32P, U32P, U32P, F32P, C.c_uint32, C.c_int,
                                          C.c_int]
        L.orc_batch_triangles.restype = C.c_uint64
        I32P = C.POINTER(C.c_int32)
        L.orc_winding_number.argtypes = [U8P, C.c_uint32, F32P, C.c_uint32, C.c_float,
                                         C.c_float, C.c_float]
        L.orc_winding_number.restype = C.c_int32
        L.orc_winding_numbers.argtypes = [U8P, C.c_uint32, F32P, C.c_uint32, F32P, C.c_uint32,
                                          C.c_float, I32P]
        L.orc_signed_area.argtypes = [U8P, C.c_uint32, F32P, C.c_uint32, C.c_float]
        L.orc_signed_area.restype = C.c_float
        L.orc_parse_svg.argtypes = [C.c_char_p, C.c_uint32, C.c_int, C.c_int, U8P, C.c_uint32,
                                    F32P, C.c_uint32, U32P, U32P, C.POINTER(C.c_int32)]
        L.orc_parse_f32.argtypes = [C.c_char_p, C.POINTER(C.c_float)]
        L.orc_parse_f32.restype = C.c_int
        _LIB = L
    return _LIB


def _ptr(a, t):
    return a.ctypes.data_as(t)


def _collect(h):
    L = lib()
    try:
        st = L.orc_result_status(h)
        nv = L.orc_result_n_vertices(h)
        ni = L.orc_result_n_indices(h)
        no = L.orc_result_n_offsets(h)
        v = np.zeros((nv, 2), np.float32)
        i = np.zeros(ni, np.uint32)
        o = np.zeros(no, np.uint32)
        L.orc_result_copy(h, _ptr(v, F32P), _ptr(i, U32P), _ptr(o, U32P))
        return st, v, i, o
    finally:
        L.orc_result_free(h)


def stroke(path, line_width=1.0, miter_limit=4.0, tolerance=0.1, start_cap=0, end_cap=0,
           line_join=0, vertex_offset=0, widths=None, attributes=None):
    """StrokeTessellator::tessellate_path on the CPU.  Returns (status, vertices, indices).
    widths: per point of the path, the width factor of the endpoint stored there
    (StrokeOptions::variable_line_width); None strokes at the fixed width.
    attributes (with widths): per point, the custom attributes (n_points x k); the result
    then has a fourth item, every vertex's interpolated attributes (n_vertices x k)."""
    v = np.ascontiguousarray(path.verbs, np.uint8)
    p = np.ascontiguousarray(path.points, np.float32)
    if widths is not None:
        w = np.ascontiguousarray(widths, np.float32)
        assert len(w) == len(p)
        a = None
        if attributes is not None:
            a = np.ascontiguousarray(attributes, np.float32).reshape(len(p), -1)
        h = lib().orc_stroke_path_vw(_ptr(v, U8P), len(v), _ptr(p, F32P), len(p),
                                     _ptr(w, F32P), line_width, miter_limit, tolerance,
                                     start_cap, end_cap, line_join, vertex_offset,
                                     _ptr(a, F32P) if a is not None else None,
                                     0 if a is None else a.shape[1])
        attrs = None
        if a is not None:
            na = lib().orc_result_n_attrs(h)
            attrs = np.zeros(na, np.float32)
            if na:
                lib().orc_result_copy_attrs(h, _ptr(attrs, F32P))
            attrs = attrs.reshape(-1, a.shape[1])
        st, vv, ii, _ = _collect(h)
        return (st, vv, ii) if a is None else (st, vv, ii, attrs)
    h = lib().orc_stroke_path(_ptr(v, U8P), len(v), _ptr(p, F32P), len(p), line_width,
                              miter_limit, tolerance, start_cap, end_cap, line_join,
                              vertex_offset)
    st, vv, ii, _ = _collect(h)
    return st, vv, ii


def fill(path, tolerance=0.1, fill_rule=0, sweep_orientation=1, handle_intersections=True,
         vertex_offset=0):
    """FillTessellator::tessellate_path on the CPU.  Returns (status, vertices, indices)."""
    v = np.ascontiguousarray(path.verbs, np.uint8)
    p = np.ascontiguousarray(path.points, np.float32)
    h = lib().orc_fill_path(_ptr(v, U8P), len(v), _ptr(p, F32P), len(p), tolerance, fill_rule,
                            sweep_orientation, int(handle_intersections), vertex_offset)
    st, vv, ii, _ = _collect(h)
    return st, vv, ii


def flatten_segments(kinds, ctrl, tolerance):
    """lyon_geom for_each_flattened_with_t per segment. kinds: 0 line, 1 quad, 2 cubic;
    ctrl: (n, 4, 2) float32.  Returns (points, offsets)."""
    k = np.ascontiguousarray(kinds, np.int32)
    c = np.ascontiguousarray(ctrl, np.float32).reshape(-1, 8)
    h = lib().orc_flatten_segments(k.ctypes.data_as(C.POINTER(C.c_int)), _ptr(c, F32P), len(k),
                                   tolerance)
    _, v, _, o = _collect(h)
    return v, o


def flatten_arcs(arcs, tolerance):
    a = np.ascontiguousarray(arcs, np.float32).reshape(-1, 7)
    h = lib().orc_flatten_arcs(_ptr(a, F32P), len(a), tolerance)
    _, v, _, o = _collect(h)
    return v, o


def arc_to_quads(arcs):
    a = np.ascontiguousarray(arcs, np.float32).reshape(-1, 7)
    h = lib().orc_arc_to_quads(_ptr(a, F32P), len(a))
    _, v, _, o = _collect(h)
    return v.reshape(-1, 3, 2), o // 3


def batch_triangles(batch, opts, is_fill, n_threads=1):
    """CPU baseline: tessellate every path of a PathBatch; returns the triangle count."""
    o = np.ascontiguousarray(opts, np.float32)
    return lib().orc_batch_triangles(_ptr(batch.verbs, U8P), _ptr(batch.points, F32P),
                                     _ptr(batch.verb_offsets, U32P),
                                     _ptr(batch.point_offsets, U32P), _ptr(o, F32P), len(batch),
                                     int(is_fill), int(n_threads))


def batch_tessellate(batch, opts, is_fill, n_threads=None):
    """Every path of a PathBatch on the host's cores (the parity reference for full-size
    batches).  opts as batch_triangles.  Returns (status (n,), vertex offsets (n+1,),
    index offsets (n+1,), vertices (N, 2) f32, indices (M,) path-local u32)."""
    import os
    L = lib()
    L.orc_batch_run.restype = C.c_void_p
    L.orc_batch_run.argtypes = [U8P, F32P, U32P, U32P, F32P, C.c_uint32, C.c_int, C.c_int]
    L.orc_batch_counts.argtypes = [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_uint64),
                                   C.POINTER(C.c_uint64)]
    L.orc_batch_copy.argtypes = [C.c_void_p, F32P, U32P]
    L.orc_batch_free.argtypes = [C.c_void_p]
    n = len(batch)
    if n_threads is None:
        n_threads = min(16, len(os.sched_getaffinity(0)))
    o = np.ascontiguousarray(opts, np.float32)
    h = L.orc_batch_run(_ptr(batch.verbs, U8P), _ptr(batch.points, F32P),
                        _ptr(batch.verb_offsets, U32P), _ptr(batch.point_offsets, U32P),
                        _ptr(o, F32P), n, int(is_fill), int(n_threads))
    try:
        st = np.zeros(n, np.int32)
        nv = np.zeros(n, np.uint64)
        ni = np.zeros(n, np.uint64)
        L.orc_batch_counts(h, st.ctypes.data_as(C.POINTER(C.c_int32)),
                           nv.ctypes.data_as(C.POINTER(C.c_uint64)),
                           ni.ctypes.data_as(C.POINTER(C.c_uint64)))
        voff = np.concatenate([[0], np.cumsum(nv)]).astype(np.int64)
        ioff = np.concatenate([[0], np.cumsum(ni)]).astype(np.int64)
        v = np.empty((int(voff[-1]), 2), np.float32)
        i = np.empty(int(ioff[-1]), np.uint32)
        L.orc_batch_copy(h, _ptr(v, F32P), _ptr(i, U32P))
    finally:
        L.orc_batch_free(h)
    return st, voff, ioff, v, i


def _box(points, pad_frac=0.01):
    p = np.asarray(points, np.float64).reshape(-1, 2)
    p = p[np.isfinite(p).all(axis=1)]
    if len(p) == 0:
        return 0.0, 0.0, 1.0, 1.0
    lo, hi = p.min(axis=0), p.max(axis=0)
    ext = max(float((hi - lo).max()), 1e-6)
    pad = ext * pad_frac
    # a square box keeps the samples isotropic
    cx, cy = (lo + hi) / 2
    h = ext / 2 + pad
    return cx - h, cy - h, cx + h, cy + h


def coverage_compare(va, ia, vb, ib, box=None, n=4096, eps=1e-4):
    """Per-sample coverage of two triangle sets (vertices (k,2) f32, local u32 indices) on
    an n x n grid; samples within eps of a triangle edge are excluded.  Returns a dict:
    mismatch, excluded, covered_a, covered_b, overlap_a, overlap_b, area_a, area_b."""
    va = np.ascontiguousarray(va, np.float32).reshape(-1, 2)
    vb = np.ascontiguousarray(vb, np.float32).reshape(-1, 2)
    ia = np.ascontiguousarray(ia, np.uint32)
    ib = np.ascontiguousarray(ib, np.uint32)
    if box is None:
        box = _box(np.concatenate([va, vb]))
    L = lib()
    L.orc_coverage_compare.argtypes = [F32P, U32P, C.c_uint64, F32P, U32P, C.c_uint64,
                                       C.c_double, C.c_double, C.c_double, C.c_double,
                                       C.c_uint32, C.c_double, C.POINTER(C.c_uint64),
                                       C.POINTER(C.c_double)]
    out = (C.c_uint64 * 8)()
    area = (C.c_double * 2)()
    L.orc_coverage_compare(_ptr(va, F32P), _ptr(ia, U32P), len(ia) // 3, _ptr(vb, F32P),
                           _ptr(ib, U32P), len(ib) // 3, *box, n, eps, out, area)
    keys = ("mismatch", "excluded", "covered_a", "covered_b", "overlap_a", "overlap_b")
    r = {k: int(out[j]) for j, k in enumerate(keys)}
    r["area_a"], r["area_b"] = float(area[0]), float(area[1])
    return r


def winding_compare(path, tolerance, fill_rule, v, i, box=None, n=4096, eps=1e-4):
    """Triangles (v, i) against the textbook winding-number coverage of the path's
    flattened edges (vertical sweep flattening).  Keys as coverage_compare, with a =
    triangles, b = winding, plus n_edges."""
    v = np.ascontiguousarray(v, np.float32).reshape(-1, 2)
    i = np.ascontiguousarray(i, np.uint32)
    pv = np.ascontiguousarray(path.verbs, np.uint8)
    pp = np.ascontiguousarray(path.points, np.float32)
    if box is None:
        box = _box(pp)
    L = lib()
    L.orc_winding_coverage_compare.argtypes = [U8P, C.c_uint32, F32P, C.c_uint32, C.c_float,
                                               C.c_int, F32P, U32P, C.c_uint64, C.c_double,
                                               C.c_double, C.c_double, C.c_double, C.c_uint32,
                                               C.c_double, C.POINTER(C.c_uint64),
                                               C.POINTER(C.c_double)]
    out = (C.c_uint64 * 8)()
    area = (C.c_double * 2)()
    L.orc_winding_coverage_compare(_ptr(pv, U8P), len(pv), _ptr(pp, F32P), len(pp), tolerance,
                                   fill_rule, _ptr(v, F32P), _ptr(i, U32P), len(i) // 3, *box,
                                   n, eps, out, area)
    keys = ("mismatch", "excluded", "covered_a", "covered_b", "overlap_a", "overlap_b",
            "n_edges")
    r = {k: int(out[j]) for j, k in enumerate(keys)}
    r["area_a"] = float(area[0])
    return r


def fill_areas(path, tolerance):
    """Exact (f64) areas of the path's fill from its region boundary, independent of any
    triangulation (oracle/refarea.cpp).  Returns (EvenOdd area, NonZero area, number of
    edge crossings)."""
    L = lib()
    L.orc_fill_area.argtypes = [U8P, C.c_uint32, F32P, C.c_uint32, C.c_float,
                                C.POINTER(C.c_double)]
    L.orc_fill_area.restype = C.c_int
    v = np.ascontiguousarray(path.verbs, np.uint8)
    p = np.ascontiguousarray(path.points, np.float32)
    out = (C.c_double * 3)()
    L.orc_fill_area(_ptr(v, U8P), len(v), _ptr(p, F32P), len(p), tolerance, out)
    return float(out[0]), float(out[1]), int(out[2])


def fill_area(path, tolerance, fill_rule):
    """(area of the fill under fill_rule, number of edge crossings)."""
    eo, nz, k = fill_areas(path, tolerance)
    return (nz if fill_rule else eo), k


def winding_numbers(path, xy, tolerance):
    """path_winding_number_at_position (hit_test.rs:21-83) of every point of xy (k, 2)."""
    v = np.ascontiguousarray(path.verbs, np.uint8)
    p = np.ascontiguousarray(path.points, np.float32)
    q = np.ascontiguousarray(xy, np.float32).reshape(-1, 2)
    out = np.zeros(len(q), np.int32)
    lib().orc_winding_numbers(_ptr(v, U8P), len(v), _ptr(p, F32P), len(p), _ptr(q, F32P),
                              len(q), tolerance, out.ctypes.data_as(C.POINTER(C.c_int32)))
    return out


def hit_test(path, point, fill_rule, tolerance):
    """hit_test_path (hit_test.rs:8-18): fill_rule 0 EvenOdd, 1 NonZero."""
    w = int(winding_numbers(path, [point], tolerance)[0])
    return (w % 2 != 0) if fill_rule == 0 else (w != 0)


def signed_area(path, tolerance):
    """approximate_signed_area (area.rs:7-54)."""
    v = np.ascontiguousarray(path.verbs, np.uint8)
    p = np.ascontiguousarray(path.points, np.float32)
    return float(np.float32(lib().orc_signed_area(_ptr(v, U8P), len(v), _ptr(p, F32P), len(p),
                                                  tolerance)))


def parse_svg(text, num_attributes=0, stop_at=None):
    """PathParser::parse (parser.rs:154-175) of one string into a Path layout.  Returns
    (verbs uint8, points float32 (n, 2), error int32[4]: kind, line, column, char)."""
    b = text.encode() if isinstance(text, str) else bytes(text)
    nv, npt = C.c_uint32(0), C.c_uint32(0)
    err = (C.c_int32 * 4)()
    sa = -1 if stop_at is None else ord(stop_at)
    cap_v = cap_p = 0
    while True:
        verbs = np.zeros(max(cap_v, 1), np.uint8)
        pts = np.zeros((max(cap_p, 1), 2), np.float32)
        lib().orc_parse_svg(b, len(b), num_attributes, sa, _ptr(verbs, U8P), cap_v,
                            _ptr(pts, F32P), cap_p, C.byref(nv), C.byref(npt), err)
        if nv.value <= cap_v and npt.value <= cap_p:
            return verbs[:nv.value], pts[:npt.value], np.array(list(err), np.int32)
        cap_v, cap_p = nv.value, npt.value


def parse_f32(text):
    """str::parse::<f32>: the float, or None for text Rust rejects."""
    v = C.c_float(0)
    ok = lib().orc_parse_f32(text.encode(), C.byref(v))
    return float(np.float32(v.value)) if ok else None


def fill_circle(center, radius, tolerance):
    """basic_shapes::fill_circle (crates/tessellation/src/basic_shapes.rs:40-205) restated
    with numpy float32 arithmetic: (vertices (n, 2) float32, indices (3 m,) local ids)."""
    f = np.float32
    cx, cy = f(center[0]), f(center[1])
    r = f(abs(f(radius)))
    if r == 0:
        return np.zeros((0, 2), np.float32), np.zeros(0, np.uint32)
    verts = [(cx + f(-1.0) * r, cy + f(0.0) * r), (cx + f(0.0) * r, cy + f(-1.0) * r),
             (cx + f(1.0) * r, cy + f(0.0) * r), (cx + f(0.0) * r, cy + f(1.0) * r)]
    tris = [(0, 3, 1), (1, 3, 2)]
    pi = f(np.pi)
    angles = [(pi, f(1.5) * pi), (f(1.5) * pi, f(2.0) * pi), (f(0.0), pi * f(0.5)),
              (pi * f(0.5), pi)]
    arc_len = f(0.5) * pi * r
    tol = min(f(tolerance), r)
    step = f(2.0) * np.sqrt(f(2.0) * tol * r - tol * tol, dtype=np.float32)
    n_seg = np.ceil(arc_len / step, dtype=np.float32)
    lg = np.log2(n_seg, dtype=np.float32)
    n = int(lg) if lg > 0 else 0

    def border(a0, a1, va, vb, k):
        if k == 0:
            return
        mid = (a0 + a1) * f(0.5)
        v = len(verts)
        verts.append((cx + np.cos(mid, dtype=np.float32) * r, cy + np.sin(mid, dtype=np.float32) * r))
        tris.append((vb, v, va))
        border(a0, mid, va, v, k - 1)
        border(mid, a1, v, vb, k - 1)

    for i in range(4):
        border(angles[i][0], angles[i][1], i, (i + 1) % 4, n)
    return np.asarray(verts, np.float32).reshape(-1, 2), np.asarray(tris, np.uint32).reshape(-1)
