"""lyon_algorithms' path validators computed by liblyg.so's HIP kernels.

Mirrors ``hit_test_path`` / ``path_winding_number_at_position``
(crates/algorithms/src/hit_test.rs:8-83) and ``approximate_signed_area``
(crates/algorithms/src/area.rs:7-19) for single paths, plus batched forms over a
:class:`~lyon_amd.batch.DeviceBatch` with results left in device memory.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as B
from .path import Path
from .tessellation import FillRule, TessellationError, _Ctx, _path_view


def _rule(fill_rule) -> int:
    r = int(fill_rule)
    if r not in (0, 1):
        raise ValueError("fill_rule must be FillRule.EvenOdd (0) or FillRule.NonZero (1)")
    return r


def _raise(st: int):
    raise TessellationError(st, B.lib().lyg_ctx_last_error(_Ctx.get()).decode())


def path_winding_numbers(points, path: Path, tolerance: float, fill_rule=FillRule.EvenOdd):
    """Winding numbers (and hit_test_path answers under ``fill_rule``) of many points
    against one path: returns (winding int32 (k,), inside bool (k,))."""
    xy = np.ascontiguousarray(points, np.float32).reshape(-1, 2)
    w = np.zeros(len(xy), np.int32)
    ins = np.zeros(len(xy), np.uint8)
    pv = _path_view(path)
    st = B.lib().lyg_path_winding_numbers(_Ctx.get(), C.byref(pv), xy.ctypes.data_as(B.F32P),
                                          len(xy), float(tolerance), _rule(fill_rule),
                                          w.ctypes.data_as(C.POINTER(C.c_int32)),
                                          ins.ctypes.data_as(B.U8P))
    if st != 0:
        _raise(st)
    return w, ins.astype(bool)


def path_winding_number_at_position(point, path: Path, tolerance: float) -> int:
    """path_winding_number_at_position (hit_test.rs:21-83)."""
    return int(path_winding_numbers([point], path, tolerance)[0][0])


def hit_test_path(point, path: Path, fill_rule, tolerance: float) -> bool:
    """hit_test_path (hit_test.rs:8-18)."""
    return bool(path_winding_numbers([point], path, tolerance, fill_rule)[1][0])


def approximate_signed_area(tolerance: float, path: Path) -> float:
    """approximate_signed_area (area.rs:7-19): the sum of the sub-paths' signed areas of
    the flattened path (f32, lyon's summation order)."""
    a = C.c_float(0.0)
    pv = _path_view(path)
    st = B.lib().lyg_approximate_signed_area(_Ctx.get(), C.byref(pv), float(tolerance),
                                             C.byref(a))
    if st != 0:
        _raise(st)
    return float(np.float32(a.value))


def winding_numbers_device(db, queries, query_paths, n: int, tolerance: float,
                           fill_rule=FillRule.EvenOdd, winding=None, inside=None):
    """Batched winding numbers: ``queries`` (device, n x, y pairs) against paths
    ``query_paths`` (device, n u32) of the DeviceBatch ``db``; ``winding`` (device, n i32)
    and optional ``inside`` (device, n u8) receive the results."""
    st = B.lib().lyg_winding_numbers(_Ctx.get(), C.byref(db.c_in()), queries.ptr,
                                     query_paths.ptr, int(n), float(tolerance),
                                     _rule(fill_rule), winding.ptr,
                                     inside.ptr if inside is not None else None)
    if st != 0:
        _raise(st)


def signed_areas_device(db, tolerance: float, areas):
    """approximate_signed_area of every path of ``db`` into ``areas`` (device, n f32)."""
    st = B.lib().lyg_signed_areas(_Ctx.get(), C.byref(db.c_in()), float(tolerance), areas.ptr)
    if st != 0:
        _raise(st)
