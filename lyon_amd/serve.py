"""Serving loop: continuous batching of single-path tessellation requests (liblyg.so's
lyg_server_*).  No lyon counterpart: each request is one ``tessellate_path`` call whose
result (vertex ids local to the request) stays in a device pool until released.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as B
from .path import Path
from .tessellation import FillOptions, StrokeOptions, TessellationError

LYG_PENDING = 21


class Server:
    def __init__(self, pool_bytes=1 << 30, max_batch_paths=1024, max_inflight=4, linger_us=0,
                 device=0, stroke_sizing=0, fill_banded_min_edges=0, max_request_bytes=0,
                 stroke_split_windows=0, max_fill_crossings=0):
        cfg = B.ServerConfigC(pool_bytes, max_batch_paths, max_inflight, linger_us, device,
                              stroke_sizing, fill_banded_min_edges, max_request_bytes,
                              stroke_split_windows, max_fill_crossings)
        self._h = B.lib().lyg_server_create(C.byref(cfg))
        if not self._h:
            raise RuntimeError("lyg_server_create failed (no usable HIP device or pool)")

    def close(self):
        if self._h:
            B.lib().lyg_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _view(path: Path):
        v = path.verbs
        p = path.points
        return B.PathViewC(v.ctypes.data_as(B.U8P), len(v), p.ctypes.data_as(B.F32P), len(p))

    def submit_fill(self, path: Path, options: FillOptions) -> int:
        t = C.c_uint64(0)
        st = B.lib().lyg_server_submit_fill(self._h, C.byref(self._view(path)),
                                            C.byref(options.to_c()), C.byref(t))
        if st != 0:
            raise TessellationError(st)
        return t.value

    def submit_stroke(self, path: Path, options: StrokeOptions) -> int:
        t = C.c_uint64(0)
        st = B.lib().lyg_server_submit_stroke(self._h, C.byref(self._view(path)),
                                              C.byref(options.to_c()), C.byref(t))
        if st != 0:
            raise TessellationError(st)
        return t.value

    def wait(self, ticket: int) -> B.ServerResultC:
        r = B.ServerResultC()
        st = B.lib().lyg_server_wait(self._h, ticket, C.byref(r))
        if st != 0:
            raise TessellationError(st)
        return r

    def done(self, ticket: int) -> bool:
        r = B.ServerResultC()
        return B.lib().lyg_server_poll(self._h, ticket, C.byref(r)) != LYG_PENDING

    def banded(self, ticket: int) -> bool:
        """Whether the request ran through the banded fill (its triangles are not lyon's)."""
        return bool(self.wait(ticket).banded)

    def result(self, ticket: int):
        """(status, vertices (n, 2) f32, indices u32) on the host; waits if needed."""
        r = self.wait(ticket)
        if r.status != 0:
            return r.status, None, None
        v = np.empty((r.n_vertices, 2), np.float32)
        i = np.empty(r.n_indices, np.uint32)
        st = B.lib().lyg_server_copy_result(self._h, ticket, v.ctypes.data_as(B.F32P),
                                            i.ctypes.data_as(B.U32P))
        if st != 0:
            raise TessellationError(st)
        return 0, v, i

    def release(self, ticket: int):
        st = B.lib().lyg_server_release(self._h, ticket)
        if st != 0:
            raise TessellationError(st)

    def stats(self) -> dict:
        s = B.ServerStatsC()
        B.lib().lyg_server_get_stats(self._h, C.byref(s))
        return {k: getattr(s, k) for k, _ in s._fields_}

    def run_stream(self, n_requests, rate_per_s, stroke_frac=0.6, min_segs=4, max_segs=65536,
                   tolerance=0.05, n_distinct=256, seed=0, fill_kind=0) -> dict:
        """Poisson request stream (lyg_server_run_stream_mixed).  fill_kind 0: the batched
        fill workload's shapes; 1: closed random paths from the strokes' generator (the
        same segment counts, kinds and coordinates); 2: closed random walks with the
        strokes' segment-count distribution."""
        s = B.StreamStatsC()
        st = B.lib().lyg_server_run_stream_mixed(self._h, seed, n_requests, rate_per_s,
                                                 stroke_frac, min_segs, max_segs, tolerance,
                                                 n_distinct, int(fill_kind), C.byref(s))
        if st != 0:
            raise TessellationError(st)
        return {k: getattr(s, k) for k, _ in s._fields_}
