"""Batched form: many paths per launch, device-resident inputs and outputs.

Device memory comes from PyTorch (plumbing only: the kernels are liblyg.so's).  The
result of ``stroke_batch`` equals calling ``StrokeTessellator.tessellate_path`` for every
path in order into one shared ``VertexBuffers`` (geometry_builder.rs:300-351): path i's
vertices start at ``vertex_offsets[i]`` and its indices are absolute ids.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as B
from .path import PathBatch
from .tessellation import TessellationError, _Ctx


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("the batched form needs a HIP device")
    return torch


class DeviceBatch:
    """A PathBatch copied to device memory once (inputs stay resident in HBM)."""

    def __init__(self, batch: PathBatch, stroke_opts=None, device="cuda:0"):
        torch = _torch()
        self.n = len(batch)
        self.verbs = torch.from_numpy(batch.verbs).to(device)
        self.points = torch.from_numpy(batch.points.reshape(-1)).to(device)
        self.verb_offsets = torch.from_numpy(batch.verb_offsets.view(np.int32)).to(device)
        self.point_offsets = torch.from_numpy(batch.point_offsets.view(np.int32)).to(device)
        self.opts = None
        self.n_opts = 0
        if stroke_opts is not None:
            raw = np.ascontiguousarray(stroke_opts).view(np.uint8).reshape(-1)
            self.opts = torch.from_numpy(raw.copy()).to(device)
            self.n_opts = len(stroke_opts)

    def c_in(self) -> B.BatchInC:
        return B.BatchInC(self.verbs.data_ptr(), self.points.data_ptr(),
                          self.verb_offsets.data_ptr(), self.point_offsets.data_ptr(), self.n)


class StrokeBatchResult:
    def __init__(self, vertices, indices, vertex_offsets, index_offsets, n_vertices, n_indices):
        self.vertices = vertices            # device tensor (capacity, 2) float32
        self.indices = indices              # device tensor (capacity,) int32 view of u32
        self.vertex_offsets = vertex_offsets  # device (n+1,) int32 view of u32
        self.index_offsets = index_offsets
        self.n_vertices = n_vertices
        self.n_indices = n_indices

    def path(self, i: int):
        """(vertices, indices) of path i as numpy arrays (indices absolute, u32)."""
        vo = self.vertex_offsets.cpu().numpy().view(np.uint32)
        io = self.index_offsets.cpu().numpy().view(np.uint32)
        v = self.vertices[int(vo[i]):int(vo[i + 1])].cpu().numpy()
        idx = self.indices[int(io[i]):int(io[i + 1])].cpu().numpy().view(np.uint32)
        return v, idx


class StrokeBatcher:
    """Reusable output arena for repeated batched strokes (no allocation per call once
    the arena is large enough)."""

    def __init__(self, device="cuda:0"):
        self.device = device
        self.vertices = None
        self.indices = None
        self.voffs = None
        self.ioffs = None

    def _ensure(self, n_paths, nv, ni):
        torch = _torch()
        if self.voffs is None or self.voffs.numel() < n_paths + 1:
            self.voffs = torch.empty(n_paths + 1, dtype=torch.int32, device=self.device)
            self.ioffs = torch.empty(n_paths + 1, dtype=torch.int32, device=self.device)
        if self.vertices is None or self.vertices.shape[0] < nv:
            self.vertices = torch.empty((max(nv, 1), 2), dtype=torch.float32, device=self.device)
        if self.indices is None or self.indices.shape[0] < ni:
            self.indices = torch.empty(max(ni, 1), dtype=torch.int32, device=self.device)

    def count(self, db: DeviceBatch):
        torch = _torch()
        nv = torch.empty(db.n, dtype=torch.int32, device=self.device)
        ni = torch.empty(db.n, dtype=torch.int32, device=self.device)
        tv = C.c_uint64(0)
        ti = C.c_uint64(0)
        cin = db.c_in()
        st = B.lib().lyg_stroke_batch_count(_Ctx.get(), C.byref(cin), db.opts.data_ptr(),
                                            db.n_opts, nv.data_ptr(), ni.data_ptr(),
                                            C.byref(tv), C.byref(ti))
        if st != 0:
            raise TessellationError(st)
        return tv.value, ti.value

    def run(self, db: DeviceBatch, vertex_offset: int = 0) -> StrokeBatchResult:
        if db.opts is None:
            raise ValueError("DeviceBatch needs stroke options")
        self._ensure(db.n, 0, 0)
        cin = db.c_in()
        for _ in range(2):
            out = B.BatchOutC(self.vertices.data_ptr(), self.indices.data_ptr(),
                              self.vertices.shape[0], self.indices.shape[0],
                              self.voffs.data_ptr(), self.ioffs.data_ptr(), vertex_offset, 0, 0)
            st = B.lib().lyg_stroke_batch(_Ctx.get(), C.byref(cin), db.opts.data_ptr(),
                                          db.n_opts, C.byref(out))
            if st == 16:
                self._ensure(db.n, out.n_vertices, out.n_indices)
                continue
            if st != 0:
                raise TessellationError(st, B.lib().lyg_ctx_last_error(_Ctx.get()).decode())
            return StrokeBatchResult(self.vertices, self.indices, self.voffs[:db.n + 1],
                                     self.ioffs[:db.n + 1], out.n_vertices, out.n_indices)
        raise RuntimeError("stroke batch did not converge")
