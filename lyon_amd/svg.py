"""SVG path data ingest computed by liblyg.so's HIP kernels.

Mirrors lyon_extra's ``PathParser`` / ``ParserOptions`` / ``ParseError``
(crates/extra/src/parser.rs:13-175): :func:`parse_path` parses one string into a
:class:`~lyon_amd.path.Path`, :func:`parse_paths` parses many strings at once on the device
into a :class:`~lyon_amd.path.PathBatch` (lyon_path's ``PathBuffer`` layout,
path_buffer.rs:14-104), and :func:`parse_paths_device` leaves the batch in device memory.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as B
from .path import Path, PathBatch
from .tessellation import TessellationError, _Ctx

ERR_NONE, ERR_NUMBER, ERR_FLAG, ERR_COMMAND, ERR_MISSING_MOVE_TO = 0, 1, 2, 3, 4
ERR_NON_ASCII, ERR_UNDECIDED_NUMBER = 5, 6
_KIND = {ERR_NUMBER: "Number", ERR_FLAG: "Flag", ERR_COMMAND: "Command",
         ERR_MISSING_MOVE_TO: "MissingMoveTo", ERR_NON_ASCII: "NonAscii",
         ERR_UNDECIDED_NUMBER: "UndecidedNumber"}


class ParserOptions:
    """ParserOptions (parser.rs:35-49)."""

    def __init__(self, num_attributes: int = 0, stop_at=None):
        self.num_attributes = int(num_attributes)
        self.stop_at = stop_at

    def _stop(self) -> int:
        if self.stop_at is None:
            return -1
        c = ord(self.stop_at)
        if c > 127:
            raise ValueError("stop_at must be an ASCII character")
        return c


class ParseError(Exception):
    """ParseError (parser.rs:13-33): kind, line, column and the character (Flag, Command,
    MissingMoveTo)."""

    def __init__(self, kind: int, line: int, column: int, char: int):
        self.kind, self.line, self.column = int(kind), int(line), int(column)
        self.char = chr(char) if char else None
        super().__init__(f"Line {line} Column {column}: {_KIND.get(kind, kind)}"
                         + (f" {self.char!r}" if self.char else ""))


def _error(rec):
    return None if int(rec[0]) == 0 else ParseError(*[int(v) for v in rec])


def _raise(st):
    raise TessellationError(st, B.lib().lyg_ctx_last_error(_Ctx.get()).decode())


def parse_path(text, options: ParserOptions = ParserOptions()):
    """PathParser::parse of one string: (Path, ParseError or None).  As with lyon's
    builder, the path holds what was parsed before an error."""
    b = text.encode() if isinstance(text, str) else bytes(text)
    nv, npt = C.c_uint32(0), C.c_uint32(0)
    err = (C.c_int32 * 4)()
    cap_v = cap_p = 0
    while True:
        verbs = np.zeros(max(cap_v, 1), np.uint8)
        pts = np.zeros((max(cap_p, 1), 2), np.float32)
        st = B.lib().lyg_parse_svg_path(_Ctx.get(), b, len(b), options.num_attributes,
                                        options._stop(), verbs.ctypes.data_as(B.U8P), cap_v,
                                        pts.ctypes.data_as(B.F32P), cap_p, C.byref(nv),
                                        C.byref(npt), err)
        if st == 16:
            cap_v, cap_p = nv.value, npt.value
            continue
        if st != 0:
            _raise(st)
        return Path(verbs[:nv.value], pts[:npt.value]), _error(err)


def pack_strings(strings):
    """Strings as one byte array and n+1 u32 offsets (the text layout of parse_paths_device)."""
    bs = [s.encode() if isinstance(s, str) else bytes(s) for s in strings]
    offs = np.zeros(len(bs) + 1, np.uint32)
    offs[1:] = np.cumsum([len(b) for b in bs], dtype=np.uint64)
    text = np.frombuffer(b"".join(bs), np.uint8) if bs else np.zeros(0, np.uint8)
    return text, offs


class DeviceSvgBatch:
    """Parsed paths left in device memory: verbs, points, verb_offsets, point_offsets
    (DeviceBuffers, the DeviceBatch layout), errors (n x 4 int32) and the totals.  With
    options attached (:meth:`with_options`) it feeds StrokeBatcher / FillBatcher directly:
    text in, triangles out, nothing on the host in between."""

    opts = None
    n_opts = 0

    def with_options(self, opts) -> "DeviceSvgBatch":
        """Attaches stroke or fill option records: one per path, or one for all."""
        from .batch import DeviceBuffer
        raw = np.ascontiguousarray(opts)
        self.opts = DeviceBuffer.from_numpy(raw.view(np.uint8).reshape(-1))
        self.n_opts = len(raw)
        return self

    def c_in(self) -> B.BatchInC:
        return B.BatchInC(self.verbs.ptr, self.points.ptr, self.verb_offsets.ptr,
                          self.point_offsets.ptr, self.n)


def parse_paths_device(text_dev, offsets_dev, n: int, options: ParserOptions = ParserOptions(),
                       out: "DeviceSvgBatch | None" = None):
    """Parses n strings held in device memory (text bytes and n+1 u32 offsets) into a
    DeviceSvgBatch; `out`, a previous result, is refilled when its buffers are large
    enough (one parse call then), else new buffers are sized by a first call."""
    from .batch import DeviceBuffer
    L = B.lib()
    ctx = _Ctx.get()
    tv, tp, ne = C.c_uint64(0), C.c_uint64(0), C.c_uint32(0)
    r = out if out is not None and out.n == n else None
    if r is not None:
        st = L.lyg_parse_svg_paths(ctx, text_dev.ptr, offsets_dev.ptr, n, options.num_attributes,
                                   options._stop(), r.verbs.ptr, r.verbs.nbytes, r.points.ptr,
                                   r.points.nbytes // 8, r.verb_offsets.ptr, r.point_offsets.ptr,
                                   r.errors.ptr, C.byref(tv), C.byref(tp), C.byref(ne))
        if st == 0:
            r.n_verbs, r.n_points, r.n_errors = tv.value, tp.value, ne.value
            return r
        if st != 16:
            _raise(st)
    r = DeviceSvgBatch()
    r.n = n
    r.verb_offsets = DeviceBuffer(4 * (n + 1))
    r.point_offsets = DeviceBuffer(4 * (n + 1))
    r.errors = DeviceBuffer(16 * max(n, 1))
    st = L.lyg_parse_svg_paths(ctx, text_dev.ptr, offsets_dev.ptr, n, options.num_attributes,
                               options._stop(), None, 0, None, 0, r.verb_offsets.ptr,
                               r.point_offsets.ptr, r.errors.ptr, C.byref(tv), C.byref(tp),
                               C.byref(ne))
    if st not in (0, 16):
        _raise(st)
    r.verbs = DeviceBuffer(max(tv.value, 1))
    r.points = DeviceBuffer(8 * max(tp.value, 1))
    st = L.lyg_parse_svg_paths(ctx, text_dev.ptr, offsets_dev.ptr, n, options.num_attributes,
                               options._stop(), r.verbs.ptr, tv.value, r.points.ptr, tp.value,
                               r.verb_offsets.ptr, r.point_offsets.ptr, r.errors.ptr,
                               C.byref(tv), C.byref(tp), C.byref(ne))
    if st != 0:
        _raise(st)
    r.n_verbs, r.n_points, r.n_errors = tv.value, tp.value, ne.value
    return r


def parse_paths(strings, options: ParserOptions = ParserOptions()):
    """Parses every string on the device: (PathBatch, [ParseError or None per string])."""
    from .batch import DeviceBuffer
    text, offs = pack_strings(strings)
    n = len(offs) - 1
    td = DeviceBuffer.from_numpy(text if len(text) else np.zeros(1, np.uint8))
    od = DeviceBuffer.from_numpy(offs)
    r = parse_paths_device(td, od, n, options)
    verbs = r.verbs.download(np.uint8, r.n_verbs)
    pts = r.points.download(np.float32, 2 * r.n_points).reshape(-1, 2)
    vo = r.verb_offsets.download(np.uint32, n + 1)
    po = r.point_offsets.download(np.uint32, n + 1)
    errs = r.errors.download(np.int32, 4 * n).reshape(-1, 4)
    return PathBatch(verbs, pts, vo, po), [_error(e) for e in errs]
