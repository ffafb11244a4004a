"""ctypes binding of liblyg.so (include/lyg.h).

The product's only compute path: if the shared library is missing, importing anything
that needs it raises — there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblyg.so")

F32P = C.POINTER(C.c_float)
U32P = C.POINTER(C.c_uint32)
U8P = C.POINTER(C.c_uint8)
I32P = C.POINTER(C.c_int32)
U64P = C.POINTER(C.c_uint64)


class StrokeOptionsC(C.Structure):
    _fields_ = [("line_width", C.c_float), ("miter_limit", C.c_float), ("tolerance", C.c_float),
                ("start_cap", C.c_uint8), ("end_cap", C.c_uint8), ("line_join", C.c_uint8),
                ("_pad", C.c_uint8)]


class FillOptionsC(C.Structure):
    _fields_ = [("tolerance", C.c_float), ("fill_rule", C.c_uint8),
                ("sweep_orientation", C.c_uint8), ("handle_intersections", C.c_uint8),
                ("_pad", C.c_uint8)]


class PathViewC(C.Structure):
    _fields_ = [("verbs", U8P), ("n_verbs", C.c_uint32), ("points", F32P),
                ("n_points", C.c_uint32)]


class OutputC(C.Structure):
    _fields_ = [("vertices", F32P), ("indices", U32P), ("vertex_capacity", C.c_uint32),
                ("index_capacity", C.c_uint32), ("vertex_offset", C.c_uint32),
                ("n_vertices", C.c_uint32), ("n_indices", C.c_uint32)]


class BatchInC(C.Structure):
    _fields_ = [("verbs", C.c_void_p), ("points", C.c_void_p), ("verb_offsets", C.c_void_p),
                ("point_offsets", C.c_void_p), ("n_paths", C.c_uint32)]


class BatchOutC(C.Structure):
    _fields_ = [("vertices", C.c_void_p), ("indices", C.c_void_p),
                ("vertex_capacity", C.c_uint64), ("index_capacity", C.c_uint64),
                ("path_vertex_offsets", C.c_void_p), ("path_index_offsets", C.c_void_p),
                ("vertex_offset", C.c_uint32), ("n_vertices", C.c_uint64),
                ("n_indices", C.c_uint64), ("path_status", C.c_void_p)]


class ServerConfigC(C.Structure):
    _fields_ = [("pool_bytes", C.c_uint64), ("max_batch_paths", C.c_uint32),
                ("max_inflight", C.c_uint32), ("linger_us", C.c_uint32), ("device", C.c_int32),
                ("stroke_sizing", C.c_uint32), ("fill_banded_min_edges", C.c_uint32),
                ("max_request_bytes", C.c_uint64), ("stroke_split_windows", C.c_uint32),
                ("max_fill_crossings", C.c_uint64)]


class ServerResultC(C.Structure):
    _fields_ = [("status", C.c_int32), ("vertices", C.c_void_p), ("indices", C.c_void_p),
                ("n_vertices", C.c_uint32), ("n_indices", C.c_uint32),
                ("latency_us", C.c_double), ("banded", C.c_int32)]


class ServerStatsC(C.Structure):
    _fields_ = [("completed", C.c_uint64), ("triangles", C.c_uint64), ("vertices", C.c_uint64),
                ("batches", C.c_uint64), ("pending", C.c_uint64),
                ("pool_high_water", C.c_uint64), ("mean_batch", C.c_double),
                ("p50_us", C.c_double), ("p99_us", C.c_double), ("max_us", C.c_double),
                ("elapsed_s", C.c_double), ("rejected", C.c_uint64)]


class StreamStatsC(C.Structure):
    _fields_ = [("requests", C.c_uint64), ("triangles", C.c_uint64), ("batches", C.c_uint64),
                ("pool_high_water", C.c_uint64), ("offered_rate", C.c_double),
                ("submit_rate", C.c_double), ("completed_rate", C.c_double),
                ("triangles_per_s", C.c_double), ("p50_us", C.c_double), ("p99_us", C.c_double),
                ("max_us", C.c_double), ("mean_batch", C.c_double), ("seconds", C.c_double),
                ("failed", C.c_uint64), ("rejected", C.c_uint64)]


class TuningC(C.Structure):
    _fields_ = [("blocks_per_cu", C.c_uint32), ("waves_per_block", C.c_uint32),
                ("lds_bytes_per_wave", C.c_uint32), ("best_ms", C.c_float),
                ("split_windows", C.c_uint32), ("edges_per_band", C.c_uint32)]


# Every symbol include/lyg.h declares, with its ctypes signature.
SIGNATURES = {
    "lyg_ctx_create": ([C.c_int], C.c_void_p),
    "lyg_ctx_destroy": ([C.c_void_p], None),
    "lyg_status_string": ([C.c_int32], C.c_char_p),
    "lyg_ctx_last_error": ([C.c_void_p], C.c_char_p),
    "lyg_ctx_enable_timing": ([C.c_void_p, C.c_int], C.c_int32),
    "lyg_ctx_kernel_times": ([C.c_void_p, F32P, F32P], C.c_int32),
    "lyg_device_alloc": ([C.c_void_p, C.c_uint64, C.POINTER(C.c_void_p)], C.c_int32),
    "lyg_device_free": ([C.c_void_p, C.c_void_p], C.c_int32),
    "lyg_copy_to_device": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64], C.c_int32),
    "lyg_copy_to_host": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64], C.c_int32),
    "lyg_ctx_synchronize": ([C.c_void_p], C.c_int32),
    "lyg_device_copy": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64], C.c_int32),
    "lyg_stroke_path": ([C.c_void_p, C.POINTER(PathViewC), C.POINTER(StrokeOptionsC),
                         C.POINTER(OutputC)], C.c_int32),
    "lyg_stroke_path_vw": ([C.c_void_p, C.POINTER(PathViewC), F32P, C.c_uint32, C.c_uint32,
                            C.POINTER(StrokeOptionsC), C.POINTER(OutputC), F32P], C.c_int32),
    "lyg_stroke_batch_vw": ([C.c_void_p, C.POINTER(BatchInC), C.c_void_p, C.c_uint32,
                             C.c_uint32, C.c_void_p, C.c_uint32, C.POINTER(BatchOutC),
                             C.c_void_p], C.c_int32),
    "lyg_fill_path": ([C.c_void_p, C.POINTER(PathViewC), C.POINTER(FillOptionsC),
                       C.POINTER(OutputC)], C.c_int32),
    "lyg_flatten_segments": ([C.c_void_p, I32P, F32P, C.c_uint32, C.c_float, F32P, C.c_uint32,
                              U32P, U32P], C.c_int32),
    "lyg_fill_rectangle": ([C.c_void_p, F32P, C.POINTER(OutputC)], C.c_int32),
    "lyg_parse_svg_paths": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32,
                             C.c_int32, C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint64,
                             C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_uint64),
                             C.POINTER(C.c_uint64), C.POINTER(C.c_uint32)], C.c_int32),
    "lyg_parse_svg_path": ([C.c_void_p, C.c_char_p, C.c_uint32, C.c_uint32, C.c_int32, U8P,
                            C.c_uint32, F32P, C.c_uint32, U32P, U32P, C.POINTER(C.c_int32)],
                           C.c_int32),
    "lyg_winding_numbers": ([C.c_void_p, C.POINTER(BatchInC), C.c_void_p, C.c_void_p, C.c_uint32,
                             C.c_float, C.c_uint8, C.c_void_p, C.c_void_p], C.c_int32),
    "lyg_signed_areas": ([C.c_void_p, C.POINTER(BatchInC), C.c_float, C.c_void_p], C.c_int32),
    "lyg_path_winding_numbers": ([C.c_void_p, C.POINTER(PathViewC), F32P, C.c_uint32, C.c_float,
                                  C.c_uint8, C.POINTER(C.c_int32), U8P], C.c_int32),
    "lyg_approximate_signed_area": ([C.c_void_p, C.POINTER(PathViewC), C.c_float,
                                     C.POINTER(C.c_float)], C.c_int32),
    "lyg_fill_circle": ([C.c_void_p, C.c_float, C.c_float, C.c_float, C.POINTER(FillOptionsC),
                         C.POINTER(OutputC)], C.c_int32),
    "lyg_flatten_segments_device": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_float,
                                     C.c_void_p, C.c_uint32, C.c_void_p, C.POINTER(C.c_uint32)],
                                    C.c_int32),
    "lyg_flatten_segments_device_noalloc": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32,
                                             C.c_float, C.c_void_p, C.c_uint32, C.c_void_p,
                                             C.POINTER(C.c_uint32)], C.c_int32),
    "lyg_flatten_noalloc_max_segments": ([], C.c_uint32),
    "lyg_flatten_arcs": ([C.c_void_p, F32P, C.c_uint32, C.c_float, F32P, C.c_uint32, U32P, U32P],
                         C.c_int32),
    "lyg_arc_to_quadratics": ([C.c_void_p, F32P, C.c_uint32, F32P, C.c_uint32, U32P, U32P],
                              C.c_int32),
    "lyg_stroke_batch": ([C.c_void_p, C.POINTER(BatchInC), C.c_void_p, C.c_uint32,
                          C.POINTER(BatchOutC)], C.c_int32),
    "lyg_fill_batch": ([C.c_void_p, C.POINTER(BatchInC), C.c_void_p, C.c_uint32,
                        C.POINTER(BatchOutC)], C.c_int32),
    "lyg_fill_batch_banded": ([C.c_void_p, C.POINTER(BatchInC), C.c_void_p, C.c_uint32,
                               C.POINTER(BatchOutC)], C.c_int32),
    "lyg_fill_path_banded": ([C.c_void_p, C.POINTER(PathViewC), C.POINTER(FillOptionsC),
                              C.POINTER(OutputC)], C.c_int32),
    "lyg_ctx_set_band_size": ([C.c_void_p, C.c_uint32, C.c_uint32], C.c_int32),
    "lyg_flattening_errors": ([C.c_void_p, C.POINTER(BatchInC), C.c_void_p, C.c_uint32, C.c_int,
                               C.c_void_p], C.c_int32),
    "lyg_gen_random_walk": ([C.c_uint64, C.c_uint32, C.c_float, U8P, F32P], C.c_int32),
    "lyg_stroke_batch_count": ([C.c_void_p, C.POINTER(BatchInC), C.c_void_p, C.c_uint32,
                                C.c_void_p, C.c_void_p, U64P, U64P], C.c_int32),
    "lyg_gen_stroke_batch": ([C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_float, C.c_int,
                              U8P, F32P, U32P, U32P, U64P, U64P], C.c_int32),
    "lyg_stroke_batch_graph": ([C.c_void_p, C.POINTER(BatchInC), C.c_void_p, C.c_uint32,
                                C.POINTER(BatchOutC), C.POINTER(C.c_void_p)], C.c_int32),
    "lyg_fill_batch_graph": ([C.c_void_p, C.POINTER(BatchInC), C.c_void_p, C.c_uint32,
                              C.POINTER(BatchOutC), C.POINTER(C.c_void_p)], C.c_int32),
    "lyg_graph_launch": ([C.c_void_p], C.c_int32),
    "lyg_graph_result": ([C.c_void_p, U64P, U64P], C.c_int32),
    "lyg_graph_destroy": ([C.c_void_p], None),
    "lyg_ctx_set_tuning": ([C.c_void_p, C.POINTER(TuningC), C.POINTER(TuningC)], C.c_int32),
    "lyg_ctx_set_stroke_staging": ([C.c_void_p, C.c_int, C.c_float], C.c_int32),
    "lyg_ctx_set_stroke_split": ([C.c_void_p, C.c_uint32], C.c_int32),
    "lyg_ctx_set_fill_lds_resident": ([C.c_void_p, C.c_int], C.c_int32),
    "lyg_ctx_stroke_split_stats": ([C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint32)],
                                   C.c_int32),
    "lyg_autotune_stroke": ([C.c_void_p, C.POINTER(BatchInC), C.c_void_p, C.c_uint32,
                             C.POINTER(BatchOutC), C.POINTER(TuningC)], C.c_int32),
    "lyg_autotune_fill": ([C.c_void_p, C.POINTER(BatchInC), C.c_void_p, C.c_uint32,
                           C.POINTER(BatchOutC), C.POINTER(TuningC)], C.c_int32),
    "lyg_autotune_banded": ([C.c_void_p, C.POINTER(BatchInC), C.c_void_p, C.c_uint32,
                             C.POINTER(BatchOutC), C.POINTER(TuningC)], C.c_int32),
    "lyg_server_create": ([C.POINTER(ServerConfigC)], C.c_void_p),
    "lyg_server_destroy": ([C.c_void_p], None),
    "lyg_server_submit_fill": ([C.c_void_p, C.POINTER(PathViewC), C.POINTER(FillOptionsC),
                                C.POINTER(C.c_uint64)], C.c_int32),
    "lyg_server_submit_stroke": ([C.c_void_p, C.POINTER(PathViewC), C.POINTER(StrokeOptionsC),
                                  C.POINTER(C.c_uint64)], C.c_int32),
    "lyg_server_poll": ([C.c_void_p, C.c_uint64, C.POINTER(ServerResultC)], C.c_int32),
    "lyg_server_wait": ([C.c_void_p, C.c_uint64, C.POINTER(ServerResultC)], C.c_int32),
    "lyg_server_copy_result": ([C.c_void_p, C.c_uint64, F32P, U32P], C.c_int32),
    "lyg_server_release": ([C.c_void_p, C.c_uint64], C.c_int32),
    "lyg_server_get_stats": ([C.c_void_p, C.POINTER(ServerStatsC)], C.c_int32),
    "lyg_server_run_stream": ([C.c_void_p, C.c_uint64, C.c_uint32, C.c_double, C.c_float,
                               C.c_uint32, C.c_uint32, C.c_float, C.c_uint32,
                               C.POINTER(StreamStatsC)], C.c_int32),
    "lyg_server_run_stream_mixed": ([C.c_void_p, C.c_uint64, C.c_uint32, C.c_double, C.c_float,
                                     C.c_uint32, C.c_uint32, C.c_float, C.c_uint32, C.c_int,
                                     C.POINTER(StreamStatsC)], C.c_int32),
    "lyg_gen_fill_batch": ([C.c_uint64, C.c_uint32, U8P, F32P, U32P, U32P, U64P, U64P],
                           C.c_int32),
    "lyg_gen_fill_options": ([C.c_uint64, C.c_uint32, C.c_float, C.POINTER(FillOptionsC)],
                             C.c_int32),
    "lyg_gen_stroke_options": ([C.c_uint64, C.c_uint32, C.c_float, C.c_float, C.c_float,
                                C.POINTER(StrokeOptionsC)], C.c_int32),
}

_lib = None


def lib():
    """Load liblyg.so once; raises RuntimeError when it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with "
                               "`python -c 'import __graft_entry__ as g; g.build()'`")
        L = C.CDLL(LIB_PATH)
        for name, (args, res) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def declared_symbols(header_path: str | None = None):
    """Names of the functions include/lyg.h declares (parsed from the header text)."""
    import re
    if header_path is None:
        header_path = os.path.join(os.path.dirname(_HERE), "include", "lyg.h")
    text = open(header_path).read()
    return sorted(set(re.findall(r"\b(lyg_[a-z0-9_]+)\s*\(", text)))
