"""lyon_amd — MI355X (gfx950) tessellation engine for lyon's hot path.

Drop-in mirror of lyon_tessellation's API (FillTessellator / StrokeTessellator /
FillOptions / StrokeOptions / VertexBuffers / BuffersBuilder) and of lyon_geom's
flattening, computed by the HIP kernels in liblyg.so.  See include/lyg.h for the C ABI.
"""
from .path import (Path, PathBatch, PathBuffer, PathBuilder, PathEvent, polygon, VERB_BEGIN, VERB_CLOSE, VERB_CUBIC,
                   VERB_END, VERB_LINE, VERB_QUADRATIC)
from .tessellation import (BuffersBuilder, FillOptions, FillRule, FillTessellator, LineCap,
                           LineJoin, Orientation, StrokeOptions, StrokeTessellator,
                           TessellationError, VertexBuffers, arc_to_quadratics, flatten_arcs,
                           flatten_segments, simple_builder)

__all__ = [
    "Path", "PathBatch", "PathBuffer", "PathBuilder", "PathEvent", "polygon", "VERB_BEGIN", "VERB_CLOSE", "VERB_CUBIC",
    "VERB_END", "VERB_LINE", "VERB_QUADRATIC", "BuffersBuilder", "FillOptions", "FillRule",
    "FillTessellator", "LineCap", "LineJoin", "Orientation", "StrokeOptions",
    "StrokeTessellator", "TessellationError", "VertexBuffers", "arc_to_quadratics",
    "flatten_arcs", "flatten_segments", "simple_builder",
]
