/*
 * lyg — MI355X (gfx950) tessellation engine for lyon's hot path.
 *
 * C ABI of liblyg.so.  Plain pointers and sizes only.  Each entry point names the
 * lyon interface it replaces (paths relative to the lyon source tree).
 *
 * Conventions shared by every call:
 *   - A path is lyon's flat storage (crates/path/src/path.rs:75-79): verbs[] (lyg_verb,
 *     order of path.rs:29-36) and points[] as x,y f32 pairs; a CLOSE verb carries the
 *     stored copy of the sub-path's first point (path.rs:537-542).
 *   - Output follows BuffersBuilder<[f32;2], u32, Positions>
 *     (crates/tessellation/src/geometry_builder.rs:300-351, 451-508): vertices are
 *     appended, and indices are absolute ids = vertex_offset + position in the vertex
 *     array of this call.  u32::MAX is reserved (VertexId::INVALID).
 *   - On error nothing is written (abort_geometry, geometry_builder.rs:478-481).
 *   - Status codes map one to one to TessellationError (crates/tessellation/src/error.rs).
 */
#ifndef LYG_H
#define LYG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Verb codes: lyon_path::path::Verb (path.rs:29-36). */
enum lyg_verb {
    LYG_VERB_LINE = 0,
    LYG_VERB_QUADRATIC = 1,
    LYG_VERB_CUBIC = 2,
    LYG_VERB_BEGIN = 3,
    LYG_VERB_CLOSE = 4,
    LYG_VERB_END = 5
};

/* lyon_path::LineCap / LineJoin / FillRule (crates/path/src/lib.rs:170-206, 262-280);
 * lyon_tessellation::Orientation (crates/tessellation/src/lib.rs:302-308). */
enum lyg_line_cap { LYG_CAP_BUTT = 0, LYG_CAP_SQUARE = 1, LYG_CAP_ROUND = 2 };
enum lyg_line_join {
    LYG_JOIN_MITER = 0,
    LYG_JOIN_MITER_CLIP = 1,
    LYG_JOIN_ROUND = 2,
    LYG_JOIN_BEVEL = 3
};
enum lyg_fill_rule { LYG_FILL_EVEN_ODD = 0, LYG_FILL_NON_ZERO = 1 };
enum lyg_orientation { LYG_ORIENT_HORIZONTAL = 0, LYG_ORIENT_VERTICAL = 1 };

/* TessellationError (error.rs:5-125) plus the runtime's own failures. */
enum lyg_status {
    LYG_OK = 0,
    LYG_ERR_TOO_MANY_VERTICES = 1,  /* GeometryBuilderError::TooManyVertices */
    LYG_ERR_INVALID_VERTEX = 2,     /* GeometryBuilderError::InvalidVertex */
    LYG_ERR_POSITION_IS_NAN = 3,    /* UnsupportedParamater::PositionIsNaN */
    LYG_ERR_TOLERANCE_IS_NAN = 4,   /* UnsupportedParamater::ToleranceIsNaN */
    LYG_ERR_INTERNAL = 5,           /* InternalError::* */
    LYG_ERR_CAPACITY = 16,          /* output arrays too small: required sizes were set */
    LYG_ERR_INVALID_ARGUMENT = 17,
    LYG_ERR_DEVICE = 18,            /* HIP runtime failure */
    LYG_ERR_UNSUPPORTED = 19,       /* option combination not implemented */
    LYG_ERR_NO_SPACE = 20,          /* device memory (working memory or pool) exhausted */
    LYG_PENDING = 21                /* serving loop: the request is not finished */
};

/* StrokeOptions (crates/tessellation/src/lib.rs:311-433); defaults: width 1, miter limit 4,
 * tolerance 0.1, Butt caps, Miter join.  variable_line_width is not supported. */
typedef struct lyg_stroke_options {
    float line_width;
    float miter_limit;
    float tolerance;
    uint8_t start_cap;
    uint8_t end_cap;
    uint8_t line_join;
    uint8_t _pad;
} lyg_stroke_options;

/* FillOptions (lib.rs:445-531); defaults: tolerance 0.1, EvenOdd, Vertical, intersections on. */
typedef struct lyg_fill_options {
    float tolerance;
    uint8_t fill_rule;
    uint8_t sweep_orientation;
    uint8_t handle_intersections;
    uint8_t _pad;
} lyg_fill_options;

/* A borrowed path: lyon_path::PathSlice (path.rs:82-87) without attributes. */
typedef struct lyg_path_view {
    const uint8_t* verbs;
    uint32_t n_verbs;
    const float* points; /* x, y pairs */
    uint32_t n_points;
} lyg_path_view;

/* VertexBuffers<[f32;2], u32> (geometry_builder.rs:262-286) seen through a BuffersBuilder. */
typedef struct lyg_output {
    float* vertices;          /* x, y pairs, capacity vertex_capacity */
    uint32_t* indices;        /* capacity index_capacity */
    uint32_t vertex_capacity;
    uint32_t index_capacity;
    uint32_t vertex_offset;   /* BuffersBuilder::with_vertex_offset */
    uint32_t n_vertices;      /* out: written (or required on LYG_ERR_CAPACITY) */
    uint32_t n_indices;       /* out */
} lyg_output;

typedef struct lyg_ctx lyg_ctx;

/* ---- context -------------------------------------------------------------------------
 * One context per host thread: owns a HIP stream, a device memory pool and launch
 * parameters.  Not thread-safe; use one context per thread.  (The reference's analogue is
 * a FillTessellator/StrokeTessellator instance, fill.rs:523-539 / stroke.rs:108-111.) */
lyg_ctx* lyg_ctx_create(int device);
void lyg_ctx_destroy(lyg_ctx* ctx);
const char* lyg_status_string(int32_t status);
/* Last HIP error text recorded by the context (empty if none). */
const char* lyg_ctx_last_error(const lyg_ctx* ctx);
/* Measurement support (no lyon counterpart): when enabled, HIP events are recorded on the
 * context's stream around two phases of every call; lyg_ctx_kernel_times returns their
 * durations for the last call in milliseconds: for a batched stroke, the sizing phase
 * (first: the staging-room kernel and its scans, or the count pass when the batch runs
 * as count + emit) and the emit pass (second); for a fill, the compaction copy (first)
 * and the first sweep round (second). */
int32_t lyg_ctx_enable_timing(lyg_ctx* ctx, int enable);
int32_t lyg_ctx_kernel_times(lyg_ctx* ctx, float* count_ms, float* emit_ms);

/* Device memory for the batched form (no lyon counterpart: lyon's buffers are host Vecs).
 * Allocations belong to the context's device; copies run on the context's stream and are
 * complete when the call returns. */
int32_t lyg_device_alloc(lyg_ctx* ctx, uint64_t bytes, void** ptr);
int32_t lyg_device_free(lyg_ctx* ctx, void* ptr);
int32_t lyg_copy_to_device(lyg_ctx* ctx, void* dst, const void* src, uint64_t bytes);
int32_t lyg_copy_to_host(lyg_ctx* ctx, void* dst, const void* src, uint64_t bytes);
/* Device-to-device copy (16-byte aligned pointers) on the context's stream, asynchronous;
 * a plain streaming kernel (also the reference kernel for measuring HBM traffic). */
int32_t lyg_device_copy(lyg_ctx* ctx, void* dst, const void* src, uint64_t bytes);
/* Waits for all work queued on the context's stream. */
int32_t lyg_ctx_synchronize(lyg_ctx* ctx);

/* ---- plain per-path drop-in (host buffers in and out) --------------------------------
 * StrokeTessellator::tessellate_path (crates/tessellation/src/stroke.rs:164-177). */
int32_t lyg_stroke_path(lyg_ctx* ctx, const lyg_path_view* path, const lyg_stroke_options* opts,
                        lyg_output* out);
/* FillTessellator::tessellate_path (crates/tessellation/src/fill.rs:628-641). */
int32_t lyg_fill_path(lyg_ctx* ctx, const lyg_path_view* path, const lyg_fill_options* opts,
                      lyg_output* out);

/* FillTessellator::tessellate_rectangle (fill.rs:653-661 -> basic_shapes.rs:12-38): four
 * vertices, two triangles; rect = min x, min y, max x, max y.  No sweep and no device work
 * (ctx may be null). */
int32_t lyg_fill_rectangle(lyg_ctx* ctx, const float* rect, lyg_output* out);
/* FillTessellator::tessellate_circle (fill.rs:663-672 -> basic_shapes.rs:40-205): four
 * quadrant vertices and recursive border arcs.  No sweep and no device work (ctx may be
 * null). */
int32_t lyg_fill_circle(lyg_ctx* ctx, float cx, float cy, float radius,
                        const lyg_fill_options* opts, lyg_output* out);

/* ---- lyon_geom flattening, array form -------------------------------------------------
 * for_each_flattened_with_t of QuadraticBezierSegment (quadratic_bezier.rs:368-391),
 * CubicBezierSegment (cubic_bezier.rs:590-614) and LineSegment (line.rs:544-550) for n
 * segments.  kinds[i]: 0 line, 1 quadratic, 2 cubic; ctrl: 8 floats per segment
 * (from, ctrl1, ctrl2, to; unused slots ignored; a line uses from and ctrl1 as its end).
 * Writes the end point of every flattened sub-segment; offsets[n+1] delimits segment i.
 * Host pointers. */
int32_t lyg_flatten_segments(lyg_ctx* ctx, const int32_t* kinds, const float* ctrl, uint32_t n,
                             float tolerance, float* out_points, uint32_t point_capacity,
                             uint32_t* offsets, uint32_t* n_points);
/* The same with device pointers (inputs, points and offsets stay in device memory; the
 * count is returned in *n_points).  Status LYG_ERR_CAPACITY (and *n_points) when
 * out_points is null or point_capacity is too small. */
int32_t lyg_flatten_segments_device(lyg_ctx* ctx, const int32_t* kinds, const float* ctrl,
                                    uint32_t n, float tolerance, float* out_points,
                                    uint32_t point_capacity, uint32_t* offsets,
                                    uint32_t* n_points);
/* No counterpart in lyon.  lyg_flatten_segments_device for callers that cannot allocate
 * (code under HIP graph capture, serving loops and harnesses that forbid hipMalloc): the
 * same kernels and results, but the per-segment counts, scan partials and totals live in a
 * fixed device array of the library (one per GPU, about 4.2 MiB, held until the process
 * ends) instead of context buffers, and the call allocates nothing.  n is limited to
 * lyg_flatten_noalloc_max_segments() (1,048,576); larger n returns LYG_ERR_CAPACITY.
 * Every call on a GPU shares that GPU's array: calls that overlap on one GPU (from
 * several contexts or threads) race and must be serialised by the caller. */
int32_t lyg_flatten_segments_device_noalloc(lyg_ctx* ctx, const int32_t* kinds,
                                            const float* ctrl, uint32_t n, float tolerance,
                                            float* out_points, uint32_t point_capacity,
                                            uint32_t* offsets, uint32_t* n_points);
uint32_t lyg_flatten_noalloc_max_segments(void);
/* Arc::for_each_flattened_with_t (arc.rs:339-368) for n arcs of 7 floats each
 * (center x, y, radii x, y, start angle, sweep angle, x rotation).  Host pointers. */
int32_t lyg_flatten_arcs(lyg_ctx* ctx, const float* arcs, uint32_t n, float tolerance,
                         float* out_points, uint32_t point_capacity, uint32_t* offsets,
                         uint32_t* n_points);
/* arc_to_quadratic_beziers_with_t (arc.rs:668-701): 6 floats (from, ctrl, to) per quadratic.
 * Host pointers. */
int32_t lyg_arc_to_quadratics(lyg_ctx* ctx, const float* arcs, uint32_t n, float* out_quads,
                              uint32_t quad_capacity, uint32_t* offsets, uint32_t* n_quads);

/* ---- batched form (device-resident buffers) ------------------------------------------
 * No lyon counterpart: equivalent to calling tessellate_path for every path in order with
 * one BuffersBuilder over a shared VertexBuffers (geometry_builder.rs:300-351).  Every
 * pointer here is device memory.  Inputs: n paths packed as verbs/points with n+1 offsets.
 * Outputs: vertices/indices plus per-path offsets (n+1 entries each) and status. */
typedef struct lyg_batch_in {
    const uint8_t* verbs;
    const float* points;
    const uint32_t* verb_offsets;
    const uint32_t* point_offsets;
    uint32_t n_paths;
} lyg_batch_in;

typedef struct lyg_batch_out {
    float* vertices;
    uint32_t* indices;
    uint64_t vertex_capacity;
    uint64_t index_capacity;
    uint32_t* path_vertex_offsets; /* n+1, device */
    uint32_t* path_index_offsets;  /* n+1, device */
    uint32_t vertex_offset;
    uint64_t n_vertices;           /* out (host field) */
    uint64_t n_indices;            /* out (host field) */
    int32_t* path_status;          /* n, device, optional: each path's lyg_status */
} lyg_batch_out;

/* opts: n_opts entries in device memory; n_opts == 1 shares one StrokeOptions. */
int32_t lyg_stroke_batch(lyg_ctx* ctx, const lyg_batch_in* in, const lyg_stroke_options* opts,
                         uint32_t n_opts, lyg_batch_out* out);
int32_t lyg_fill_batch(lyg_ctx* ctx, const lyg_batch_in* in, const lyg_fill_options* opts,
                       uint32_t n_opts, lyg_batch_out* out);

/* ---- variable-width stroke ---------------------------------------------------------------
 * StrokeTessellator::tessellate_path with StrokeOptions::variable_line_width(width_attribute)
 * (stroke.rs:653-767, 1380-1512, 1809-1866, 2060-2128): every endpoint carries n_attributes
 * custom attributes (Path::builder_with_attributes, path.rs:115-118), given per point of
 * the path(s) — n_points x n_attributes floats, the rows of control points unread — and
 * its half width is line_width * attributes[width_attribute] * 0.5; curves interpolate the
 * widths of their endpoints.  Positions agree with lyon's within float rounding of asin /
 * cos / sin; status as lyg_stroke_path / lyg_stroke_batch, plus LYG_ERR_NO_SPACE for a path
 * with a curve of more than 4,194,304 flattened points.  One GPU lane strokes one path.
 * out_attributes (optional, n_attributes floats per output vertex, vertex_capacity rows):
 * StrokeVertex::interpolated_attributes (stroke.rs:2728-2746) of every vertex — its
 * source endpoint's row, or on a curve a*(1-t) + b*t of the curve's two endpoint rows. */
int32_t lyg_stroke_path_vw(lyg_ctx* ctx, const lyg_path_view* path, const float* attributes,
                           uint32_t n_attributes, uint32_t width_attribute,
                           const lyg_stroke_options* opts, lyg_output* out,
                           float* out_attributes);
/* attributes and out_attributes in device memory, per point of the batch (lyg_batch_in's
 * points) and per output vertex. */
int32_t lyg_stroke_batch_vw(lyg_ctx* ctx, const lyg_batch_in* in, const float* attributes,
                            uint32_t n_attributes, uint32_t width_attribute,
                            const lyg_stroke_options* opts, uint32_t n_opts,
                            lyg_batch_out* out, float* out_attributes);

/* ---- banded fill (no lyon counterpart for the method) ---------------------------------
 * The fill of FillTessellator::tessellate_path (fill.rs:628-641) computed for large paths
 * across the whole GPU: the flattened edges (as lyon's event queue builder flattens them,
 * event_queue.rs:661-900) are cut into horizontal bands swept independently into
 * trapezoids.  The covered region is lyon's (EvenOdd / NonZero over the same edges); the
 * triangles and vertices are not lyon's (vertices are not shared, no monotone spans), and
 * sweep_orientation / handle_intersections do not change the result.  Errors:
 * LYG_ERR_TOLERANCE_IS_NAN, LYG_ERR_POSITION_IS_NAN, LYG_ERR_INTERNAL (a band denser than
 * 4096 edges), LYG_ERR_CAPACITY (sizes set).  Host buffers as lyg_fill_path. */
int32_t lyg_fill_path_banded(lyg_ctx* ctx, const lyg_path_view* path,
                             const lyg_fill_options* opts, lyg_output* out);
/* Batched form over device buffers (as lyg_fill_batch). */
int32_t lyg_fill_batch_banded(lyg_ctx* ctx, const lyg_batch_in* in,
                              const lyg_fill_options* opts, uint32_t n_opts,
                              lyg_batch_out* out);
/* No lyon counterpart: how far each path's flattened curves stray from the true curves.
 * For every chord of every flattened curve (the f32 points as the tessellator computes
 * them: stroke_rule 0 the fill's flattening, event_queue.rs:758-900, curves turned down the
 * sweep; 1 the stroker's, stroke.rs:2755-2779, quadratics with a sharp turn split first),
 * the largest distance from the curve, evaluated in f64 from its f32 control points over
 * the chord's parameter range, to the chord; path_error[i] (device, n_paths) gets
 * path i's largest (0 for paths without curves).  tolerance: device, 1 or n_paths values.
 * Lyon's Wang count keeps this within the tolerance up to the truncation of its last
 * segment (quadratic_bezier.rs:375-390). */
int32_t lyg_flattening_errors(lyg_ctx* ctx, const lyg_batch_in* in, const float* tolerance,
                              uint32_t n_tol, int stroke_rule, float* path_error);
/* Band size of the banded fill: about edges_per_band edge starts per band (default 256),
 * at most max_bands bands per path (default 16384). */
int32_t lyg_ctx_set_band_size(lyg_ctx* ctx, uint32_t edges_per_band, uint32_t max_bands);

/* ---- lyon_algorithms path validators ---------------------------------------------------
 * path_winding_number_at_position / hit_test_path (crates/algorithms/src/hit_test.rs:8-83)
 * and approximate_signed_area (crates/algorithms/src/area.rs:7-19), bit-exact with lyon's
 * f32 arithmetic.  Paths are as lyon's builders make them (every sub-path opens with
 * Begin).  tolerance must be >= f32::EPSILON^2 (for_each_flattened's own precondition,
 * quadratic_bezier.rs:344), else LYG_ERR_TOLERANCE_IS_NAN.  fill_rule: 0 EvenOdd,
 * 1 NonZero (FillRule::is_in, path/src/lib.rs:267-280).
 *
 * Batched, device pointers: query i is the point queries[2i], queries[2i+1] against path
 * query_paths[i] of the batch (an index past the batch gives winding 0); writes winding[i]
 * and, when inside is not null, inside[i] = hit_test_path's answer. */
int32_t lyg_winding_numbers(lyg_ctx* ctx, const lyg_batch_in* in, const float* queries,
                            const uint32_t* query_paths, uint32_t n_queries, float tolerance,
                            uint8_t fill_rule, int32_t* winding, uint8_t* inside);
/* approximate_signed_area of every path of a device batch into areas[n_paths] (device). */
int32_t lyg_signed_areas(lyg_ctx* ctx, const lyg_batch_in* in, float tolerance, float* areas);
/* One path and n query points, host pointers (inside may be null). */
int32_t lyg_path_winding_numbers(lyg_ctx* ctx, const lyg_path_view* path, const float* xy,
                                 uint32_t n, float tolerance, uint8_t fill_rule,
                                 int32_t* winding, uint8_t* inside);
/* approximate_signed_area of one path, host pointers. */
int32_t lyg_approximate_signed_area(lyg_ctx* ctx, const lyg_path_view* path, float tolerance,
                                    float* area);

/* ---- SVG path data ingest ---------------------------------------------------------------
 * lyon_extra's PathParser::parse (crates/extra/src/parser.rs:154-462) into lyon_path's Path
 * layout (BuilderImpl, crates/path/src/path.rs:516-611): one path per string, all strings
 * parsed at once on the device, the paths written end to end as a batch (the layout of
 * lyon_path's PathBuffer, path_buffer.rs:14-104, with n+1 offsets).  Numbers convert as
 * Rust's str::parse::<f32> does (correctly rounded); arcs become quadratics as
 * SvgArc::to_arc + Arc::for_each_quadratic_bezier_with_t make them (arc.rs:58-146, 668-701).
 *
 * Per path, errors[4*i .. 4*i+3] = kind, line, column, character (ParseError's fields,
 * parser.rs:13-33); a path with an error keeps what the parser had built before it, as
 * lyon's builder does.  num_attributes: numbers read after every endpoint (the extended
 * syntax) and dropped, since this path layout has no attributes.  stop_at: an ASCII
 * character that ends the parse (ParserOptions::stop_at) or -1.  Only ASCII text: a
 * non-ASCII byte reached by the parser ends the path with LYG_SVG_ERR_NON_ASCII. */
enum {
    LYG_SVG_OK = 0,
    LYG_SVG_ERR_NUMBER = 1,           /* ParseError::Number */
    LYG_SVG_ERR_FLAG = 2,             /* ParseError::Flag */
    LYG_SVG_ERR_COMMAND = 3,          /* ParseError::Command */
    LYG_SVG_ERR_MISSING_MOVE_TO = 4,  /* ParseError::MissingMoveTo */
    LYG_SVG_ERR_NON_ASCII = 5,        /* not in lyon: non-ASCII input */
    LYG_SVG_ERR_UNDECIDED_NUMBER = 6  /* no longer returned (numbers of any length are
                                         rounded correctly); kept for the ABI */
};
/* Device pointers: text[text_offsets[i] .. text_offsets[i+1]) is string i.  verb_offsets /
 * point_offsets (n+1) and errors (4n, may be null) are written; verbs / points are written
 * when their capacities (in verbs and in points) hold n_verbs / n_points, else
 * LYG_ERR_CAPACITY with the required sizes set.  n_errors: paths with an error. */
int32_t lyg_parse_svg_paths(lyg_ctx* ctx, const uint8_t* text, const uint32_t* text_offsets,
                            uint32_t n, uint32_t num_attributes, int32_t stop_at,
                            uint8_t* verbs, uint64_t verb_capacity, float* points,
                            uint64_t point_capacity, uint32_t* verb_offsets,
                            uint32_t* point_offsets, int32_t* errors, uint64_t* n_verbs,
                            uint64_t* n_points, uint32_t* n_errors);
/* One string, host pointers (the lyon call: PathParser::parse into a Path builder). */
int32_t lyg_parse_svg_path(lyg_ctx* ctx, const char* text, uint32_t len,
                           uint32_t num_attributes, int32_t stop_at, uint8_t* verbs,
                           uint32_t verb_capacity, float* points, uint32_t point_capacity,
                           uint32_t* n_verbs, uint32_t* n_points, int32_t* error);

/* Size query for the batched stroke: per-path vertex/index counts into device arrays
 * (n entries each) and totals.  Same kernels as lyg_stroke_batch's first pass. */
int32_t lyg_stroke_batch_count(lyg_ctx* ctx, const lyg_batch_in* in,
                               const lyg_stroke_options* opts, uint32_t n_opts,
                               uint32_t* path_nv, uint32_t* path_ni, uint64_t* total_nv,
                               uint64_t* total_ni);

/* ---- HIP graph form of the batched stroke (no lyon counterpart) -------------------------
 * Captures one lyg_stroke_batch (counter reset, count kernel, offset scans, emit kernel)
 * bound to these device buffers; lyg_graph_launch replays it with no host synchronisation
 * inside (the emit kernel skips a path that does not fit the output capacity and reports
 * it); lyg_graph_result waits and returns the totals and status. */
typedef struct lyg_graph lyg_graph;
int32_t lyg_stroke_batch_graph(lyg_ctx* ctx, const lyg_batch_in* in,
                               const lyg_stroke_options* opts, uint32_t n_opts,
                               const lyg_batch_out* out, lyg_graph** graph);
/* The batched fill's first sweep round as HIP graphs (no lyon counterpart): the round's
 * plan (size classes, placement forms, staging and working memory) is made at capture
 * time from the batch's event bounds and intersection estimates, as lyg_fill_batch's first
 * round makes it; a launch replays the sweep launches (on the context's class streams, at
 * once), the offset scans and the copy (each path's range checked against the output
 * capacity on the device) with no host synchronisation inside, and writes
 * out->path_status when it is set.  lyg_graph_result waits and returns lyg_fill_batch's
 * statuses, plus LYG_ERR_NO_SPACE when a path outgrew the room planned at capture (its
 * data changed since; lyg_fill_batch sweeps it again with more room).  The graph owns its
 * working, staging and status memory; out's arrays must outlive it. */
int32_t lyg_fill_batch_graph(lyg_ctx* ctx, const lyg_batch_in* in,
                             const lyg_fill_options* opts, uint32_t n_opts,
                             const lyg_batch_out* out, lyg_graph** graph);
int32_t lyg_graph_launch(lyg_graph* graph);
int32_t lyg_graph_result(lyg_graph* graph, uint64_t* n_vertices, uint64_t* n_indices);
void lyg_graph_destroy(lyg_graph* graph);

/* ---- autotuner (no lyon counterpart) -----------------------------------------------------
 * Runs the batch under each candidate launch shape (stroke: resident blocks per CU of the
 * persistent grid, then windows per part of split paths; fill: waves per block and LDS
 * bytes per wave for the sweep's hot arrays; banded fill: edges per band), keeps the
 * fastest in the context for later calls and reports it. */
typedef struct lyg_tuning {
    uint32_t blocks_per_cu;
    uint32_t waves_per_block;
    uint32_t lds_bytes_per_wave;
    float best_ms;
    uint32_t split_windows;    /* stroke: windows per part (lyg_ctx_set_stroke_split) */
    uint32_t edges_per_band;   /* banded fill (lyg_ctx_set_band_size) */
} lyg_tuning;
/* Sets the launch shapes explicitly (either pointer may be null), e.g. a tuning kept from an
 * earlier run. */
int32_t lyg_ctx_set_tuning(lyg_ctx* ctx, const lyg_tuning* stroke, const lyg_tuning* fill);
/* How lyg_stroke_batch sizes its output (no lyon counterpart).  enabled = 1 (default): one
 * emit pass into per-path staging room (room_scale times the computed room; a path that
 * outgrows it is emitted again at its final place), then a compaction copy; enabled = 0:
 * a count pass and an emit pass.  Both give the same output. */
int32_t lyg_ctx_set_stroke_staging(lyg_ctx* ctx, int enabled, float room_scale);
/* Intra-path parallelism of the staged stroke (no lyon counterpart).  windows_per_part > 0:
 * a path of one sub-path with at least 2 * windows_per_part windows of 64 steps (sub-path
 * start, line ends, flattened curve points) is cut into parts of that many windows, walked
 * by different waves at once; each part first re-walks the window before it to rebuild
 * the walk's state, and a check per path confirms it matches the state the previous part
 * ended in (else the path is walked again whole).  The output is the same bytes as the
 * whole-path walk.  Default 64 (parts of 4,096 steps); 0: every path is walked by one
 * wave. */
int32_t lyg_ctx_set_stroke_split(lyg_ctx* ctx, uint32_t windows_per_part);
/* Placement of the batched fill sweep's working arrays (no lyon counterpart).  enabled
 * (default): the first round sweeps each path whose estimated active edges, spans, pending
 * intersections and monotone nodes fit 150 KiB in a form that keeps them all in LDS (a
 * path that outgrows them is swept again in the global-memory form); 0: every path starts
 * with its arrays in LDS and moves them to global memory as they grow.  Both give the
 * same output.  enabled > 1 divides the LDS-resident capacities by that factor, so that
 * paths outgrow them and take the fallback (for tests). */
int32_t lyg_ctx_set_fill_lds_resident(lyg_ctx* ctx, int enabled);
/* Of the last split batched stroke on this context: the items walked (whole paths and
 * parts) and the paths walked again whole because a part's prologue did not rebuild the
 * state or a part outgrew its room (no lyon counterpart). */
int32_t lyg_ctx_stroke_split_stats(lyg_ctx* ctx, uint64_t* items, uint32_t* walked_again);
int32_t lyg_autotune_stroke(lyg_ctx* ctx, const lyg_batch_in* in, const lyg_stroke_options* opts,
                            uint32_t n_opts, lyg_batch_out* out, lyg_tuning* chosen);
int32_t lyg_autotune_fill(lyg_ctx* ctx, const lyg_batch_in* in, const lyg_fill_options* opts,
                          uint32_t n_opts, lyg_batch_out* out, lyg_tuning* chosen);
int32_t lyg_autotune_banded(lyg_ctx* ctx, const lyg_batch_in* in, const lyg_fill_options* opts,
                            uint32_t n_opts, lyg_batch_out* out, lyg_tuning* chosen);

/* ---- serving loop (no lyon counterpart) -------------------------------------------------
 * Continuous batching of single-path requests on one GPU.  Requests (host paths, copied at
 * submit) are admitted into batches while their worst-case output fits the unreserved part
 * of a device pool; up to max_inflight batches run at once, each on its own context and
 * stream, so requests arriving while others run are batched as soon as a worker and pool
 * room free up.  Results stay in the pool (device pointers, vertex ids local to the
 * request, as tessellate_path into an empty VertexBuffers) until released.  Every call is
 * thread-safe. */
typedef struct lyg_server lyg_server;
typedef struct lyg_server_config {
    uint64_t pool_bytes;       /* device pool holding the results */
    uint32_t max_batch_paths;  /* requests per batch */
    uint32_t max_inflight;     /* batches on the device at once */
    uint32_t linger_us;        /* wait this long for a short queue to grow (0: none) */
    int32_t device;
    uint32_t stroke_sizing;    /* 0: one staged emit pass (as lyg_stroke_batch), 1: count pass
                                  then emit (lyg_ctx_set_stroke_staging(ctx, 0, ..)) */
    uint32_t fill_banded_min_edges; /* fills flattening to at least this many edges run through
                                       the banded fill (lyon's coverage, other triangles); 0:
                                       never */
    uint64_t max_request_bytes; /* a request whose worst-case output exceeds this completes at
                                   once with LYG_ERR_NO_SPACE (0: the pool's size); a fill's
                                   worst case counts its estimated crossings */
    uint32_t stroke_split_windows; /* lyg_ctx_set_stroke_split of every worker (0: the
                                      library's default) */
    uint64_t max_fill_crossings;   /* a fill whose estimated crossings exceed this completes at
                                      once with LYG_ERR_NO_SPACE (0: no limit) */
} lyg_server_config;
typedef struct lyg_server_result {
    int32_t status;            /* the request's lyg_status */
    const float* vertices;     /* device pointer (x, y pairs); valid until released */
    const uint32_t* indices;   /* device pointer */
    uint32_t n_vertices, n_indices;
    double latency_us;         /* submit to completion */
    int32_t banded;            /* 1: a fill that ran through the banded fill (lyon's
                                  coverage, its own triangles); 0: lyon's triangles */
} lyg_server_result;
typedef struct lyg_server_stats {
    uint64_t completed, triangles, vertices, batches, pending, pool_high_water;
    double mean_batch;         /* requests per batch */
    double p50_us, p99_us, max_us;  /* submit-to-completion latency over completed requests */
    double elapsed_s;          /* first submit to last completion */
    uint64_t rejected;         /* requests refused at admission (worst case above the limit) */
} lyg_server_stats;
lyg_server* lyg_server_create(const lyg_server_config* cfg);
void lyg_server_destroy(lyg_server* server);
int32_t lyg_server_submit_fill(lyg_server* server, const lyg_path_view* path,
                               const lyg_fill_options* opts, uint64_t* ticket);
int32_t lyg_server_submit_stroke(lyg_server* server, const lyg_path_view* path,
                                 const lyg_stroke_options* opts, uint64_t* ticket);
/* LYG_PENDING until the request is done, then LYG_OK with the result filled in. */
int32_t lyg_server_poll(lyg_server* server, uint64_t ticket, lyg_server_result* out);
int32_t lyg_server_wait(lyg_server* server, uint64_t ticket, lyg_server_result* out);
/* Copies a finished result to host arrays (n_vertices pairs, n_indices ids). */
int32_t lyg_server_copy_result(lyg_server* server, uint64_t ticket, float* vertices,
                               uint32_t* indices);
/* Frees the request and, with the last request of its batch, the batch's pool space. */
int32_t lyg_server_release(lyg_server* server, uint64_t ticket);
int32_t lyg_server_get_stats(lyg_server* server, lyg_server_stats* out);
/* Serving measurement: n_requests with Poisson arrivals at rate_per_s, each a stroke with
 * probability stroke_frac (one path of log-uniform [min_segs, max_segs] segments, 25% line
 * / 25% quadratic / 50% cubic, points in [0, 4096]^2, the batched stroke workload's
 * options) or else a fill (a path of the batched fill workload, EvenOdd/NonZero 50/50);
 * tolerance for both.  Requests draw from n_distinct pre-generated paths of each kind;
 * every result is released when done.  Latency percentiles and the mean batch cover this
 * stream's requests only. */
typedef struct lyg_stream_stats {
    uint64_t requests, triangles, batches, pool_high_water;
    double offered_rate, submit_rate, completed_rate;  /* requests/s */
    double triangles_per_s, p50_us, p99_us, max_us, mean_batch, seconds;
    uint64_t failed;           /* requests that completed with an error status (rejected
                                  ones included) */
    uint64_t rejected;         /* of which refused at admission */
} lyg_stream_stats;
int32_t lyg_server_run_stream(lyg_server* server, uint64_t seed, uint32_t n_requests,
                              double rate_per_s, float stroke_frac, uint32_t min_segs,
                              uint32_t max_segs, float tolerance, uint32_t n_distinct,
                              lyg_stream_stats* out);
/* The same with a choice of fill requests: fill_kind 0 = the batched fill workload's stars
 * and cubic loops (as lyg_server_run_stream); 1 = closed paths from the stroke requests'
 * generator (log-uniform segment counts in [min_segs, max_segs], lines / quadratics /
 * cubics 25/25/50, points uniform in [0, 4096]^2); 2 = closed random walks (the single
 * large fill's generator) of a log-uniform edge count in [min_segs, max_segs] rescaled
 * onto [0, 4096]^2. */
int32_t lyg_server_run_stream_mixed(lyg_server* server, uint64_t seed, uint32_t n_requests,
                                    double rate_per_s, float stroke_frac, uint32_t min_segs,
                                    uint32_t max_segs, float tolerance, uint32_t n_distinct,
                                    int fill_kind, lyg_stream_stats* out);

/* ---- synthetic inputs ----------------------------------------------------------------
 * Seeded generators for the benchmark workloads (SplitMix64 -> PCG32).  Host memory.
 * Call once with null arrays to get sizes, then again with arrays of those sizes. */
int32_t lyg_gen_stroke_batch(uint64_t seed, uint32_t n_paths, uint32_t min_segs,
                             uint32_t max_segs, float coord_max, int closed, uint8_t* verbs,
                             float* points, uint32_t* verb_offsets, uint32_t* point_offsets,
                             uint64_t* n_verbs, uint64_t* n_points);
/* Per-path StrokeOptions of the batched stroke workload: width uniform in [w_min, w_max],
 * join uniform over {Miter, Round, Bevel}, one cap uniform over {Butt, Square, Round},
 * miter limit 4.  `out` has n_paths entries (host memory). */
int32_t lyg_gen_stroke_options(uint64_t seed, uint32_t n_paths, float tolerance, float w_min,
                               float w_max, lyg_stroke_options* out);

/* The batched fill workload: closed paths in [0, 4096]^2, even paths self-intersecting
 * stars (5-512 points at unsorted random angles, radii R*U[0.2,1], R in [16, 512]), odd
 * paths loops of 4-256 cubics scaled to flatten to 16-2048 edges at tolerance 0.05.
 * Same two-call protocol. */
int32_t lyg_gen_fill_batch(uint64_t seed, uint32_t n_paths, uint8_t* verbs, float* points,
                           uint32_t* verb_offsets, uint32_t* point_offsets, uint64_t* n_verbs,
                           uint64_t* n_points);
/* The single large fill workload: a closed polyline of n vertices from a seeded 2D random
 * walk (steps N(0,1) per axis, Box-Muller on PCG32, f32), rescaled per axis to
 * [0, extent]^2.  verbs: n + 1 entries (Begin, Lines, Close), points: n + 1 pairs. */
int32_t lyg_gen_random_walk(uint64_t seed, uint32_t n, float extent, uint8_t* verbs,
                            float* points);
/* Per-path FillOptions of the batched fill workload: EvenOdd/NonZero 50/50, vertical
 * sweep, intersections handled. */
int32_t lyg_gen_fill_options(uint64_t seed, uint32_t n_paths, float tolerance,
                             lyg_fill_options* out);

#ifdef __cplusplus
}
#endif
#endif /* LYG_H */
