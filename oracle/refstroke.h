// ORACLE — test infrastructure only (see refgeom.h).
//
// CPU restatement of lyon_tessellation's fixed-width StrokeTessellator
// (crates/tessellation/src/stroke.rs) writing through BuffersBuilder semantics
// (geometry_builder.rs:300-351, 451-508) into VertexBuffers<[f32;2], u32>.
// Variable line width (stroke.rs:1380-1512) is not part of this restatement.
#pragma once
#include "refgeom.h"
#include "refpath.h"

namespace refstroke {
using namespace refgeom;

enum { SIDE_POSITIVE = 0, SIDE_NEGATIVE = 1 };
enum LineCap { CAP_BUTT = 0, CAP_SQUARE = 1, CAP_ROUND = 2 };
enum LineJoin { JOIN_MITER = 0, JOIN_MITER_CLIP = 1, JOIN_ROUND = 2, JOIN_BEVEL = 3 };
static const uint32_t INVALID_ID = 0xffffffffu;

struct StrokeOptions {
    float line_width = 1.0f;
    float miter_limit = 4.0f;
    float tolerance = 0.1f;
    int start_cap = CAP_BUTT, end_cap = CAP_BUTT, line_join = JOIN_MITER;
};

enum Status { OK = 0, TOO_MANY_VERTICES = 1 };

// BuffersBuilder<[f32;2], u32, Positions> over a caller-owned VertexBuffers.
struct Output {
    std::vector<V2>* vertices;
    std::vector<uint32_t>* indices;
    uint32_t vertex_offset = 0;
    size_t first_vertex = 0, first_index = 0;
    // custom attributes per point (n_attrs each) and, when set, the interpolated
    // attributes of every vertex
    const float* attrs = nullptr;
    uint32_t n_attrs = 0;
    std::vector<float>* attr_out = nullptr;
    void begin_geometry() {
        first_vertex = vertices->size();
        first_index = indices->size();
    }
    // add_vertex_impl (geometry_builder.rs:338-350); u32::MAX is VertexId::INVALID.
    bool add_vertex(V2 p, uint32_t* id) {
        uint64_t v = (uint64_t)vertices->size() + vertex_offset;
        if (v > 0xfffffffeull) return false;
        vertices->push_back(p);
        *id = (uint32_t)v;
        return true;
    }
    void add_triangle(uint32_t a, uint32_t b, uint32_t c) {
        indices->push_back(a);
        indices->push_back(b);
        indices->push_back(c);
    }
    void abort_geometry() {
        vertices->resize(first_vertex);
        indices->resize(first_index);
        if (attr_out) attr_out->resize(first_vertex * n_attrs);
    }
};

struct SidePoints {
    V2 prev{NAN, NAN}, next{NAN, NAN};
    bool has_single = false;
    V2 single_vertex{0, 0};
    uint32_t prev_vertex = INVALID_ID, next_vertex = INVALID_ID;
};

// EndpointData (stroke.rs:320-352); vertex sources are omitted because they never
// influence positions or topology (advancement does, through the variable-width joins).
// VertexSource (stroke.rs / lib.rs): an endpoint (a point index) or a point at t on the
// edge between two endpoints.
struct Source {
    uint32_t from = 0xffffffffu, to = 0xffffffffu;
    float t = 0.0f;
    bool edge = false;
};

struct EndpointData {
    V2 position{0, 0};
    float half_width = NAN;
    float advancement = NAN;
    Source src;
    int line_join = JOIN_MITER;
    SidePoints side_points[2];
    bool fold[2] = {false, false};
    bool is_flattening_step = false;
};

// StrokeVertexData (stroke.rs:2655-2666): only what StrokeVertex::position reads.
struct VertexData {
    V2 position_on_path{0, 0};
    float half_width = 0;
    V2 normal{0, 0};
    int side = SIDE_NEGATIVE;
    Source src;
};

// StrokeVertex::position (stroke.rs:2676-2678); with attributes to interpolate, also
// StrokeVertex::interpolated_attributes (stroke.rs:2728-2746) into out.attr_out.
inline bool emit_vertex(Output& out, const VertexData& v, uint32_t* id) {
    V2 pos = v.position_on_path + v.normal * v.half_width;
    if (!out.add_vertex(pos, id)) return false;
    if (out.attr_out) {
        const float* a = out.attrs + (size_t)v.src.from * out.n_attrs;
        const float* b = out.attrs + (size_t)v.src.to * out.n_attrs;
        for (uint32_t k = 0; k < out.n_attrs; ++k)
            out.attr_out->push_back(v.src.edge ? a[k] * (1.0f - v.src.t) + b[k] * v.src.t : a[k]);
    }
    return true;
}

// PointBuffer (stroke.rs:2371-2457)
struct PointBuffer {
    EndpointData points[3];
    size_t start = 0, count_ = 0;
    void push(const EndpointData& p) {
        if (count_ < 3) {
            points[count_] = p;
            count_ += 1;
            return;
        }
        points[start] = p;
        start += 1;
        if (start == 3) start = 0;
    }
    void clear() { count_ = start = 0; }
    size_t count() const { return count_; }
    EndpointData& get(size_t idx) { return points[(idx + start) % 3]; }
    EndpointData& last() { return get(count_ - 1); }
    // last_two_mut returns (second to last, last)
    EndpointData& second_last() { return points[(start + count_ - 2) % 3]; }
    EndpointData& last_of_two() { return points[(start + count_ - 1) % 3]; }
    // replace_last (stroke.rs:2400-2406): the slot last written
    void replace_last(const EndpointData& p) { points[(start == 0 ? count_ : start) - 1] = p; }
};

int stroke_path(const refpath::PathView& path, const StrokeOptions& opts, Output& out);
// StrokeOptions::variable_line_width: widths[i] is the line-width factor attribute of the
// endpoint stored at point i (control points' entries are not read).
// With out.attr_out set, the interpolated custom attributes of every vertex go there;
// out.attrs is the attribute table, point_base the row of the path's first point.
int stroke_path_vw(const refpath::PathView& path, const StrokeOptions& opts,
                   const float* widths, Output& out, uint32_t point_base = 0);
bool find_sharp_turn_export(float fx, float fy, float cx, float cy, float tx, float ty,
                            float* t_out);
uint32_t arc_subdivisions_export(float diff, float radius, float tolerance);

}  // namespace refstroke
