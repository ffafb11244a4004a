// Direct constructions of lyon's basic fill shapes (crates/tessellation/src/basic_shapes.rs):
// FillTessellator::tessellate_rectangle and tessellate_circle do not run the sweep; they
// emit a fixed vertex/triangle pattern.  The handful of vertices is computed here on the
// host with f32 arithmetic in lyon's operation order; sqrt/log2/cos/sin are the platform
// libm's single-precision functions, which is what Rust's f32 methods call.
#include <cmath>
#include <cstdint>
#include <vector>

#include "ctx.h"

namespace {

constexpr float PI_F = 3.14159265358979323846f;

struct ShapeOut {
    std::vector<float> v;  // x, y pairs
    std::vector<uint32_t> t;  // local vertex ids, three per triangle
    uint32_t vertex(float x, float y) {
        v.push_back(x);
        v.push_back(y);
        return (uint32_t)(v.size() / 2 - 1);
    }
    void tri(uint32_t a, uint32_t b, uint32_t c) {
        t.push_back(a);
        t.push_back(b);
        t.push_back(c);
    }
};

// Writes the shape into the caller's arrays as BuffersBuilder would (ids vertex_offset +
// local id, geometry_builder.rs:338-350): the two-phase capacity protocol of lyg_output.
int32_t write_out(const ShapeOut& s, lyg_output* out) {
    const uint64_t nv = s.v.size() / 2, ni = s.t.size();
    out->n_vertices = (uint32_t)nv;
    out->n_indices = (uint32_t)ni;
    if (nv > 0 && (uint64_t)out->vertex_offset + nv - 1 > 0xfffffffeull) {
        out->n_vertices = out->n_indices = 0;
        return LYG_ERR_TOO_MANY_VERTICES;
    }
    if (nv > out->vertex_capacity || ni > out->index_capacity) return LYG_ERR_CAPACITY;
    if ((nv && !out->vertices) || (ni && !out->indices)) return LYG_ERR_INVALID_ARGUMENT;
    for (uint64_t k = 0; k < 2 * nv; ++k) out->vertices[k] = s.v[k];
    for (uint64_t k = 0; k < ni; ++k) out->indices[k] = s.t[k] + out->vertex_offset;
    return LYG_OK;
}

// Rust's `f as u32`: truncation, saturating, NaN to 0.
uint32_t as_u32(float f) {
    if (!(f > 0.0f)) return 0u;
    if (f >= 4294967296.0f) return 0xffffffffu;
    return (uint32_t)f;
}

// fill_border_radius (basic_shapes.rs:160-205): pre-order recursion over the arc.
void border_radius(ShapeOut& s, float cx, float cy, float a0, float a1, float radius,
                   uint32_t va, uint32_t vb, uint32_t n) {
    if (n == 0) return;
    const float mid = (a0 + a1) * 0.5f;
    const float nx = std::cos(mid), ny = std::sin(mid);
    const uint32_t v = s.vertex(cx + nx * radius, cy + ny * radius);
    s.tri(vb, v, va);
    border_radius(s, cx, cy, a0, mid, radius, va, v, n - 1);
    border_radius(s, cx, cy, mid, a1, radius, v, vb, n - 1);
}

}  // namespace

extern "C" {

// FillTessellator::tessellate_rectangle -> basic_shapes::fill_rectangle
// (crates/tessellation/src/basic_shapes.rs:12-38).  rect: min x, min y, max x, max y.
int32_t lyg_fill_rectangle(lyg_ctx* c, const float* rect, lyg_output* out) {
    (void)c;
    if (!rect || !out) return LYG_ERR_INVALID_ARGUMENT;
    ShapeOut s;
    const uint32_t a = s.vertex(rect[0], rect[1]);  // min
    const uint32_t b = s.vertex(rect[0], rect[3]);  // bottom_left
    const uint32_t cc = s.vertex(rect[2], rect[3]);  // bottom_right = max
    const uint32_t d = s.vertex(rect[2], rect[1]);  // top_right
    s.tri(a, b, cc);
    s.tri(a, cc, d);
    return write_out(s, out);
}

// FillTessellator::tessellate_circle -> basic_shapes::fill_circle (basic_shapes.rs:40-124),
// with its own circle_flattening_step (basic_shapes.rs:137-146).
int32_t lyg_fill_circle(lyg_ctx* c, float cx, float cy, float radius,
                        const lyg_fill_options* opts, lyg_output* out) {
    (void)c;
    if (!opts || !out) return LYG_ERR_INVALID_ARGUMENT;
    out->n_vertices = out->n_indices = 0;
    radius = std::fabs(radius);
    if (radius == 0.0f) return LYG_OK;
    ShapeOut s;
    const uint32_t v0 = s.vertex(cx + -1.0f * radius, cy + 0.0f * radius);  // left
    const uint32_t v1 = s.vertex(cx + 0.0f * radius, cy + -1.0f * radius);  // up
    const uint32_t v2 = s.vertex(cx + 1.0f * radius, cy + 0.0f * radius);   // right
    const uint32_t v3 = s.vertex(cx + 0.0f * radius, cy + 1.0f * radius);   // down
    s.tri(v0, v3, v1);
    s.tri(v1, v3, v2);
    const float angles[4][2] = {{PI_F, 1.5f * PI_F},
                                {1.5f * PI_F, 2.0f * PI_F},
                                {0.0f, PI_F * 0.5f},
                                {PI_F * 0.5f, PI_F}};
    const float arc_len = 0.5f * PI_F * radius;
    const float tol = std::fmin(opts->tolerance, radius);
    const float step = 2.0f * std::sqrt(2.0f * tol * radius - tol * tol);
    const float num_segments = std::ceil(arc_len / step);
    const uint32_t n = as_u32(std::log2(num_segments));
    // 4 + 4 (2^n - 1) vertices: beyond u32 ids BuffersBuilder fails with TooManyVertices
    if (n >= 30) return LYG_ERR_TOO_MANY_VERTICES;
    const uint32_t v[4] = {v0, v1, v2, v3};
    for (int i = 0; i < 4; ++i)
        border_radius(s, cx, cy, angles[i][0], angles[i][1], radius, v[i], v[(i + 1) % 4], n);
    return write_out(s, out);
}

}  // extern "C"
