// ORACLE — test infrastructure only (see refgeom.h).  Loaded through ctypes by tests/,
// by __graft_entry__.smoke() and by bench.py's cpu_baseline leg; never by lyon_amd/.
//
// extern "C" face of the CPU restatement: stroke, fill and the lyon_geom flattening
// calls, each producing VertexBuffers<[f32;2],u32> / point lists in a result handle.
#include <atomic>
#include <cstring>
#include <chrono>
#include <cstdlib>
#include <thread>
#include <vector>

#include "refgeom.h"
#include "refpath.h"
#include "refstroke.h"
#include "reffill.h"

using namespace refgeom;

struct orc_result {
    std::vector<V2> vertices;
    std::vector<uint32_t> indices;
    std::vector<uint32_t> offsets;  // flatten: per-segment point offsets
    std::vector<float> attrs;       // variable-width stroke: interpolated vertex attributes
    int status = 0;
};

extern "C" {

void orc_result_free(orc_result* r) { delete r; }
int orc_result_status(const orc_result* r) { return r->status; }
uint64_t orc_result_n_vertices(const orc_result* r) { return r->vertices.size(); }
uint64_t orc_result_n_indices(const orc_result* r) { return r->indices.size(); }
uint64_t orc_result_n_offsets(const orc_result* r) { return r->offsets.size(); }
void orc_result_copy(const orc_result* r, float* vertices, uint32_t* indices, uint32_t* offsets) {
    if (vertices && !r->vertices.empty())
        std::memcpy(vertices, r->vertices.data(), r->vertices.size() * sizeof(V2));
    if (indices && !r->indices.empty())
        std::memcpy(indices, r->indices.data(), r->indices.size() * 4);
    if (offsets && !r->offsets.empty())
        std::memcpy(offsets, r->offsets.data(), r->offsets.size() * 4);
}

// StrokeTessellator::tessellate_path into BuffersBuilder::new(..).with_vertex_offset(vo).
orc_result* orc_stroke_path(const uint8_t* verbs, uint32_t n_verbs, const float* points,
                            uint32_t n_points, float line_width, float miter_limit,
                            float tolerance, int start_cap, int end_cap, int line_join,
                            uint32_t vertex_offset) {
    orc_result* r = new orc_result();
    refpath::PathView pv{verbs, n_verbs, points, n_points};
    refstroke::StrokeOptions o;
    o.line_width = line_width;
    o.miter_limit = miter_limit;
    o.tolerance = tolerance;
    o.start_cap = start_cap;
    o.end_cap = end_cap;
    o.line_join = line_join;
    refstroke::Output out{&r->vertices, &r->indices, vertex_offset};
    r->status = refstroke::stroke_path(pv, o, out);
    return r;
}

// The same with StrokeOptions::variable_line_width: widths[i] is the width factor of the
// endpoint stored at point i (tessellate_with_ids_vw, stroke.rs:653-767).
size_t orc_result_n_attrs(const orc_result* r) { return r->attrs.size(); }
void orc_result_copy_attrs(const orc_result* r, float* out) {
    if (!r->attrs.empty()) std::memcpy(out, r->attrs.data(), r->attrs.size() * 4);
}

// attrs (optional): n_points x n_attrs custom attributes; each vertex's interpolated
// attributes (StrokeVertex::interpolated_attributes) are then kept in the result.
orc_result* orc_stroke_path_vw(const uint8_t* verbs, uint32_t n_verbs, const float* points,
                               uint32_t n_points, const float* widths, float line_width,
                               float miter_limit, float tolerance, int start_cap, int end_cap,
                               int line_join, uint32_t vertex_offset, const float* attrs,
                               uint32_t n_attrs) {
    orc_result* r = new orc_result();
    refpath::PathView pv{verbs, n_verbs, points, n_points};
    refstroke::StrokeOptions o;
    o.line_width = line_width;
    o.miter_limit = miter_limit;
    o.tolerance = tolerance;
    o.start_cap = start_cap;
    o.end_cap = end_cap;
    o.line_join = line_join;
    refstroke::Output out{&r->vertices, &r->indices, vertex_offset};
    if (attrs && n_attrs) {
        out.attrs = attrs;
        out.n_attrs = n_attrs;
        out.attr_out = &r->attrs;
    }
    r->status = refstroke::stroke_path_vw(pv, o, widths, out);
    return r;
}

// FillTessellator::tessellate_path into BuffersBuilder (fill.rs:628-641).
orc_result* orc_fill_path(const uint8_t* verbs, uint32_t n_verbs, const float* points,
                          uint32_t n_points, float tolerance, int fill_rule,
                          int sweep_orientation, int handle_intersections,
                          uint32_t vertex_offset) {
    orc_result* r = new orc_result();
    refpath::PathView pv{verbs, n_verbs, points, n_points};
    reffill::FillOptions o;
    o.tolerance = tolerance;
    o.fill_rule = fill_rule;
    o.sweep_orientation = sweep_orientation;
    o.handle_intersections = handle_intersections != 0;
    reffill::Output out{&r->vertices, &r->indices, vertex_offset};
    r->status = reffill::fill_path(pv, o, out);
    return r;
}

// The flattened polylines of every curve of a path, as lyon_geom's
// for_each_flattened_with_t emits them: kind 0 line (1 point), 1 quadratic,
// 2 cubic; `offsets` gets one entry per segment plus a final total.
orc_result* orc_flatten_segments(const int* kinds, const float* ctrl, uint32_t n, float tol) {
    orc_result* r = new orc_result();
    r->offsets.push_back(0);
    for (uint32_t i = 0; i < n; ++i) {
        const float* c = ctrl + 8 * i;
        V2 p0{c[0], c[1]}, p1{c[2], c[3]}, p2{c[4], c[5]}, p3{c[6], c[7]};
        auto cb = [&](V2, V2 to, float, float) { r->vertices.push_back(to); };
        if (kinds[i] == 0) {
            r->vertices.push_back(p1);
        } else if (kinds[i] == 1) {
            quad_flatten_with_t(p0, p1, p2, tol, cb);
        } else {
            cubic_flatten_with_t(p0, p1, p2, p3, tol, cb);
        }
        r->offsets.push_back((uint32_t)r->vertices.size());
    }
    return r;
}

// Arc::for_each_flattened_with_t for n arcs given as (cx, cy, rx, ry, start, sweep, xrot).
orc_result* orc_flatten_arcs(const float* arcs, uint32_t n, float tol) {
    orc_result* r = new orc_result();
    r->offsets.push_back(0);
    for (uint32_t i = 0; i < n; ++i) {
        const float* a = arcs + 7 * i;
        refgeom::Arc arc{V2{a[0], a[1]}, V2{a[2], a[3]}, a[4], a[5], a[6]};
        arc_flatten_with_t(arc, tol, [&](V2, V2 to, float, float) { r->vertices.push_back(to); });
        r->offsets.push_back((uint32_t)r->vertices.size());
    }
    return r;
}

// arc_to_quadratic_beziers_with_t: vertices hold (from, ctrl, to) triples.
orc_result* orc_arc_to_quads(const float* arcs, uint32_t n) {
    orc_result* r = new orc_result();
    r->offsets.push_back(0);
    for (uint32_t i = 0; i < n; ++i) {
        const float* a = arcs + 7 * i;
        refgeom::Arc arc{V2{a[0], a[1]}, V2{a[2], a[3]}, a[4], a[5], a[6]};
        arc_to_quads_with_t(arc, [&](V2 f, V2 c, V2 t, float, float) {
            r->vertices.push_back(f);
            r->vertices.push_back(c);
            r->vertices.push_back(t);
        });
        r->offsets.push_back((uint32_t)r->vertices.size());
    }
    return r;
}

// ---------------------------------------------------------------------------------
// CPU baseline: stroke or fill many paths given in a packed batch, single thread or
// a pool of worker threads over paths (the "rayon over paths" analogue), and return
// the triangle count.  Output goes to per-worker reserved VertexBuffers.
// ---------------------------------------------------------------------------------
struct BatchJob {
    const uint8_t* verbs;
    const float* points;
    const uint32_t* verb_off;
    const uint32_t* point_off;
    const float* opts;  // stroke: width, miter, tol, start_cap, end_cap, join (6 floats)
                        // fill: tol, rule (2 floats)
    int is_fill;
};

static uint64_t run_range(const BatchJob& j, uint32_t a, uint32_t b) {
    std::vector<V2> vs;
    std::vector<uint32_t> is;
    vs.reserve(1 << 16);
    is.reserve(1 << 18);
    uint64_t tris = 0;
    for (uint32_t p = a; p < b; ++p) {
        vs.clear();
        is.clear();
        refpath::PathView pv{j.verbs + j.verb_off[p], j.verb_off[p + 1] - j.verb_off[p],
                             j.points + 2 * (size_t)j.point_off[p],
                             j.point_off[p + 1] - j.point_off[p]};
        if (j.is_fill) {
            reffill::FillOptions o;
            o.tolerance = j.opts[2 * p];
            o.fill_rule = (int)j.opts[2 * p + 1];
            reffill::Output out{&vs, &is, 0};
            reffill::fill_path(pv, o, out);
        } else {
            const float* s = j.opts + 6 * (size_t)p;
            refstroke::StrokeOptions o;
            o.line_width = s[0];
            o.miter_limit = s[1];
            o.tolerance = s[2];
            o.start_cap = (int)s[3];
            o.end_cap = (int)s[4];
            o.line_join = (int)s[5];
            refstroke::Output out{&vs, &is, 0};
            refstroke::stroke_path(pv, o, out);
        }
        tris += is.size() / 3;
    }
    return tris;
}

uint64_t orc_batch_triangles(const uint8_t* verbs, const float* points, const uint32_t* verb_off,
                             const uint32_t* point_off, const float* opts, uint32_t n_paths,
                             int is_fill, int n_threads) {
    BatchJob j{verbs, points, verb_off, point_off, opts, is_fill};
    if (n_threads <= 1) return run_range(j, 0, n_paths);
    std::vector<std::thread> ts;
    std::vector<uint64_t> res(n_threads, 0);
    std::atomic<uint32_t> next{0};
    for (int t = 0; t < n_threads; ++t) {
        ts.emplace_back([&, t] {
            for (;;) {
                uint32_t p = next.fetch_add(16);
                if (p >= n_paths) break;
                uint32_t e = p + 16 < n_paths ? p + 16 : n_paths;
                res[t] += run_range(j, p, e);
            }
        });
    }
    uint64_t tot = 0;
    for (int t = 0; t < n_threads; ++t) {
        ts[t].join();
        tot += res[t];
    }
    return tot;
}

// Every path of a batch tessellated on n_threads host threads, each path's output kept
// (path-local ids): orc_batch_counts gives the per-path statuses and sizes,
// orc_batch_copy the outputs back to back in path order.
struct BatchOut {
    std::vector<int> status;
    std::vector<std::vector<V2>> v;
    std::vector<std::vector<uint32_t>> i;
};

BatchOut* orc_batch_run(const uint8_t* verbs, const float* points, const uint32_t* verb_off,
                        const uint32_t* point_off, const float* opts, uint32_t n_paths,
                        int is_fill, int n_threads) {
    BatchOut* r = new BatchOut();
    r->status.assign(n_paths, 0);
    r->v.resize(n_paths);
    r->i.resize(n_paths);
    std::atomic<uint32_t> next{0};
    auto work = [&] {
        for (;;) {
            const uint32_t a = next.fetch_add(8);
            if (a >= n_paths) break;
            const uint32_t b = a + 8 < n_paths ? a + 8 : n_paths;
            for (uint32_t p = a; p < b; ++p) {
                refpath::PathView pv{verbs + verb_off[p], verb_off[p + 1] - verb_off[p],
                                     points + 2 * (size_t)point_off[p],
                                     point_off[p + 1] - point_off[p]};
                if (is_fill) {
                    reffill::FillOptions o;
                    o.tolerance = opts[2 * p];
                    o.fill_rule = (int)opts[2 * p + 1];
                    reffill::Output out{&r->v[p], &r->i[p], 0};
                    r->status[p] = reffill::fill_path(pv, o, out);
                } else {
                    const float* q = opts + 6 * (size_t)p;
                    refstroke::StrokeOptions o;
                    o.line_width = q[0];
                    o.miter_limit = q[1];
                    o.tolerance = q[2];
                    o.start_cap = (int)q[3];
                    o.end_cap = (int)q[4];
                    o.line_join = (int)q[5];
                    refstroke::Output out{&r->v[p], &r->i[p], 0};
                    r->status[p] = refstroke::stroke_path(pv, o, out);
                }
            }
        }
    };
    std::vector<std::thread> ts;
    for (int t = 0; t < (n_threads > 1 ? n_threads : 1); ++t) ts.emplace_back(work);
    for (auto& t : ts) t.join();
    return r;
}

void orc_batch_counts(const BatchOut* r, int32_t* status, uint64_t* nv, uint64_t* ni) {
    for (size_t p = 0; p < r->status.size(); ++p) {
        status[p] = r->status[p];
        nv[p] = r->v[p].size();
        ni[p] = r->i[p].size();
    }
}

void orc_batch_copy(const BatchOut* r, float* vertices, uint32_t* indices) {
    size_t ov = 0, oi = 0;
    for (size_t p = 0; p < r->status.size(); ++p) {
        for (const V2& q : r->v[p]) {
            vertices[2 * ov] = q.x;
            vertices[2 * ov + 1] = q.y;
            ++ov;
        }
        std::memcpy(indices + oi, r->i[p].data(), 4 * r->i[p].size());
        oi += r->i[p].size();
    }
}

void orc_batch_free(BatchOut* r) { delete r; }

// stroke.rs:3178-3212 find_sharp_turn: returns 1 and the split t when the quadratic turns
// sharply, else 0.
int orc_find_sharp_turn(float fx, float fy, float cx, float cy, float tx, float ty, float* t) {
    *t = 0.0f;
    return refstroke::find_sharp_turn_export(fx, fy, cx, cy, tx, ty, t) ? 1 : 0;
}
// stroke.rs:2002-2004 / 2488-2490: round join / cap subdivision depth.
uint32_t orc_arc_subdivisions(float diff, float radius, float tolerance) {
    return refstroke::arc_subdivisions_export(diff, radius, tolerance);
}

}  // extern "C"

extern "C" {
// euclid Rotation2D::transform_point applied in place (used to replay lyon's
// test_path_with_rotations, fill_tests.rs:134-150): angle accumulates in f32 by the caller.
void orc_rotate_points(float* pts, uint32_t n, float angle) {
    for (uint32_t i = 0; i < n; ++i) {
        V2 p = refgeom::rotate(angle, V2{pts[2 * i], pts[2 * i + 1]});
        pts[2 * i] = p.x;
        pts[2 * i + 1] = p.y;
    }
}
}
