// ORACLE — test infrastructure only (see refgeom.h).
//
// StrokeTessellator restated from crates/tessellation/src/stroke.rs: the fixed-width
// stroker and the variable-width one (StrokeOptions::variable_line_width).
// Function names follow the reference; each cites the lines it restates.
#include "refstroke.h"

#include <algorithm>
#include <utility>

namespace refstroke {

static const float PI_F = 3.14159265358979323846f;

// math_utils.rs:16-40
static V2 compute_normal(V2 v1, V2 v2_) {
    const float epsilon = 1e-4f;
    V2 n1 = v2(-v1.y, v1.x);
    V2 v12 = v1 + v2_;
    if (square_length(v12) < epsilon) return v2(0.0f, 0.0f);
    V2 t = normalize(v12);
    V2 n = v2(-t.y, t.x);
    float inv_len = dot(n, n1);
    if (std::fabs(inv_len) < epsilon) return n1;
    return n / inv_len;
}

// stroke.rs:2749-2753
static float circle_flattening_step(float radius, float tolerance) {
    tolerance = fminr(tolerance, radius);
    return 2.0f * std::acos((radius - tolerance) / radius);
}

// Rust `f32 as u32`: saturating, NaN -> 0.
static uint32_t f32_as_u32(float v) {
    if (!(v > 0.0f)) return 0;
    if (v >= 4294967296.0f) return 0xffffffffu;
    return (uint32_t)v;
}

// Round-join / round-cap subdivision count: (|diff| / step).ceil().log2().round() as u32
// (stroke.rs:2002-2004, 2488-2490).
static uint32_t arc_subdivisions(float diff, float radius, float tolerance) {
    float step = circle_flattening_step(radius, tolerance);
    float num_segments = std::ceil(std::fabs(diff) / step);
    return f32_as_u32(std::round(std::log2(num_segments)));
}

// stroke.rs:3178-3212
static bool find_sharp_turn(V2 from, V2 ctrl, V2 to, float* t_out) {
    V2 baseline = to - from;
    V2 v = ctrl - from;
    V2 n = v2(-baseline.y, baseline.x);
    float v_dot_b = dot(v, baseline);
    float v_dot_n = dot(v, n);
    V2 long_axis;
    if ((v_dot_b >= 0.0f && v_dot_b <= dot(baseline, baseline)) ||
        std::fabs(v_dot_n) * 2.0f >= std::fabs(v_dot_b)) {
        if (square_length(baseline) * 30.0f > square_length(v)) return false;
        long_axis = v;
    } else {
        long_axis = baseline;
    }
    float rot = -angle_from_x_axis(long_axis);
    V2 rctrl = rotate(rot, v);
    V2 rto = rotate(rot, baseline);
    return quad_local_x_extremum_t(v2(0.0f, 0.0f), rctrl, rto, t_out);
}

// stroke.rs:2320-2322
static bool miter_limit_is_exceeded(V2 normal, float miter_limit) {
    return square_length(normal) > miter_limit * miter_limit * 4.0f;
}

// stroke.rs:2279-2315
static void get_clip_intersections(V2 previous_normal, V2 next_normal, V2 normal,
                                   float clip_distance, V2* i1, V2* i2) {
    V2 clip_point = normalize(normal) * clip_distance;
    V2 clip_vec = tangent(normal);
    if (!line_intersection_f64(clip_point, clip_vec, previous_normal, tangent(previous_normal), i1))
        *i1 = normal;
    if (!line_intersection_f64(clip_point, clip_vec, next_normal, tangent(next_normal), i2))
        *i2 = normal;
}

namespace {

struct Builder {
    StrokeOptions options;
    Output& output;
    VertexData vertex;
    PointBuffer point_buffer;
    EndpointData firsts[2];
    int n_firsts = 0;
    float square_merge_threshold;
    bool may_need_empty_cap = false;
    bool error = false;
    bool vw = false;  // variable_line_width is set
    float sub_path_start_advancement = 0.0f;

    // StrokeBuilderImpl::new (stroke.rs:591-631)
    Builder(const StrokeOptions& o, Output& out) : options(o), output(out) {
        output.begin_geometry();
        square_merge_threshold = fmaxr(
            fminr(options.tolerance * options.tolerance * 0.5f,
                  options.line_width * options.line_width * 0.05f),
            1e-8f);
        vertex.half_width = options.line_width * 0.5f;
    }

    bool add(uint32_t* id) { return emit_vertex(output, vertex, id); }

    // stroke.rs:1318-1320
    bool points_are_too_close(V2 p0, V2 p1) const {
        return square_length(p0 - p1) < square_merge_threshold;
    }

    // stroke.rs:1869-1904
    void add_edge_triangles(const EndpointData& p0, const EndpointData& p1) {
        uint32_t p0_neg = p0.side_points[SIDE_NEGATIVE].next_vertex;
        uint32_t p0_pos = p0.side_points[SIDE_POSITIVE].next_vertex;
        uint32_t p1_neg = p1.side_points[SIDE_NEGATIVE].prev_vertex;
        uint32_t p1_pos = p1.side_points[SIDE_POSITIVE].prev_vertex;
        if (p0.fold[SIDE_POSITIVE]) p0_neg = p0.side_points[SIDE_POSITIVE].prev_vertex;
        if (p0.fold[SIDE_NEGATIVE]) p0_pos = p0.side_points[SIDE_NEGATIVE].prev_vertex;
        if (p1.fold[SIDE_POSITIVE]) p1_neg = p1.side_points[SIDE_POSITIVE].next_vertex;
        if (p1.fold[SIDE_NEGATIVE]) p1_pos = p1.side_points[SIDE_NEGATIVE].next_vertex;
        if (p0_neg == p1_pos) return;
        if (p0_neg != p0_pos && p0_pos != p1_pos) output.add_triangle(p0_neg, p0_pos, p1_pos);
        if (p0_neg != p1_neg && p1_pos != p1_neg) output.add_triangle(p0_neg, p1_pos, p1_neg);
    }

    // stroke.rs:2612-2653
    bool tessellate_arc(float a0, float a1, uint32_t va, uint32_t vb, uint32_t n) {
        if (n == 0) return true;
        float mid = (a0 + a1) * 0.5f;
        vertex.normal = v2(std::cos(mid), std::sin(mid));
        uint32_t id;
        if (!add(&id)) return false;
        output.add_triangle(va, id, vb);
        if (!tessellate_arc(a0, mid, va, id, n - 1)) return false;
        return tessellate_arc(mid, a1, id, vb, n - 1);
    }

    // stroke.rs:1968-2021
    bool tessellate_round_join(EndpointData& join, int side) {
        V2 center = join.position;
        float radius = join.half_width;
        V2 start_normal = join.side_points[side].prev - center;
        V2 end_normal = join.side_points[side].next - center;
        uint32_t start_vertex = join.side_points[side].prev_vertex;
        uint32_t end_vertex = join.side_points[side].next_vertex;
        float angle_sign = side == SIDE_NEGATIVE ? 1.0f : -1.0f;
        float start_angle = angle_from_x_axis(start_normal);
        float diff = angle_to(start_angle, angle_from_x_axis(end_normal));
        if (diff * angle_sign < 0.0f) diff = angle_sign * (2.0f * PI_F - std::fabs(diff));
        float end_angle = start_angle + diff;
        if (side == SIDE_NEGATIVE) {
            std::swap(start_angle, end_angle);
            std::swap(start_vertex, end_vertex);
        }
        uint32_t nsub = arc_subdivisions(diff, radius, options.tolerance);
        vertex.side = side;
        return tessellate_arc(start_angle, end_angle, start_vertex, end_vertex, nsub);
    }

    // stroke.rs:1907-1965
    bool tessellate_join(EndpointData& join) {
        bool needs[2] = {
            !join.side_points[SIDE_POSITIVE].has_single && !join.fold[SIDE_NEGATIVE],
            !join.side_points[SIDE_NEGATIVE].has_single && !join.fold[SIDE_POSITIVE],
        };
        SidePoints& P = join.side_points[SIDE_POSITIVE];
        SidePoints& N = join.side_points[SIDE_NEGATIVE];
        if (!join.fold[SIDE_POSITIVE] && !join.fold[SIDE_NEGATIVE]) {
            if (needs[0] && needs[1]) {
                output.add_triangle(P.prev_vertex, P.next_vertex, N.next_vertex);
                output.add_triangle(P.prev_vertex, N.next_vertex, N.prev_vertex);
            } else if (!needs[0] && needs[1]) {
                output.add_triangle(N.prev_vertex, P.prev_vertex, N.next_vertex);
            } else if (needs[0] && !needs[1]) {
                output.add_triangle(N.prev_vertex, P.prev_vertex, P.next_vertex);
            }
        }
        for (int side = 0; side < 2; ++side) {
            if (!needs[side]) continue;
            if (join.line_join == JOIN_ROUND) {
                if (!tessellate_round_join(join, side)) return false;
            }
        }
        return true;
    }

    // stroke.rs:2024-2055
    bool add_join_base_vertices(EndpointData& join, int side) {
        vertex.side = side;
        SidePoints& sp = join.side_points[side];
        if (sp.has_single) {
            vertex.normal = (sp.single_vertex - join.position) / join.half_width;
            uint32_t id;
            if (!add(&id)) return false;
            sp.prev_vertex = id;
            sp.next_vertex = id;
        } else {
            vertex.normal = (sp.prev - join.position) / join.half_width;
            uint32_t a;
            if (!add(&a)) return false;
            vertex.normal = (sp.next - join.position) / join.half_width;
            uint32_t b;
            if (!add(&b)) return false;
            sp.prev_vertex = a;
            sp.next_vertex = b;
        }
        return true;
    }

    // stroke.rs:1643-1733
    void compute_join_side_positions_fixed_width(const EndpointData& prev, EndpointData& join,
                                                 const EndpointData& next, float miter_limit) {
        V2 prev_tangent = join.position - prev.position;
        V2 next_tangent = next.position - join.position;
        float prev_length = length(prev_tangent);
        float next_length = length(next_tangent);
        prev_tangent = prev_tangent / prev_length;
        next_tangent = next_tangent / next_length;

        V2 normal = compute_normal(prev_tangent, next_tangent);
        int front_side;
        V2 front_normal;
        if (cross(prev_tangent, next_tangent) >= 0.0f) {
            front_side = SIDE_NEGATIVE;
            front_normal = -normal;
        } else {
            front_side = SIDE_POSITIVE;
            front_normal = normal;
        }
        int back_side = 1 - front_side;

        V2 extruded_normal = front_normal * vertex.half_width;
        bool unclipped_miter =
            (join.line_join == JOIN_MITER || join.line_join == JOIN_MITER_CLIP) &&
            !miter_limit_is_exceeded(front_normal, miter_limit);

        bool fold = false;
        bool angle_is_sharp = dot(next_tangent, prev_tangent) < 0.0f;
        if (!unclipped_miter && angle_is_sharp) {
            float d_next = dot(extruded_normal, -next_tangent) - next_length;
            float d_prev = dot(extruded_normal, prev_tangent) - prev_length;
            if (fminr(d_next, d_prev) > 0.0f || square_length(normal) < 1e-5f) {
                join.fold[front_side] = true;
                fold = true;
            }
        }

        V2 n0 = v2(-prev_tangent.y, prev_tangent.x) * vertex.half_width;
        V2 n1 = v2(-next_tangent.y, next_tangent.x) * vertex.half_width;
        join.side_points[SIDE_POSITIVE].prev = join.position + n0;
        join.side_points[SIDE_POSITIVE].next = join.position + n1;
        join.side_points[SIDE_NEGATIVE].prev = join.position - n0;
        join.side_points[SIDE_NEGATIVE].next = join.position - n1;

        if (!fold) {
            V2 miter_pos[2] = {join.position + normal * vertex.half_width,
                               join.position - normal * vertex.half_width};
            join.side_points[back_side].has_single = true;
            join.side_points[back_side].single_vertex = miter_pos[back_side];
            if (unclipped_miter) {
                join.side_points[front_side].has_single = true;
                join.side_points[front_side].single_vertex = miter_pos[front_side];
            } else if (join.line_join == JOIN_MITER_CLIP) {
                V2 a = join.side_points[front_side].prev - join.position;
                V2 b = join.side_points[front_side].next - join.position;
                V2 pn, nn;
                get_clip_intersections(a, b, front_normal, miter_limit * vertex.half_width, &pn,
                                       &nn);
                join.side_points[front_side].prev = join.position + pn;
                join.side_points[front_side].next = join.position + nn;
            }
        }
    }

    // stroke.rs:1743-1806.  Returns false on a builder error; *skipped mirrors Ok(true).
    bool flattened_step(EndpointData& prev, EndpointData& join, EndpointData& next,
                        bool* skipped = nullptr) {
        if (skipped) *skipped = false;
        V2 prev_edge = join.position - prev.position;
        float prev_length = length(prev_edge);
        V2 prev_tangent = prev_edge / prev_length;
        V2 next_edge = next.position - join.position;
        float next_length = length(next_edge);
        V2 next_tangent = next_edge / next_length;
        V2 normal = compute_normal(prev_tangent, next_tangent);
        if (std::isnan(join.advancement)) join.advancement = prev.advancement + prev_length;
        if (std::isnan(next.advancement)) next.advancement = join.advancement + next_length;

        V2 p0 = join.position + normal * vertex.half_width;
        join.side_points[SIDE_POSITIVE].prev = p0;
        join.side_points[SIDE_POSITIVE].next = p0;
        join.side_points[SIDE_POSITIVE].has_single = true;
        join.side_points[SIDE_POSITIVE].single_vertex = p0;
        V2 p1 = join.position - normal * vertex.half_width;
        join.side_points[SIDE_NEGATIVE].prev = p1;
        join.side_points[SIDE_NEGATIVE].next = p1;
        join.side_points[SIDE_NEGATIVE].has_single = true;
        join.side_points[SIDE_NEGATIVE].single_vertex = p1;

        V2 v0 = p0 - prev.side_points[SIDE_POSITIVE].next;
        V2 v1 = p1 - prev.side_points[SIDE_NEGATIVE].next;
        if (dot(prev_edge, v0) < 0.0f && dot(prev_edge, v1) < 0.0f) {
            if (skipped) *skipped = true;
            return true;
        }

        vertex.normal = normal;
        vertex.side = SIDE_POSITIVE;
        uint32_t pos_vertex, neg_vertex;
        if (!add(&pos_vertex)) return false;
        vertex.normal = -normal;
        vertex.side = SIDE_NEGATIVE;
        if (!add(&neg_vertex)) return false;
        join.side_points[SIDE_POSITIVE].prev_vertex = pos_vertex;
        join.side_points[SIDE_POSITIVE].next_vertex = pos_vertex;
        join.side_points[SIDE_NEGATIVE].prev_vertex = neg_vertex;
        join.side_points[SIDE_NEGATIVE].next_vertex = neg_vertex;
        return true;
    }

    // stroke.rs:1521-1632.  Returns false on error; *added mirrors the Ok(bool).
    bool step_impl(EndpointData next, bool* added) {
        size_t count = point_buffer.count();
        if (count > 0) {
            if (points_are_too_close(point_buffer.last().position, next.position)) {
                if (count == 1) may_need_empty_cap = true;
                *added = false;
                return true;
            }
            if (count == 1) {
                EndpointData& first = point_buffer.last();
                V2 edge = next.position - first.position;
                float len = length(edge);
                V2 t = edge / len;
                V2 n = v2(-t.y, t.x) * next.half_width;
                first.side_points[SIDE_POSITIVE].next = first.position + n;
                first.side_points[SIDE_NEGATIVE].next = first.position - n;
            }
        }
        if (count > 1) {
            EndpointData& prev = point_buffer.second_last();
            EndpointData& join = point_buffer.last_of_two();
            vertex.position_on_path = join.position;
            vertex.half_width = join.half_width;
            bool fast_path = false;
            if (join.is_flattening_step) {
                V2 v0 = join.position - prev.position;
                V2 v1 = next.position - join.position;
                fast_path = dot(v0, v1) > 0.0f;
            }
            if (fast_path) {
                join.line_join = JOIN_MITER;
                if (!flattened_step(prev, join, next)) return false;
            } else {
                compute_join_side_positions_fixed_width(prev, join, next, options.miter_limit);
                if (!add_join_base_vertices(join, SIDE_NEGATIVE)) return false;
                if (!add_join_base_vertices(join, SIDE_POSITIVE)) return false;
            }
            if (count > 2) add_edge_triangles(prev, join);
            if (!tessellate_join(join)) return false;
            if (count == 2) {
                firsts[0] = prev;
                firsts[1] = join;
                n_firsts = 2;
            }
        }
        point_buffer.push(next);
        *added = true;
        return true;
    }

    // stroke.rs:1809-1836
    static void compute_edge_attachment_positions(EndpointData& p0, EndpointData& p1) {
        V2 edge = p1.position - p0.position;
        float d = length(edge);
        float edge_angle = angle_from_x_axis(edge);
        // sin(vwidth_angle) = (hw1 - hw0) / d; NaN (|.| > 1) counts as 0
        float sin_vwidth_angle = (p1.half_width - p0.half_width) / d;
        float vwidth_angle = std::asin(sin_vwidth_angle);
        if (std::isnan(vwidth_angle)) vwidth_angle = 0.0f;
        compute_side_attachment_positions(p0, p1, edge_angle, vwidth_angle, SIDE_POSITIVE);
        compute_side_attachment_positions(p0, p1, edge_angle, vwidth_angle, SIDE_NEGATIVE);
        if (std::isnan(p1.advancement)) p1.advancement = p0.advancement + d;
    }

    // stroke.rs:1838-1866 (side_sign: stroke.rs:2324-2330)
    static void compute_side_attachment_positions(EndpointData& p0, EndpointData& p1,
                                                  float edge_angle, float vwidth_angle,
                                                  int side) {
        float nl = side == SIDE_NEGATIVE ? -1.0f : 1.0f;
        float normal_angle = edge_angle + nl * (PI_F * 0.5f + vwidth_angle);
        V2 normal = v2(std::cos(normal_angle), std::sin(normal_angle));
        p0.side_points[side].next = p0.position + normal * p0.half_width;
        p1.side_points[side].prev = p1.position + normal * p1.half_width;
    }

    // stroke.rs:2060-2128
    static void compute_join_side_positions(const EndpointData& prev, EndpointData& join,
                                            const EndpointData& next, float miter_limit,
                                            int side) {
        float sign = side == SIDE_NEGATIVE ? -1.0f : 1.0f;
        V2 v0 = normalize(join.side_points[side].prev - prev.side_points[side].next);
        V2 v1 = normalize(next.side_points[side].prev - join.side_points[side].next);
        bool inward = cross(v0, v1) * sign > 0.0f;
        bool forward = dot(v0, v1) > 0.0f;
        V2 normal = compute_normal(v0, v1) * sign;
        V2 path_v0 = normalize(join.position - prev.position);
        V2 path_v1 = normalize(next.position - join.position);
        bool normal_same_side = dot(v0 + v1, path_v0 + path_v1) >= 0.0f;
        bool angle_is_sharp = inward && !forward && normal_same_side;
        if (angle_is_sharp) {
            V2 extruded_normal = normal * join.half_width;
            float prev_length = join.advancement - prev.advancement;
            float next_length = next.advancement - join.advancement;
            float d_next = dot(extruded_normal, v1) - next_length;
            float d_prev = dot(extruded_normal, -v0) - prev_length;
            if (fminr(d_next, d_prev) > 0.0f || square_length(normal) < 1e-5f)
                join.fold[side] = true;
        }
        bool concave = inward && normal_same_side && !join.fold[side];
        if (concave || ((join.line_join == JOIN_MITER || join.line_join == JOIN_MITER_CLIP) &&
                        !miter_limit_is_exceeded(normal, miter_limit))) {
            join.side_points[side].has_single = true;
            join.side_points[side].single_vertex = join.position + normal * join.half_width;
        } else if (join.line_join == JOIN_MITER_CLIP) {
            V2 n0 = join.side_points[side].prev - join.position;
            V2 n1 = join.side_points[side].next - join.position;
            V2 pn, nn;
            get_clip_intersections(n0, n1, normal, miter_limit * join.half_width, &pn, &nn);
            join.side_points[side].prev = join.position + pn;
            join.side_points[side].next = join.position + nn;
        }
    }

    // stroke.rs:1380-1512 (the variable-width step)
    bool step_impl_vw(EndpointData next, bool* added) {
        size_t count = point_buffer.count();
        if (count > 0 && points_are_too_close(point_buffer.last().position, next.position)) {
            if (count == 1) may_need_empty_cap = true;
            *added = false;
            return true;
        }
        if (count > 0) {
            EndpointData& join = point_buffer.last();
            if (!join.is_flattening_step || !next.is_flattening_step)
                compute_edge_attachment_positions(join, next);
        }
        bool skip = false;
        if (count > 1) {
            EndpointData& prev = point_buffer.second_last();
            EndpointData& join = point_buffer.last_of_two();
            vertex.src = join.src;
            vertex.position_on_path = join.position;
            vertex.half_width = join.half_width;
            bool fast_path = false;
            if (join.is_flattening_step) {
                V2 v0 = join.position - prev.position;
                V2 v1 = next.position - join.position;
                fast_path = dot(v0, v1) > 0.0f;
            }
            if (fast_path) {
                join.line_join = JOIN_MITER;
                if (!flattened_step(prev, join, next, &skip)) return false;
            } else {
                compute_join_side_positions(prev, join, next, options.miter_limit, SIDE_POSITIVE);
                compute_join_side_positions(prev, join, next, options.miter_limit, SIDE_NEGATIVE);
                // no folding when the other side is concave
                if (join.side_points[SIDE_POSITIVE].has_single) join.fold[SIDE_NEGATIVE] = false;
                if (join.side_points[SIDE_NEGATIVE].has_single) join.fold[SIDE_POSITIVE] = false;
                if (!add_join_base_vertices(join, SIDE_NEGATIVE)) return false;
                if (!add_join_base_vertices(join, SIDE_POSITIVE)) return false;
            }
            if (!skip) {
                if (count > 2) add_edge_triangles(prev, join);
                if (!tessellate_join(join)) return false;
                if (count == 2) {
                    firsts[0] = prev;
                    firsts[1] = join;
                    n_firsts = 2;
                }
            }
        }
        if (skip) point_buffer.replace_last(next);
        else point_buffer.push(next);
        *added = true;
        return true;
    }

    bool step_any(const EndpointData& next, bool* added) {
        return vw ? step_impl_vw(next, added) : step_impl(next, added);
    }

    void step(const EndpointData& next) {
        bool added;
        if (!step_any(next, &added)) error = true;
    }

    EndpointData make_point(V2 pos, bool is_flattening_step) const {
        EndpointData e;
        e.position = pos;
        e.half_width = options.line_width * 0.5f;
        e.line_join = options.line_join;
        e.is_flattening_step = is_flattening_step;
        return e;
    }

    // stroke.rs:2526-2560
    bool tessellate_empty_square_cap(V2 position) {
        vertex.position_on_path = position;
        uint32_t a, b, c, d;
        vertex.normal = v2(1.0f, 1.0f);
        vertex.side = SIDE_NEGATIVE;
        if (!add(&a)) return false;
        vertex.normal = v2(1.0f, -1.0f);
        vertex.side = SIDE_POSITIVE;
        if (!add(&b)) return false;
        vertex.normal = v2(-1.0f, -1.0f);
        vertex.side = SIDE_POSITIVE;
        if (!add(&c)) return false;
        vertex.normal = v2(-1.0f, 1.0f);
        vertex.side = SIDE_NEGATIVE;
        if (!add(&d)) return false;
        output.add_triangle(a, b, c);
        output.add_triangle(a, c, d);
        return true;
    }

    // stroke.rs:2459-2524
    bool tessellate_round_cap(V2 center, float radius, V2 start_normal, uint32_t start_vertex,
                              uint32_t end_vertex, V2 edge_normal, bool is_start) {
        if (radius < options.tolerance) return true;
        int first_side = (is_start ^ (cross(edge_normal, start_normal) >= 0.0f)) ? SIDE_POSITIVE
                                                                                  : SIDE_NEGATIVE;
        float start_angle = angle_from_x_axis(start_normal);
        float diff = angle_to(start_angle, angle_from_x_axis(edge_normal));
        float mid_angle = start_angle + diff;
        float end_angle = mid_angle + diff;
        uint32_t nsub = arc_subdivisions(diff, radius, options.tolerance);
        vertex.position_on_path = center;
        vertex.half_width = radius;
        vertex.side = first_side;
        vertex.normal = normalize(edge_normal);
        uint32_t mid_vertex;
        if (!add(&mid_vertex)) return false;
        output.add_triangle(start_vertex, mid_vertex, end_vertex);
        if (!tessellate_arc(start_angle, mid_angle, start_vertex, mid_vertex, nsub)) return false;
        vertex.side = 1 - first_side;
        return tessellate_arc(mid_angle, end_angle, mid_vertex, end_vertex, nsub);
    }

    // stroke.rs:2562-2609
    bool tessellate_empty_round_cap(V2 center) {
        float radius = vertex.half_width;
        vertex.position_on_path = center;
        uint32_t left_id, right_id;
        vertex.normal = v2(-1.0f, 0.0f);
        vertex.side = SIDE_POSITIVE;
        if (!add(&left_id)) return false;
        vertex.normal = v2(1.0f, 0.0f);
        vertex.side = SIDE_NEGATIVE;
        if (!add(&right_id)) return false;
        if (!tessellate_round_cap(center, radius, v2(-1.0f, 0.0f), left_id, right_id,
                                  v2(0.0f, 1.0f), true))
            return false;
        return tessellate_round_cap(center, radius, v2(1.0f, 0.0f), right_id, left_id,
                                    v2(0.0f, -1.0f), false);
    }

    // stroke.rs:1282-1316
    bool tessellate_empty_cap() {
        EndpointData& p = point_buffer.get(0);
        vertex.src = p.src;
        vertex.half_width = p.half_width;
        if (options.start_cap == CAP_SQUARE) return tessellate_empty_square_cap(p.position);
        if (options.start_cap == CAP_ROUND) return tessellate_empty_round_cap(p.position);
        return true;
    }

    // stroke.rs:2130-2206
    bool tessellate_last_edge(const EndpointData& p0, EndpointData& p1, bool is_first_edge) {
        V2 v = p1.position - p0.position;
        p1.advancement = p0.advancement + length(v);
        vertex.src = p1.src;
        vertex.position_on_path = p1.position;
        vertex.half_width = p1.half_width;
        for (int side = 0; side < 2; ++side) {
            V2 side_position = p1.side_points[side].prev;
            bool has_clip = options.end_cap != CAP_ROUND;
            float clip = options.end_cap == CAP_SQUARE ? p1.half_width : 0.0f;
            if (has_clip) {
                V2 normal = normalize(p1.position - p0.position);
                V2 inter;
                if (!line_intersection_f64(p1.position + normal * clip, tangent(normal),
                                           side_position,
                                           side_position - p0.side_points[side].next, &inter))
                    inter = p1.side_points[side].prev;
                p1.side_points[side].prev = inter;
            }
            vertex.side = side;
            vertex.normal = (p1.side_points[side].prev - p1.position) / p1.half_width;
            uint32_t id;
            if (!add(&id)) return false;
            p1.side_points[side].prev_vertex = id;
        }
        if (!is_first_edge) add_edge_triangles(p0, p1);
        if (options.end_cap == CAP_ROUND) {
            return tessellate_round_cap(p1.position, p1.half_width,
                                        p1.side_points[SIDE_POSITIVE].prev - p1.position,
                                        p1.side_points[SIDE_POSITIVE].prev_vertex,
                                        p1.side_points[SIDE_NEGATIVE].prev_vertex, v, false);
        }
        return true;
    }

    // stroke.rs:2208-2276
    bool tessellate_first_edge(EndpointData& first, const EndpointData& second) {
        vertex.src = first.src;
        vertex.position_on_path = first.position;
        vertex.half_width = first.half_width;
        for (int side = 0; side < 2; ++side) {
            V2 side_position = first.side_points[side].next;
            bool has_clip = options.start_cap != CAP_ROUND;
            float clip = options.start_cap == CAP_SQUARE ? first.half_width : 0.0f;
            if (has_clip) {
                V2 normal = normalize(first.position - second.position);
                V2 inter;
                if (!line_intersection_f64(first.position + normal * clip, tangent(normal),
                                           side_position,
                                           side_position - second.side_points[side].prev,
                                           &inter))
                    inter = first.side_points[side].next;
                side_position = inter;
            }
            vertex.side = side;
            vertex.normal = (side_position - first.position) / first.half_width;
            uint32_t id;
            if (!add(&id)) return false;
            first.side_points[side].next_vertex = id;
        }
        add_edge_triangles(first, second);
        if (options.start_cap == CAP_ROUND) {
            return tessellate_round_cap(first.position, first.half_width,
                                        first.side_points[SIDE_NEGATIVE].next - first.position,
                                        first.side_points[SIDE_NEGATIVE].next_vertex,
                                        first.side_points[SIDE_POSITIVE].next_vertex,
                                        first.position - second.position, true);
        }
        return true;
    }

    // stroke.rs:1322-1378
    bool end_with_caps() {
        size_t count = point_buffer.count();
        if (may_need_empty_cap && count == 1) return tessellate_empty_cap();
        if (count >= 2) {
            EndpointData p0 = point_buffer.second_last();
            EndpointData p1 = point_buffer.last_of_two();
            if (!vw) {  // the fixed-width step leaves the last edge's side points unset
                V2 t = normalize(p1.position - p0.position);
                V2 n = v2(-t.y, t.x) * p1.half_width;
                p1.side_points[SIDE_POSITIVE].prev = p1.position + n;
                p1.side_points[SIDE_NEGATIVE].prev = p1.position - n;
            }
            bool is_first = count == 2;
            if (!tessellate_last_edge(p0, p1, is_first)) return false;
            sub_path_start_advancement = p1.advancement;
            if (count > 2) {
                p0 = firsts[0];
                p1 = firsts[1];
            }
            if (!tessellate_first_edge(p0, p1)) return false;
        }
        return true;
    }

    // stroke.rs:1206-1279
    bool close() {
        if (point_buffer.count() == 1) {
            if (!tessellate_empty_cap()) return false;
        }
        if (point_buffer.count() <= 2) return true;
        EndpointData p = firsts[0];
        p.advancement = NAN;  // recomputed, not the sub-path start's
        bool segment_added;
        if (!step_any(p, &segment_added)) return false;
        if (!segment_added) point_buffer.last().position = p.position;
        if (n_firsts >= 2) {
            EndpointData p2 = firsts[1];
            bool dummy;
            if (!step_any(p2, &dummy)) return false;
            EndpointData& p0 = point_buffer.second_last();
            EndpointData& p1 = point_buffer.last_of_two();
            vertex.src = p0.src;
            vertex.position_on_path = p0.position;
            vertex.half_width = p0.half_width;
            for (int side = 0; side < 2; ++side) {
                vertex.side = side;
                V2 sp = p0.side_points[side].has_single ? p0.side_points[side].single_vertex
                                                        : p0.side_points[side].next;
                vertex.normal = (sp - p0.position) / p0.half_width;
                uint32_t id;
                if (!add(&id)) return false;
                p0.side_points[side].next_vertex = id;
            }
            add_edge_triangles(p0, p1);
        }
        return true;
    }

    // stroke.rs:1179-1193
    void end(bool close_) {
        may_need_empty_cap = may_need_empty_cap || (close_ && point_buffer.count() == 1);
        bool ok = (close_ && point_buffer.count() > 2) ? close() : end_with_caps();
        if (!ok) error = true;
        point_buffer.clear();
        n_firsts = 0;
    }

    // stroke.rs:2755-2779 + quadratic_bezier_to_fw (:1108-1143)
    void quadratic_to(V2 from, V2 ctrl, V2 to) {
        float t_split;
        float tol = options.tolerance;
        if (find_sharp_turn(from, ctrl, to, &t_split)) {
            V2 a[3], b[3];
            quad_split(from, ctrl, to, t_split, a, b);
            quad_flatten_with_t(a[0], a[1], a[2], tol, [&](V2, V2 p, float, float t1) {
                step(make_point(p, t1 != 1.0f));
            });
            quad_flatten_with_t(b[0], b[1], b[2], tol, [&](V2, V2 p, float, float t1) {
                step(make_point(p, t1 != 1.0f));
            });
        } else {
            quad_flatten_with_t(from, ctrl, to, tol,
                                [&](V2, V2 p, float, float t1) { step(make_point(p, t1 != 1.0f)); });
        }
    }

    // quadratic_bezier_to with widths (stroke.rs:997-1030) over flatten_quad
    // (stroke.rs:2755-2779): the half width at each flattened point interpolates the
    // endpoints' widths at its t on the whole curve.
    void quadratic_to_vw(V2 from, V2 ctrl, V2 to, float w0, float w1, uint32_t id0,
                         uint32_t id1) {
        float t_split;
        float tol = options.tolerance;
        auto at = [&](V2 p, float t, bool flat) {
            EndpointData e = make_point(p, flat);
            e.half_width = (w0 * (1.0f - t) + w1 * t) * 0.5f;
            e.src = edge_src(id0, id1, t);
            step(e);
        };
        if (find_sharp_turn(from, ctrl, to, &t_split)) {
            V2 a[3], b[3];
            quad_split(from, ctrl, to, t_split, a, b);
            quad_flatten_with_t(a[0], a[1], a[2], tol, [&](V2, V2 p, float, float t1) {
                at(p, t1 * t_split, t1 != 1.0f);
            });
            quad_flatten_with_t(b[0], b[1], b[2], tol, [&](V2, V2 p, float, float t1) {
                at(p, t_split + t1 * (1.0f - t_split), t1 != 1.0f);
            });
        } else {
            quad_flatten_with_t(from, ctrl, to, tol,
                                [&](V2, V2 p, float, float t1) { at(p, t1, t1 != 1.0f); });
        }
    }

    // cubic_bezier_to with widths (stroke.rs:1032-1067)
    void cubic_to_vw(V2 from, V2 c1, V2 c2, V2 to, float w0, float w1, uint32_t id0,
                     uint32_t id1) {
        cubic_flatten_with_t(from, c1, c2, to, options.tolerance,
                             [&](V2, V2 p, float, float t1) {
                                 EndpointData e = make_point(p, t1 != 1.0f);
                                 e.half_width = (w0 * (1.0f - t1) + w1 * t1) * 0.5f;
                                 e.src = edge_src(id0, id1, t1);
                                 step(e);
                             });
    }

    // quadratic_bezier_to / cubic_bezier_to (stroke.rs:1005-1013, 1041-1050): the
    // endpoint at t == 1, else the edge at t
    static Source edge_src(uint32_t id0, uint32_t id1, float t) {
        Source r;
        r.from = t == 1.0f ? id1 : id0;
        r.to = id1;
        r.t = t;
        r.edge = t != 1.0f;
        return r;
    }
    static Source endpoint_src(uint32_t id) {
        Source r;
        r.from = r.to = id;
        return r;
    }

    // cubic_bezier_to_fw (stroke.rs:1145-1177)
    void cubic_to(V2 from, V2 c1, V2 c2, V2 to) {
        cubic_flatten_with_t(from, c1, c2, to, options.tolerance,
                             [&](V2, V2 p, float, float t1) { step(make_point(p, t1 != 1.0f)); });
    }
};

}  // namespace

// StrokeTessellator::tessellate_path -> tessellate_fw (stroke.rs:164-177, 872-952)
int stroke_path(const refpath::PathView& path, const StrokeOptions& opts, Output& out) {
    Builder b(opts, out);
    refpath::Iter it(path);
    refpath::Event e;
    V2 current{NAN, NAN};
    while (it.next(e)) {
        switch (e.kind) {
            case refpath::EV_BEGIN:
                current = e.to;
                b.may_need_empty_cap = false;
                b.step(b.make_point(e.to, false));
                break;
            case refpath::EV_LINE:
                current = e.to;
                b.step(b.make_point(e.to, false));
                break;
            case refpath::EV_QUAD: {
                V2 from = current;
                current = e.to;
                b.quadratic_to(from, e.ctrl1, e.to);
                break;
            }
            case refpath::EV_CUBIC: {
                V2 from = current;
                current = e.to;
                b.cubic_to(from, e.ctrl1, e.ctrl2, e.to);
                break;
            }
            case refpath::EV_END:
                b.end(e.close);
                break;
        }
        if (b.error) {
            out.abort_geometry();
            return TOO_MANY_VERTICES;
        }
    }
    return OK;
}

// StrokeTessellator::tessellate_path -> tessellate_with_ids_vw (stroke.rs:653-767): each
// endpoint's half width is line_width * its width factor * 0.5; curves interpolate the
// full widths of their two endpoints.
int stroke_path_vw(const refpath::PathView& path, const StrokeOptions& opts,
                   const float* widths, Output& out, uint32_t point_base) {
    Builder b(opts, out);
    b.vw = true;
    refpath::Iter it(path);
    refpath::Event e;
    V2 current{NAN, NAN};
    float cur_w = NAN;  // width factor of the current endpoint
    uint32_t cur_id = 0;  // its point index (+ point_base: the ids of the attribute rows)
    auto width_at = [&](uint32_t pi) { return pi < path.n_points ? widths[pi] : NAN; };
    while (it.next(e)) {
        const uint32_t pi = it.pi - 1;  // the event's endpoint (Begin, Line, Quad, Cubic)
        switch (e.kind) {
            case refpath::EV_BEGIN: {
                current = e.to;
                cur_w = width_at(pi);
                b.may_need_empty_cap = false;
                EndpointData p = b.make_point(e.to, false);
                p.half_width = opts.line_width * cur_w * 0.5f;
                p.advancement = b.sub_path_start_advancement;
                cur_id = point_base + pi;
                p.src = Builder::endpoint_src(cur_id);
                b.step(p);
                break;
            }
            case refpath::EV_LINE: {
                current = e.to;
                cur_w = width_at(pi);
                EndpointData p = b.make_point(e.to, false);
                p.half_width = opts.line_width * cur_w * 0.5f;
                cur_id = point_base + pi;
                p.src = Builder::endpoint_src(cur_id);
                b.step(p);
                break;
            }
            case refpath::EV_QUAD: {
                V2 from = current;
                current = e.to;
                const float w0 = opts.line_width * cur_w;
                cur_w = width_at(pi);
                const uint32_t id0 = cur_id;
                cur_id = point_base + pi;
                b.quadratic_to_vw(from, e.ctrl1, e.to, w0, opts.line_width * cur_w, id0, cur_id);
                break;
            }
            case refpath::EV_CUBIC: {
                V2 from = current;
                current = e.to;
                const float w0 = opts.line_width * cur_w;
                cur_w = width_at(pi);
                const uint32_t id0 = cur_id;
                cur_id = point_base + pi;
                b.cubic_to_vw(from, e.ctrl1, e.ctrl2, e.to, w0, opts.line_width * cur_w, id0,
                              cur_id);
                break;
            }
            case refpath::EV_END:
                b.end(e.close);
                break;
        }
        if (b.error) {
            out.abort_geometry();
            return TOO_MANY_VERTICES;
        }
    }
    return OK;
}

// Exposed for the tests that pin these rules against the geometry they encode.
bool find_sharp_turn_export(float fx, float fy, float cx, float cy, float tx, float ty,
                            float* t_out) {
    return find_sharp_turn(v2(fx, fy), v2(cx, cy), v2(tx, ty), t_out);
}
uint32_t arc_subdivisions_export(float diff, float radius, float tolerance) {
    return arc_subdivisions(diff, radius, tolerance);
}

}  // namespace refstroke
