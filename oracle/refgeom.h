// ORACLE — test infrastructure only. Nothing in the product (lyon_amd/) may include,
// link or call this file; only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg use it, as the checker.
//
// CPU restatement of the f32/f64 vector arithmetic that lyon's hot path gets from
// euclid 0.22.9 (Cargo.lock:603-605, not vendored) and of lyon_geom's curve
// flattening.  Every operation keeps the operand order of the Rust source so that
// results are bit-identical under IEEE-754 with no FMA contraction
// (build with -ffp-contract=off, no -ffast-math).
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace refgeom {

struct V2 {
    float x, y;
};
inline V2 v2(float x, float y) { return V2{x, y}; }
inline V2 operator+(V2 a, V2 b) { return V2{a.x + b.x, a.y + b.y}; }
inline V2 operator-(V2 a, V2 b) { return V2{a.x - b.x, a.y - b.y}; }
inline V2 operator-(V2 a) { return V2{-a.x, -a.y}; }
inline V2 operator*(V2 a, float s) { return V2{a.x * s, a.y * s}; }
inline V2 operator/(V2 a, float s) { return V2{a.x / s, a.y / s}; }
inline bool operator==(V2 a, V2 b) { return a.x == b.x && a.y == b.y; }
inline bool operator!=(V2 a, V2 b) { return !(a == b); }
// euclid Vector2D::dot / cross / square_length / length / normalize
inline float dot(V2 a, V2 b) { return a.x * b.x + a.y * b.y; }
inline float cross(V2 a, V2 b) { return a.x * b.y - a.y * b.x; }
inline float square_length(V2 a) { return a.x * a.x + a.y * a.y; }
inline float length(V2 a) { return std::sqrt(square_length(a)); }
inline V2 normalize(V2 a) { return a / length(a); }
// euclid Point2D::lerp: one_t * a + t * b per component.
inline V2 lerp(V2 a, V2 b, float t) {
    float one_t = 1.0f - t;
    return V2{one_t * a.x + t * b.x, one_t * a.y + t * b.y};
}
// lyon_geom utils.rs:15-17 tangent(v) = (-v.y, v.x)
inline V2 tangent(V2 v) { return V2{-v.y, v.x}; }

// Rust f32::max / f32::min (IEEE maxNum/minNum: a NaN operand yields the other).
inline float fmaxr(float a, float b) { return std::fmax(a, b); }
inline float fminr(float a, float b) { return std::fmin(a, b); }

// euclid Trig::fast_atan2 (euclid 0.22.9 trig.rs), the polynomial atan2 approximation
// published at math.stackexchange.com/questions/1098487 (answer 1105038); used by
// Vector2D::angle_from_x_axis.
inline float fast_atan2(float y, float x) {
    const float PI = 3.14159265358979323846f;
    const float FRAC_PI_2 = 1.57079632679489661923f;
    float x_abs = std::fabs(x);
    float y_abs = std::fabs(y);
    float a = fminr(x_abs, y_abs) / fmaxr(x_abs, y_abs);
    float s = a * a;
    float result = ((-0.0464964749f * s + 0.15931422f) * s - 0.327622764f) * s * a + a;
    if (y_abs > x_abs) result = FRAC_PI_2 - result;
    if (x < 0.0f) result = PI - result;
    if (y < 0.0f) result = -result;
    return result;
}
inline float angle_from_x_axis(V2 v) { return fast_atan2(v.y, v.x); }
// euclid Angle::angle_to: d = (to - self) % 2pi; (2d % 2pi) - d
inline float angle_to(float self_r, float to_r) {
    const float PI = 3.14159265358979323846f;
    float two = 1.0f + 1.0f;
    float mx = PI * two;
    float d = std::fmod(to_r - self_r, mx);
    return std::fmod(two * d, mx) - d;
}
// euclid Rotation2D::transform_point with (sin, cos) = angle.sin_cos()
inline V2 rotate(float angle, V2 p) {
    float s = std::sin(angle), c = std::cos(angle);
    return V2{p.x * c - p.y * s, p.y * c + p.x * s};
}

// f64 mirror for Line::intersection (line.rs:562-578) and LineSegment<f64>.
struct D2 {
    double x, y;
};
inline D2 d2(V2 v) { return D2{(double)v.x, (double)v.y}; }
inline D2 operator+(D2 a, D2 b) { return D2{a.x + b.x, a.y + b.y}; }
inline D2 operator-(D2 a, D2 b) { return D2{a.x - b.x, a.y - b.y}; }
inline D2 operator*(D2 a, double s) { return D2{a.x * s, a.y * s}; }
inline double crossd(D2 a, D2 b) { return a.x * b.y - a.y * b.x; }

// Line::intersection (line.rs:562-578) evaluated on the f64 cast of two f32 lines,
// result cast back to f32.  Returns false when |det| <= 1e-8 (f64 EPSILON).
inline bool line_intersection_f64(V2 p1, V2 v1, V2 p2, V2 v2_, V2* out) {
    D2 sp = d2(p1), sv = d2(v1), op = d2(p2), ov = d2(v2_);
    double det = crossd(sv, ov);
    if (std::fabs(det) <= 1e-8) return false;
    double inv_det = 1.0 / det;
    D2 self_p2 = sp + sv;
    D2 other_p2 = op + ov;
    double a = crossd(sp, self_p2);
    double b = crossd(op, other_p2);
    out->x = (float)((b * sv.x - a * ov.x) * inv_det);
    out->y = (float)((b * sv.y - a * ov.y) * inv_det);
    return true;
}

// ---------------------------------------------------------------------------------
// lyon_geom flattening
// ---------------------------------------------------------------------------------

// quadratic_bezier.rs:943-974 / cubic_bezier.rs:1369-1403: the shared tail that turns
// err4 into a (possibly non-integer) segment count.
inline float wang_tail(float err4) {
    static const float LUT[24] = {1.0f,      16.0f,     81.0f,     256.0f,    625.0f,
                                  1296.0f,   2401.0f,   4096.0f,   6561.0f,   10000.0f,
                                  14641.0f,  20736.0f,  28561.0f,  38416.0f,  50625.0f,
                                  65536.0f,  83521.0f,  104976.0f, 130321.0f, 160000.0f,
                                  194481.0f, 234256.0f, 279841.0f, 331776.0f};
    if (err4 <= 331776.0f) {
        for (int i = 0; i < 24; ++i) {
            if (err4 <= LUT[i]) return (float)(i + 1);
        }
    }
    return fmaxr(std::sqrt(std::sqrt(err4)), 1.0f);
}

// quadratic_bezier.rs:943-974
inline float quad_wang(V2 from, V2 ctrl, V2 to, float tol) {
    V2 l = ((from - ctrl * 2.0f) + to) * 2.0f;
    float d = 1.0f / (8.0f * tol);
    float err4 = dot(l, l) * d * d;
    return wang_tail(err4);
}

// cubic_bezier.rs:1369-1403
inline float cubic_wang(V2 from, V2 c1, V2 c2, V2 to, float tol) {
    V2 v1 = ((from - c1 * 2.0f) + c2) * 6.0f;
    V2 v2_ = ((c1 - c2 * 2.0f) + to) * 6.0f;
    float l = fmaxr(dot(v1, v1), dot(v2_, v2_));
    float d = 1.0f / (8.0f * tol);
    float err4 = l * d * d;
    return wang_tail(err4);
}

// num_traits ToPrimitive::to_u32 on f32 followed by unwrap_or(1): truncation when the
// value is in (-1, 2^32), otherwise (NaN, inf, out of range) 1.
inline uint32_t to_u32_or_1(float n) {
    if (!(n > -1.0f) || !(n < 4294967296.0f)) return 1u;
    return (uint32_t)n;
}

// Calls cb(to, t0, t1) for every flattened segment, following
// QuadraticBezierSegment::for_each_flattened_with_t (quadratic_bezier.rs:368-391).
template <class F>
inline void quad_flatten_with_t(V2 from, V2 ctrl, V2 to, float tol, F&& cb) {
    float n = quad_wang(from, ctrl, to, tol);
    float step = 1.0f / n;
    float t0 = 0.0f;
    // polynomial_form (:891-900)
    V2 a0 = from, a1 = (ctrl - from) * 2.0f, a2 = (from + to) - ctrl * 2.0f;
    uint32_t cnt = to_u32_or_1(n) - 1;
    V2 f = from;
    for (uint32_t i = 0; i < cnt; ++i) {
        float t1 = t0 + step;
        // Horner (:928-938)
        V2 v = a0;
        float tt = t1;
        v = v + a1 * tt;
        tt = tt * t1;
        v = v + a2 * tt;
        cb(f, v, t0, t1);
        f = v;
        t0 = t1;
    }
    cb(f, to, t0, 1.0f);
}

// CubicBezierSegment::for_each_flattened_with_t (cubic_bezier.rs:590-614)
template <class F>
inline void cubic_flatten_with_t(V2 from, V2 c1, V2 c2, V2 to, float tol, F&& cb) {
    float n = cubic_wang(from, c1, c2, to, tol);
    float step = 1.0f / n;
    float t0 = 0.0f;
    // polynomial_form (:1318-1325)
    V2 a0 = from;
    V2 a1 = (c1 - from) * 3.0f;
    V2 a2 = (from * 3.0f - c1 * 6.0f) + c2 * 3.0f;
    V2 a3 = (to - from) + (c1 - c2) * 3.0f;
    uint32_t cnt = to_u32_or_1(n) - 1;
    V2 f = from;
    for (uint32_t i = 0; i < cnt; ++i) {
        float t1 = t0 + step;
        // Horner (:1352-1364)
        V2 v = a0;
        float tt = t1;
        v = v + a1 * tt;
        tt = tt * t1;
        v = v + a2 * tt;
        tt = tt * t1;
        v = v + a3 * tt;
        cb(f, v, t0, t1);
        f = v;
        t0 = t1;
    }
    cb(f, to, t0, 1.0f);
}

// QuadraticBezierSegment::sample (Bernstein form, quadratic_bezier.rs:35-41)
inline V2 quad_sample(V2 from, V2 ctrl, V2 to, float t) {
    float t2 = t * t;
    float one_t = 1.0f - t;
    float one_t2 = one_t * one_t;
    return (from * one_t2 + ctrl * 2.0f * one_t * t) + to * t2;
}

// QuadraticBezierSegment::split (quadratic_bezier.rs:208-223)
inline void quad_split(V2 from, V2 ctrl, V2 to, float t, V2 out_a[3], V2 out_b[3]) {
    V2 sp = quad_sample(from, ctrl, to, t);
    out_a[0] = from;
    out_a[1] = lerp(from, ctrl, t);
    out_a[2] = sp;
    out_b[0] = sp;
    out_b[1] = lerp(ctrl, to, t);
    out_b[2] = to;
}

// QuadraticBezierSegment::local_x_extremum_t (quadratic_bezier.rs:180-191)
inline bool quad_local_x_extremum_t(V2 from, V2 ctrl, V2 to, float* t_out) {
    float div = from.x - 2.0f * ctrl.x + to.x;
    if (div == 0.0f) return false;
    float t = (from.x - ctrl.x) / div;
    if (t > 0.0f && t < 1.0f) {
        *t_out = t;
        return true;
    }
    return false;
}

// ---------------------------------------------------------------------------------
// Arcs (arc.rs)
// ---------------------------------------------------------------------------------
struct Arc {
    V2 center, radii;
    float start_angle, sweep_angle, x_rotation;
};
// sample_ellipse (arc.rs:741-746)
inline V2 sample_ellipse(V2 radii, float x_rotation, float angle) {
    return rotate(x_rotation, V2{radii.x * std::cos(angle), radii.y * std::sin(angle)});
}
// Arc::sample / from / to (arc.rs:194-197, 227-234)
inline V2 arc_sample(const Arc& a, float t) {
    float angle = a.start_angle + a.sweep_angle * t;
    return a.center + sample_ellipse(a.radii, a.x_rotation, angle);
}
// Arc::after_split (arc.rs:286-295)
inline Arc arc_after_split(const Arc& a, float t) {
    float split_angle = a.sweep_angle * t;
    Arc r = a;
    r.start_angle = a.start_angle + split_angle;
    r.sweep_angle = a.sweep_angle - split_angle;
    return r;
}
// Arc::flattening_step (arc.rs:372-388)
inline float arc_flattening_step(const Arc& a, float tol) {
    float r = length(arc_sample(a, 0.0f) - a.center);
    float ang = 2.0f * std::acos((r - tol) / r);
    float result = fminr(ang / std::fabs(a.sweep_angle), 1.0f);
    if (result < 1e-4f) return 1.0f;
    return result;
}
// Arc::for_each_flattened_with_t (arc.rs:339-368)
template <class F>
inline void arc_flatten_with_t(const Arc& arc, float tol, F&& cb) {
    Arc iter = arc;
    float t0 = 0.0f;
    V2 from = arc_sample(arc, 0.0f);
    for (;;) {
        float step = arc_flattening_step(iter, tol);
        if (step >= 1.0f) break;
        iter = arc_after_split(iter, step);
        float t1 = t0 + step * (1.0f - t0);
        V2 to = arc_sample(iter, 0.0f);
        cb(from, to, t0, t1);
        from = to;
        t0 = t1;
    }
    cb(from, arc_sample(arc, 1.0f), t0, 1.0f);
}
// Arc::tangent_at_angle (arc.rs:510-516)
inline V2 arc_tangent_at_angle(const Arc& a, float angle) {
    return rotate(a.x_rotation, V2{-a.radii.x * std::sin(angle), a.radii.y * std::cos(angle)});
}
// arc_to_quadratic_beziers_with_t (arc.rs:668-701)
template <class F>
inline void arc_to_quads_with_t(const Arc& arc, F&& cb) {
    const float PI = 3.14159265358979323846f;
    const float FRAC_PI_4 = 0.785398163397448309616f;
    float sign = std::copysign(1.0f, arc.sweep_angle);
    if (std::isnan(arc.sweep_angle)) sign = arc.sweep_angle;
    float sweep = fminr(std::fabs(arc.sweep_angle), PI * 2.0f);
    float n_steps = std::ceil(sweep / FRAC_PI_4);
    float step = sweep / n_steps * sign;
    float t0 = 0.0f;
    float dt = 1.0f / n_steps;
    float alpha = std::tan(step * 0.5f);
    int n = (int)n_steps;
    for (int i = 0; i < n; ++i) {
        float a1 = arc.start_angle + step * (float)i;
        float a2 = arc.start_angle + step * (float)(i + 1);
        V2 v1 = sample_ellipse(arc.radii, arc.x_rotation, a1);
        V2 v2_ = sample_ellipse(arc.radii, arc.x_rotation, a2);
        V2 from = arc.center + v1;
        V2 to = arc.center + v2_;
        V2 ctrl = from + arc_tangent_at_angle(arc, a1) * alpha;
        float t1 = (i + 1 == n) ? 1.0f : t0 + dt;
        cb(from, ctrl, to, t0, t1);
        t0 = t1;
    }
}

}  // namespace refgeom
