// Arcs on the device: lyon_geom's Arc (center parameterization, crates/geom/src/arc.rs) and
// its quadratic approximation.  Shared by the flattening kernels and the SVG path parser.
#pragma once
#include "lyg_math.h"
#include "stroke_core.h"

namespace lyg {

struct ArcD {
    vec2 center, radii;
    float start, sweep, xrot;
};

// Arc::sample (arc.rs:194-197) with sample_ellipse (arc.rs:741-746)
__device__ __forceinline__ vec2 arc_sample(const ArcD& a, float t) {
    float angle = a.start + a.sweep * t;
    vec2 e = mk(a.radii.x * t_cos(angle), a.radii.y * t_sin(angle));
    return add(a.center, rotate_v(a.xrot, e));
}

// Number of quadratics arc_to_quadratic_beziers_with_t (arc.rs:668-701) makes.
__device__ __forceinline__ int arc_quad_count(const ArcD& arc) {
    const float PI = 3.14159265358979323846f;
    const float FRAC_PI_4 = 0.785398163397448309616f;
    float sweep = ffmin(ffabs(arc.sweep), PI * 2.0f);
    float n_steps = __builtin_ceilf(sweep / FRAC_PI_4);
    return (n_steps >= 0.0f && n_steps < 2147483647.0f) ? (int)n_steps : 0;
}

// arc_to_quadratic_beziers_with_t (arc.rs:668-701): f(k, from, ctrl, to) per quadratic.
template <class F>
__device__ __forceinline__ void arc_to_quads_each(const ArcD& arc, F&& f) {
    const float PI = 3.14159265358979323846f;
    const float FRAC_PI_4 = 0.785398163397448309616f;
    float sign = __builtin_copysignf(1.0f, arc.sweep);
    if (__builtin_isnan(arc.sweep)) sign = arc.sweep;
    float sweep = ffmin(ffabs(arc.sweep), PI * 2.0f);
    float n_steps = __builtin_ceilf(sweep / FRAC_PI_4);
    int ns = (n_steps >= 0.0f && n_steps < 2147483647.0f) ? (int)n_steps : 0;
    float step = sweep / n_steps * sign;
    float alpha = (float)tan((double)(step * 0.5f));
    for (int k = 0; k < ns; ++k) {
        float a1 = arc.start + step * (float)k;
        float a2 = arc.start + step * (float)(k + 1);
        vec2 v1 = rotate_v(arc.xrot, mk(arc.radii.x * t_cos(a1), arc.radii.y * t_sin(a1)));
        vec2 v2 = rotate_v(arc.xrot, mk(arc.radii.x * t_cos(a2), arc.radii.y * t_sin(a2)));
        vec2 from = add(arc.center, v1);
        vec2 to = add(arc.center, v2);
        // Arc::tangent_at_angle (arc.rs:510-516)
        vec2 tan_ = rotate_v(arc.xrot, mk(-arc.radii.x * t_sin(a1), arc.radii.y * t_cos(a1)));
        vec2 ctrl = add(from, mul(tan_, alpha));
        f(k, from, ctrl, to);
    }
}

}  // namespace lyg
