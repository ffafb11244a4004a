// Device-side geometry of lyon's fixed-width stroker: joins, caps, round arcs.
//
// The join/cap rules must produce the same vertices and triangles, in the same order,
// as crates/tessellation/src/stroke.rs, so the arithmetic below keeps the reference's
// formulas and operation order (each function cites the lines it must agree with).
// What differs is the execution model: a join is a pure function of four positions
// (the kept points before and after it) plus the id base the wave scan assigned it,
// so any lane can compute any join; vertices and indices go through a Sink that
// either counts (first pass) or writes at scan-assigned offsets (second pass).
#pragma once
#include "lyg_math.h"
#include "lyg_types.h"
#include "wave.h"

namespace lyg {

enum { SP = 0, SN = 1 };  // Side::Positive / Side::Negative slots (stroke.rs:25-26)
static constexpr uint32_t INVALID_ID = 0xffffffffu;

// Transcendentals: evaluated in f64 and rounded once to f32.
__device__ __forceinline__ float t_acos(float x) { return (float)acos((double)x); }
__device__ __forceinline__ float t_cos(float x) { return (float)cos((double)x); }
__device__ __forceinline__ float t_sin(float x) { return (float)sin((double)x); }
__device__ __forceinline__ float t_log2(float x) { return (float)log2((double)x); }

// euclid Trig::fast_atan2 (euclid 0.22.9), used by Vector2D::angle_from_x_axis.
__device__ __forceinline__ float fast_atan2(float y, float x) {
    const float PI = 3.14159265358979323846f;
    const float FRAC_PI_2 = 1.57079632679489661923f;
    float x_abs = ffabs(x), y_abs = ffabs(y);
    float a = ffmin(x_abs, y_abs) / ffmax(x_abs, y_abs);
    float s = a * a;
    float r = ((-0.0464964749f * s + 0.15931422f) * s - 0.327622764f) * s * a + a;
    if (y_abs > x_abs) r = FRAC_PI_2 - r;
    if (x < 0.0f) r = PI - r;
    if (y < 0.0f) r = -r;
    return r;
}
__device__ __forceinline__ float angle_of(vec2 v) { return fast_atan2(v.y, v.x); }
// fmodf(x, y) for y > 0: exact results without the general reduction when |x| < 2y
// (|x| < y: x itself; y < |x| < 2y: x - y with x's sign, exact by Sterbenz's lemma).
__device__ __forceinline__ float fmod_near(float x, float y) {
    const float ax = ffabs(x);
    if (ax < y) return x;
    if (ax > y && ax < 2.0f * y) return x - __builtin_copysignf(y, x);
    return fmodf(x, y);
}
// euclid Angle::angle_to
__device__ __forceinline__ float angle_to(float from, float to) {
    const float PI = 3.14159265358979323846f;
    float two = 1.0f + 1.0f;
    float mx = PI * two;
    float d = fmod_near(to - from, mx);
    return fmod_near(two * d, mx) - d;
}
// euclid Rotation2D::transform_vector
__device__ __forceinline__ vec2 rotate_v(float angle, vec2 p) {
    float s = t_sin(angle), c = t_cos(angle);
    return mk(p.x * c - p.y * s, p.y * c + p.x * s);
}

// math_utils.rs:16-40
__device__ __forceinline__ vec2 compute_normal(vec2 v1, vec2 v2) {
    const float eps = 1e-4f;
    vec2 n1 = mk(-v1.y, v1.x);
    vec2 v12 = add(v1, v2);
    if (sqlen(v12) < eps) return mk(0.0f, 0.0f);
    vec2 t = normalize(v12);
    vec2 n = mk(-t.y, t.x);
    float inv_len = dot(n, n1);
    if (ffabs(inv_len) < eps) return n1;
    return divs(n, inv_len);
}

// Line::intersection (line.rs:562-578) on f64 casts of two f32 lines.
__device__ __forceinline__ bool line_isect(vec2 p1, vec2 v1, vec2 p2, vec2 v2, vec2* out) {
    double spx = p1.x, spy = p1.y, svx = v1.x, svy = v1.y;
    double opx = p2.x, opy = p2.y, ovx = v2.x, ovy = v2.y;
    double det = svx * ovy - svy * ovx;
    if (fabs(det) <= 1e-8) return false;
    double inv_det = 1.0 / det;
    double s2x = spx + svx, s2y = spy + svy;
    double o2x = opx + ovx, o2y = opy + ovy;
    double a = spx * s2y - spy * s2x;
    double b = opx * o2y - opy * o2x;
    out->x = (float)((b * svx - a * ovx) * inv_det);
    out->y = (float)((b * svy - a * ovy) * inv_det);
    return true;
}

// quadratic_bezier.rs:35-41, 180-191, 208-223
__device__ __forceinline__ vec2 quad_sample(vec2 from, vec2 ctrl, vec2 to, float t) {
    float t2 = t * t;
    float one_t = 1.0f - t;
    float one_t2 = one_t * one_t;
    return add(add(mul(from, one_t2), mul(mul(mul(ctrl, 2.0f), one_t), t)), mul(to, t2));
}

// stroke.rs:3178-3212: split parameter of a quadratic with a sharp turn, or -1.
__device__ __forceinline__ float find_sharp_turn(vec2 from, vec2 ctrl, vec2 to) {
    vec2 baseline = sub(to, from);
    vec2 v = sub(ctrl, from);
    vec2 n = mk(-baseline.y, baseline.x);
    float v_dot_b = dot(v, baseline);
    float v_dot_n = dot(v, n);
    vec2 long_axis;
    if ((v_dot_b >= 0.0f && v_dot_b <= dot(baseline, baseline)) ||
        ffabs(v_dot_n) * 2.0f >= ffabs(v_dot_b)) {
        if (sqlen(baseline) * 30.0f > sqlen(v)) return -1.0f;
        long_axis = v;
    } else {
        long_axis = baseline;
    }
    float rot = -angle_of(long_axis);
    vec2 rc = rotate_v(rot, v);
    vec2 rt = rotate_v(rot, baseline);
    // local_x_extremum_t with from = (0, 0)
    float div = 0.0f - 2.0f * rc.x + rt.x;
    if (div == 0.0f) return -1.0f;
    float t = (0.0f - rc.x) / div;
    if (t > 0.0f && t < 1.0f) return t;
    return -1.0f;
}

struct SParams {
    float hw, tol, miter_limit, thr;
    int start_cap, end_cap, join;
    float arc_step;          // circle flattening step of every round join and cap of the path
    const float* nsub_thr;   // subdivision thresholds (nsub_thresholds_kernel) or null
};
__device__ __forceinline__ float circle_step(float radius, float tol);
__device__ __forceinline__ SParams make_params(const lyg_stroke_options& o,
                                               const float* nsub_thr = nullptr) {
    SParams p;
    p.hw = o.line_width * 0.5f;
    p.tol = o.tolerance;
    p.miter_limit = o.miter_limit;
    // StrokeBuilderImpl::new (stroke.rs:603-605)
    p.thr = ffmax(ffmin(o.tolerance * o.tolerance * 0.5f, o.line_width * o.line_width * 0.05f),
                  1e-8f);
    p.start_cap = o.start_cap;
    p.end_cap = o.end_cap;
    p.join = o.line_join;
    // every arc of a path has radius half_width: its flattening step is a per-path constant
    const bool round = o.line_join == LYG_JOIN_ROUND || o.start_cap == LYG_CAP_ROUND ||
                       o.end_cap == LYG_CAP_ROUND;
    p.arc_step = round ? circle_step(p.hw, p.tol) : 0.0f;
    p.nsub_thr = nsub_thr;
    return p;
}

// Output sink: counts in the first pass, writes at scan offsets in the second.  Writes
// are bounds-checked against the item's path range; a write outside it is dropped and
// flagged (it would mean the two passes disagree).
struct Sink {
    uint32_t vid;       // id of the next vertex
    uint32_t nv, ni;    // vertices / triangles produced so far
    float* vout;        // emit: vertex array (index = id - vid0)
    uint32_t vid0;
    uint32_t* iout;     // emit: first index slot of this item
    bool emit;
    uint32_t vlim;      // emit: exclusive vertex index limit of the path
    uint32_t ileft;     // emit: index slots left in the path from iout
    uint32_t* err;      // emit: error flag
};
__device__ __forceinline__ uint32_t sink_vertex(Sink& s, vec2 p) {
    uint32_t id = s.vid++;
    if (s.emit) {
        uint32_t idx = id - s.vid0;
        if (idx < s.vlim) {
            float2* o = reinterpret_cast<float2*>(s.vout) + idx;
            *o = make_float2(p.x, p.y);
        } else if (s.err) {
            atomicOr(s.err, 1u);
        }
    }
    s.nv++;
    return id;
}
__device__ __forceinline__ void sink_tri(Sink& s, uint32_t a, uint32_t b, uint32_t c) {
    if (s.emit) {
        if (3u * s.ni + 3u <= s.ileft) {
            uint32_t* o = s.iout + 3 * s.ni;
            o[0] = a;
            o[1] = b;
            o[2] = c;
        } else if (s.err) {
            atomicOr(s.err, 2u);
        }
    }
    s.ni++;
}
// StrokeVertex::position (stroke.rs:2676-2678): position_on_path + normal * half_width.
__device__ __forceinline__ uint32_t sink_sv(Sink& s, vec2 on_path, vec2 normal, float hw) {
    return sink_vertex(s, add(on_path, mul(normal, hw)));
}

// EndpointData (stroke.rs:320-352) restricted to what positions and topology use.
struct EPD {
    vec2 pos;
    vec2 prev[2], next[2], single[2];
    uint32_t pv[2], nv[2];
    uint8_t has_single[2], fold[2];
    uint8_t flat, line_join;
};
__device__ __forceinline__ void epd_init(EPD& e, vec2 pos, bool flat, int join) {
    e.pos = pos;
    const float nan = __builtin_nanf("");
    for (int s = 0; s < 2; ++s) {
        e.prev[s] = mk(nan, nan);
        e.next[s] = mk(nan, nan);
        e.single[s] = mk(0.0f, 0.0f);
        e.pv[s] = INVALID_ID;
        e.nv[s] = INVALID_ID;
        e.has_single[s] = 0;
        e.fold[s] = 0;
    }
    e.flat = flat;
    e.line_join = (uint8_t)join;
}

// Compact per-point layout code published to the next lane: which vertices a point owns.
enum : uint32_t {
    LC_FAST = 1u,        // flattened_step fast path
    LC_WEIRD = 2u,       // fast path that emitted nothing (stroke.rs:1786-1788)
    LC_SINGLE_P = 4u,
    LC_SINGLE_N = 8u,
    LC_FOLD_P = 16u,
    LC_FOLD_N = 32u,
    LC_FIRST = 64u,      // sub-path start point: side points only, no vertices
};
__device__ __forceinline__ uint32_t layout_of(const EPD& j, bool fast, bool weird) {
    uint32_t c = 0;
    if (fast) c |= LC_FAST;
    if (weird) c |= LC_WEIRD;
    if (j.has_single[SP]) c |= LC_SINGLE_P;
    if (j.has_single[SN]) c |= LC_SINGLE_N;
    if (j.fold[SP]) c |= LC_FOLD_P;
    if (j.fold[SN]) c |= LC_FOLD_N;
    return c;
}
// Vertex ids of a point from its layout code and the id of its first vertex: mirrors the
// order of flattened_step (POS, NEG) and add_join_base_vertices (NEG then POS).
__device__ __forceinline__ void ids_of(uint32_t code, uint32_t base, uint32_t pv[2],
                                       uint32_t nv[2]) {
    if ((code & LC_FIRST) || (code & LC_WEIRD)) {
        pv[0] = pv[1] = nv[0] = nv[1] = INVALID_ID;
        return;
    }
    if (code & LC_FAST) {
        pv[SP] = nv[SP] = base;
        pv[SN] = nv[SN] = base + 1;
        return;
    }
    uint32_t b = base;
    pv[SN] = b;
    nv[SN] = (code & LC_SINGLE_N) ? b : b + 1;
    b = nv[SN] + 1;
    pv[SP] = b;
    nv[SP] = (code & LC_SINGLE_P) ? b : b + 1;
}

// stroke.rs:1869-1904
__device__ __forceinline__ void add_edge_triangles(const EPD& p0, const EPD& p1, Sink& s) {
    uint32_t p0_neg = p0.nv[SN], p0_pos = p0.nv[SP];
    uint32_t p1_neg = p1.pv[SN], p1_pos = p1.pv[SP];
    if (p0.fold[SP]) p0_neg = p0.pv[SP];
    if (p0.fold[SN]) p0_pos = p0.pv[SN];
    if (p1.fold[SP]) p1_neg = p1.nv[SP];
    if (p1.fold[SN]) p1_pos = p1.nv[SN];
    if (p0_neg == p1_pos) return;
    if (p0_neg != p0_pos && p0_pos != p1_pos) sink_tri(s, p0_neg, p0_pos, p1_pos);
    if (p0_neg != p1_neg && p1_pos != p1_neg) sink_tri(s, p0_neg, p1_pos, p1_neg);
}

// circle_flattening_step (stroke.rs:2749-2753).
__device__ __forceinline__ float circle_step(float radius, float tol) {
    float t = ffmin(tol, radius);
    return 2.0f * t_acos((radius - t) / radius);
}
// round(log2(x)) as the subdivision count of stroke.rs:2002-2004 / 2488-2490 (0 when not
// positive, saturating).
__device__ __forceinline__ uint32_t nsub_log2(float x) {
    float r = __builtin_roundf(t_log2(x));
    if (!(r > 0.0f)) return 0u;
    if (r >= 4294967296.0f) return 0xffffffffu;
    return (uint32_t)r;
}
// Subdivision count of an arc of `diff` radians with radius and tolerance as given.
__device__ __forceinline__ uint32_t arc_subdivisions(float diff, float radius, float tol) {
    float num_segments = __builtin_ceilf(ffabs(diff) / circle_step(radius, tol));
    return nsub_log2(num_segments);
}
// The same for an arc of the path's own radius (half_width), from the per-path step and,
// when present, the threshold table: round(log2(x)) for x in [2^k, 2^(k+1)) is k or k + 1
// and grows with x, so it is k + (x >= thr[k]), where thr[k] was found by bisection with
// the very same log2 (nsub_thresholds_kernel).
__device__ __forceinline__ uint32_t path_arc_subdivisions(float diff, const SParams& p) {
    const float x = __builtin_ceilf(ffabs(diff) / p.arc_step);
    if (p.nsub_thr && x >= 1.0f && x < 2147483648.0f) {
        const int k = (int)((f2u(x) >> 23) & 0xffu) - 127;
        return (uint32_t)k + (x >= p.nsub_thr[k] ? 1u : 0u);
    }
    return nsub_log2(x);
}

// tessellate_arc (stroke.rs:2612-2653): pre-order recursion made iterative.
__device__ __noinline__ void tessellate_arc(float a0, float a1, uint32_t va, uint32_t vb,
                                               uint32_t n, vec2 on_path, float hw, Sink& s) {
    if (n == 0) return;
    if (n > 30) n = 30;
    struct Fr {
        float a0, a1;
        uint32_t va, vb, d;
    };
    Fr st[32];
    int sp = 0;
    st[sp++] = Fr{a0, a1, va, vb, n};
    while (sp > 0) {
        Fr f = st[--sp];
        if (f.d == 0) continue;
        float mid = (f.a0 + f.a1) * 0.5f;
        uint32_t id = sink_sv(s, on_path, mk(t_cos(mid), t_sin(mid)), hw);
        sink_tri(s, f.va, id, f.vb);
        st[sp++] = Fr{mid, f.a1, id, f.vb, f.d - 1};
        st[sp++] = Fr{f.a0, mid, f.va, id, f.d - 1};
    }
}

// Number of vertices (= triangles) tessellate_arc produces for n subdivisions.
__device__ __forceinline__ uint32_t arc_count(uint32_t n) {
    if (n > 30) n = 30;
    return (1u << n) - 1u;
}

// Angles and subdivision count of one side's round join: tessellate_round_join
// (stroke.rs:1968-2021) up to the tessellate_arc call.
struct RoundArc {
    float a0, a1;
    uint32_t nsub;
    bool swap;  // NEG side: start/end vertices swapped
};
__device__ __forceinline__ RoundArc round_join_arc(const EPD& j, int side, const SParams& p) {
    const float PI = 3.14159265358979323846f;
    vec2 center = j.pos;
    vec2 start_normal = sub(j.prev[side], center);
    vec2 end_normal = sub(j.next[side], center);
    float angle_sign = side == SN ? 1.0f : -1.0f;
    float start_angle = angle_of(start_normal);
    float diff = angle_to(start_angle, angle_of(end_normal));
    if (diff * angle_sign < 0.0f) diff = angle_sign * (2.0f * PI - ffabs(diff));
    float end_angle = start_angle + diff;
    RoundArc r;
    r.swap = side == SN;
    r.a0 = r.swap ? end_angle : start_angle;
    r.a1 = r.swap ? start_angle : end_angle;
    r.nsub = path_arc_subdivisions(diff, p);
    return r;
}

// Everything about one join that does not depend on vertex ids.
struct JInfo {
    EPD x;           // the join point (side points, singles, folds, effective line_join)
    vec2 normal;     // flattened_step normal (fast path)
    bool fast, weird;
    bool needs[2];   // side_needs_join of tessellate_join
    RoundArc arc[2]; // round join arcs (valid when needs[side] and Round)
};

// A round-join arc left for the wave to emit cooperatively (emit_arcs): the same
// vertices and triangles tessellate_arc would write, at the ids and triangle slots the
// sink had reached when the arc came up.
struct ArcJob {
    float a0, a1, cx, cy;
    uint32_t va, vb, n, first_id;
    uint32_t* tri;   // index slots of the arc's first triangle
    uint32_t ileft;  // index slots left in the path from `tri`
    uint32_t pad;
};
struct ArcJobs {  // at most one arc per side of a lane's join
    ArcJob sp, sn;
    bool has_sp = false, has_sn = false;
};

// One side's round join arc (stroke.rs:1957-1962); `side` is a constant so the join
// state is never indexed at run time.  With `jobs` the arc is recorded and the sink
// advanced past it; without, it is emitted here.
template <int side>
__device__ __forceinline__ void round_join_side(const JInfo& J, const SParams& p, Sink& s,
                                                ArcJobs* jobs) {
    if (!J.needs[side]) return;
    const uint32_t pv = J.x.pv[side], nv = J.x.nv[side];
    const bool swap = J.arc[side].swap;
    if (jobs) {
        uint32_t n = J.arc[side].nsub;
        if (n == 0) return;  // tessellate_arc emits nothing
        if (n > 30) n = 30;
        const uint32_t cnt = (1u << n) - 1u;
        ArcJob a;
        a.a0 = J.arc[side].a0;
        a.a1 = J.arc[side].a1;
        a.cx = J.x.pos.x;
        a.cy = J.x.pos.y;
        a.va = swap ? nv : pv;
        a.vb = swap ? pv : nv;
        a.n = n;
        a.first_id = s.vid;
        a.tri = s.iout + 3u * s.ni;
        a.ileft = s.ileft > 3u * s.ni ? s.ileft - 3u * s.ni : 0u;
        a.pad = 0;
        if (side == SP) {
            jobs->sp = a;
            jobs->has_sp = true;
        } else {
            jobs->sn = a;
            jobs->has_sn = true;
        }
        s.vid += cnt;
        s.nv += cnt;
        s.ni += cnt;
        return;
    }
    Sink s2 = s;  // the call takes its address: keep the caller's sink in registers
    tessellate_arc(J.arc[side].a0, J.arc[side].a1, swap ? nv : pv, swap ? pv : nv,
                   J.arc[side].nsub, J.x.pos, p.hw, s2);
    s = s2;
}

// tessellate_join (stroke.rs:1907-1965) with the arcs precomputed in `J`.
__device__ __forceinline__ void tessellate_join(const JInfo& J, const SParams& p, Sink& s,
                                                ArcJobs* jobs = nullptr) {
    const EPD& j = J.x;
    bool needs_p = J.needs[SP], needs_n = J.needs[SN];
    if (!j.fold[SP] && !j.fold[SN]) {
        if (needs_p && needs_n) {
            sink_tri(s, j.pv[SP], j.nv[SP], j.nv[SN]);
            sink_tri(s, j.pv[SP], j.nv[SN], j.pv[SN]);
        } else if (!needs_p && needs_n) {
            sink_tri(s, j.pv[SN], j.pv[SP], j.nv[SN]);
        } else if (needs_p && !needs_n) {
            sink_tri(s, j.pv[SN], j.pv[SP], j.nv[SP]);
        }
    }
    if (j.line_join == LYG_JOIN_ROUND) {
        round_join_side<SP>(J, p, s, jobs);
        round_join_side<SN>(J, p, s, jobs);
    }
}

// get_clip_intersections (stroke.rs:2279-2315)
__device__ __forceinline__ void clip_intersections(vec2 pn, vec2 nn, vec2 normal, float dist,
                                                   vec2* i1, vec2* i2) {
    vec2 cp = mul(normalize(normal), dist);
    vec2 cv = perp(normal);
    if (!line_isect(cp, cv, pn, perp(pn), i1)) *i1 = normal;
    if (!line_isect(cp, cv, nn, perp(nn), i2)) *i2 = normal;
}

// compute_join_side_positions_fixed_width (stroke.rs:1643-1733) on `join`, given the
// previous and next kept positions.
__device__ __forceinline__ void join_side_positions(vec2 prevp, EPD& join, vec2 nextp,
                                                    const SParams& p) {
    vec2 pt = sub(join.pos, prevp);
    vec2 nt = sub(nextp, join.pos);
    float pl = vlen(pt), nl = vlen(nt);
    pt = divs(pt, pl);
    nt = divs(nt, nl);
    vec2 normal = compute_normal(pt, nt);
    // front side: SN when turning left.  Slots are only ever indexed by constants (a
    // runtime index would keep the join state in private memory).
    const bool fsn = cross(pt, nt) >= 0.0f;
    vec2 fnormal = fsn ? neg(normal) : normal;
    vec2 extruded = mul(fnormal, p.hw);
    bool unclipped_miter =
        (join.line_join == LYG_JOIN_MITER || join.line_join == LYG_JOIN_MITER_CLIP) &&
        !(sqlen(fnormal) > p.miter_limit * p.miter_limit * 4.0f);
    bool fold = false;
    bool sharp = dot(nt, pt) < 0.0f;
    if (!unclipped_miter && sharp) {
        float d_next = dot(extruded, neg(nt)) - nl;
        float d_prev = dot(extruded, pt) - pl;
        if (ffmin(d_next, d_prev) > 0.0f || sqlen(normal) < 1e-5f) {
            if (fsn) join.fold[SN] = 1;
            else join.fold[SP] = 1;
            fold = true;
        }
    }
    vec2 n0 = mul(mk(-pt.y, pt.x), p.hw);
    vec2 n1 = mul(mk(-nt.y, nt.x), p.hw);
    join.prev[SP] = add(join.pos, n0);
    join.next[SP] = add(join.pos, n1);
    join.prev[SN] = sub(join.pos, n0);
    join.next[SN] = sub(join.pos, n1);
    if (!fold) {
        const vec2 mp_p = add(join.pos, mul(normal, p.hw));
        const vec2 mp_n = sub(join.pos, mul(normal, p.hw));
        // the back side always gets its single vertex, the front one for a miter
        if (fsn || unclipped_miter) {
            join.has_single[SP] = 1;
            join.single[SP] = mp_p;
        }
        if (!fsn || unclipped_miter) {
            join.has_single[SN] = 1;
            join.single[SN] = mp_n;
        }
        if (!unclipped_miter && join.line_join == LYG_JOIN_MITER_CLIP) {
            const vec2 fprev = fsn ? join.prev[SN] : join.prev[SP];
            const vec2 fnext = fsn ? join.next[SN] : join.next[SP];
            vec2 a = sub(fprev, join.pos);
            vec2 b = sub(fnext, join.pos);
            vec2 i1, i2;
            clip_intersections(a, b, fnormal, p.miter_limit * p.hw, &i1, &i2);
            const vec2 np = add(join.pos, i1), nn = add(join.pos, i2);
            if (fsn) {
                join.prev[SN] = np;
                join.next[SN] = nn;
            } else {
                join.prev[SP] = np;
                join.next[SP] = nn;
            }
        }
    }
}

// Side points, singles and folds of the join at J.x (whose pos/flat/line_join are set),
// given the previous and next kept positions: the geometric half of the `count > 1`
// branch of fixed_width_step_impl (stroke.rs:1558-1604), i.e. flattened_step's side
// points (stroke.rs:1743-1784) or compute_join_side_positions_fixed_width.
__device__ __forceinline__ void join_state(vec2 prevp, JInfo& J, vec2 nextp, const SParams& p) {
    EPD& join = J.x;
    J.fast = false;
    J.weird = false;
    if (join.flat) J.fast = dot(sub(join.pos, prevp), sub(nextp, join.pos)) > 0.0f;
    if (J.fast) {
        join.line_join = LYG_JOIN_MITER;
        vec2 pe = sub(join.pos, prevp);
        float pl = vlen(pe);
        vec2 pt = divs(pe, pl);
        vec2 ne = sub(nextp, join.pos);
        float nl = vlen(ne);
        vec2 nt = divs(ne, nl);
        vec2 normal = compute_normal(pt, nt);
        J.normal = normal;
        vec2 p0 = add(join.pos, mul(normal, p.hw));
        vec2 p1 = sub(join.pos, mul(normal, p.hw));
        join.prev[SP] = join.next[SP] = join.single[SP] = p0;
        join.prev[SN] = join.next[SN] = join.single[SN] = p1;
        join.has_single[SP] = join.has_single[SN] = 1;
    } else {
        join_side_positions(prevp, join, nextp, p);
    }
    J.needs[SP] = !join.has_single[SP] && !join.fold[SN];
    J.needs[SN] = !join.has_single[SN] && !join.fold[SP];
    for (int side = 0; side < 2; ++side) {
        J.arc[side].nsub = 0;
        J.arc[side].a0 = J.arc[side].a1 = 0.0f;
        J.arc[side].swap = false;
        if (join.line_join == LYG_JOIN_ROUND && J.needs[side]) J.arc[side] = round_join_arc(join, side, p);
    }
}

// flattened_step's fold-back test (stroke.rs:1784-1788): in the fast path a join whose
// two new side points both lie behind the previous point's emits nothing.
__device__ __forceinline__ void join_weird(JInfo& J, vec2 prevp, vec2 prev_next_p,
                                           vec2 prev_next_n) {
    if (!J.fast) return;
    vec2 pe = sub(J.x.pos, prevp);
    vec2 v0 = sub(J.x.single[SP], prev_next_p);
    vec2 v1 = sub(J.x.single[SN], prev_next_n);
    J.weird = dot(pe, v0) < 0.0f && dot(pe, v1) < 0.0f;
}

__device__ __forceinline__ uint32_t join_code(const JInfo& J) {
    return layout_of(J.x, J.fast, J.weird);
}

// Vertices and triangles the join emits, from the layout codes alone: ids only matter
// through equality, which the codes fix (ids of different points never coincide unless
// both are VertexId::INVALID).
__device__ __forceinline__ void join_counts(const JInfo& J, uint32_t pcode, bool with_edge,
                                            uint32_t* nv, uint32_t* nt) {
    uint32_t v = 0, t = 0;
    uint32_t code = join_code(J);
    if (J.fast) {
        v = J.weird ? 0u : 2u;
    } else {
        v = (J.x.has_single[SN] ? 1u : 2u) + (J.x.has_single[SP] ? 1u : 2u);
    }
    if (with_edge) {
        EPD a, b;
        a.fold[SP] = (pcode & LC_FOLD_P) ? 1 : 0;
        a.fold[SN] = (pcode & LC_FOLD_N) ? 1 : 0;
        b.fold[SP] = J.x.fold[SP];
        b.fold[SN] = J.x.fold[SN];
        ids_of(pcode, 0x40000000u, a.pv, a.nv);
        ids_of(code, 0x80000000u, b.pv, b.nv);
        Sink c{0, 0, 0, nullptr, 0, nullptr, false, 0, 0, nullptr};
        add_edge_triangles(a, b, c);
        t += c.ni;
    }
    if (!J.x.fold[SP] && !J.x.fold[SN]) {
        if (J.needs[SP] && J.needs[SN]) t += 2;
        else if (J.needs[SP] || J.needs[SN]) t += 1;
    }
    if (J.x.line_join == LYG_JOIN_ROUND) {
        for (int side = 0; side < 2; ++side) {
            if (!J.needs[side]) continue;
            uint32_t a = arc_count(J.arc[side].nsub);
            v += a;
            t += a;
        }
    }
    *nv = v;
    *nt = t;
}

// One side's join vertices (add_join_base_vertices, stroke.rs:2024-2055): the ids of
// its prev/next vertices (equal when the side has a single vertex).
template <int side>
__device__ __forceinline__ void join_base_vertices(const EPD& join, const SParams& p, Sink& s,
                                                   uint32_t& a, uint32_t& b) {
    if (join.has_single[side]) {
        a = b = sink_sv(s, join.pos, divs(sub(join.single[side], join.pos), p.hw), p.hw);
    } else {
        a = sink_sv(s, join.pos, divs(sub(join.prev[side], join.pos), p.hw), p.hw);
        b = sink_sv(s, join.pos, divs(sub(join.next[side], join.pos), p.hw), p.hw);
    }
}

// The emitting half: join vertices (flattened_step's POS/NEG pair or
// add_join_base_vertices NEG then POS, stroke.rs:2024-2055), the edge triangles from
// `prev` when with_edge (stroke.rs:1616-1618) and tessellate_join.  Fills J.x's ids.
// The ids are computed as values and stored once after the branches: stores to the same
// slots in both branches would be merged into stores through a selected address, which
// keeps the join state in private memory.
__device__ __forceinline__ void join_emit(JInfo& J, const EPD& prev, bool with_edge,
                                          const SParams& p, Sink& s, ArcJobs* jobs = nullptr) {
    EPD& join = J.x;
    uint32_t pp = INVALID_ID, np = INVALID_ID, pn = INVALID_ID, nn = INVALID_ID;
    if (J.fast) {
        if (!J.weird) {
            pp = np = sink_sv(s, join.pos, J.normal, p.hw);
            pn = nn = sink_sv(s, join.pos, neg(J.normal), p.hw);
        }
    } else {
        join_base_vertices<SN>(join, p, s, pn, nn);
        join_base_vertices<SP>(join, p, s, pp, np);
    }
    join.pv[SP] = pp;
    join.nv[SP] = np;
    join.pv[SN] = pn;
    join.nv[SN] = nn;
    if (with_edge) add_edge_triangles(prev, join, s);
    tessellate_join(J, p, s, jobs);
}

// Emits the arcs recorded by the wave's lanes (ArcJobs) with all 64 lanes: node g of the
// concatenated arcs goes to lane g mod 64.  Node p of an arc of depth n is found by
// walking down from the root in tessellate_arc's pre-order (the node, its lower half,
// its upper half): the same midpoints, ids first_id + p and triangles (va, id, vb) as
// the serial recursion.  `jl` (128 jobs) and `xl`, `cl` (64 each) are the wave's LDS;
// `s` is the calling lane's sink (every lane of the wave writes into the same path).
__device__ __forceinline__ void emit_arcs(const ArcJobs& aj, const Sink& s, float hw, ArcJob* jl,
                                          uint32_t* xl, uint32_t* cl) {
    const int l = lane_id();
    const uint32_t c0 = aj.has_sp ? (1u << aj.sp.n) - 1u : 0u;
    const uint32_t c1 = aj.has_sn ? (1u << aj.sn.n) - 1u : 0u;
    const uint32_t mine = c0 + c1;
    const uint32_t incl = incl_scan(mine);
    const uint32_t total = readlane_u(incl, WAVE - 1);
    if (total == 0) return;
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();  // earlier readers of the LDS slots are done
    if (c0) jl[2 * l] = aj.sp;
    if (c1) jl[2 * l + 1] = aj.sn;
    xl[l] = incl - mine;
    cl[l] = c0;
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
    for (uint32_t base = 0; base < total; base += WAVE) {
        const uint32_t g = base + (uint32_t)l;
        if (g >= total) continue;
        // owner lane: the last lane whose first node is <= g
        int lo = 0, hi = WAVE - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (xl[mid] <= g) lo = mid;
            else hi = mid - 1;
        }
        uint32_t r = g - xl[lo];
        const uint32_t oc0 = cl[lo];
        const ArcJob a = r < oc0 ? jl[2 * lo] : jl[2 * lo + 1];
        const uint32_t p = r < oc0 ? r : r - oc0;
        float a0 = a.a0, a1 = a.a1;
        uint32_t va = a.va, vb = a.vb, idx = 0, size = (1u << a.n) - 1u;
        float mid = (a0 + a1) * 0.5f;
        while (idx != p) {
            const uint32_t half = (size - 1u) >> 1;
            const uint32_t cur = a.first_id + idx;
            if (p <= idx + half) {
                a1 = mid;
                vb = cur;
                idx += 1;
            } else {
                a0 = mid;
                va = cur;
                idx += 1 + half;
            }
            size = half;
            mid = (a0 + a1) * 0.5f;
        }
        Sink w = s;  // this lane's copy of the path's sink, placed at the node
        w.vid = a.first_id + p;
        const uint32_t id = sink_sv(w, mk(a.cx, a.cy), mk(t_cos(mid), t_sin(mid)), hw);
        if (3u * p + 3u <= a.ileft) {
            uint32_t* o = a.tri + 3u * p;
            o[0] = va;
            o[1] = id;
            o[2] = vb;
        } else if (s.err) {
            atomicOr(s.err, 2u);
        }
    }
}

// The whole `count > 1` branch of fixed_width_step_impl (stroke.rs:1558-1625) for the
// join at `join` (previous kept point `prev`, next kept position `nextp`).
__device__ __forceinline__ uint32_t process_join(const EPD& prev, EPD& join, vec2 nextp,
                                                 bool with_edge, const SParams& p, Sink& s) {
    JInfo J;
    J.x = join;
    join_state(prev.pos, J, nextp, p);
    join_weird(J, prev.pos, prev.next[SP], prev.next[SN]);
    join_emit(J, prev, with_edge, p, s);
    join = J.x;
    return join_code(J);
}

// EPD of an already processed point from what its lane published.
__device__ __forceinline__ void epd_from_pub(EPD& e, vec2 pos, vec2 next_p, vec2 next_n,
                                             uint32_t code, uint32_t base, int join) {
    epd_init(e, pos, false, join);
    e.next[SP] = next_p;
    e.next[SN] = next_n;
    ids_of(code, base, e.pv, e.nv);
    e.fold[SP] = (code & LC_FOLD_P) ? 1 : 0;
    e.fold[SN] = (code & LC_FOLD_N) ? 1 : 0;
}

// Side points of the sub-path's first point once the second one is known
// (stroke.rs:1538-1556).
__device__ __forceinline__ void first_point_sides(EPD& first, vec2 nextp, float hw) {
    vec2 edge = sub(nextp, first.pos);
    float len = vlen(edge);
    vec2 t = divs(edge, len);
    vec2 n = mul(mk(-t.y, t.x), hw);
    first.next[SP] = add(first.pos, n);
    first.next[SN] = sub(first.pos, n);
}

// tessellate_round_cap (stroke.rs:2459-2524)
__device__ __forceinline__ void round_cap(vec2 center, float radius, vec2 start_normal,
                                          uint32_t start_v, uint32_t end_v, vec2 edge_normal,
                                          bool is_start, const SParams& p, Sink& s) {
    if (radius < p.tol) return;
    float start_angle = angle_of(start_normal);
    float diff = angle_to(start_angle, angle_of(edge_normal));
    float mid_angle = start_angle + diff;
    float end_angle = mid_angle + diff;
    // radius is half_width at every call (caps of the path): the per-path step applies
    uint32_t nsub = radius == p.hw ? path_arc_subdivisions(diff, p) : arc_subdivisions(diff, radius, p.tol);
    uint32_t mid = sink_sv(s, center, normalize(edge_normal), radius);
    sink_tri(s, start_v, mid, end_v);
    tessellate_arc(start_angle, mid_angle, start_v, mid, nsub, center, radius, s);
    tessellate_arc(mid_angle, end_angle, mid, end_v, nsub, center, radius, s);
}

// tessellate_last_edge (stroke.rs:2130-2206)
__device__ __forceinline__ void last_edge(const EPD& p0, EPD& p1, bool is_first,
                                          const SParams& p, Sink& s) {
    vec2 v = sub(p1.pos, p0.pos);
    for (int side = 0; side < 2; ++side) {
        vec2 side_position = p1.prev[side];
        if (p.end_cap != LYG_CAP_ROUND) {
            float clip = p.end_cap == LYG_CAP_SQUARE ? p.hw : 0.0f;
            vec2 normal = normalize(sub(p1.pos, p0.pos));
            vec2 inter;
            if (!line_isect(add(p1.pos, mul(normal, clip)), perp(normal), side_position,
                            sub(side_position, p0.next[side]), &inter))
                inter = p1.prev[side];
            p1.prev[side] = inter;
        }
        p1.pv[side] = sink_sv(s, p1.pos, divs(sub(p1.prev[side], p1.pos), p.hw), p.hw);
    }
    if (!is_first) add_edge_triangles(p0, p1, s);
    if (p.end_cap == LYG_CAP_ROUND)
        round_cap(p1.pos, p.hw, sub(p1.prev[SP], p1.pos), p1.pv[SP], p1.pv[SN], v, false, p, s);
}

// tessellate_first_edge (stroke.rs:2208-2276)
__device__ __forceinline__ void first_edge(EPD& first, const EPD& second, const SParams& p,
                                           Sink& s) {
    for (int side = 0; side < 2; ++side) {
        vec2 side_position = first.next[side];
        if (p.start_cap != LYG_CAP_ROUND) {
            float clip = p.start_cap == LYG_CAP_SQUARE ? p.hw : 0.0f;
            vec2 normal = normalize(sub(first.pos, second.pos));
            vec2 inter;
            if (!line_isect(add(first.pos, mul(normal, clip)), perp(normal), side_position,
                            sub(side_position, second.prev[side]), &inter))
                inter = first.next[side];
            side_position = inter;
        }
        first.nv[side] = sink_sv(s, first.pos, divs(sub(side_position, first.pos), p.hw), p.hw);
    }
    add_edge_triangles(first, second, s);
    if (p.start_cap == LYG_CAP_ROUND)
        round_cap(first.pos, p.hw, sub(first.next[SN], first.pos), first.nv[SN], first.nv[SP],
                  sub(first.pos, second.pos), true, p, s);
}

// tessellate_empty_square_cap / tessellate_empty_round_cap (stroke.rs:2526-2609)
__device__ __forceinline__ void empty_cap(vec2 pos, const SParams& p, Sink& s) {
    if (p.start_cap == LYG_CAP_SQUARE) {
        uint32_t a = sink_sv(s, pos, mk(1.0f, 1.0f), p.hw);
        uint32_t b = sink_sv(s, pos, mk(1.0f, -1.0f), p.hw);
        uint32_t c = sink_sv(s, pos, mk(-1.0f, -1.0f), p.hw);
        uint32_t d = sink_sv(s, pos, mk(-1.0f, 1.0f), p.hw);
        sink_tri(s, a, b, c);
        sink_tri(s, a, c, d);
    } else if (p.start_cap == LYG_CAP_ROUND) {
        uint32_t l = sink_sv(s, pos, mk(-1.0f, 0.0f), p.hw);
        uint32_t r = sink_sv(s, pos, mk(1.0f, 0.0f), p.hw);
        round_cap(pos, p.hw, mk(-1.0f, 0.0f), l, r, mk(0.0f, 1.0f), true, p, s);
        round_cap(pos, p.hw, mk(1.0f, 0.0f), r, l, mk(0.0f, -1.0f), false, p, s);
    }
}

// Everything a sub-path's end needs, gathered from the window or the carry.
struct EndInput {
    uint32_t m1;        // kept points in the sub-path
    bool close, mnec;
    vec2 pm1, pm;       // last two kept points P_{m-1}, P_m
    bool fm;            // is_flattening_step of P_m
    vec2 pm1_next[2];   // side "next" points of P_{m-1} after its join (m1 >= 3)
    uint32_t pm1_code;  // layout code of the join at P_{m-1} (m1 >= 3)
    uint32_t pm1_base;  // id base of the join at P_{m-1} (m1 >= 3)
    vec2 f0, f1, f2;    // first three kept points of the sub-path
    bool ff1;
    uint32_t f1_base;   // id base of the join at P_1 (m1 >= 3)
};

// The sub-path's first point once the second is known (stroke.rs:1538-1556).
__device__ __forceinline__ void first_point_state(EPD& e, vec2 p0, vec2 p1, const SParams& p) {
    epd_init(e, p0, false, p.join);
    first_point_sides(e, p1, p.hw);
}

// P_1 as stored in firsts[1]: its join processed with P_0 before and P_2 after.
__device__ __forceinline__ void second_point_state(EPD& e, const EndInput& in, const SParams& p) {
    EPD f0;
    first_point_state(f0, in.f0, in.f1, p);
    epd_init(e, in.f1, in.ff1, p.join);
    Sink ids{in.f1_base, 0, 0, nullptr, in.f1_base, nullptr, false, 0, 0, nullptr};
    process_join(f0, e, in.f2, false, p, ids);
}

// End of a sub-path: StrokeBuilderImpl::end (stroke.rs:1179-1193) with close()
// (stroke.rs:1206-1279) and end_with_caps() (stroke.rs:1322-1378).
__device__ __noinline__ void end_subpath(const EndInput& in, const SParams& p, Sink& s) {
    bool mnec = in.mnec || (in.close && in.m1 == 1);
    if (in.close && in.m1 >= 3) {
        // Point buffer holds (..., P_{m-1}, P_m): a = second to last, b = last.
        EPD a, b;
        epd_from_pub(a, in.pm1, in.pm1_next[SP], in.pm1_next[SN], in.pm1_code, in.pm1_base,
                     p.join);
        epd_init(b, in.pm, in.fm, p.join);
        // step(firsts[0])
        if (sqlen(sub(b.pos, in.f0)) < p.thr) {
            b.pos = in.f0;
        } else {
            process_join(a, b, in.f0, true, p, s);
            a = b;
            epd_init(b, in.f0, false, p.join);
        }
        // step(firsts[1])
        EPD f1;
        second_point_state(f1, in, p);
        if (!(sqlen(sub(b.pos, f1.pos)) < p.thr)) {
            process_join(a, b, f1.pos, true, p, s);
            a = b;
            b = f1;
        }
        // Re-create the first two vertices of the edge towards the second endpoint.
        for (int side = 0; side < 2; ++side) {
            vec2 sp_ = a.has_single[side] ? a.single[side] : a.next[side];
            a.nv[side] = sink_sv(s, a.pos, divs(sub(sp_, a.pos), p.hw), p.hw);
        }
        add_edge_triangles(a, b, s);
        return;
    }
    if (mnec && in.m1 == 1) {
        empty_cap(in.f0, p, s);
        return;
    }
    if (in.m1 >= 2) {
        EPD p0, p1;
        if (in.m1 == 2) {
            first_point_state(p0, in.pm1, in.pm, p);
        } else {
            epd_from_pub(p0, in.pm1, in.pm1_next[SP], in.pm1_next[SN], in.pm1_code,
                         in.pm1_base, p.join);
        }
        epd_init(p1, in.pm, in.fm, p.join);
        vec2 t = normalize(sub(p1.pos, p0.pos));
        vec2 n = mul(mk(-t.y, t.x), p.hw);
        p1.prev[SP] = add(p1.pos, n);
        p1.prev[SN] = sub(p1.pos, n);
        last_edge(p0, p1, in.m1 == 2, p, s);
        if (in.m1 > 2) {
            first_point_state(p0, in.f0, in.f1, p);
            second_point_state(p1, in, p);
        }
        first_edge(p0, p1, p, s);
    }
}

}  // namespace lyg
