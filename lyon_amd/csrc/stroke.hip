// Batched fixed-width stroke tessellation for gfx950.
//
// Replaces StrokeTessellator::tessellate_path (crates/tessellation/src/stroke.rs:164-177,
// 872-952) for many paths at once.  One wave (64 lanes) owns one path at a time and
// pulls path ids from a device work queue (persistent kernel).  A path is walked as a
// stream of "steps" — the points lyon hands to fixed_width_step_impl: sub-path starts,
// line ends and every flattened curve point — in windows of 64 steps, one per lane:
//
//   1. decode up to 64 verbs into LDS (one lane per verb: flattening counts, curve
//      polynomials, sharp-turn splits of quadratics), wave prefix sum -> step offsets;
//   2. each lane evaluates its step (binary search of its verb, closed-form t chain,
//      Horner), then the "points too close" filter runs as a ballot — and only when
//      a lane trips it, as a 64-iteration serial replay of stroke.rs:1530-1536;
//   3. every kept point produces the join of the kept point before it; neighbours come
//      from ds_bpermute shuffles (or the carry from the previous window), ids from a
//      wave prefix sum of the per-join vertex/triangle counts;
//   4. END/CLOSE steps run the sub-path end logic in their own lane.
//
// The count pass (EMIT = false) stores per-path vertex/index counts; the emit pass
// recomputes everything and writes at the offsets the host scan produced, so global
// memory sees one read of the path and one write of the geometry.
#include <hip/hip_runtime.h>

#include "lyg_math.h"
#include "lyg_types.h"
#include "stroke_core.h"
#include "wave.h"

namespace lyg {

static constexpr int WAVES_PER_BLOCK = 4;

enum : uint32_t {
    K_BEGIN = 1u,
    K_LINE = 2u,
    K_QUAD = 3u,
    K_CUBIC = 4u,
    K_END = 5u,
};
enum : uint32_t { F_BEGIN = 1u, F_END = 2u, F_CLOSE = 4u, F_FLAT = 8u };
static constexpr uint32_t NO_CHAIN = 0xffffffffu;

struct VRec {
    vec2 a0, a1, a2, a3;  // curve A polynomial form (quadratic: a3 unused)
    vec2 b0, b1, b2;      // curve B (second half of a split quadratic)
    vec2 toA, toB;        // exact last points; BEGIN/LINE keep their point in toA
    float sA, sB;         // flattening step 1/n
    uint32_t nA, nB;      // flattened point counts
    uint32_t off;         // first step of this verb in the chunk
    uint32_t kind;        // K_* | (close << 8)
};

// What a lane publishes about the kept point before it (for the next lane's join).
struct Pub {
    vec2 next[2];
    uint32_t code;
    uint32_t base;
};

// Cross-window state of the path being walked (uniform across the wave).
struct Carry {
    // last three items, most recent first
    vec2 pos[3];
    uint32_t flags[3];
    uint32_t isp[3];
    Pub pub[3];
    vec2 last_kept;
    uint32_t sp_cnt;
    bool mnec;
    bool prev_end;  // last step was an END
    // first three kept points of the current sub-path
    vec2 f0, f1, f2;
    bool ff1;
    uint32_t f1_base;
    uint32_t run_v, run_i;  // vertices / triangles emitted so far in this path
};

// Slot j of three values, selected field by field on values (a select between struct
// lvalues would become a select of addresses and pin the carry in private memory).
__device__ __forceinline__ uint32_t sel3(uint32_t a, uint32_t b, uint32_t c, int j) {
    uint32_t r = j == 1 ? b : c;
    return j == 0 ? a : r;
}
__device__ __forceinline__ float sel3f(float a, float b, float c, int j) {
    float r = j == 1 ? b : c;
    return j == 0 ? a : r;
}
__device__ __forceinline__ vec2 sel3(vec2 a, vec2 b, vec2 c, int j) {
    return mk(sel3f(a.x, b.x, c.x, j), sel3f(a.y, b.y, c.y, j));
}
__device__ __forceinline__ Pub sel3(const Pub& a, const Pub& b, const Pub& c, int j) {
    Pub r;
    r.next[0] = sel3(a.next[0], b.next[0], c.next[0], j);
    r.next[1] = sel3(a.next[1], b.next[1], c.next[1], j);
    r.code = sel3(a.code, b.code, c.code, j);
    r.base = sel3(a.base, b.base, c.base, j);
    return r;
}

// Value of lane `src`, or of carry slot -src-1 when src < 0.  The carry slots are read
// with constant indexes so the carry stays in registers.
__device__ __forceinline__ vec2 gather_v(vec2 v, int src, const vec2 c[3]) {
    vec2 s = shfl_v(v, src < 0 ? 0 : src);
    if (src >= 0) return s;
    return sel3(c[0], c[1], c[2], -src - 1);
}
__device__ __forceinline__ uint32_t gather_u(uint32_t v, int src, const uint32_t c[3]) {
    uint32_t s = shfl_u(v, src < 0 ? 0 : src);
    if (src >= 0) return s;
    return sel3(c[0], c[1], c[2], -src - 1);
}
__device__ __forceinline__ uint32_t gather_u3(uint32_t v, int src, uint32_t c0, uint32_t c1,
                                              uint32_t c2) {
    uint32_t s = shfl_u(v, src < 0 ? 0 : src);
    if (src >= 0) return s;
    return sel3(c0, c1, c2, -src - 1);
}
__device__ __forceinline__ Pub gather_pub(const Pub& v, int src, const Pub c[3]) {
    Pub r;
    int s = src < 0 ? 0 : src;
    r.next[0] = shfl_v(v.next[0], s);
    r.next[1] = shfl_v(v.next[1], s);
    r.code = shfl_u(v.code, s);
    r.base = shfl_u(v.base, s);
    if (src < 0) r = sel3(c[0], c[1], c[2], -src - 1);
    return r;
}

__device__ __forceinline__ vec2 ldp(const float* pts, uint32_t n, uint32_t i) {
    if (i >= n) return mk(__builtin_nanf(""), __builtin_nanf(""));
    const float2 v = reinterpret_cast<const float2*>(pts)[i];
    return mk(v.x, v.y);
}

__device__ __forceinline__ vec2 horner(const VRec& r, bool b, bool cubic, float t) {
    vec2 v = b ? r.b0 : r.a0;
    float tt = t;
    v = add(v, mul(b ? r.b1 : r.a1, tt));
    tt = tt * t;
    v = add(v, mul(b ? r.b2 : r.a2, tt));
    if (cubic) {
        tt = tt * t;
        v = add(v, mul(r.a3, tt));
    }
    return v;
}

// Decode verb `vi` of the path into rec.  `p` is the index of its first point.
__device__ __forceinline__ void decode_verb(uint8_t verb, const float* pts, uint32_t npts,
                                            uint32_t p, float tol, VRec& r) {
    const float nan = __builtin_nanf("");
    r.nA = 0;
    r.nB = 0;
    r.sA = r.sB = 0.0f;
    r.a0 = r.a1 = r.a2 = r.a3 = mk(0.0f, 0.0f);
    r.b0 = r.b1 = r.b2 = mk(0.0f, 0.0f);
    r.toA = r.toB = mk(nan, nan);
    vec2 from = p > 0 ? ldp(pts, npts, p - 1) : mk(nan, nan);
    switch (verb) {
        case LYG_VERB_BEGIN:
            r.kind = K_BEGIN;
            r.toA = ldp(pts, npts, p);
            r.nA = 1;
            break;
        case LYG_VERB_LINE:
            r.kind = K_LINE;
            r.toA = ldp(pts, npts, p);
            r.nA = 1;
            break;
        case LYG_VERB_QUADRATIC: {
            r.kind = K_QUAD;
            vec2 ctrl = ldp(pts, npts, p), to = ldp(pts, npts, p + 1);
            float ts = find_sharp_turn(from, ctrl, to);
            vec2 A0 = from, A1 = ctrl, A2 = to;
            if (ts > 0.0f) {
                // QuadraticBezierSegment::split (quadratic_bezier.rs:208-223)
                vec2 sp = quad_sample(from, ctrl, to, ts);
                A1 = lerp(from, ctrl, ts);
                A2 = sp;
                vec2 B0 = sp, B1 = lerp(ctrl, to, ts), B2 = to;
                float nb = quad_wang(B0, B1, B2, tol);
                r.sB = 1.0f / nb;
                r.nB = seg_count(nb);
                r.b0 = B0;
                r.b1 = mul(sub(B1, B0), 2.0f);
                r.b2 = sub(add(B0, B2), mul(B1, 2.0f));
                r.toB = B2;
            }
            float na = quad_wang(A0, A1, A2, tol);
            r.sA = 1.0f / na;
            r.nA = seg_count(na);
            r.a0 = A0;
            r.a1 = mul(sub(A1, A0), 2.0f);
            r.a2 = sub(add(A0, A2), mul(A1, 2.0f));
            r.toA = A2;
            break;
        }
        case LYG_VERB_CUBIC: {
            r.kind = K_CUBIC;
            vec2 c1 = ldp(pts, npts, p), c2 = ldp(pts, npts, p + 1), to = ldp(pts, npts, p + 2);
            float n = cubic_wang(from, c1, c2, to, tol);
            r.sA = 1.0f / n;
            r.nA = seg_count(n);
            // polynomial_form (cubic_bezier.rs:1318-1325)
            r.a0 = from;
            r.a1 = mul(sub(c1, from), 3.0f);
            r.a2 = add(sub(mul(from, 3.0f), mul(c1, 6.0f)), mul(c2, 3.0f));
            r.a3 = add(sub(to, from), mul(sub(c1, c2), 3.0f));
            r.toA = to;
            break;
        }
        case LYG_VERB_CLOSE:
            r.kind = K_END | (1u << 8);
            r.nA = 1;
            break;
        default:  // END
            r.kind = K_END;
            r.nA = 1;
            break;
    }
}

__device__ __forceinline__ uint32_t verb_points(uint8_t v) {
    switch (v) {
        case LYG_VERB_BEGIN:
        case LYG_VERB_LINE:
        case LYG_VERB_CLOSE:
            return 1;
        case LYG_VERB_QUADRATIC:
            return 2;
        case LYG_VERB_CUBIC:
            return 3;
        default:
            return 0;
    }
}

// The carry at the start of a path.
__device__ __forceinline__ void carry_init(Carry& c) {
    const float nan = __builtin_nanf("");
    for (int k = 0; k < 3; ++k) {
        c.pos[k] = mk(nan, nan);
        c.flags[k] = 0;
        c.isp[k] = 0;
        c.pub[k].next[0] = c.pub[k].next[1] = mk(nan, nan);
        c.pub[k].code = LC_FIRST;
        c.pub[k].base = INVALID_ID;
    }
    c.last_kept = mk(nan, nan);
    c.sp_cnt = 0;
    c.mnec = false;
    c.prev_end = true;
    c.f0 = c.f1 = c.f2 = mk(nan, nan);
    c.ff1 = false;
    c.f1_base = INVALID_ID;
    c.run_v = 0;
    c.run_i = 0;
}

// The carry a part of a split path starts its prologue from: inside the sub-path (three
// or more points kept before it), nothing known about the last items, the sub-path's
// first points from the path's head walk.  Ids that refer to the head's vertices carry
// the marker bit so the compaction resolves them against the path instead of the part.
__device__ __forceinline__ void carry_mid_path(Carry& c, const StrokeHeads& H) {
    carry_init(c);
    for (int k = 0; k < 3; ++k) c.isp[k] = 3;
    c.sp_cnt = 3;
    c.prev_end = false;
    c.f0 = mk(H.f0.x, H.f0.y);
    c.f1 = mk(H.f1.x, H.f1.y);
    c.f2 = mk(H.f2.x, H.f2.y);
    c.ff1 = H.ff1 != 0;
    c.f1_base = H.f1_base == INVALID_ID ? INVALID_ID : (ID_HEAD_MARK | H.f1_base);
}

// The part of the carry the rest of a walk depends on, for checking that a part's
// prologue rebuilt the state the previous part ended in.  Counts of kept points matter
// only up to 3 (joins need two points before them, the sub-path end distinguishes 1, 2
// and more), so they are stored clamped; ids are relative to the part's first vertex.
__device__ __forceinline__ void carry_store(const Carry& c, CarryRec* r) {
    CarryRec x;
    for (int k = 0; k < 3; ++k) {
        x.pos[k] = F2{c.pos[k].x, c.pos[k].y};
        x.next[k][0] = F2{c.pub[k].next[0].x, c.pub[k].next[0].y};
        x.next[k][1] = F2{c.pub[k].next[1].x, c.pub[k].next[1].y};
        x.flags[k] = c.flags[k];
        x.isp[k] = min(c.isp[k], 3u);
        x.code[k] = c.pub[k].code;
        x.base[k] = c.pub[k].base;
    }
    x.last_kept = F2{c.last_kept.x, c.last_kept.y};
    x.sp = min(c.sp_cnt, 3u);
    x.prev_end = c.prev_end ? 1u : 0u;
    x.mnec = c.mnec ? 1u : 0u;
    x.pad = 0;
    *r = x;
}

// A range of one path's walk.  Steps are addressed as (verb chunk k, step w within the
// chunk's steps), w a multiple of the window.  Windows before (ke, we) are a prologue:
// walked to rebuild the carry, nothing emitted and their counts dropped.  The range
// ends before (k1, w1); k1 = the number of chunks ends it with the path.
struct WalkRange {
    uint32_t k0, w0, ke, we, k1, w1;
    uint32_t pt0;  // points consumed by the verbs before chunk k0
};

// Walks [R.k0, R.k1) of one path with the whole wave from the carry `c`, which it
// updates; c.run_v / c.run_i count the vertices / triangles from the range's emitted
// start.  EMIT: output to [v0, v_end) / [i0, i_end) of the output arrays.  `spec`: where
// the carry is stored when the prologue ends (null when there is none).
template <bool EMIT, bool RANGED = true>
__device__ __forceinline__ void stroke_walk(const StrokeLaunch& L, uint32_t gpath, const WalkRange& R,
                            Carry& c, uint32_t path_v0, uint32_t path_v_end, uint32_t path_i0,
                            uint32_t path_i_end, CarryRec* spec, VRec* recs, ArcJob* arc_jobs,
                            uint32_t* arc_first, uint32_t* arc_c0) {
    const int l = lane_id();
    const uint32_t vb = L.in.verb_off[gpath], ve = L.in.verb_off[gpath + 1];
    const uint32_t pb = L.in.point_off[gpath], pe = L.in.point_off[gpath + 1];
    const uint8_t* verbs = L.in.verbs + vb;
    const float* pts = L.in.points + 2ull * pb;
    const uint32_t nverbs = ve - vb, npts = pe - pb;
    const lyg_stroke_options o = L.opts[L.opts_stride ? gpath : 0];
    const SParams P = make_params(o, L.nsub_thr);
    const float nan = __builtin_nanf("");
    bool started = !RANGED || (R.k0 == R.ke && R.w0 == R.we);

    uint32_t pt_carry = RANGED ? R.pt0 : 0u;
    for (uint32_t v0 = RANGED ? R.k0 * WAVE : 0u; v0 < nverbs; v0 += WAVE) {
        const uint32_t kch = v0 / WAVE;
        if (RANGED && (kch > R.k1 || (kch == R.k1 && R.w1 == 0))) break;
        // ---- 1. decode a chunk of verbs ------------------------------------------------
        const uint32_t nch = min((uint32_t)WAVE, nverbs - v0);
        uint8_t verb = LYG_VERB_END;
        uint32_t np = 0;
        if ((uint32_t)l < nch) {
            verb = verbs[v0 + l];
            np = verb_points(verb);
        }
        uint32_t np_incl = incl_scan(np);
        uint32_t first_pt = pt_carry + np_incl - np;
        pt_carry += readlane_u(np_incl, WAVE - 1);
        VRec r;
        uint32_t cnt = 0;
        if ((uint32_t)l < nch) {
            decode_verb(verb, pts, npts, first_pt, P.tol, r);
            cnt = r.nA + r.nB;
        }
        uint32_t cnt_incl = incl_scan(cnt);
        r.off = cnt_incl - cnt;
        const uint32_t nsteps = readlane_u(cnt_incl, WAVE - 1);
        wave_sync();
        if ((uint32_t)l < nch) recs[l] = r;
        wave_sync();

        // ---- 2. windows of 64 steps ----------------------------------------------------
        // t chain of the curve half the previous window ended in (key, steps, units)
        uint32_t ch_key = NO_CHAIN, ch_j = 0;
        uint64_t ch_T = 0;
        for (uint32_t w0 = RANGED && kch == R.k0 ? R.w0 : 0u; w0 < nsteps; w0 += WAVE) {
            if (RANGED && kch == R.k1 && w0 >= R.w1) break;
            const bool silent = RANGED && (kch < R.ke || (kch == R.ke && w0 < R.we));
            if (!silent && !started) {
                // the prologue ends: ids from here on count from the range's start
                started = true;
                for (int k = 0; k < 3; ++k)
                    if (c.pub[k].base != INVALID_ID) c.pub[k].base -= c.run_v;
                c.run_v = 0;
                c.run_i = 0;
                if (spec && l == 0) carry_store(c, spec);
            }
            const uint32_t g = w0 + (uint32_t)l;
            const bool valid = g < nsteps;
            vec2 pos = mk(nan, nan);
            uint32_t flags = 0;
            uint32_t my_key = NO_CHAIN, my_j = 0;  // the t chain this lane evaluated
            uint64_t my_T = 0;
            if (valid) {
                // last verb whose offset <= g
                int lo = 0, hi = (int)nch - 1;
                while (lo < hi) {
                    int mid = (lo + hi + 1) >> 1;
                    if (recs[mid].off <= g) lo = mid;
                    else hi = mid - 1;
                }
                const VRec& vr = recs[lo];
                uint32_t k = g - vr.off;
                uint32_t kind = vr.kind & 0xffu;
                if (kind == K_BEGIN) {
                    pos = vr.toA;
                    flags = F_BEGIN;
                } else if (kind == K_LINE) {
                    pos = vr.toA;
                } else if (kind == K_END) {
                    flags = F_END | ((vr.kind >> 8) ? F_CLOSE : 0u);
                } else {
                    const bool cubic = kind == K_CUBIC;
                    const bool inA = k < vr.nA;
                    const uint32_t kk = inA ? k : k - vr.nA;
                    if (kk + 1 == (inA ? vr.nA : vr.nB)) {
                        pos = inA ? vr.toA : vr.toB;
                    } else {
                        // t after kk + 1 steps, continuing the previous window's chain
                        // when this curve half carries on from it
                        const float sstep = inA ? vr.sA : vr.sB;
                        const uint32_t key = ((uint32_t)lo << 1) | (inA ? 0u : 1u);
                        const uint32_t j = kk + 1;
                        const bool cont = key == ch_key && ch_j < j;
                        float t;
                        uint64_t T;
                        if (!cont && j <= (uint32_t)WAVE) {
                            // a curve half that starts in this window: lyon's serial sum
                            t = 0.0f;
                            for (uint32_t q = 0; q < j; ++q) t = t + sstep;
                            T = tchain_units_of(sstep, t);
                        } else {
                            T = tchain_units_from(tchain_mant(sstep), j, cont ? ch_j : 0u,
                                                  cont ? ch_T : 0ull);
                            t = tchain_value(sstep, T);
                        }
                        pos = horner(vr, !inA, inA && cubic, t);
                        flags = F_FLAT;
                        my_key = key;
                        my_j = j;
                        my_T = T;
                    }
                }
            }
            const bool isB = (flags & F_BEGIN) != 0;
            const bool isE = (flags & F_END) != 0;
            const uint64_t endm = ballot(valid && isE);
            const uint64_t begm = ballot(valid && isB);
            // a new sub-path starts at a BEGIN and right after an END
            const uint64_t restart = begm | (endm << 1) | (c.prev_end ? 1ull : 0ull);

            // ---- keep filter (stroke.rs:1530-1536) ----
            vec2 prev_pos = shfl_v(pos, l > 0 ? l - 1 : 0);
            if (l == 0) prev_pos = c.last_kept;
            bool starts = (restart >> l) & 1ull;
            bool close_pred = valid && !isE && !starts && (sqlen(sub(prev_pos, pos)) < P.thr);
            uint64_t keptm;
            bool my_mnec;
            const uint64_t validm = ballot(valid);
            if (ballot(close_pred) == 0ull) {
                keptm = validm & ~endm;
                // mnec only changes on a drop; a restart clears it.
                uint64_t rb = restart & mask_le(l);
                my_mnec = rb ? false : c.mnec;
            } else {
                // Serial replay of the filter over the window.
                vec2 lk = c.last_kept;
                uint32_t cntk = c.sp_cnt;
                bool mn = c.mnec;
                bool after_end = c.prev_end;
                bool my_kept = false;
                my_mnec = false;
                const int nvalid = __builtin_popcountll(validm);
                for (int i = 0; i < nvalid; ++i) {
                    vec2 pi = mk(readlane_f(pos.x, i), readlane_f(pos.y, i));
                    uint32_t fi = readlane_u(flags, i);
                    bool kept_i = false;
                    if (fi & F_END) {
                        if (l == i) my_mnec = mn;
                        after_end = true;
                        cntk = 0;
                    } else {
                        if (fi & F_BEGIN) mn = false;
                        if ((fi & F_BEGIN) || after_end || cntk == 0) {
                            kept_i = true;
                            lk = pi;
                            cntk = 1;
                        } else if (sqlen(sub(lk, pi)) < P.thr) {
                            if (cntk == 1) mn = true;
                        } else {
                            kept_i = true;
                            lk = pi;
                            cntk += 1;
                        }
                        after_end = false;
                        if (l == i) my_mnec = mn;
                    }
                    if (l == i) my_kept = kept_i;
                }
                keptm = ballot(valid && my_kept);
            }
            const bool kept = (keptm >> l) & 1ull;
            const bool item = valid && (kept || isE);
            const uint64_t itemm = keptm | endm;

            // index of this item within its sub-path (kept points before it)
            uint32_t isp;
            {
                uint64_t rb = restart & mask_le(l);
                if (rb) {
                    int b = highest_bit(rb);
                    isp = __builtin_popcountll(keptm & mask_lt(l) & ~mask_lt(b));
                } else {
                    isp = c.sp_cnt + __builtin_popcountll(keptm & mask_lt(l));
                }
            }

            // previous items: lane index, or -1/-2/-3 for the carry slots
            int p1 = highest_bit(itemm & mask_lt(l));
            if (p1 < 0) p1 = -1;
            int p2s = shfl_i(p1, p1 < 0 ? 0 : p1);
            int p2 = p1 >= 0 ? p2s : p1 - 1;
            int p3s = shfl_i(p1, p2 < 0 ? 0 : p2);
            int p3 = p2 >= 0 ? p3s : p2 - 1;
            (void)p3;

            const vec2 X = gather_v(pos, p1, c.pos);
            const uint32_t Xf = gather_u(flags, p1, c.flags);
            const vec2 PP = gather_v(pos, p2, c.pos);

            // firsts of this item's sub-path (only a sub-path end reads them)
            vec2 F0 = c.f0, F1 = c.f1, F2 = c.f2;
            bool FF1 = c.ff1;
            int f2lane = -1;
            if (endm != 0) {
                uint64_t rb = restart & mask_le(l);
                int rbl = rb ? highest_bit(rb) : -1;
                uint64_t m0 = ballot(kept && isp == 0);
                uint64_t m1 = ballot(kept && isp == 1);
                uint64_t m2 = ballot(kept && isp == 2);
                int s0 = highest_bit(m0 & mask_lt(l));
                int s1 = highest_bit(m1 & mask_lt(l));
                int s2 = highest_bit(m2 & mask_lt(l));
                vec2 g0 = shfl_v(pos, s0 < 0 ? 0 : s0);
                vec2 g1 = shfl_v(pos, s1 < 0 ? 0 : s1);
                uint32_t gf1 = shfl_u(flags, s1 < 0 ? 0 : s1);
                vec2 g2 = shfl_v(pos, s2 < 0 ? 0 : s2);
                if (s0 >= 0 && s0 >= rbl) F0 = g0;
                if (s1 >= 0 && s1 >= rbl) {
                    F1 = g1;
                    FF1 = (gf1 & F_FLAT) != 0;
                }
                if (s2 >= 0 && s2 >= rbl) {
                    F2 = g2;
                    f2lane = s2;
                }
            }

            // ---- 3. per-item work -----------------------------------------------------
            // A kept point with isp >= 2 owns the join at the kept point X before it
            // (previous kept PP, next = this point); isp == 1 owns the side points of the
            // sub-path's first point; END owns the sub-path end.
            const bool is_join = item && kept && isp >= 2;
            Pub pub;
            pub.next[0] = pub.next[1] = mk(nan, nan);
            pub.code = LC_FIRST;
            pub.base = INVALID_ID;
            JInfo J;
            if (item && kept && isp == 1) {
                EPD f;
                epd_init(f, X, false, P.join);
                first_point_sides(f, pos, P.hw);
                pub.next[0] = f.next[0];
                pub.next[1] = f.next[1];
            }
            if (is_join) {
                epd_init(J.x, X, (Xf & F_FLAT) != 0, P.join);
                join_state(PP, J, pos, P);
                pub.next[0] = J.x.next[0];
                pub.next[1] = J.x.next[1];
            }
            // the fast path's fold-back test reads the previous point's side points
            Pub prv = gather_pub(pub, p1, c.pub);
            if (is_join) {
                join_weird(J, PP, prv.next[0], prv.next[1]);
                pub.code = join_code(J);
            }
            const uint32_t fake_base = 0x80000000u + ((uint32_t)l << 24);
            if (is_join) pub.base = fake_base;
            // only the code and the base of the published record changed since
            prv.code = gather_u3(pub.code, p1, c.pub[0].code, c.pub[1].code, c.pub[2].code);
            prv.base = gather_u3(pub.base, p1, c.pub[0].base, c.pub[1].base, c.pub[2].base);

            // END: everything the sub-path end needs
            EndInput ein;
            ein.m1 = isp;
            ein.close = (flags & F_CLOSE) != 0;
            ein.mnec = my_mnec;
            ein.pm = X;
            ein.fm = (Xf & F_FLAT) != 0;
            ein.pm1 = PP;
            ein.pm1_next[0] = prv.next[0];
            ein.pm1_next[1] = prv.next[1];
            ein.pm1_code = prv.code;
            ein.pm1_base = prv.base;
            ein.f0 = F0;
            ein.f1 = F1;
            ein.f2 = F2;
            ein.ff1 = FF1;
            if (endm != 0) {
                uint32_t f1b = shfl_u(pub.base, f2lane < 0 ? 0 : f2lane);
                ein.f1_base = f2lane >= 0 ? f1b : c.f1_base;
            }

            uint32_t nv = 0, nt = 0;
            if (is_join) join_counts(J, prv.code, isp >= 3, &nv, &nt);
            if (item && isE) {
                // the call takes addresses: it gets copies made here, so the
                // originals never live in private memory
                Sink sk{fake_base, 0, 0, nullptr, 0, nullptr, false, 0, 0, nullptr};
                const EndInput e2 = ein;
                const SParams P2 = P;
                end_subpath(e2, P2, sk);
                nv = sk.nv;
                nt = sk.ni;
            }

            // ---- 4. ids: wave prefix sums ----
            const uint32_t nv_incl = incl_scan(nv);
            const uint32_t nt_incl = incl_scan(nt);
            const uint32_t my_vbase = c.run_v + nv_incl - nv;  // path-relative
            const uint32_t my_tbase = c.run_i + nt_incl - nt;
            // ids: vertex_offset + position in the batch, or in the path with local_ids
            const uint32_t id_v0 = L.local_ids ? 0u : path_v0;
            const uint32_t real_base = EMIT ? (L.vertex_offset + id_v0 + my_vbase) : my_vbase;
            if (is_join) pub.base = real_base;
            prv.base = gather_u3(pub.base, p1, c.pub[0].base, c.pub[1].base, c.pub[2].base);
            if (endm != 0) {
                uint32_t f1b = shfl_u(pub.base, f2lane < 0 ? 0 : f2lane);
                ein.f1_base = f2lane >= 0 ? f1b : c.f1_base;
            }
            ein.pm1_base = prv.base;

            if (EMIT && !silent) {
                // vertex id -> slot: id - (vertex_offset + id_v0 - path_v0), modulo 2^32
                Sink se{real_base, 0, 0, L.out_vertices, L.vertex_offset + id_v0 - path_v0, nullptr, true,
                        path_v_end, 0, L.err};
                const uint64_t islot = (uint64_t)path_i0 + 3ull * my_tbase;
                se.iout = L.out_indices + islot;
                se.ileft = islot < path_i_end ? (uint32_t)(path_i_end - islot) : 0u;
                ArcJobs aj;  // round-join arcs, emitted by the whole wave below
                if (is_join) {
                    EPD pj;
                    epd_from_pub(pj, PP, prv.next[0], prv.next[1], prv.code, prv.base, P.join);
                    join_emit(J, pj, isp >= 3, P, se, &aj);
                }
                if (item && isE) {
                    Sink s2 = se;
                    const EndInput e2 = ein;
                    const SParams P2 = P;
                    end_subpath(e2, P2, s2);
                    se = s2;
                }
                emit_arcs(aj, se, P.hw, arc_jobs, arc_first, arc_c0);
            }

            // ---- 5. carry to the next window ----
            {
                const uint32_t tot_v = readlane_u(nv_incl, WAVE - 1);
                const uint32_t tot_t = readlane_u(nt_incl, WAVE - 1);
                // last three items
                uint64_t im = itemm;
                int lanes[3] = {-1, -1, -1};
                for (int k = 0; k < 3; ++k) {
                    int h = highest_bit(im);
                    lanes[k] = h;
                    if (h >= 0) im &= ~(1ull << h);
                }
                vec2 npos[3];
                uint32_t nflags[3], nisp[3];
                Pub npub[3];
                int taken = 0;
                for (int k = 0; k < 3; ++k) {
                    int s = lanes[k] < 0 ? 0 : lanes[k];
                    vec2 pv_ = readlane_v(pos, s);
                    uint32_t fv = readlane_u(flags, s);
                    uint32_t iv = readlane_u(isp, s);
                    Pub pb_;
                    pb_.next[0] = readlane_v(pub.next[0], s);
                    pb_.next[1] = readlane_v(pub.next[1], s);
                    pb_.code = readlane_u(pub.code, s);
                    pb_.base = readlane_u(pub.base, s);
                    if (lanes[k] >= 0) {
                        npos[k] = pv_;
                        nflags[k] = fv;
                        nisp[k] = iv;
                        npub[k] = pb_;
                        taken++;
                    }
                }
                vec2 cpos[3];
                uint32_t cfl[3], cisp[3];
                Pub cpub[3];
                for (int k = 0; k < 3; ++k) {
                    if (k < taken) {
                        cpos[k] = npos[k];
                        cfl[k] = nflags[k];
                        cisp[k] = nisp[k];
                        cpub[k] = npub[k];
                    } else {
                        // constant-index selects keep the carry in registers
                        const int j = k - taken;
                        cpos[k] = sel3(c.pos[0], c.pos[1], c.pos[2], j);
                        cfl[k] = sel3(c.flags[0], c.flags[1], c.flags[2], j);
                        cisp[k] = sel3(c.isp[0], c.isp[1], c.isp[2], j);
                        cpub[k] = sel3(c.pub[0], c.pub[1], c.pub[2], j);
                    }
                }
                for (int k = 0; k < 3; ++k) {
                    c.pos[k] = cpos[k];
                    c.flags[k] = cfl[k];
                    c.isp[k] = cisp[k];
                    c.pub[k] = cpub[k];
                }
                // keep-filter state
                int lk = highest_bit(keptm);
                vec2 lkp = readlane_v(pos, lk < 0 ? 0 : lk);
                const int lastv = highest_bit(validm);
                const bool last_is_end = lastv >= 0 && ((endm >> lastv) & 1ull);
                // sub-path counters after the window
                uint64_t rball = begm | (endm << 1) | (c.prev_end ? 1ull : 0ull);
                uint64_t rbw = rball & mask_le(lastv);
                uint32_t newcnt;
                bool newm;
                // mnec of the last lane, plus the effect of the last lane itself
                bool last_mnec = __builtin_bit_cast(int, readlane_u((uint32_t)my_mnec, lastv < 0 ? 0 : lastv)) != 0;
                if (last_is_end) {
                    newcnt = 0;
                    newm = false;
                } else if (rbw) {
                    int b = highest_bit(rbw);
                    newcnt = __builtin_popcountll(keptm & mask_le(lastv) & ~mask_lt(b));
                    newm = last_mnec;
                } else {
                    newcnt = c.sp_cnt + __builtin_popcountll(keptm);
                    newm = last_mnec;
                }
                // the t chain of the last step, for the next window
                {
                    const int src = lastv < 0 ? 0 : lastv;
                    ch_key = lastv < 0 ? NO_CHAIN : readlane_u(my_key, src);
                    ch_j = readlane_u(my_j, src);
                    ch_T = (uint64_t)readlane_u((uint32_t)my_T, src) |
                           ((uint64_t)readlane_u((uint32_t)(my_T >> 32), src) << 32);
                }
                if (lk >= 0) c.last_kept = lkp;
                if (last_is_end) c.last_kept = mk(nan, nan);
                // firsts of the (possibly new) current sub-path
                {
                    int rbl = rbw ? highest_bit(rbw) : -1;
                    uint64_t m0 = ballot(kept && isp == 0);
                    uint64_t m1 = ballot(kept && isp == 1);
                    uint64_t m2 = ballot(kept && isp == 2);
                    int s0 = highest_bit(m0), s1 = highest_bit(m1), s2 = highest_bit(m2);
                    vec2 g0 = readlane_v(pos, s0 < 0 ? 0 : s0);
                    vec2 g1 = readlane_v(pos, s1 < 0 ? 0 : s1);
                    uint32_t gf1 = readlane_u(flags, s1 < 0 ? 0 : s1);
                    vec2 g2 = readlane_v(pos, s2 < 0 ? 0 : s2);
                    uint32_t gb2 = readlane_u(pub.base, s2 < 0 ? 0 : s2);
                    if (s0 >= 0 && s0 >= rbl) c.f0 = g0;
                    if (s1 >= 0 && s1 >= rbl) {
                        c.f1 = g1;
                        c.ff1 = (gf1 & F_FLAT) != 0;
                    }
                    if (s2 >= 0 && s2 >= rbl) {
                        c.f2 = g2;
                        c.f1_base = gb2;
                    }
                }
                c.sp_cnt = newcnt;
                c.mnec = newm;
                c.prev_end = last_is_end;
                c.run_v += tot_v;
                c.run_i += tot_t;
            }
        }
        wave_sync();
    }
}

// Walks one whole path with the whole wave.  Returns (vertices, triangles).
template <bool EMIT>
__device__ void stroke_one_path(const StrokeLaunch& L, uint32_t gpath, VRec* recs,
                                ArcJob* arc_jobs, uint32_t* arc_first, uint32_t* arc_c0,
                                uint32_t* out_nv, uint32_t* out_ni) {
    uint32_t path_v0 = 0, path_i0 = 0, path_v_end = 0, path_i_end = 0;
    if (EMIT) {
        path_v0 = L.vtx_off[gpath];
        path_i0 = L.idx_off[gpath];
        path_v_end = L.vtx_off[gpath + 1];
        path_i_end = L.idx_off[gpath + 1];
        // a path that does not fit the output arrays is skipped and reported
        if (path_v_end > L.vertex_capacity || path_i_end > L.index_capacity) {
            if (lane_id() == 0 && L.err) atomicOr(L.err, 2u);
            return;
        }
    }
    Carry c;
    carry_init(c);
    const WalkRange R{0, 0, 0, 0, 0xffffffffu, 0, 0};
    stroke_walk<EMIT, false>(L, gpath, R, c, path_v0, path_v_end, path_i0, path_i_end, nullptr,
                             recs, arc_jobs, arc_first, arc_c0);
    *out_nv = c.run_v;
    *out_ni = c.run_i;
}

// One item of the split walk: a whole path, or a part of a long one (after its prologue),
// into its own staging room with item-local ids; stores its exact counts and, for parts,
// the carries the check compares.
template <bool EMIT>
__device__ void stroke_item(const StrokeLaunch& L, uint32_t item, VRec* recs, ArcJob* arc_jobs,
                            uint32_t* arc_first, uint32_t* arc_c0) {
    const StrokeItem it = L.items[item];
    const WalkRange R{it.k0, it.w0, it.ke, it.we, it.k1, it.w1, it.pt0};
    const bool first = (it.flags & ITEM_FIRST) != 0;
    if (EMIT && ((uint64_t)it.v_off + it.v_room > L.vertex_capacity ||
                 (uint64_t)it.i_off + it.i_room > L.index_capacity)) {
        if (lane_id() == 0 && L.err) atomicOr(L.err, 2u);
        return;
    }
    Carry c;
    if (first) carry_init(c);
    else carry_mid_path(c, L.heads[it.path]);
    stroke_walk<EMIT>(L, it.path, R, c, it.v_off, it.v_off + it.v_room, it.i_off,
                      it.i_off + it.i_room, first ? nullptr : L.item_spec + item, recs, arc_jobs,
                      arc_first, arc_c0);
    if (lane_id() == 0) {
        L.items[item].nv = c.run_v;
        L.items[item].ni = 3u * c.run_i;
        if (it.flags & ITEM_SPLIT) carry_store(c, L.item_end + item);
    }
}

// Fetches the next path id for the whole wave.  Every lane executes the atomic (lane 0
// adds 1, the others 0) so there is no divergent region around it: a divergent
// `if (lane == 0) atomicAdd` followed by readlane lets the control-flow structurizer park
// lane 0 outside the loop while the other lanes re-read a stale id.
__device__ __forceinline__ uint32_t next_path(uint32_t* counter) {
    uint32_t mine = atomicAdd(counter, lane_id() == 0 ? 1u : 0u);
    return readlane_u(mine, 0);
}

// Waves per SIMD: the walk is latency bound (one wave's window is a long dependent
// chain of shuffles, LDS reads and scans), so occupancy beats spill-free code.  Measured
// on the 10,000-path stroke workload: count pass 11.4 ms at 1 wave/SIMD, 7.2 at 2, 6.0 at
// 4; emit pass 16.0 ms at 1, 11.1 at 2, 15.7 at 4 (its larger state spills too much).
template <bool EMIT, bool ITEMS>
__global__ void __launch_bounds__(WAVE* WAVES_PER_BLOCK)
    LYG_WAVES_PER_EU(EMIT ? 2 : 4) stroke_kernel(StrokeLaunch L) {
    __shared__ VRec recs_all[WAVES_PER_BLOCK][WAVE];
    // the emit pass's cooperative arcs: two jobs per lane, first node and side split
    constexpr int AW = EMIT ? WAVES_PER_BLOCK : 1;
    __shared__ ArcJob arcs_all[AW][EMIT ? 2 * WAVE : 1];
    __shared__ uint32_t arc_first_all[AW][EMIT ? WAVE : 1], arc_c0_all[AW][EMIT ? WAVE : 1];
    const int w = threadIdx.x / WAVE;
    VRec* recs = recs_all[w];
    ArcJob* arc_jobs = arcs_all[EMIT ? w : 0];
    uint32_t* arc_first = arc_first_all[EMIT ? w : 0];
    uint32_t* arc_c0 = arc_c0_all[EMIT ? w : 0];
    const int l = lane_id();
    uint32_t pid = next_path(L.work_counter);
    if constexpr (ITEMS) {
        // parts of split paths and whole paths, as listed by stroke_cut_kernel
        const uint32_t n_work =
            L.n_extra_dev ? L.in.n_paths + (uint32_t)min(*L.n_extra_dev,
                                                          (unsigned long long)L.cap_extra)
                          : L.n_items;
        while (pid < n_work) {
            stroke_item<EMIT>(L, pid, recs, arc_jobs, arc_first, arc_c0);
            pid = next_path(L.work_counter);
        }
        return;
    }
    const uint32_t n_list = L.n_list_dev ? *L.n_list_dev : L.in.n_paths;
    while (pid < n_list) {
        const uint32_t gp = L.path_list ? L.path_list[pid] : L.in.first_path + pid;
        uint32_t nv = 0, ni = 0;
        stroke_one_path<EMIT>(L, gp, recs, arc_jobs, arc_first, arc_c0, &nv, &ni);
        if (!EMIT || L.store_counts) {
            // every lane holds the totals; lane 0's store is the only one kept
            uint32_t* dst_v = L.path_nv + gp;
            uint32_t* dst_i = L.path_ni + gp;
            if (l == 0) {
                *dst_v = nv;
                *dst_i = 3u * ni;
            }
        }
        pid = next_path(L.work_counter);
    }
}

// thr[k], k = 0..30: the smallest float x in [2^k, 2^(k+1)) whose round(log2(x)) is k + 1
// (2^(k+1) if none), by bisection over the binade's bit patterns with the log2 the
// subdivision count uses; the count of x in that binade is then k + (x >= thr[k]).
__global__ void nsub_thresholds_kernel(float* thr) {
    const int k = (int)threadIdx.x;
    if (k > 30) return;
    uint32_t lo = (uint32_t)(127 + k) << 23, hi = (uint32_t)(128 + k) << 23;
    while (lo < hi) {
        const uint32_t mid = lo + (hi - lo) / 2;
        if (__builtin_roundf(t_log2(u2f(mid))) >= (float)(k + 1)) hi = mid;
        else lo = mid + 1;
    }
    thr[k] = u2f(lo);
}

// ---- staged single pass -----------------------------------------------------------------
// Room for one path's staged output: two vertices and two triangles per flattened point
// (the fast path of every curve point), plus one full join per verb (four vertices and the
// round-join arcs on both sides) and caps per sub-path.  A join inside a curve that is not
// a fast-path join (a cusp) is not counted: such a path can outgrow its room, which the
// compaction detects from the exact counts; it is then emitted again on its own.
__global__ void __launch_bounds__(WAVE) stroke_stage_caps_kernel(StrokeLaunch L, uint32_t n,
                                                                 float room_scale,
                                                                 uint32_t* cap_v, uint32_t* cap_i) {
    // one wave per path, one lane per verb
    const uint32_t path = blockIdx.x;
    if (path >= n) return;
    const int l = lane_id();
    const lyg_stroke_options o = L.opts[L.opts_stride ? path : 0];
    const uint32_t vb = L.in.verb_off[path], ve = L.in.verb_off[path + 1];
    const uint32_t pb = L.in.point_off[path], pe = L.in.point_off[path + 1];
    const float* pts = L.in.points + 2ull * pb;
    const uint32_t npts = pe - pb;
    const float tol = o.tolerance;
    uint32_t steps = 0, subs = 0, pt_carry = 0;
    for (uint32_t v0 = vb; v0 < ve; v0 += WAVE) {
        const uint32_t v = v0 + (uint32_t)l;
        const uint8_t verb = v < ve ? L.in.verbs[v] : (uint8_t)LYG_VERB_END;
        const uint32_t np = v < ve ? verb_points(verb) : 0u;
        const uint32_t incl = incl_scan(np);
        const uint32_t first = pt_carry + incl - np;  // this verb's first point
        pt_carry += readlane_u(incl, WAVE - 1);
        if (v < ve) {
            // the current point is the point stored before the verb's own points
            const vec2 cur = ldp(pts, npts, first > 0 ? first - 1 : 0);
            const vec2 a = ldp(pts, npts, first), b = ldp(pts, npts, first + 1),
                       c = ldp(pts, npts, first + 2);
            uint32_t add = 2;
            if (verb == LYG_VERB_BEGIN) {
                subs += 1;
                add = 1;
            } else if (verb == LYG_VERB_QUADRATIC) {
                add = (tol > 0.0f ? min(seg_count(quad_wang(cur, a, b, tol)), 1u << 24) : 1u) + 2;
            } else if (verb == LYG_VERB_CUBIC) {
                add = (tol > 0.0f ? min(seg_count(cubic_wang(cur, a, b, c, tol)), 1u << 24) : 1u) + 1;
            }
            steps = min(steps + add, 1u << 25);  // 64 lanes' sum stays below 2^32
        }
    }
    // wave sums
    const uint64_t verbs = ve - vb;
    const uint64_t steps_all = readlane_u(incl_scan(steps), WAVE - 1);
    const uint64_t subs_all = readlane_u(incl_scan(subs), WAVE - 1);
    if (l != 0) return;
    const bool round = o.line_join == LYG_JOIN_ROUND || o.start_cap == LYG_CAP_ROUND ||
                       o.end_cap == LYG_CAP_ROUND;
    const float hw = o.line_width * 0.5f;
    uint64_t A = 0;
    if (round && tol > 0.0f && hw > 0.0f)
        A = arc_count(arc_subdivisions(2.0f * 3.14159265358979323846f, hw, tol));
    const double f = room_scale > 0.0f ? (double)room_scale : 1.0;
    const uint64_t cv = (uint64_t)(f * (double)(2 * steps_all + verbs * (4 + 2 * A) +
                                                subs_all * (8 + 4 * A) + 64));
    const uint64_t ct = (uint64_t)(f * (double)(2 * steps_all + verbs * (6 + 2 * A) +
                                                subs_all * (10 + 4 * A) + 64));
    cap_v[path] = (uint32_t)(cv < 0x7fffffffull ? cv : 0x7fffffffull);
    cap_i[path] = (uint32_t)(3 * ct < 0x7fffffffull ? 3 * ct : 0x7fffffffull);
}

// ---- split walks: plan, check, compaction ------------------------------------------------
// Steps of one verb chunk (64 verbs from v0): every lane decodes its verb as the walk does;
// *my_off / *my_cnt: the lane's verb's first step in the chunk and its steps.
__device__ __forceinline__ uint32_t chunk_steps(const uint8_t* verbs, uint32_t nverbs,
                                                uint32_t v0, const float* pts, uint32_t npts,
                                                uint32_t* pt_carry, float tol, uint8_t* my_verb,
                                                uint32_t* my_off, uint32_t* my_cnt) {
    const int l = lane_id();
    const uint32_t nch = min((uint32_t)WAVE, nverbs - v0);
    uint8_t verb = LYG_VERB_END;
    uint32_t np = 0;
    if ((uint32_t)l < nch) {
        verb = verbs[v0 + l];
        np = verb_points(verb);
    }
    const uint32_t np_incl = incl_scan(np);
    const uint32_t first_pt = *pt_carry + np_incl - np;
    *pt_carry += readlane_u(np_incl, WAVE - 1);
    uint32_t cnt = 0;
    if ((uint32_t)l < nch) {
        VRec r;
        decode_verb(verb, pts, npts, first_pt, tol, r);
        cnt = r.nA + r.nB;
    }
    *my_verb = (uint32_t)l < nch ? verb : (uint8_t)0xff;
    const uint32_t incl = incl_scan(cnt);
    *my_off = incl - cnt;
    *my_cnt = cnt;
    return readlane_u(incl, WAVE - 1);
}

// Staging room of `steps` steps with `verbs` verb boundaries (joins), as the bound of
// stroke_stage_caps_kernel: two vertices and triangles per step (the fast path), a full
// join with its round arcs per boundary, caps per sub-path.
__device__ __forceinline__ void item_room(uint64_t steps, uint64_t verbs, uint64_t subs,
                                          uint64_t A, float room_scale, uint32_t* rv,
                                          uint32_t* ri) {
    const double f = room_scale > 0.0f ? (double)room_scale : 1.0;
    const uint64_t cv = (uint64_t)(f * (double)(2 * steps + verbs * (4 + 2 * A) +
                                                subs * (8 + 4 * A) + 64));
    const uint64_t ct = (uint64_t)(f * (double)(2 * steps + verbs * (6 + 2 * A) +
                                                subs * (10 + 4 * A) + 64));
    *rv = (uint32_t)(cv < 0x7fffffffull ? cv : 0x7fffffffull);
    *ri = (uint32_t)(3 * ct < 0x7fffffffull ? 3 * ct : 0x7fffffffull);
}

// The cutting of one path into items.  A path of one sub-path with at least
// 2 * split_windows windows is cut into parts of split_windows windows (part j starts at
// window j * split_windows of the path; its prologue is the window before it); every
// other path is one item.  PLAN: counts the parts beyond the first and the staging room
// of all of them (*extra, *rv, *ri).  Otherwise: writes the items, item 0 at index
// `path`, parts 1.. from extra_first, staging rooms from v_base / i_base on (the rooms
// are recomputed exactly as planned).  One wave.
template <bool PLAN>
__device__ void cut_path(const StrokeLaunch& L, uint32_t path, uint32_t split_windows,
                         float room_scale, uint32_t* extra, uint32_t* rv, uint32_t* ri,
                         uint32_t n_extra_planned, uint32_t extra_first, uint32_t v_base,
                         uint32_t i_base) {
    const int l = lane_id();
    const uint32_t vb = L.in.verb_off[path], ve = L.in.verb_off[path + 1];
    const uint32_t pb = L.in.point_off[path], pe = L.in.point_off[path + 1];
    const uint8_t* verbs = L.in.verbs + vb;
    const float* pts = L.in.points + 2ull * pb;
    const uint32_t nverbs = ve - vb, npts = pe - pb;
    const lyg_stroke_options o = L.opts[L.opts_stride ? path : 0];
    const float tol = o.tolerance;
    const bool round = o.line_join == LYG_JOIN_ROUND || o.start_cap == LYG_CAP_ROUND ||
                       o.end_cap == LYG_CAP_ROUND;
    const float hw = o.line_width * 0.5f;
    uint64_t A = 0;
    if (round && tol > 0.0f && hw > 0.0f)
        A = arc_count(arc_subdivisions(2.0f * 3.14159265358979323846f, hw, tol));
    // parts of split_windows windows; with PLAN the eligibility is decided at the end
    const bool cut = PLAN ? split_windows > 0 : n_extra_planned > 0;
    uint64_t steps = 0, windows = 0, sum_v = 0, sum_i = 0;
    uint32_t n_begin = 0, n_end = 0, pt_carry = 0, open_j = 0;
    bool end_last = false;
    // the last four windows seen (chunk, first step, points before the chunk, steps), for
    // prologues: a prologue goes back over whole windows until it holds PRO_STEPS steps,
    // so the last items it carries over are all computed from points it walked
    constexpr uint32_t PRO_STEPS = 32;
    uint32_t rk[4] = {0, 0, 0, 0}, rw[4] = {0, 0, 0, 0}, rp[4] = {0, 0, 0, 0},
             rn[4] = {0, 0, 0, 0};
    uint32_t n_seen = 0;
    StrokeItem cur{};
    cur.path = path;
    cur.flags = ITEM_FIRST | ITEM_SPLIT;
    uint64_t cur_steps = 0, cur_verbs = 0;
    auto close_item = [&]() {
        // joins: the verbs that start in the part, the one in progress at its start, and
        // room for one more (a cusp inside a curve) per verb
        uint32_t pv, pi;
        item_room(cur_steps, 2 * (cur_verbs + 1), 1, A, room_scale, &pv, &pi);
        if (!PLAN && l == 0) {
            cur.v_room = pv;
            cur.i_room = pi;
            cur.v_off = v_base + (uint32_t)sum_v;
            cur.i_off = i_base + (uint32_t)sum_i;
            L.items[open_j == 0 ? path : extra_first + open_j - 1] = cur;
        }
        sum_v += pv;
        sum_i += pi;
    };
    if (PLAN || cut) {
        for (uint32_t v0 = 0; v0 < nverbs; v0 += WAVE) {
            const uint32_t k = v0 / WAVE;
            const uint32_t pt_k = pt_carry;
            uint8_t mv;
            uint32_t off, cnt;
            const uint32_t ns =
                chunk_steps(verbs, nverbs, v0, pts, npts, &pt_carry, tol, &mv, &off, &cnt);
            const bool has = mv != (uint8_t)0xff;
            steps += ns;
            n_begin += __popcll(ballot(mv == LYG_VERB_BEGIN));
            const uint64_t em = ballot(mv == LYG_VERB_END || mv == LYG_VERB_CLOSE);
            n_end += __popcll(em);
            if (em && v0 + (uint32_t)highest_bit(em) == nverbs - 1) end_last = true;
            if (!cut) continue;
            const uint32_t nw = (ns + WAVE - 1) / WAVE;
            uint32_t part_s = 0;  // first step of the open part in this chunk
            for (uint32_t w = 0; w < nw; ++w) {
                const uint64_t gw = windows + w;
                if (gw > 0 && gw % split_windows == 0) {
                    // the open part ends here; the next starts with its prologue
                    const uint32_t s0 = w * WAVE;
                    cur_verbs += __popcll(ballot(has && off >= part_s && off < s0));
                    cur.k1 = k;
                    cur.w1 = s0;
                    close_item();
                    ++open_j;
                    // the prologue: back over the windows seen until PRO_STEPS steps
                    uint32_t q = 0, got = 0;
                    const uint32_t avail = n_seen < 4 ? n_seen : 4u;
                    for (; q + 1 < avail && got + rn[q] < PRO_STEPS; ++q) got += rn[q];
                    cur = StrokeItem{};
                    cur.path = path;
                    cur.flags = ITEM_SPLIT;
                    cur.k0 = rk[q];
                    cur.w0 = rw[q];
                    cur.pt0 = rp[q];
                    cur.ke = k;
                    cur.we = s0;
                    cur_steps = 0;
                    cur_verbs = 0;
                    part_s = s0;
                }
                const uint32_t wn = min((uint32_t)WAVE, ns - w * WAVE);
                cur_steps += wn;
                // newest window first in the ring
                for (int r = 3; r > 0; --r) {
                    rk[r] = rk[r - 1];
                    rw[r] = rw[r - 1];
                    rp[r] = rp[r - 1];
                    rn[r] = rn[r - 1];
                }
                rk[0] = k;
                rw[0] = w * WAVE;
                rp[0] = pt_k;
                rn[0] = wn;
                ++n_seen;
            }
            cur_verbs += __popcll(ballot(has && off >= part_s));
            windows += nw;
        }
        if (cut) {
            cur.k1 = 0xffffffffu;
            cur.w1 = 0;
            close_item();
        }
    }
    if (PLAN) {
        const bool one_sub = nverbs > 0 && verbs[0] == LYG_VERB_BEGIN && n_begin == 1 &&
                             n_end == 1 && end_last;
        const bool split = cut && one_sub && tol > 0.0f && windows >= 2ull * split_windows;
        uint32_t wv, wi;
        item_room(steps, nverbs, n_begin, A, room_scale, &wv, &wi);
        *extra = split ? open_j : 0u;
        *rv = split ? (uint32_t)min<uint64_t>(sum_v, 0x7fffffffull) : wv;
        *ri = split ? (uint32_t)min<uint64_t>(sum_i, 0x7fffffffull) : wi;
    } else if (!cut && l == 0) {
        StrokeItem it{};
        it.path = path;
        it.flags = ITEM_FIRST;
        it.k1 = 0xffffffffu;
        it.v_room = *rv;
        it.i_room = *ri;
        it.v_off = v_base;
        it.i_off = i_base;
        L.items[path] = it;
    }
}

// One wave per path: the plan (parts beyond the first, staging rooms) into per-path
// arrays, which scans turn into item and staging offsets.
__global__ void __launch_bounds__(WAVE) stroke_plan_kernel(StrokeLaunch L, uint32_t n_paths,
                                                           uint32_t split_windows,
                                                           float room_scale, uint32_t* extra,
                                                           uint32_t* rv, uint32_t* ri) {
    for (uint32_t path = blockIdx.x; path < n_paths; path += gridDim.x) {
        uint32_t e, v, i;
        cut_path<true>(L, path, split_windows, room_scale, &e, &v, &i, 0, 0, 0, 0);
        if (lane_id() == 0) {
            extra[path] = e;
            rv[path] = v;
            ri[path] = i;
        }
    }
}

// One wave per path: the items, at the offsets the scans gave (extra_off: parts beyond
// the first of the paths before; v_off / i_off: staging), and the path's item range.
__global__ void __launch_bounds__(WAVE) stroke_cut_kernel(StrokeLaunch L, uint32_t n_paths,
                                                          uint32_t split_windows,
                                                          float room_scale,
                                                          const uint32_t* extra,
                                                          const uint32_t* extra_off,
                                                          const uint32_t* rv, const uint32_t* ri,
                                                          const uint32_t* v_off,
                                                          const uint32_t* i_off,
                                                          uint32_t cap_extra, uint32_t* err) {
    for (uint32_t path = blockIdx.x; path < n_paths; path += gridDim.x) {
        uint32_t e = extra[path];
        const uint32_t ef = n_paths + extra_off[path];
        if (e && extra_off[path] + e > cap_extra) {
            // beyond the item arrays (graph replay with more parts than captured): the
            // path is not cut; the replay reports it
            if (lane_id() == 0 && err) atomicOr(err, 4u);
            e = 0;
        }
        uint32_t v = rv[path], i = ri[path];
        cut_path<false>(L, path, split_windows, room_scale, nullptr, &v, &i, e, ef, v_off[path],
                        i_off[path]);
        if (lane_id() == 0) {
            StrokeHeads H;
            H.f0 = H.f1 = H.f2 = F2{0.0f, 0.0f};
            H.ff1 = 0;
            H.f1_base = INVALID_ID;
            H.n_items = e + 1;
            H.extra_first = e ? ef : 0u;
            H.redo = 0;
            L.heads[path] = H;
        }
    }
}

// One wave per split path: the sub-path's first points, from its first window walked
// alone (a path whose first window does not keep three points is marked to be walked
// again whole).
__global__ void __launch_bounds__(WAVE) stroke_heads_kernel(StrokeLaunch L, uint32_t n_paths) {
    __shared__ VRec recs[WAVE];
    for (uint32_t path = blockIdx.x; path < n_paths; path += gridDim.x) {
        if (L.heads[path].n_items < 2) continue;
        Carry c;
        carry_init(c);
        const WalkRange R{0, 0, 0, 0, 0, (uint32_t)WAVE, 0};
        stroke_walk<false>(L, path, R, c, 0, 0, 0, 0, nullptr, recs, nullptr, nullptr, nullptr);
        if (lane_id() == 0) {
            StrokeHeads* H = L.heads + path;
            if (c.sp_cnt >= 3 && !c.prev_end) {
                H->f0 = F2{c.f0.x, c.f0.y};
                H->f1 = F2{c.f1.x, c.f1.y};
                H->f2 = F2{c.f2.x, c.f2.y};
                H->ff1 = c.ff1 ? 1u : 0u;
                H->f1_base = c.f1_base;
            } else {
                H->redo = 2;  // the parts cannot be trusted
            }
        }
    }
}

// Whether the carry the previous part ended with (ids relative to that part, which made
// nv_prev vertices) is the carry the next part's prologue rebuilt.
__device__ __forceinline__ bool carry_match(const CarryRec& e, uint32_t nv_prev,
                                            const CarryRec& s) {
    auto same = [](F2 a, F2 b) {
        return __float_as_uint(a.x) == __float_as_uint(b.x) &&
               __float_as_uint(a.y) == __float_as_uint(b.y);
    };
    bool ok = same(e.last_kept, s.last_kept) && e.sp == s.sp && e.prev_end == s.prev_end &&
              (e.sp >= 2 || e.mnec == s.mnec);
    for (int k = 0; k < 3; ++k) {
        const uint32_t eb = e.base[k] == INVALID_ID ? INVALID_ID : e.base[k] - nv_prev;
        ok = ok && same(e.pos[k], s.pos[k]) && same(e.next[k][0], s.next[k][0]) &&
             same(e.next[k][1], s.next[k][1]) && e.flags[k] == s.flags[k] &&
             e.isp[k] == s.isp[k] && e.code[k] == s.code[k] && eb == s.base[k];
    }
    return ok;
}

// One wave per path: checks its parts' carries, places the parts in the path (vertex and
// index bases), stores the path's counts, and marks it to be emitted again whole when a
// part's prologue went wrong (its counts then come from a whole walk here) or a part
// outgrew its staging room.
__global__ void __launch_bounds__(WAVE) stroke_items_verify_kernel(StrokeLaunch L,
                                                                   uint32_t n_paths) {
    __shared__ VRec recs[WAVE];
    const int l = lane_id();
    for (uint32_t path = blockIdx.x; path < n_paths; path += gridDim.x) {
        const uint32_t n = L.heads[path].n_items;
        const uint32_t ef = L.heads[path].extra_first;
        uint32_t vrun = 0, irun = 0;
        bool mism = L.heads[path].redo != 0, over = false;
        for (uint32_t j0 = 0; j0 < n; j0 += WAVE) {
            const uint32_t j = j0 + (uint32_t)l;
            const bool valid = j < n;
            const uint32_t q = j == 0 ? path : ef + j - 1;
            uint32_t nv = 0, ni = 0;
            bool o = false, m = false;
            if (valid) {
                const StrokeItem& it = L.items[q];
                nv = it.nv;
                ni = it.ni;
                o = nv > it.v_room || ni > it.i_room;
                if (j > 0) {
                    const uint32_t qp = j == 1 ? path : ef + j - 2;
                    m = !carry_match(L.item_end[qp], L.items[qp].nv, L.item_spec[q]);
                }
            }
            const uint32_t vi = incl_scan(nv), ii = incl_scan(ni);
            if (valid) {
                L.items[q].vbase = vrun + vi - nv;
                L.items[q].ibase = irun + ii - ni;
            }
            vrun += readlane_u(vi, WAVE - 1);
            irun += readlane_u(ii, WAVE - 1);
            mism = mism || ballot(m) != 0;
            over = over || ballot(o) != 0;
        }
        if (mism) {
            Carry c;
            carry_init(c);
            const WalkRange R{0, 0, 0, 0, 0xffffffffu, 0, 0};
            stroke_walk<false, false>(L, path, R, c, 0, 0, 0, 0, nullptr, recs, nullptr, nullptr,
                                      nullptr);
            vrun = c.run_v;
            irun = 3u * c.run_i;
        }
        if (l == 0) {
            L.path_nv[path] = vrun;
            L.path_ni[path] = irun;
            // 1: a part's prologue did not rebuild the state; 3: a part outgrew its room
            L.heads[path].redo = mism ? 1u : (over ? 3u : 0u);
        }
    }
}

// One block of 256 threads per item (grid-stride): copies the item's staged vertices to
// its place in its path's final range and its indices with ids made final (item-local
// less ITEM_ID_BIAS + the item's base in the path + the path's id base; head-marked ids +
// the path's id base); a path marked for a whole walk is listed once instead.
__global__ void __launch_bounds__(256) stroke_items_compact_kernel(StrokeCompact C,
                                                                   const StrokeItem* items,
                                                                   const StrokeHeads* heads,
                                                                   uint32_t n_items) {
    if (C.n_extra_dev)
        n_items = C.n_paths + (uint32_t)min(*C.n_extra_dev, (unsigned long long)C.cap_extra);
    for (uint32_t q = blockIdx.x; q < n_items; q += gridDim.x) {
        const StrokeItem it = items[q];
        const uint32_t p = it.path;
        if (heads[p].redo) {
            if (threadIdx.x == 0 && (it.flags & ITEM_FIRST))
                C.fallback[atomicAdd(C.n_fallback, 1u)] = p;
            continue;
        }
        const uint32_t v0 = C.vtx_off[p] + it.vbase, i0 = C.idx_off[p] + it.ibase;
        if ((uint64_t)v0 + it.nv > C.out_vcap || (uint64_t)i0 + it.ni > C.out_icap) {
            if (threadIdx.x == 0 && C.err) atomicOr(C.err, 2u);
            continue;
        }
        const uint2* __restrict__ sv = reinterpret_cast<const uint2*>(C.stage_v) + it.v_off;
        uint2* __restrict__ dv = reinterpret_cast<uint2*>(C.out_vertices) + v0;
        const uint32_t nv = it.nv, ni = it.ni;
        uint32_t j = threadIdx.x;
        for (; j + 768 < nv; j += 1024) {
            const uint2 a = sv[j], b = sv[j + 256], c = sv[j + 512], d = sv[j + 768];
            dv[j] = a;
            dv[j + 256] = b;
            dv[j + 512] = c;
            dv[j + 768] = d;
        }
        for (; j < nv; j += 256) dv[j] = sv[j];
        const uint32_t pbase = C.vertex_offset + (C.local_ids ? 0u : C.vtx_off[p]);
        const uint32_t add = pbase + it.vbase - ITEM_ID_BIAS;
        const uint32_t* __restrict__ si = C.stage_i + it.i_off;
        uint32_t* __restrict__ di = C.out_indices + i0;
        for (j = threadIdx.x; j < ni; j += 256) {
            const uint32_t x = si[j];
            di[j] = (x & ID_HEAD_MASK) == ID_HEAD_MARK ? (x - ID_HEAD_MARK) + pbase : x + add;
        }
    }
}

// One block of 256 threads per path (grid-stride over paths): copies the path's staged vertices and
// indices (adding its id base) to the final offsets, or lists the path for a second emit
// when its exact counts exceed its staging room.
__global__ void __launch_bounds__(256) stroke_compact_kernel(StrokeCompact C) {
    for (uint32_t p = blockIdx.x; p < C.n_paths; p += gridDim.x) {
        const uint32_t nv = C.path_nv[p], ni = C.path_ni[p];
        const uint32_t sv0 = C.stage_voff[p], si0 = C.stage_ioff[p];
        if (nv > C.stage_voff[p + 1] - sv0 || ni > C.stage_ioff[p + 1] - si0) {
            if (threadIdx.x == 0) C.fallback[atomicAdd(C.n_fallback, 1u)] = p;
            continue;
        }
        const uint32_t v0 = C.vtx_off[p], i0 = C.idx_off[p];
        const uint2* __restrict__ sv = reinterpret_cast<const uint2*>(C.stage_v) + sv0;
        uint2* __restrict__ dv = reinterpret_cast<uint2*>(C.out_vertices) + v0;
        // four loads in flight per thread, then the four stores
        uint32_t j = threadIdx.x;
        for (; j + 768 < nv; j += 1024) {
            const uint2 a = sv[j], b = sv[j + 256], c = sv[j + 512], d = sv[j + 768];
            dv[j] = a;
            dv[j + 256] = b;
            dv[j + 512] = c;
            dv[j + 768] = d;
        }
        for (; j < nv; j += 256) dv[j] = sv[j];
        const uint32_t base = C.vertex_offset + (C.local_ids ? 0u : v0);
        const uint32_t* __restrict__ si = C.stage_i + si0;
        uint32_t* __restrict__ di = C.out_indices + i0;
        j = threadIdx.x;
        for (; j + 768 < ni; j += 1024) {
            const uint32_t a = si[j], b = si[j + 256], c = si[j + 512], d = si[j + 768];
            di[j] = a + base;
            di[j + 256] = b + base;
            di[j + 512] = c + base;
            di[j + 768] = d + base;
        }
        for (; j < ni; j += 256) di[j] = si[j] + base;
    }
}

// ---- variable-width stroke (StrokeOptions::variable_line_width) ---------------------------
// lyon's second stroker (stroke.rs:653-767, 1380-1512): every endpoint carries a width
// factor (a custom attribute), curves interpolate it, and each edge gets its own side lines
// (tangent to its two end circles), from which every join finds its sides.  Its steps read
// the side lines of the edges on both sides of a join, so the walk is serial: one lane
// strokes one path, in a count pass (sizes), then an emit pass at the scanned offsets,
// through the same vertex/triangle sink, joins' arcs and caps' arithmetic as the
// fixed-width kernel.
namespace vw {

static constexpr uint32_t MAX_CURVE_POINTS = 1u << 22;  // per curve; more: LYG_ERR_NO_SPACE

struct Side {
    vec2 prev, next, single;
    uint32_t prev_v, next_v;
    bool has_single;
};
// EndpointData (stroke.rs:320-352) without the vertex source
struct Pt {
    vec2 pos;
    float hw, adv;
    int join;
    Side sp[2];
    bool fold[2];
    bool flat;
    // VertexSource: endpoint (point index sf) or the point at st on the edge sf -> sto
    uint32_t sf, sto;
    float st;
    bool sedge;
};

__device__ __forceinline__ void pt_init(Pt& e, vec2 pos, float hw, int join, bool flat) {
    const float nan = __builtin_nanf("");
    e.sf = e.sto = 0;
    e.st = 0.0f;
    e.sedge = false;
    e.pos = pos;
    e.hw = hw;
    e.adv = nan;
    e.join = join;
    e.flat = flat;
    for (int k = 0; k < 2; ++k) {
        e.sp[k].prev = e.sp[k].next = mk(nan, nan);
        e.sp[k].single = mk(0.0f, 0.0f);
        e.sp[k].prev_v = e.sp[k].next_v = INVALID_ID;
        e.sp[k].has_single = false;
        e.fold[k] = false;
    }
}

struct Stroker {
    float lw, tol, miter_limit, merge2;
    uint32_t start_cap, end_cap, line_join;
    Sink* s;
    vec2 on_path;  // StrokeVertexData: position_on_path, half_width
    float vhw;
    Pt buf[3];     // PointBuffer (stroke.rs:2371-2457), oldest first: buf[count - 1] is last
    uint32_t count;
    Pt firsts[2];
    uint32_t n_firsts;
    bool may_need_empty_cap;
    float sub_path_start_adv;
    uint32_t status;
    // interpolated attributes (emit pass, when asked for): the current vertex source and
    // the first vertex id not yet given its row
    const float* attrs;
    uint32_t n_attrs;
    float* attr_out;  // row of the path's first vertex
    uint32_t sf, sto, filled;
    float st;
    bool sedge;
};

// StrokeVertex::interpolated_attributes (stroke.rs:2728-2746) for the vertices emitted
// since the last source change, then the new source (where lyon sets vertex.src).
__device__ __forceinline__ void flush_attrs(Stroker& S) {
    if (!S.attr_out) return;
    const Sink& k = *S.s;
    for (uint32_t v = S.filled; v < k.vid; ++v) {
        const uint32_t row = v - k.vid0;
        if (row >= k.vlim) break;
        const float* a = S.attrs + (size_t)S.sf * S.n_attrs;
        const float* b = S.attrs + (size_t)S.sto * S.n_attrs;
        float* o = S.attr_out + (size_t)row * S.n_attrs;
        for (uint32_t q = 0; q < S.n_attrs; ++q)
            o[q] = S.sedge ? a[q] * (1.0f - S.st) + b[q] * S.st : a[q];
    }
    S.filled = k.vid;
}
__device__ __forceinline__ void set_src(Stroker& S, const Pt& p) {
    flush_attrs(S);
    S.sf = p.sf;
    S.sto = p.sto;
    S.st = p.st;
    S.sedge = p.sedge;
}

// Slots are addressed with constant indices only (the step is instantiated per count), so
// the points stay in registers.
template <uint32_t C>
__device__ __forceinline__ void push_n(Stroker& S, const Pt& p) {
    if (C < 3) {
        S.buf[C < 3 ? C : 2] = p;
        S.count = C + 1;
    } else {
        S.buf[0] = S.buf[1];
        S.buf[1] = S.buf[2];
        S.buf[2] = p;
    }
}
__device__ __forceinline__ uint32_t vtx(Stroker& S, vec2 normal) {
    return sink_sv(*S.s, S.on_path, normal, S.vhw);
}

// stroke.rs:1869-1904
__device__ __forceinline__ void edge_tris(Stroker& S, const Pt& p0, const Pt& p1) {
    uint32_t p0n = p0.sp[SN].next_v, p0p = p0.sp[SP].next_v;
    uint32_t p1n = p1.sp[SN].prev_v, p1p = p1.sp[SP].prev_v;
    if (p0.fold[SP]) p0n = p0.sp[SP].prev_v;
    if (p0.fold[SN]) p0p = p0.sp[SN].prev_v;
    if (p1.fold[SP]) p1n = p1.sp[SP].next_v;
    if (p1.fold[SN]) p1p = p1.sp[SN].next_v;
    if (p0n == p1p) return;
    if (p0n != p0p && p0p != p1p) sink_tri(*S.s, p0n, p0p, p1p);
    if (p0n != p1n && p1p != p1n) sink_tri(*S.s, p0n, p1p, p1n);
}

// stroke.rs:1968-2021
__device__ __forceinline__ void round_join(Stroker& S, Pt& j, int side) {
    const float PI = 3.14159265358979323846f;
    vec2 sn = sub(j.sp[side].prev, j.pos), en = sub(j.sp[side].next, j.pos);
    uint32_t sv = j.sp[side].prev_v, ev = j.sp[side].next_v;
    const float sign = side == SN ? 1.0f : -1.0f;
    float a0 = angle_of(sn);
    float diff = angle_to(a0, angle_of(en));
    if (diff * sign < 0.0f) diff = sign * (2.0f * PI - ffabs(diff));
    float a1 = a0 + diff;
    if (side == SN) {
        float ta = a0;
        a0 = a1;
        a1 = ta;
        uint32_t tv = sv;
        sv = ev;
        ev = tv;
    }
    tessellate_arc(a0, a1, sv, ev, arc_subdivisions(diff, j.hw, S.tol), S.on_path, S.vhw, *S.s);
}

// stroke.rs:1907-1965
__device__ __forceinline__ void join_tris(Stroker& S, Pt& j) {
    const bool need0 = !j.sp[SP].has_single && !j.fold[SN];
    const bool need1 = !j.sp[SN].has_single && !j.fold[SP];
    const Side &P = j.sp[SP], &N = j.sp[SN];
    if (!j.fold[SP] && !j.fold[SN]) {
        if (need0 && need1) {
            sink_tri(*S.s, P.prev_v, P.next_v, N.next_v);
            sink_tri(*S.s, P.prev_v, N.next_v, N.prev_v);
        } else if (!need0 && need1) {
            sink_tri(*S.s, N.prev_v, P.prev_v, N.next_v);
        } else if (need0 && !need1) {
            sink_tri(*S.s, N.prev_v, P.prev_v, P.next_v);
        }
    }
    if (j.join == LYG_JOIN_ROUND) {
        if (need0) round_join(S, j, SP);
        if (need1) round_join(S, j, SN);
    }
}

// stroke.rs:2024-2055
__device__ __forceinline__ void base_vertices(Stroker& S, Pt& j, int side) {
    Side& sp = j.sp[side];
    if (sp.has_single) {
        uint32_t id = vtx(S, divs(sub(sp.single, j.pos), j.hw));
        sp.prev_v = sp.next_v = id;
    } else {
        uint32_t a = vtx(S, divs(sub(sp.prev, j.pos), j.hw));
        uint32_t b = vtx(S, divs(sub(sp.next, j.pos), j.hw));
        sp.prev_v = a;
        sp.next_v = b;
    }
}

// stroke.rs:1743-1806; true: the join folds back and is skipped
__device__ __forceinline__ bool flattened_step(Stroker& S, Pt& prev, Pt& j, Pt& next) {
    vec2 pe = sub(j.pos, prev.pos);
    float pl = vlen(pe);
    vec2 pt = divs(pe, pl);
    vec2 ne = sub(next.pos, j.pos);
    float nl = vlen(ne);
    vec2 nt = divs(ne, nl);
    vec2 n = compute_normal(pt, nt);
    if (j.adv != j.adv) j.adv = prev.adv + pl;
    if (next.adv != next.adv) next.adv = j.adv + nl;
    vec2 p0 = add(j.pos, mul(n, S.vhw));
    vec2 p1 = sub(j.pos, mul(n, S.vhw));
    j.sp[SP].prev = j.sp[SP].next = j.sp[SP].single = p0;
    j.sp[SP].has_single = true;
    j.sp[SN].prev = j.sp[SN].next = j.sp[SN].single = p1;
    j.sp[SN].has_single = true;
    vec2 v0 = sub(p0, prev.sp[SP].next), v1 = sub(p1, prev.sp[SN].next);
    if (dot(pe, v0) < 0.0f && dot(pe, v1) < 0.0f) return true;
    uint32_t a = vtx(S, n);
    uint32_t b = vtx(S, neg(n));
    j.sp[SP].prev_v = j.sp[SP].next_v = a;
    j.sp[SN].prev_v = j.sp[SN].next_v = b;
    return false;
}

// stroke.rs:1809-1866: side lines of the edge p0 -> p1, tangent to both end circles
__device__ __forceinline__ void attachments(Pt& p0, Pt& p1) {
    const float PI = 3.14159265358979323846f;
    vec2 e = sub(p1.pos, p0.pos);
    float d = vlen(e);
    float ea = angle_of(e);
    float va = (float)asin((double)((p1.hw - p0.hw) / d));
    if (va != va) va = 0.0f;
    for (int side = 0; side < 2; ++side) {
        float nl = side == SN ? -1.0f : 1.0f;
        float na = ea + nl * (PI * 0.5f + va);
        vec2 n = mk(t_cos(na), t_sin(na));
        p0.sp[side].next = add(p0.pos, mul(n, p0.hw));
        p1.sp[side].prev = add(p1.pos, mul(n, p1.hw));
    }
    if (p1.adv != p1.adv) p1.adv = p0.adv + d;
}

// stroke.rs:2060-2128
__device__ __forceinline__ void join_side(const Pt& prev, Pt& j, const Pt& next, float ml,
                                          int side) {
    const float sign = side == SN ? -1.0f : 1.0f;
    vec2 v0 = normalize(sub(j.sp[side].prev, prev.sp[side].next));
    vec2 v1 = normalize(sub(next.sp[side].prev, j.sp[side].next));
    bool inward = cross(v0, v1) * sign > 0.0f;
    bool forward = dot(v0, v1) > 0.0f;
    vec2 n = mul(compute_normal(v0, v1), sign);
    vec2 pv0 = normalize(sub(j.pos, prev.pos));
    vec2 pv1 = normalize(sub(next.pos, j.pos));
    bool same_side = dot(add(v0, v1), add(pv0, pv1)) >= 0.0f;
    if (inward && !forward && same_side) {
        vec2 en = mul(n, j.hw);
        float pl = j.adv - prev.adv;
        float nl = next.adv - j.adv;
        float dn = dot(en, v1) - nl;
        float dp = dot(en, neg(v0)) - pl;
        if (ffmin(dn, dp) > 0.0f || sqlen(n) < 1e-5f) j.fold[side] = true;
    }
    bool concave = inward && same_side && !j.fold[side];
    bool miter = j.join == LYG_JOIN_MITER || j.join == LYG_JOIN_MITER_CLIP;
    if (concave || (miter && !(sqlen(n) > ml * ml * 4.0f))) {
        j.sp[side].has_single = true;
        j.sp[side].single = add(j.pos, mul(n, j.hw));
    } else if (j.join == LYG_JOIN_MITER_CLIP) {
        vec2 n0 = sub(j.sp[side].prev, j.pos), n1 = sub(j.sp[side].next, j.pos);
        vec2 i0, i1;
        clip_intersections(n0, n1, n, ml * j.hw, &i0, &i1);
        j.sp[side].prev = add(j.pos, i0);
        j.sp[side].next = add(j.pos, i1);
    }
}

// step_impl (stroke.rs:1380-1512) with C points buffered; returns whether the point was
// added
template <uint32_t C>
__device__ __forceinline__ bool step_n(Stroker& S, Pt next) {
    if (C > 0) {
        Pt& lp = S.buf[C > 0 ? C - 1 : 0];
        vec2 d = sub(lp.pos, next.pos);
        if (sqlen(d) < S.merge2) {
            if (C == 1) S.may_need_empty_cap = true;
            return false;
        }
        if (!lp.flat || !next.flat) attachments(lp, next);
    }
    bool skip = false;
    if (C > 1) {
        Pt& prev = S.buf[C > 1 ? C - 2 : 0];
        Pt& j = S.buf[C > 1 ? C - 1 : 0];
        set_src(S, j);
        S.on_path = j.pos;
        S.vhw = j.hw;
        bool fast = false;
        if (j.flat) fast = dot(sub(j.pos, prev.pos), sub(next.pos, j.pos)) > 0.0f;
        if (fast) {
            j.join = LYG_JOIN_MITER;
            skip = flattened_step(S, prev, j, next);
        } else {
            join_side(prev, j, next, S.miter_limit, SP);
            join_side(prev, j, next, S.miter_limit, SN);
            if (j.sp[SP].has_single) j.fold[SN] = false;
            if (j.sp[SN].has_single) j.fold[SP] = false;
            base_vertices(S, j, SN);
            base_vertices(S, j, SP);
        }
        if (!skip) {
            if (C > 2) edge_tris(S, prev, j);
            join_tris(S, j);
            if (C == 2) {
                S.firsts[0] = prev;
                S.firsts[1] = j;
                S.n_firsts = 2;
            }
        }
    }
    if (skip) S.buf[C > 0 ? C - 1 : 0] = next;  // replace_last
    else push_n<C>(S, next);
    return true;
}
__device__ __forceinline__ bool step(Stroker& S, const Pt& next) {
    switch (S.count) {
        case 0: return step_n<0>(S, next);
        case 1: return step_n<1>(S, next);
        case 2: return step_n<2>(S, next);
        default: return step_n<3>(S, next);
    }
}

// stroke.rs:2459-2524
__device__ __forceinline__ void round_cap(Stroker& S, vec2 center, float radius, vec2 sn,
                                       uint32_t sv, uint32_t ev, vec2 edge_n, bool is_start) {
    if (radius < S.tol) return;
    float a0 = angle_of(sn);
    float diff = angle_to(a0, angle_of(edge_n));
    float mid = a0 + diff;
    float a1 = mid + diff;
    uint32_t nsub = arc_subdivisions(diff, radius, S.tol);
    S.on_path = center;
    S.vhw = radius;
    uint32_t mv = vtx(S, normalize(edge_n));
    sink_tri(*S.s, sv, mv, ev);
    tessellate_arc(a0, mid, sv, mv, nsub, center, radius, *S.s);
    tessellate_arc(mid, a1, mv, ev, nsub, center, radius, *S.s);
}

// stroke.rs:1282-1316, 2526-2609
__device__ __forceinline__ void empty_cap(Stroker& S) {
    Pt& p = S.buf[0];
    set_src(S, p);
    S.vhw = p.hw;
    S.on_path = p.pos;
    if (S.start_cap == LYG_CAP_SQUARE) {
        uint32_t a = vtx(S, mk(1.0f, 1.0f));
        uint32_t b = vtx(S, mk(1.0f, -1.0f));
        uint32_t c = vtx(S, mk(-1.0f, -1.0f));
        uint32_t d = vtx(S, mk(-1.0f, 1.0f));
        sink_tri(*S.s, a, b, c);
        sink_tri(*S.s, a, c, d);
    } else if (S.start_cap == LYG_CAP_ROUND) {
        const vec2 center = p.pos;
        const float r = S.vhw;
        uint32_t l = vtx(S, mk(-1.0f, 0.0f));
        uint32_t rr = vtx(S, mk(1.0f, 0.0f));
        round_cap(S, center, r, mk(-1.0f, 0.0f), l, rr, mk(0.0f, 1.0f), true);
        round_cap(S, center, r, mk(1.0f, 0.0f), rr, l, mk(0.0f, -1.0f), false);
    }
}

// stroke.rs:2130-2206
__device__ __forceinline__ void last_edge(Stroker& S, const Pt& p0, Pt& p1, bool is_first) {
    vec2 v = sub(p1.pos, p0.pos);
    p1.adv = p0.adv + vlen(v);
    set_src(S, p1);
    S.on_path = p1.pos;
    S.vhw = p1.hw;
    for (int side = 0; side < 2; ++side) {
        vec2 sp = p1.sp[side].prev;
        if (S.end_cap != LYG_CAP_ROUND) {
            float clip = S.end_cap == LYG_CAP_SQUARE ? p1.hw : 0.0f;
            vec2 n = normalize(sub(p1.pos, p0.pos));
            vec2 inter;
            if (!line_isect(add(p1.pos, mul(n, clip)), perp(n), sp, sub(sp, p0.sp[side].next),
                            &inter))
                inter = p1.sp[side].prev;
            p1.sp[side].prev = inter;
        }
        p1.sp[side].prev_v = vtx(S, divs(sub(p1.sp[side].prev, p1.pos), p1.hw));
    }
    if (!is_first) edge_tris(S, p0, p1);
    if (S.end_cap == LYG_CAP_ROUND)
        round_cap(S, p1.pos, p1.hw, sub(p1.sp[SP].prev, p1.pos), p1.sp[SP].prev_v,
                  p1.sp[SN].prev_v, v, false);
}

// stroke.rs:2208-2276
__device__ __forceinline__ void first_edge(Stroker& S, Pt& f, const Pt& second) {
    set_src(S, f);
    S.on_path = f.pos;
    S.vhw = f.hw;
    for (int side = 0; side < 2; ++side) {
        vec2 sp = f.sp[side].next;
        if (S.start_cap != LYG_CAP_ROUND) {
            float clip = S.start_cap == LYG_CAP_SQUARE ? f.hw : 0.0f;
            vec2 n = normalize(sub(f.pos, second.pos));
            vec2 inter;
            if (!line_isect(add(f.pos, mul(n, clip)), perp(n), sp,
                            sub(sp, second.sp[side].prev), &inter))
                inter = f.sp[side].next;
            sp = inter;
        }
        f.sp[side].next_v = vtx(S, divs(sub(sp, f.pos), f.hw));
    }
    edge_tris(S, f, second);
    if (S.start_cap == LYG_CAP_ROUND)
        round_cap(S, f.pos, f.hw, sub(f.sp[SN].next, f.pos), f.sp[SN].next_v, f.sp[SP].next_v,
                  sub(f.pos, second.pos), true);
}

// stroke.rs:1322-1378
__device__ __forceinline__ void end_with_caps(Stroker& S) {
    if (S.may_need_empty_cap && S.count == 1) {
        empty_cap(S);
        return;
    }
    if (S.count >= 2) {
        Pt p0 = S.count == 2 ? S.buf[0] : S.buf[1];
        Pt p1 = S.count == 2 ? S.buf[1] : S.buf[2];
        last_edge(S, p0, p1, S.count == 2);
        S.sub_path_start_adv = p1.adv;
        if (S.count > 2) {
            p0 = S.firsts[0];
            p1 = S.firsts[1];
        }
        first_edge(S, p0, p1);
    }
}

// stroke.rs:1206-1279
__device__ __forceinline__ void close_sub(Stroker& S) {
    if (S.count == 1) empty_cap(S);
    if (S.count <= 2) return;
    Pt p = S.firsts[0];
    p.adv = __builtin_nanf("");
    if (!step(S, p)) S.buf[2].pos = p.pos;
    if (S.n_firsts >= 2) {
        step(S, S.firsts[1]);
        // here count is 3: both steps kept or replaced a point after the two firsts
        Pt& p0 = S.buf[1];
        Pt& p1 = S.buf[2];
        set_src(S, p0);
        S.on_path = p0.pos;
        S.vhw = p0.hw;
        for (int side = 0; side < 2; ++side) {
            vec2 sp = p0.sp[side].has_single ? p0.sp[side].single : p0.sp[side].next;
            p0.sp[side].next_v = vtx(S, divs(sub(sp, p0.pos), p0.hw));
        }
        edge_tris(S, p0, p1);
    }
}

// stroke.rs:1179-1193
__device__ __forceinline__ void end_sub(Stroker& S, bool close) {
    S.may_need_empty_cap = S.may_need_empty_cap || (close && S.count == 1);
    if (close && S.count > 2) close_sub(S);
    else end_with_caps(S);
    S.count = 0;
    S.n_firsts = 0;
}

// One flattened curve piece (half `b` of a split quadratic, or the whole curve): its
// points with the half width interpolated at their t on the whole curve
// (t_whole = t0 + t * t_scale), as quadratic_bezier_to / cubic_bezier_to do.
__device__ __forceinline__ void curve_piece(Stroker& S, const VRec& r, bool b, bool cubic,
                                            float t0, float t_scale, float w0, float w1,
                                            uint32_t id0, uint32_t id1) {
    const uint32_t n = b ? r.nB : r.nA;
    const float stp = b ? r.sB : r.sA;
    if (n > MAX_CURVE_POINTS) {
        S.status = LYG_ERR_NO_SPACE;
        return;
    }
    float t = 0.0f;
    for (uint32_t k = 0; k < n; ++k) {
        vec2 pos;
        float t1;
        if (k + 1 == n) {
            pos = b ? r.toB : r.toA;
            t1 = 1.0f;
        } else {
            t = t + stp;
            t1 = t;
            pos = horner(r, b, cubic && !b, t);
        }
        const float tw = t_scale == 1.0f && t0 == 0.0f ? t1 : t0 + t1 * t_scale;
        Pt e;
        pt_init(e, pos, (w0 * (1.0f - tw) + w1 * tw) * 0.5f, (int)S.line_join, t1 != 1.0f);
        // the endpoint at t == 1, else the edge (stroke.rs:1005-1013, 1041-1050)
        e.sedge = tw != 1.0f;
        e.sf = e.sedge ? id0 : id1;
        e.sto = id1;
        e.st = tw;
        step(S, e);
    }
}

}  // namespace vw


// One lane per path, the first L.paths_per_wave lanes of each one-wave block taking paths
// (grid-stride): few paths per wave spread the serial walks over every SIMD and keep the
// lanes of a wave from waiting on each other's branches.  Counting (EMIT false) or writing
// at the scanned offsets.
template <bool EMIT>
__global__ void __launch_bounds__(64) stroke_vw_kernel(StrokeVWLaunch L) {
    using namespace vw;
    const uint32_t ppw = L.paths_per_wave;
    if (threadIdx.x >= ppw) return;
    for (uint32_t path = blockIdx.x * ppw + threadIdx.x; path < L.in.n_paths;
         path += gridDim.x * ppw) {
        const lyg_stroke_options o = L.opts[L.opts_stride ? path : 0];
        const uint32_t vb = L.in.verb_off[path], ve = L.in.verb_off[path + 1];
        const uint32_t pb = L.in.point_off[path], pe = L.in.point_off[path + 1];
        const float* pts = L.in.points + 2ull * pb;
        const uint32_t npts = pe - pb;
        Sink sk;
        sk.nv = sk.ni = 0;
        sk.emit = EMIT;
        sk.err = L.err;
        if (EMIT) {
            sk.vid = sk.vid0 = L.vertex_offset + L.vtx_off[path];
            sk.vout = L.out_vertices + 2ull * L.vtx_off[path];
            sk.iout = L.out_indices + L.idx_off[path];
            sk.vlim = L.vtx_off[path + 1] - L.vtx_off[path];
            sk.ileft = L.idx_off[path + 1] - L.idx_off[path];
            if (L.path_status[path] != 0) continue;  // the count pass failed it: no output
        } else {
            sk.vid = sk.vid0 = 0;
            sk.vout = nullptr;
            sk.iout = nullptr;
            sk.vlim = sk.ileft = 0;
        }
        Stroker S;
        S.lw = o.line_width;
        S.tol = o.tolerance;
        S.miter_limit = o.miter_limit;
        S.merge2 = ffmax(ffmin(o.tolerance * o.tolerance * 0.5f,
                               o.line_width * o.line_width * 0.05f),
                         1e-8f);  // StrokeBuilderImpl::new (stroke.rs:591-631)
        S.start_cap = o.start_cap;
        S.end_cap = o.end_cap;
        S.line_join = o.line_join;
        S.s = &sk;
        S.on_path = mk(0.0f, 0.0f);
        S.vhw = o.line_width * 0.5f;
        S.count = 0;
        S.n_firsts = 0;
        S.may_need_empty_cap = false;
        S.sub_path_start_adv = 0.0f;
        S.status = 0;
        S.attrs = L.attrs;
        S.n_attrs = L.n_attrs;
        S.attr_out = EMIT && L.out_attrs ? L.out_attrs + (size_t)L.vtx_off[path] * L.n_attrs
                                         : nullptr;
        S.sf = S.sto = 0;
        S.st = 0.0f;
        S.sedge = false;
        S.filled = sk.vid;
        uint32_t cur_id = 0;  // point index (in the batch) of the current endpoint
        auto width = [&](uint32_t i) {
            return i < npts ? L.widths[(size_t)(pb + i) * L.widths_stride]
                            : __builtin_nanf("");
        };
        float cur_w = __builtin_nanf("");
        uint32_t p = 0;
        for (uint32_t v = vb; v < ve && S.status == 0; ++v) {
            const uint8_t verb = L.in.verbs[v];
            VRec r;
            decode_verb(verb, pts, npts, p, o.tolerance, r);
            if (verb == LYG_VERB_BEGIN) {
                cur_w = width(p);
                S.may_need_empty_cap = false;
                Pt e;
                pt_init(e, r.toA, o.line_width * cur_w * 0.5f, (int)o.line_join, false);
                e.adv = S.sub_path_start_adv;
                cur_id = pb + p;
                e.sf = e.sto = cur_id;
                step(S, e);
                p += 1;
            } else if (verb == LYG_VERB_LINE) {
                cur_w = width(p);
                Pt e;
                pt_init(e, r.toA, o.line_width * cur_w * 0.5f, (int)o.line_join, false);
                cur_id = pb + p;
                e.sf = e.sto = cur_id;
                step(S, e);
                p += 1;
            } else if (verb == LYG_VERB_QUADRATIC) {
                const float w0 = o.line_width * cur_w;
                cur_w = width(p + 1);
                const float w1 = o.line_width * cur_w;
                vec2 from = p > 0 ? ldp(pts, npts, p - 1) : mk(__builtin_nanf(""), __builtin_nanf(""));
                const float ts = find_sharp_turn(from, ldp(pts, npts, p), ldp(pts, npts, p + 1));
                const uint32_t id0 = cur_id;
                cur_id = pb + p + 1;
                if (ts > 0.0f) {  // flatten_quad's two halves (stroke.rs:2755-2779)
                    curve_piece(S, r, false, false, 0.0f, ts, w0, w1, id0, cur_id);
                    if (S.status == 0)
                        curve_piece(S, r, true, false, ts, 1.0f - ts, w0, w1, id0, cur_id);
                } else {
                    curve_piece(S, r, false, false, 0.0f, 1.0f, w0, w1, id0, cur_id);
                }
                p += 2;
            } else if (verb == LYG_VERB_CUBIC) {
                const float w0 = o.line_width * cur_w;
                cur_w = width(p + 2);
                const uint32_t id0 = cur_id;
                cur_id = pb + p + 2;
                curve_piece(S, r, false, true, 0.0f, 1.0f, w0, o.line_width * cur_w, id0,
                            cur_id);
                p += 3;
            } else {
                const bool close = verb == LYG_VERB_CLOSE;
                end_sub(S, close);
                if (close) p += 1;
            }
        }
        flush_attrs(S);
        if (!EMIT) {
            const bool ok = S.status == 0;
            L.path_status[path] = S.status;
            L.path_nv[path] = ok ? sk.nv : 0u;
            L.path_ni[path] = ok ? 3u * sk.ni : 0u;
        }
    }
}

}  // namespace lyg

extern "C" {

// Launch helpers used by the host runtime (lyg_api.cpp).
hipError_t lyg_launch_stroke_stage_caps(const lyg::StrokeLaunch* L, uint32_t n, float room_scale,
                                        uint32_t* cap_v, uint32_t* cap_i, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(lyg::stroke_stage_caps_kernel, dim3(n), dim3(lyg::WAVE), 0, s, *L, n,
                       room_scale, cap_v, cap_i);
    return hipGetLastError();
}
hipError_t lyg_launch_nsub_thresholds(float* thr, hipStream_t s) {
    hipLaunchKernelGGL(lyg::nsub_thresholds_kernel, dim3(1), dim3(64), 0, s, thr);
    return hipGetLastError();
}
hipError_t lyg_launch_stroke_compact(const lyg::StrokeCompact* C, int grid, hipStream_t s) {
    if (C->n_paths == 0) return hipSuccess;
    hipLaunchKernelGGL(lyg::stroke_compact_kernel, dim3(grid), dim3(256), 0, s, *C);
    return hipGetLastError();
}
hipError_t lyg_launch_stroke_plan(const lyg::StrokeLaunch* L, uint32_t n_paths,
                                  uint32_t split_windows, float room_scale, uint32_t* extra,
                                  uint32_t* rv, uint32_t* ri, int grid, hipStream_t s) {
    if (n_paths == 0) return hipSuccess;
    hipLaunchKernelGGL(lyg::stroke_plan_kernel, dim3(grid), dim3(lyg::WAVE), 0, s, *L, n_paths,
                       split_windows, room_scale, extra, rv, ri);
    return hipGetLastError();
}
hipError_t lyg_launch_stroke_cut(const lyg::StrokeLaunch* L, uint32_t n_paths,
                                 uint32_t split_windows, float room_scale, const uint32_t* extra,
                                 const uint32_t* extra_off, const uint32_t* rv,
                                 const uint32_t* ri, const uint32_t* v_off, const uint32_t* i_off,
                                 uint32_t cap_extra, uint32_t* err, int grid, hipStream_t s) {
    if (n_paths == 0) return hipSuccess;
    hipLaunchKernelGGL(lyg::stroke_cut_kernel, dim3(grid), dim3(lyg::WAVE), 0, s, *L, n_paths,
                       split_windows, room_scale, extra, extra_off, rv, ri, v_off, i_off,
                       cap_extra, err);
    return hipGetLastError();
}
hipError_t lyg_launch_stroke_heads(const lyg::StrokeLaunch* L, uint32_t n_paths, int grid,
                                   hipStream_t s) {
    if (n_paths == 0) return hipSuccess;
    hipLaunchKernelGGL(lyg::stroke_heads_kernel, dim3(grid), dim3(lyg::WAVE), 0, s, *L, n_paths);
    return hipGetLastError();
}
hipError_t lyg_launch_stroke_items_verify(const lyg::StrokeLaunch* L, uint32_t n_paths, int grid,
                                          hipStream_t s) {
    if (n_paths == 0) return hipSuccess;
    hipLaunchKernelGGL(lyg::stroke_items_verify_kernel, dim3(grid), dim3(lyg::WAVE), 0, s, *L,
                       n_paths);
    return hipGetLastError();
}
hipError_t lyg_launch_stroke_items_compact(const lyg::StrokeCompact* C,
                                           const lyg::StrokeItem* items,
                                           const lyg::StrokeHeads* heads, uint32_t n_items,
                                           int grid, hipStream_t s) {
    if (n_items == 0) return hipSuccess;
    hipLaunchKernelGGL(lyg::stroke_items_compact_kernel, dim3(grid), dim3(256), 0, s, *C, items,
                       heads, n_items);
    return hipGetLastError();
}
hipError_t lyg_launch_stroke(const lyg::StrokeLaunch* L, int emit, int grid, hipStream_t stream) {
    dim3 block(lyg::WAVE * lyg::WAVES_PER_BLOCK);
    if (L->items) {
        if (!emit) return hipErrorInvalidValue;  // items are emitted only
        hipLaunchKernelGGL((lyg::stroke_kernel<true, true>), dim3(grid), block, 0, stream, *L);
    } else if (emit) {
        hipLaunchKernelGGL((lyg::stroke_kernel<true, false>), dim3(grid), block, 0, stream, *L);
    } else {
        hipLaunchKernelGGL((lyg::stroke_kernel<false, false>), dim3(grid), block, 0, stream, *L);
    }
    return hipGetLastError();
}

hipError_t lyg_launch_stroke_vw(const lyg::StrokeVWLaunch* L, int emit, int grid,
                                hipStream_t s) {
    if (L->in.n_paths == 0) return hipSuccess;
    if (emit)
        hipLaunchKernelGGL(lyg::stroke_vw_kernel<true>, dim3(grid), dim3(64), 0, s, *L);
    else
        hipLaunchKernelGGL(lyg::stroke_vw_kernel<false>, dim3(grid), dim3(64), 0, s, *L);
    return hipGetLastError();
}

}  // extern "C"
