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

// Walks one path with the whole wave.  Returns (vertices, triangles) in the count pass.
template <bool EMIT>
__device__ void stroke_one_path(const StrokeLaunch& L, uint32_t path, VRec* recs,
                                ArcJob* arc_jobs, uint32_t* arc_first, uint32_t* arc_c0,
                                uint32_t* out_nv, uint32_t* out_ni) {
    const int l = lane_id();
    const uint32_t gpath = path;
    const uint32_t vb = L.in.verb_off[gpath], ve = L.in.verb_off[gpath + 1];
    const uint32_t pb = L.in.point_off[gpath], pe = L.in.point_off[gpath + 1];
    const uint8_t* verbs = L.in.verbs + vb;
    const float* pts = L.in.points + 2ull * pb;
    const uint32_t nverbs = ve - vb, npts = pe - pb;
    const lyg_stroke_options o = L.opts[L.opts_stride ? gpath : 0];
    const SParams P = make_params(o, L.nsub_thr);

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

    uint32_t pt_carry = 0;
    for (uint32_t v0 = 0; v0 < nverbs; v0 += WAVE) {
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
        for (uint32_t w0 = 0; w0 < nsteps; w0 += WAVE) {
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

            if (EMIT) {
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
    *out_nv = c.run_v;
    *out_ni = c.run_i;
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
template <bool EMIT>
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
    while (pid < L.in.n_paths) {
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
hipError_t lyg_launch_stroke(const lyg::StrokeLaunch* L, int emit, int grid, hipStream_t stream) {
    dim3 block(lyg::WAVE * lyg::WAVES_PER_BLOCK);
    if (emit)
        hipLaunchKernelGGL(lyg::stroke_kernel<true>, dim3(grid), block, 0, stream, *L);
    else
        hipLaunchKernelGGL(lyg::stroke_kernel<false>, dim3(grid), block, 0, stream, *L);
    return hipGetLastError();
}

}  // extern "C"
