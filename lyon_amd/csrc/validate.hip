// lyon_algorithms' path validators for gfx950, over the batched path layout:
//
//   * winding numbers / hit tests of query points (path_winding_number_at_position and
//     hit_test_path, crates/algorithms/src/hit_test.rs:8-117): one wave per query point.
//     The lanes take the path's verbs 64 at a time; a DPP scan of the points each verb
//     consumes gives every lane its verb's points, a running maximum gives the latest
//     Begin (the sub-path's first point, which End and Close connect back to).  Each lane
//     walks the flattened points of its own segment and counts the crossings left of the
//     query; the count is an integer, so the wave sum is exact in any order.
//   * approximate signed areas (approximate_signed_area, crates/algorithms/src/area.rs:7-54):
//     lyon's serial f32 sum of cross products in event order; one wave per path computes the
//     terms 64 at a time and adds them in order.
//
// Paths are as lyon's builders make them: every sub-path opens with Begin, so the current
// point of a verb is the last point consumed before it (path.rs:843-909).  Curves are
// flattened as for_each_flattened / Flattened do (quadratic_bezier.rs:337-359, 975-1031;
// cubic_bezier.rs:563-582, 1405-1454): t = step, step + step, ..., the last point `to`.
#include <hip/hip_runtime.h>

#include "lyg_math.h"
#include "lyg_types.h"
#include "wave.h"

namespace lyg {

namespace {

constexpr int VAL_WAVES = 4;

// Points each verb consumes (path.rs:29-36): Line 1, Quadratic 2, Cubic 3, Begin 1,
// Close 1 (the stored first point), End 0.
__device__ __forceinline__ uint32_t points_of(uint32_t verb) {
    return verb < 5u ? (verb == 2u ? 3u : verb == 1u ? 2u : 1u) : 0u;
}

// One path's points; reading past them gives NaN, as PointIter::next does.
struct PathPts {
    const float* p;
    uint32_t n;
    __device__ __forceinline__ vec2 at(uint32_t i) const {
        if (i < n) return mk(p[2ull * i], p[2ull * i + 1]);
        return mk(__builtin_nanf(""), __builtin_nanf(""));
    }
};

// hit_test.rs:85-117 test_segment; LineSegment::sample is euclid's lerp.
__device__ __forceinline__ int seg_winding(vec2 P, vec2 a, vec2 b) {
    const float min_y = ffmin(a.y, b.y), max_y = ffmax(a.y, b.y);
    if (min_y > P.y || max_y <= P.y || ffmin(a.x, b.x) > P.x) return 0;
    if (a.y == b.y) return 0;
    const float d = b.y - a.y;
    const float t = (P.y - a.y) / d;
    const float x = lerp(a, b, t).x;
    if (x > P.x) return 0;
    return d > 0.0f ? 1 : -1;
}

// Flattened points of a quadratic (kind 1) or cubic (kind 2) in lyon's order; f(p) is
// called for every point after `from`, the last one being `to`.
template <class F>
__device__ __forceinline__ void for_each_flat_point(int kind, vec2 from, vec2 c1, vec2 c2,
                                                    vec2 to, float tol, F&& f) {
    vec2 a1, a2, a3 = mk(0.0f, 0.0f);
    float n;
    if (kind == 1) {
        n = quad_wang(from, c1, c2, tol);  // c2 holds the quadratic's end point
        a1 = mul(sub(c1, from), 2.0f);
        a2 = sub(add(from, c2), mul(c1, 2.0f));
    } else {
        n = cubic_wang(from, c1, c2, to, tol);
        a1 = mul(sub(c1, from), 3.0f);
        a2 = add(sub(mul(from, 3.0f), mul(c1, 6.0f)), mul(c2, 3.0f));
        a3 = add(sub(to, from), mul(sub(c1, c2), 3.0f));
    }
    const float step = 1.0f / n;
    const uint32_t cnt = seg_count(n) - 1u;
    float t = 0.0f;
    for (uint32_t k = 0; k < cnt; ++k) {
        t = t + step;
        vec2 v = from;
        float tt = t;
        v = add(v, mul(a1, tt));
        tt = tt * t;
        v = add(v, mul(a2, tt));
        if (kind == 2) {
            tt = tt * t;
            v = add(v, mul(a3, tt));
        }
        f(v);
    }
    f(to);
}

}  // namespace

__global__ void __launch_bounds__(WAVE* VAL_WAVES) winding_kernel(ValidateLaunch L) {
    const int l = lane_id();
    const uint32_t stride = gridDim.x * VAL_WAVES;
    for (uint32_t q = blockIdx.x * VAL_WAVES + threadIdx.x / WAVE; q < L.n_queries; q += stride) {
        const uint32_t pid = L.query_path[q];
        int32_t w = 0;
        if (pid < L.n_paths) {
            const vec2 P = mk(L.queries[2ull * q], L.queries[2ull * q + 1]);
            const uint32_t vb = L.verb_off[pid], nv = L.verb_off[pid + 1] - vb;
            const uint32_t pb = L.point_off[pid];
            const PathPts pts{L.points + 2ull * pb, L.point_off[pid + 1] - pb};
            uint32_t pc = 0;   // points consumed by the verbs before this window
            uint32_t beg = 0;  // 1 + point index of the latest Begin so far, 0 before any
            for (uint32_t c = 0; c < nv; c += WAVE) {
                const uint32_t i = c + (uint32_t)l;
                const uint32_t verb = i < nv ? (uint32_t)L.verbs[vb + i] : 0xffu;
                const uint32_t cons = points_of(verb);
                const uint32_t incl = incl_scan(cons);
                const uint32_t pi = pc + incl - cons;  // this verb's first point
                uint32_t b = incl_max_u32(verb == 3u ? pi + 1u : 0u);
                b = b > beg ? b : beg;
                if (verb != 3u && verb != 0xffu) {
                    const vec2 from = pts.at(pi - 1u);
                    if (verb == 0u) {
                        w += seg_winding(P, from, pts.at(pi));
                    } else if (verb == 1u || verb == 2u) {
                        const vec2 c1 = pts.at(pi), c2 = pts.at(pi + 1u);
                        const vec2 to = verb == 1u ? c2 : pts.at(pi + 2u);
                        // fast_bounding_range_y (quadratic_bezier.rs:561-566,
                        // cubic_bezier.rs:906-920)
                        float lo = ffmin(from.y, c1.y), hi = ffmax(from.y, c1.y);
                        if (verb == 2u) {
                            lo = ffmin(lo, c2.y);
                            hi = ffmax(hi, c2.y);
                        }
                        lo = ffmin(lo, to.y);
                        hi = ffmax(hi, to.y);
                        if (!(lo > P.y || hi < P.y)) {
                            vec2 prev = from;
                            for_each_flat_point((int)verb, from, c1, c2, to, L.tolerance,
                                                [&](vec2 p) {
                                                    w += seg_winding(P, prev, p);
                                                    prev = p;
                                                });
                        }
                    } else {
                        // End / Close: the edge from the current point back to the first
                        w += seg_winding(P, from, pts.at(b - 1u));
                    }
                }
                pc += readlane_u(incl, WAVE - 1);
                beg = readlane_u(b, WAVE - 1);
            }
        }
        w = (int32_t)readlane_u(incl_scan((uint32_t)w), WAVE - 1);
        if (l == 0) {
            L.winding[q] = w;
            if (L.inside) L.inside[q] = L.fill_rule == 0 ? (w % 2 != 0) : (w != 0);
        }
    }
}

// approximate_signed_area with one wave per path.  lyon's result is a serial f32 sum in
// event order, so the additions stay serial, but everything before them is spread over the
// lanes: the verbs are read 64 at a time; between two boundary verbs (Begin, End, Close) the
// flattened points of a run of segments are computed 1,024 at a time, 16 consecutive ones
// per lane (t from the closed form once, then lyon's serial sum), each lane forms the cross
// products of its points with the points before them, and the terms are added to the
// running sum in point order through readlane.  Begin / End / Close
// are applied in order between the runs, with area.rs's rules for ill-formed paths.
constexpr int AREA_WAVES = 4;
constexpr int AREA_RUN = 16;  // consecutive points per lane

__global__ void __launch_bounds__(WAVE* AREA_WAVES) signed_area_kernel(ValidateLaunch L) {
    __shared__ uint32_t pre_all[AREA_WAVES][WAVE + 1];  // term prefix of the window's verbs
    __shared__ uint32_t pi_all[AREA_WAVES][WAVE];       // first point of each verb
    __shared__ uint32_t vb_all[AREA_WAVES][WAVE];       // verb codes
    const int l = lane_id();
    const uint32_t w = threadIdx.x / WAVE;
    uint32_t* pre = pre_all[w];
    uint32_t* pis = pi_all[w];
    uint32_t* vbs = vb_all[w];
    const uint32_t stride = gridDim.x * AREA_WAVES;
    for (uint32_t pid = blockIdx.x * AREA_WAVES + w; pid < L.n_paths; pid += stride) {
        const uint32_t vb = L.verb_off[pid], nv = L.verb_off[pid + 1] - vb;
        const uint32_t pb = L.point_off[pid];
        const PathPts pts{L.points + 2ull * pb, L.point_off[pid + 1] - pb};
        float area = 0.0f, da = 0.0f;
        vec2 first = mk(0.0f, 0.0f), cur = mk(0.0f, 0.0f), v0c = mk(0.0f, 0.0f);
        bool in_sub = false, stop = false;
        uint32_t pc = 0;
        for (uint32_t c = 0; c < nv && !stop; c += WAVE) {
            const uint32_t i = c + (uint32_t)l;
            const uint32_t verb = i < nv ? (uint32_t)L.verbs[vb + i] : 0xffu;
            const uint32_t cons = points_of(verb);
            const uint32_t incl = incl_scan(cons);
            const uint32_t pi = pc + incl - cons;
            // flattened points of each segment verb (Line 1, curves Wang's count)
            uint32_t cnt = 0;
            if (verb == 0u) {
                cnt = 1;
            } else if (verb == 1u) {
                cnt = seg_count(quad_wang(pts.at(pi - 1u), pts.at(pi), pts.at(pi + 1u), L.tolerance));
            } else if (verb == 2u) {
                cnt = seg_count(cubic_wang(pts.at(pi - 1u), pts.at(pi), pts.at(pi + 1u),
                                           pts.at(pi + 2u), L.tolerance));
            }
            const uint32_t tincl = incl_scan(cnt);
            __builtin_amdgcn_wave_barrier();
            pre[l + 1] = tincl;
            if (l == 0) pre[0] = 0;
            pis[l] = pi;
            vbs[l] = verb;
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
            __builtin_amdgcn_wave_barrier();
            const uint64_t bmask = ballot(verb >= 3u);  // boundaries and lanes past the path
            int j = 0;
            while (j < WAVE && !stop) {
                const uint32_t vj = vbs[j];
                if (vj >= 3u) {
                    if (vj == 0xffu) break;
                    if (vj == 3u) {
                        // Begin: a sub-path starts (one inside a sub-path ends the sum)
                        if (in_sub) {
                            stop = true;
                            break;
                        }
                        first = pts.at(pis[j]);
                        cur = first;
                        da = 0.0f;
                        v0c = mk(0.0f, 0.0f);
                        in_sub = true;
                    } else {
                        // End / Close: the closing term, then the sub-path's area
                        if (!in_sub) {
                            stop = true;
                            break;
                        }
                        da = da + cross(v0c, sub(cur, first));
                        area = area + da * 0.5f;
                        in_sub = false;
                    }
                    ++j;
                    continue;
                }
                // a run of segment verbs j .. j2 - 1
                const uint64_t rest = bmask & ~mask_lt(j);
                const int j2 = rest ? __builtin_ctzll(rest) : WAVE;
                if (!in_sub) {
                    stop = true;
                    break;
                }
                const uint32_t r0 = pre[j], r1 = pre[j2];
                for (uint32_t q0 = r0; q0 < r1; q0 += WAVE * AREA_RUN) {
                    // lane l: points [g0, gend) of the run, in order
                    const uint32_t g0 = q0 + (uint32_t)l * AREA_RUN;
                    const uint32_t gend = g0 < r1 ? (g0 + AREA_RUN < r1 ? g0 + AREA_RUN : r1) : g0;
                    vec2 v1s[AREA_RUN];
                    vec2 p_last = cur, v1_last = mk(0.0f, 0.0f);
#pragma unroll
                    for (int u = 0; u < AREA_RUN; ++u) v1s[u] = mk(0.0f, 0.0f);
                    if (g0 < gend) {
                        int lo = j, hi = j2 - 1;  // last verb of the run with pre <= g0
                        while (lo < hi) {
                            const int mid = (lo + hi + 1) >> 1;
                            if (pre[mid] <= g0) lo = mid;
                            else hi = mid - 1;
                        }
                        int my = lo;
                        int kind = 0;
                        vec2 a0 = mk(0.0f, 0.0f), a1 = a0, a2 = a0, a3 = a0, to = a0;
                        float step = 1.0f, t = 0.0f;
                        uint32_t n_pts = 1, k = 0;
                        auto setup = [&](int v, uint32_t kk) {
                            const uint32_t sp = pis[v];
                            kind = (int)vbs[v];
                            k = kk;
                            if (kind == 0) {
                                to = pts.at(sp);
                                n_pts = 1;
                                return;
                            }
                            const vec2 f = pts.at(sp - 1u), c1 = pts.at(sp), c2 = pts.at(sp + 1u);
                            float nf;
                            if (kind == 1) {
                                nf = quad_wang(f, c1, c2, L.tolerance);
                                a1 = mul(sub(c1, f), 2.0f);
                                a2 = sub(add(f, c2), mul(c1, 2.0f));
                                to = c2;
                            } else {
                                const vec2 c3 = pts.at(sp + 2u);
                                nf = cubic_wang(f, c1, c2, c3, L.tolerance);
                                a1 = mul(sub(c1, f), 3.0f);
                                a2 = add(sub(mul(f, 3.0f), mul(c1, 6.0f)), mul(c2, 3.0f));
                                a3 = add(sub(c3, f), mul(sub(c1, c2), 3.0f));
                                to = c3;
                            }
                            a0 = f;
                            step = 1.0f / nf;
                            n_pts = seg_count(nf);
                            // t of point kk + 1: lyon's serial sum (closed form deep inside)
                            t = 0.0f;
                            if (kk + 1 < n_pts) {
                                if (kk + 1 <= (uint32_t)WAVE) {
                                    for (uint32_t q = 0; q <= kk; ++q) t = t + step;
                                } else {
                                    t = tchain(step, kk + 1);
                                }
                            }
                        };
                        setup(my, g0 - pre[my]);
#pragma unroll
                        for (int u = 0; u < AREA_RUN; ++u) {
                            const uint32_t g = g0 + (uint32_t)u;
                            if (g < gend) {
                                if (g >= pre[my + 1]) {
                                    ++my;
                                    setup(my, 0);
                                }
                                vec2 p;
                                if (k + 1 >= n_pts) {
                                    p = to;
                                } else {
                                    vec2 v = a0;
                                    float tt = t;
                                    v = add(v, mul(a1, tt));
                                    tt = tt * t;
                                    v = add(v, mul(a2, tt));
                                    if (kind == 2) {
                                        tt = tt * t;
                                        v = add(v, mul(a3, tt));
                                    }
                                    p = v;
                                }
                                ++k;
                                t = t + step;
                                v1s[u] = sub(p, first);
                                v1_last = v1s[u];
                                p_last = p;
                            }
                        }
                    }
                    // the point before the lane's first: the previous lane's last (its run is
                    // full whenever this one has points), or the carry for lane 0
                    vec2 v0 = shfl_v(v1s[AREA_RUN - 1], l == 0 ? 0 : l - 1);
                    if (l == 0) v0 = v0c;
                    float terms[AREA_RUN];
#pragma unroll
                    for (int u = 0; u < AREA_RUN; ++u) {
                        terms[u] = cross(v0, v1s[u]);
                        v0 = v1s[u];
                    }
                    // lyon's serial sum, in point order
                    const uint32_t m = r1 - q0 < (uint32_t)(WAVE * AREA_RUN) ? r1 - q0
                                                                             : (uint32_t)(WAVE * AREA_RUN);
                    const uint32_t full = m / AREA_RUN, rem = m % AREA_RUN;
                    for (uint32_t ln = 0; ln < full; ++ln) {
#pragma unroll
                        for (int u = 0; u < AREA_RUN; ++u) da = da + readlane_f(terms[u], (int)ln);
                    }
                    if (rem) {
#pragma unroll
                        for (int u = 0; u < AREA_RUN; ++u)
                            if ((uint32_t)u < rem) da = da + readlane_f(terms[u], (int)full);
                    }
                    const int last_lane = (int)((m - 1) / AREA_RUN);
                    v0c = readlane_v(v1_last, last_lane);
                    cur = readlane_v(p_last, last_lane);
                }
                j = j2;
            }
            pc += readlane_u(incl, WAVE - 1);
            __builtin_amdgcn_wave_barrier();
        }
        if (l == 0) L.areas[pid] = area;
    }
}

}  // namespace lyg

extern "C" {

hipError_t lyg_launch_winding(const lyg::ValidateLaunch* L, hipStream_t s) {
    if (L->n_queries == 0) return hipSuccess;
    uint32_t blocks = (L->n_queries + lyg::VAL_WAVES - 1) / lyg::VAL_WAVES;
    if (blocks > 65536u) blocks = 65536u;
    hipLaunchKernelGGL(lyg::winding_kernel, dim3(blocks), dim3(lyg::WAVE * lyg::VAL_WAVES), 0, s, *L);
    return hipGetLastError();
}

hipError_t lyg_launch_signed_area(const lyg::ValidateLaunch* L, hipStream_t s) {
    if (L->n_paths == 0) return hipSuccess;
    uint32_t blocks = (L->n_paths + lyg::AREA_WAVES - 1) / lyg::AREA_WAVES;
    if (blocks > 65536u) blocks = 65536u;
    hipLaunchKernelGGL(lyg::signed_area_kernel, dim3(blocks), dim3(lyg::WAVE * lyg::AREA_WAVES), 0,
                       s, *L);
    return hipGetLastError();
}


}  // extern "C"
