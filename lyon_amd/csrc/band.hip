// Banded fill: the fill of large paths spread over the whole GPU.
//
// lyon's FillTessellator sweeps one path in one sequential loop (fill.rs:836-867), so a
// path of a million edges keeps one core busy for a minute.  Here the path's flattened
// edges (flattened as lyon's event-queue builder flattens them: curves turned to run down
// the sweep, event_queue.rs:758-900, every sub-path closed, :661-682) are cut into
// horizontal bands of about equal numbers of edge starts, and every band is swept on its
// own:
//
//   1. per verb: points consumed, edges produced (Wang counts of the turned curves),
//      then the edges themselves, each oriented top to bottom with winding +1/-1;
//   2. per path: a histogram of edge tops gives band boundaries; every edge is listed in
//      the bands it overlaps;
//   3. per band (one workgroup): every pair of band edges that cross strictly inside both
//      is found with exact orientation signs in f64, sorted edge lists by x range
//      keeping the candidate pairs few; the crossing's height is taken in f64;
//   4. per band (one workgroup): the band's events (edge starts, edge ends, crossings)
//      sorted by height, and the order of the edges entering at the band's top;
//   5. per band (one wave): a sweep keeping the active edges in left-to-right order in
//      LDS with the winding number of every gap between neighbours; every event rebuilds
//      the few positions it touches (sorted by x half-way to the next event, so order is
//      restored even where rounding bends the exact crossing order), and every gap whose
//      pair of edges changes closes a trapezoid, written when the fill rule puts it
//      inside;
//   6. trapezoids to triangles (two, one at a crossing point, none when empty).
//
// Coverage is the path's (EvenOdd / NonZero over the same flattened edges); the triangles
// are not lyon's (its monotone spans differ) and vertices are not shared between
// trapezoids.
#include <hip/hip_runtime.h>

#include "lyg_math.h"
#include "lyg_types.h"
#include "wave.h"
#include "band_params.h"
#include "stroke_core.h"

namespace lyg {
namespace band {

static constexpr uint32_t M_MAX = 4096;    // edges of one band (LDS arrays)
static constexpr uint32_t E_LDS = 4096;    // events sorted in LDS (more: in global memory)
static constexpr uint32_t BLK_MAX = 32;    // positions one event may rebuild
static constexpr uint32_t NO_POS = 0xffffu;

// band status bits
static constexpr uint32_t BS_DENSE = 1;     // more than M_MAX edges in the band
static constexpr uint32_t BS_RECORDS = 2;   // trapezoid room exceeded (retry with more)
static constexpr uint32_t BS_BLOCK = 4;     // an event touched more than BLK_MAX positions
static constexpr uint32_t BS_NAN = 8;       // a NaN position in the path

// Params: band_params.h


__device__ __forceinline__ uint32_t okey(float f) {
    if (f == 0.0f) f = 0.0f;
    uint32_t u = f2u(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float okey_inv(uint32_t k) {
    uint32_t u = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
    return u2f(u);
}
__device__ __forceinline__ bool after(vec2 a, vec2 b) {
    return a.y > b.y || (a.y == b.y && a.x > b.x);
}
__device__ __forceinline__ bool inside(uint32_t rule, int32_t w) {
    return rule == 0 ? ((w & 1) != 0) : (w != 0);
}
__device__ __forceinline__ uint32_t path_of(const uint32_t* off, uint32_t n, uint32_t i) {
    uint32_t lo = 0, hi = n;  // largest p with off[p] <= i
    while (hi - lo > 1) {
        uint32_t m = (lo + hi) >> 1;
        if (off[m] <= i) lo = m;
        else hi = m;
    }
    return lo;
}
__device__ __forceinline__ const lyg_fill_options& opt_of(const Params& P, uint32_t p) {
    return P.opts[P.opts_stride ? p : 0];
}

// x of the edge's line at height y, in f64 (exact at its end points).
__device__ __forceinline__ double xat(float4 g, double y) {
    if (y == (double)g.y) return g.x;
    if (y == (double)g.w) return g.z;
    return (double)g.x + (y - (double)g.y) * (((double)g.z - (double)g.x) / ((double)g.w - (double)g.y));
}
__device__ __forceinline__ double orient(double px, double py, double qx, double qy, double rx,
                                         double ry) {
    return (qx - px) * (ry - py) - (qy - py) * (rx - px);
}

// ---- 1. verbs --------------------------------------------------------------------------
// Points each verb consumes (path.rs:843-909): Begin 1, Line 1, Quadratic 2, Cubic 3,
// Close 1 (the stored copy of the first point), End 0.
__global__ void verb_info_kernel(Params P) {
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= P.n_verbs) return;
    const uint8_t k = P.verbs[v];
    uint32_t n = 0;
    switch (k) {
        case LYG_VERB_BEGIN: case LYG_VERB_LINE: case LYG_VERB_CLOSE: n = 1; break;
        case LYG_VERB_QUADRATIC: n = 2; break;
        case LYG_VERB_CUBIC: n = 3; break;
        default: n = 0;
    }
    P.v_npts[v] = n;
    P.v_isbeg[v] = k == LYG_VERB_BEGIN ? 1u : 0u;
    P.v_path[v] = path_of(P.verb_off, P.n_paths, v);
}

__global__ void begin_list_kernel(Params P) {
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= P.n_verbs) return;
    if (P.verbs[v] == LYG_VERB_BEGIN) P.begin_verb[P.v_begord[v]] = v;
}

struct Seg {
    uint8_t kind;  // 0 none, 1 line, 2 quadratic, 3 cubic
    vec2 from, c1, c2, to;
    bool nan;
};

// The segment verb v contributes to the event queue builder (begin/line/quadratic/cubic/
// end of event_queue.rs:661-900, driven by Iter::next): its start is the point before its
// own, and End / Close close the sub-path with a line to its first point.
__device__ Seg seg_of(const Params& P, uint32_t v) {
    Seg s;
    s.kind = 0;
    s.nan = false;
    const uint8_t k = P.verbs[v];
    const uint32_t p = P.v_path[v];
    const uint32_t pb = P.point_off[p] + (P.v_pbase[v] - P.v_pbase[P.verb_off[p]]);
    if (k == LYG_VERB_BEGIN) return s;
    // no current point (a path not starting with Begin), or points missing at the end
    const uint32_t pend = P.point_off[p + 1];
    const uint32_t need = k == LYG_VERB_LINE ? 1u : (k == LYG_VERB_QUADRATIC ? 2u : (k == LYG_VERB_CUBIC ? 3u : 0u));
    if (pb == P.point_off[p] || pb > pend || pend - pb < need) return s;
    const float* pt = P.points;
    s.from = mk(pt[2 * (pb - 1)], pt[2 * (pb - 1) + 1]);
    if (k == LYG_VERB_END || k == LYG_VERB_CLOSE) {
        const uint32_t ord = P.v_begord[v];  // Begin verbs before v
        if (ord == 0) return s;
        const uint32_t bv = P.begin_verb[ord - 1];
        if (bv < P.verb_off[p]) return s;
        const uint32_t fb = P.point_off[p] + (P.v_pbase[bv] - P.v_pbase[P.verb_off[p]]);
        if (fb >= pend) return s;
        s.to = mk(pt[2 * fb], pt[2 * fb + 1]);
        s.kind = 1;
    } else if (k == LYG_VERB_LINE) {
        s.to = mk(pt[2 * pb], pt[2 * pb + 1]);
        s.kind = 1;
    } else if (k == LYG_VERB_QUADRATIC) {
        s.c1 = mk(pt[2 * pb], pt[2 * pb + 1]);
        s.to = mk(pt[2 * pb + 2], pt[2 * pb + 3]);
        s.kind = 2;
    } else if (k == LYG_VERB_CUBIC) {
        s.c1 = mk(pt[2 * pb], pt[2 * pb + 1]);
        s.c2 = mk(pt[2 * pb + 2], pt[2 * pb + 3]);
        s.to = mk(pt[2 * pb + 4], pt[2 * pb + 5]);
        s.kind = 3;
    }
    // non-finite positions are rejected (lyon checks NaN; an infinite one has no fill)
    auto bad = [](vec2 q) { return !(ffabs(q.x) <= 3.4028235e38f) || !(ffabs(q.y) <= 3.4028235e38f); };
    s.nan = bad(s.from) || bad(s.to) || (s.kind >= 2 && bad(s.c1)) || (s.kind == 3 && bad(s.c2));
    return s;
}

// A curve is flattened after being turned to run down the sweep (event_queue.rs:758-900):
// the turned control points, in lyon's order.
__device__ __forceinline__ void turned(const Seg& s, vec2& f, vec2& a, vec2& b, vec2& t, bool& sw) {
    f = s.from;
    a = s.c1;
    b = s.c2;
    t = s.to;
    sw = after(s.from, s.to);
    if (sw) {
        f = s.to;
        t = s.from;
        if (s.kind == 3) {
            a = s.c2;
            b = s.c1;
        }
    }
}

__global__ void edge_count_kernel(Params P) {
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= P.n_verbs) return;
    Seg s = seg_of(P, v);
    uint32_t n = 0;
    if (s.nan) {
        atomicOr(&P.p_status[P.v_path[v]], BS_NAN);
    } else if (s.kind == 1) {
        n = 1;
    } else if (s.kind >= 2) {
        const float tol = opt_of(P, P.v_path[v]).tolerance;
        vec2 f, a, b, t;
        bool sw;
        turned(s, f, a, b, t, sw);
        n = seg_count(s.kind == 3 ? cubic_wang(f, a, b, t, tol) : quad_wang(f, a, t, tol));
    }
    P.v_nedge[v] = n;
}

__device__ __forceinline__ void put_edge(const Params& P, uint32_t e, uint32_t path, vec2 a,
                                         vec2 b, int w, uint32_t* ymin, uint32_t* ymax) {
    // oriented top (first in sweep order) to bottom; horizontal and empty edges add
    // nothing to a fill and get winding 0
    if (after(a, b)) {
        vec2 x = a;
        a = b;
        b = x;
        w = -w;
    }
    P.geo[e] = make_float4(a.x, a.y, b.x, b.y);
    const bool flat = a.y == b.y;
    P.wind[e] = flat ? 0 : (int8_t)w;
    P.e_path[e] = path;
    if (!flat) {
        *ymin = min(*ymin, okey(a.y));
        *ymax = max(*ymax, okey(b.y));
    }
}

__global__ void edge_emit_kernel(Params P) {
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t ymin = 0xffffffffu, ymax = 0u;
    uint32_t path = 0;
    if (v < P.n_verbs && P.v_nedge[v] > 0) {
        path = P.v_path[v];
        Seg s = seg_of(P, v);
        uint32_t e = P.v_ebase[v];
        if (s.kind == 1) {
            put_edge(P, e, path, s.from, s.to, 1, &ymin, &ymax);
        } else {
            const float tol = opt_of(P, path).tolerance;
            vec2 f, a, b, t;
            bool sw;
            turned(s, f, a, b, t, sw);
            const bool cubic = s.kind == 3;
            // polynomial form, Wang count and serial t (quadratic_bezier.rs:368-391, 891-974;
            // cubic_bezier.rs:590-614, 1318-1403)
            float n = cubic ? cubic_wang(f, a, b, t, tol) : quad_wang(f, a, t, tol);
            float step = 1.0f / n;
            uint32_t cnt = seg_count(n);
            vec2 a0 = f, a1, a2, a3 = mk(0.0f, 0.0f);
            if (cubic) {
                a1 = mul(sub(a, f), 3.0f);
                a2 = add(sub(mul(f, 3.0f), mul(a, 6.0f)), mul(b, 3.0f));
                a3 = add(sub(t, f), mul(sub(a, b), 3.0f));
            } else {
                a1 = mul(sub(a, f), 2.0f);
                a2 = sub(add(f, t), mul(a, 2.0f));
            }
            float t0 = 0.0f;
            vec2 lf = f;
            const int w = sw ? -1 : 1;
            for (uint32_t i = 0; i < cnt; ++i) {
                vec2 lt;
                if (i + 1 == cnt) {
                    lt = t;
                } else {
                    float t1 = t0 + step;
                    vec2 q = a0;
                    float tt = t1;
                    q = add(q, mul(a1, tt));
                    tt = tt * t1;
                    q = add(q, mul(a2, tt));
                    if (cubic) {
                        tt = tt * t1;
                        q = add(q, mul(a3, tt));
                    }
                    lt = q;
                    t0 = t1;
                }
                put_edge(P, e + i, path, lf, lt, w, &ymin, &ymax);
                lf = lt;
            }
        }
    }
    // per-path y range: the lanes of a wave mostly share a path
    const uint32_t p0 = readlane_u(path, 0);
    const bool same = ballot(path == p0 || ymin == 0xffffffffu) == ~0ull;
    if (same) {
        uint32_t mn = wave_min_u32(ymin);
        uint32_t mx = ~wave_min_u32(~ymax);
        if (lane_id() == 0 && mn != 0xffffffffu) {
            atomicMin(&P.p_yr[2 * p0], mn);
            atomicMax(&P.p_yr[2 * p0 + 1], mx);
        }
    } else if (ymin != 0xffffffffu) {
        atomicMin(&P.p_yr[2 * path], ymin);
        atomicMax(&P.p_yr[2 * path + 1], ymax);
    }
}

// ---- 2. bands --------------------------------------------------------------------------
__global__ void path_bands_kernel(Params P, const uint32_t* p_nedge_edges) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P.n_paths) return;
    const uint32_t ne = p_nedge_edges[p + 1] - p_nedge_edges[p];
    uint64_t nb = ((uint64_t)ne + P.edges_per_band - 1) / P.edges_per_band;
    if (P.p_mult) nb *= P.p_mult[p];  // paths whose bands came out too dense: more bands
    nb = max((uint64_t)1, min(nb, (uint64_t)P.max_bands));
    if (P.p_yr[2 * p] == 0xffffffffu || P.p_status[p] != 0) nb = 1;  // nothing to fill
    P.p_nband[p] = (uint32_t)nb;
}

__global__ void hist_kernel(Params P) {
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= P.n_edges || P.wind[e] == 0) return;
    const uint32_t p = P.e_path[e];
    const uint32_t nb = P.p_nband[p];
    if (nb <= 1) return;
    const uint32_t nbin = 64 * nb;
    const float y0 = okey_inv(P.p_yr[2 * p]), y1 = okey_inv(P.p_yr[2 * p + 1]);
    const double span = (double)y1 - (double)y0;
    const double f = ((double)P.geo[e].y - (double)y0) / span * (double)nbin;
    const uint32_t bin = f >= 0.0 ? (f < (double)(nbin - 1) ? (uint32_t)f : nbin - 1) : 0u;
    atomicAdd(&P.hist[P.p_hist_off[p] + bin], 1u);
}

__device__ __forceinline__ float bin_y(float y0, float y1, uint32_t i, uint32_t nbin) {
    if (i == 0) return y0;
    if (i >= nbin) return y1;
    return (float)((double)y0 + ((double)y1 - (double)y0) * (double)i / (double)nbin);
}

// One block per path: band boundaries at the bins where the running count of edge tops
// passes each multiple of (tops / bands); bands of zero height stay empty.
__global__ void __launch_bounds__(256) bounds_kernel(Params P) {
    const uint32_t p = blockIdx.x;
    const uint32_t nb = P.p_nband[p];
    const uint32_t b0 = P.p_band_off[p];
    const float y0 = okey_inv(P.p_yr[2 * p]), y1 = okey_inv(P.p_yr[2 * p + 1]);
    const bool empty = P.p_yr[2 * p] == 0xffffffffu;
    if (nb <= 1) {
        if (threadIdx.x == 0) {
            P.b_ytop[b0] = empty ? 0.0f : y0;
            P.b_ybot[b0] = empty ? 0.0f : y1;
            P.b_path[b0] = p;
        }
        return;
    }
    const uint32_t nbin = 64 * nb;
    const uint32_t* h = P.hist + P.p_hist_off[p];
    __shared__ uint32_t part[256];
    const uint32_t t = threadIdx.x;
    const uint32_t per = (nbin + 255) / 256;
    const uint32_t lo = t * per, hi = min(nbin, lo + per);
    uint32_t s = 0;
    for (uint32_t i = lo; i < hi; ++i) s += h[i];
    part[t] = s;
    __syncthreads();
    for (uint32_t o = 1; o < 256; o <<= 1) {
        uint32_t x = t >= o ? part[t - o] : 0;
        __syncthreads();
        part[t] += x;
        __syncthreads();
    }
    const uint64_t total = part[255];
    // band k (1 .. nb-1) starts at the first bin i whose running count before it reaches
    // T_k = ceil(k * total / nb); bands never reached start at the path's bottom (empty)
    for (uint32_t k = 1 + t; k < nb; k += 256) P.b_ytop[b0 + k] = y1;
    __syncthreads();
    uint64_t run = part[t] - s;
    for (uint32_t i = lo; i < hi; ++i) {
        const uint64_t ci = run;  // tops in bins before i
        if (i > 0 && total > 0) {
            const uint64_t cprev = ci - h[i - 1];
            const uint64_t kmin = (cprev * nb) / total + 1;
            const uint64_t kmax = (ci * nb) / total;
            for (uint64_t k = kmin; k <= kmax && k < nb; ++k) P.b_ytop[b0 + k] = bin_y(y0, y1, i, nbin);
        }
        run += h[i];
    }
    __syncthreads();
    if (t == 0) P.b_ytop[b0] = y0;
    __syncthreads();
    for (uint32_t k = t; k < nb; k += 256) {
        P.b_ybot[b0 + k] = (k + 1 < nb) ? P.b_ytop[b0 + k + 1] : y1;
        P.b_path[b0 + k] = p;
    }
}

// Bands of path p overlapped by an edge [ya, yb] (ya < yb): k with ytop_k < yb and
// ybot_k > ya.
__device__ __forceinline__ void band_range(const Params& P, uint32_t p, float ya, float yb,
                                           uint32_t& k0, uint32_t& k1) {
    const uint32_t b0 = P.p_band_off[p], nb = P.p_band_off[p + 1] - b0;
    // first band whose bottom is above ya... i.e. the last band with ytop <= ya
    uint32_t lo = 0, hi = nb;
    while (hi - lo > 1) {
        uint32_t m = (lo + hi) >> 1;
        if (P.b_ytop[b0 + m] <= ya) lo = m;
        else hi = m;
    }
    k0 = b0 + lo;
    k1 = k0;
    while (k1 + 1 < b0 + nb && P.b_ytop[k1 + 1] < yb) ++k1;
}

__global__ void band_count_kernel(Params P) {
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= P.n_edges || P.wind[e] == 0) return;
    const uint32_t p = P.e_path[e];
    if (P.p_status[p] != 0) return;
    const float4 g = P.geo[e];
    uint32_t k0, k1;
    band_range(P, p, g.y, g.w, k0, k1);
    for (uint32_t k = k0; k <= k1; ++k)
        if (P.b_ytop[k] < P.b_ybot[k]) atomicAdd(&P.b_nedge[k], 1u);
}

__global__ void band_fill_kernel(Params P) {
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= P.n_edges || P.wind[e] == 0) return;
    const uint32_t p = P.e_path[e];
    if (P.p_status[p] != 0) return;
    const float4 g = P.geo[e];
    uint32_t k0, k1;
    band_range(P, p, g.y, g.w, k0, k1);
    for (uint32_t k = k0; k <= k1; ++k)
        if (P.b_ytop[k] < P.b_ybot[k]) {
            const uint32_t slot = atomicAdd(&P.b_fill[k], 1u);
            P.band_edges[P.b_edge_off[k] + slot] = e;
        }
}

// ---- sorting helpers (bitonic, one workgroup) --------------------------------------------
template <class T, class Less>
__device__ void block_bitonic(T* a, uint32_t n_pow2, Less less) {
    for (uint32_t k = 2; k <= n_pow2; k <<= 1) {
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t i = threadIdx.x; i < n_pow2; i += blockDim.x) {
                const uint32_t l = i ^ j;
                if (l > i) {
                    const bool up = (i & k) == 0;
                    T x = a[i], y = a[l];
                    if (up ? less(y, x) : less(x, y)) {
                        a[i] = y;
                        a[l] = x;
                    }
                }
            }
            __syncthreads();
        }
    }
}
__device__ __forceinline__ uint32_t pow2_at_least(uint32_t n) {
    uint32_t p = 1;
    while (p < n) p <<= 1;
    return p;
}

// ---- 3. crossings ------------------------------------------------------------------------
// Exact crossing test of band-local edges a < b (proper crossings only: strictly inside
// both); *y gets the crossing height.
__device__ __forceinline__ bool crossing(float4 A, float4 B, double* y) {
    const double d1 = orient(A.x, A.y, A.z, A.w, B.x, B.y);
    const double d2 = orient(A.x, A.y, A.z, A.w, B.z, B.w);
    if (!((d1 > 0 && d2 < 0) || (d1 < 0 && d2 > 0))) return false;
    const double d3 = orient(B.x, B.y, B.z, B.w, A.x, A.y);
    const double d4 = orient(B.x, B.y, B.z, B.w, A.z, A.w);
    if (!((d3 > 0 && d4 < 0) || (d3 < 0 && d4 > 0))) return false;
    const double t = d3 / (d3 - d4);
    double yy = (double)A.y + t * ((double)A.w - (double)A.y);
    const double lo = fmax((double)A.y, (double)B.y), hi = fmin((double)A.w, (double)B.w);
    yy = fmin(fmax(yy, lo), hi);
    *y = yy;
    return true;
}

// One block of 256 per band: the band's edge list sorted by edge id (band-local ids are
// then reproducible), the band's copy of its edges' geometry and windings, and its counts
// of edges entering at the top, starting inside and ending inside.
__global__ void __launch_bounds__(256) band_sort_kernel(Params P, int8_t* bw) {
    const uint32_t b = blockIdx.x;
    const uint32_t m = P.b_nedge[b];
    const uint32_t off = P.b_edge_off[b];
    __shared__ uint64_t key[M_MAX];
    __shared__ uint32_t c_start, c_end, c_init;
    if (threadIdx.x == 0) {
        c_start = c_end = c_init = 0;
    }
    if (m > M_MAX) {
        if (threadIdx.x == 0) {
            P.b_status[b] = BS_DENSE;
            atomicOr(&P.p_status[P.b_path[b]], BS_DENSE);
            P.b_nev[b] = 0;
            P.b_ninit[b] = 0;
        }
        return;
    }
    const float ytop = P.b_ytop[b], ybot = P.b_ybot[b];
    const uint32_t n2 = pow2_at_least(max(m, 1u));
    for (uint32_t i = threadIdx.x; i < n2; i += blockDim.x)
        key[i] = i < m ? (uint64_t)P.band_edges[off + i] : ~0ull;
    __syncthreads();
    block_bitonic(key, n2, [](uint64_t x, uint64_t y) { return x < y; });
    uint32_t ns = 0, nd = 0, ni = 0;
    for (uint32_t i = threadIdx.x; i < m; i += blockDim.x) {
        const uint32_t e = (uint32_t)key[i];
        const float4 g = P.geo[e];
        P.band_edges[off + i] = e;
        P.bgeo[off + i] = g;
        bw[off + i] = P.wind[e];
        if (g.y >= ytop) ++ns;
        else ++ni;
        if (g.w < ybot) ++nd;
    }
    atomicAdd(&c_start, ns);
    atomicAdd(&c_end, nd);
    atomicAdd(&c_init, ni);
    __syncthreads();
    if (threadIdx.x == 0) {
        P.b_nev[b] = c_start + c_end;  // crossings are added once counted
        P.b_ninit[b] = c_init;
    }
}

// One block of 256 per band.  EMIT = 0 counts the band's crossings, EMIT = 1 writes them.
template <int EMIT>
__global__ void __launch_bounds__(256) cross_kernel(Params P) {
    const uint32_t b = blockIdx.x;
    const uint32_t m = P.b_nedge[b];
    const uint32_t off = P.b_edge_off[b];
    if (m == 0 || P.b_status[b] != 0) {
        if (!EMIT && threadIdx.x == 0) P.b_ncross[b] = 0;
        return;
    }
    __shared__ uint64_t key[M_MAX];
    __shared__ float xhi[M_MAX];
    __shared__ uint32_t cnt;
    const double ytop = P.b_ytop[b], ybot = P.b_ybot[b];
    const uint32_t n2 = pow2_at_least(m);
    if (threadIdx.x == 0) cnt = 0;
    // x range of every edge over the band (rounded outwards), sorted by its low end
    for (uint32_t i = threadIdx.x; i < n2; i += blockDim.x) {
        if (i < m) {
            const float4 g = P.bgeo[off + i];
            const double ya = fmax((double)g.y, ytop), yb = fmin((double)g.w, ybot);
            const double xa = xat(g, ya), xb = xat(g, yb);
            const double lo = fmin(xa, xb), hi = fmax(xa, xb);
            float flo = (float)lo, fhi = (float)hi;
            if ((double)flo > lo) flo = okey_inv(okey(flo) - 1u);  // next float down
            if ((double)fhi < hi) fhi = okey_inv(okey(fhi) + 1u);  // next float up
            xhi[i] = fhi;
            key[i] = ((uint64_t)okey(flo) << 32) | i;
        } else {
            key[i] = ~0ull;
        }
    }
    __syncthreads();
    block_bitonic(key, n2, [](uint64_t x, uint64_t y) { return x < y; });
    const uint32_t base = EMIT ? P.b_cross_off[b] : 0;
    uint32_t mine = 0;
    for (uint32_t i = threadIdx.x; i < m; i += blockDim.x) {
        const uint32_t ia = (uint32_t)key[i];
        const float4 A = P.bgeo[off + ia];
        const float ha = xhi[ia];
        for (uint32_t j = i + 1; j < m; ++j) {
            const uint64_t kj = key[j];
            if (okey_inv((uint32_t)(kj >> 32)) > ha) break;
            const uint32_t ib = (uint32_t)kj;
            const float4 B = P.bgeo[off + ib];
            if (!(B.y < A.w && A.y < B.w)) continue;
            // band-local ids follow edge ids: the lower id is always the first operand,
            // so a pair gets the same height in every band that holds both
            double y;
            if (!(ia < ib ? crossing(A, B, &y) : crossing(B, A, &y))) continue;
            if (y < ytop || y >= ybot) continue;
            if (EMIT) {
                const uint32_t slot = atomicAdd(&cnt, 1u);
                if (slot < P.b_ncross[b]) {  // the count pass saw the same pairs
                    P.cross_y[base + slot] = y;
                    P.cross_ab[base + slot] = min(ia, ib) | (max(ia, ib) << 16);
                }
            } else {
                ++mine;
            }
        }
    }
    if (!EMIT) {
        if (mine) atomicAdd(&cnt, mine);
        __syncthreads();
        if (threadIdx.x == 0) P.b_ncross[b] = cnt;
    }
}

// ---- 4. events -----------------------------------------------------------------------------
struct EvRec {
    double y;
    float x;
    uint32_t k;  // kind << 30 | ids
};
static constexpr uint32_t EV_CROSS = 0u, EV_END = 1u, EV_START = 2u;
__device__ __forceinline__ bool ev_less(const EvRec& a, const EvRec& b) {
    if (a.y != b.y) return a.y < b.y;
    if (a.x != b.x) return a.x < b.x;
    return a.k < b.k;  // crossings, then ends, then starts; then ids
}

__global__ void event_count_kernel(Params P) {
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= P.n_bands) return;
    const bool ok = P.b_status[b] == 0;
    const uint32_t n = ok ? P.b_nev[b] + P.b_ncross[b] : 0;
    const uint32_t ninit = ok ? P.b_ninit[b] : 0;
    P.b_nev[b] = n;
    // room for the band's trapezoids: every gap that closes was opened at the top or by an
    // event, and an event opens at most (positions it rebuilds + 1) gaps, 3 in the usual
    // cases (a crossing, a vertex with two edges)
    P.b_rec_cap[b] = ok && P.b_nedge[b] ? ninit + 3 * n + 4 : 0;
}

// One block of 256 per band: the band's events sorted (in LDS when they fit), and the
// edges entering at the top sorted by x half-way down to the first event below the top.
__global__ void __launch_bounds__(256) event_sort_kernel(Params P, EvRec* scratch) {
    const uint32_t b = blockIdx.x;
    if (P.b_status[b] != 0 || P.b_nedge[b] == 0) return;
    const uint32_t m = P.b_nedge[b];
    const uint32_t off = P.b_edge_off[b];
    const uint32_t ne = P.b_nev[b];
    const uint32_t eoff = P.b_ev_off[b];
    const float ytop = P.b_ytop[b], ybot = P.b_ybot[b];
    __shared__ EvRec lds[E_LDS];
    __shared__ uint32_t c_ev, c_init;
    __shared__ double first_below;
    const uint32_t n2 = ne ? pow2_at_least(ne) : 0;
    const bool in_lds = n2 <= E_LDS;
    EvRec* ev = in_lds ? lds : scratch + 2ull * eoff;  // 2 * eoff: room for the padding
    if (threadIdx.x == 0) {
        c_ev = 0;
        c_init = 0;
        first_below = (double)ybot;
    }
    __syncthreads();
    const uint32_t nc = P.b_ncross[b], coff = P.b_cross_off[b];
    for (uint32_t i = threadIdx.x; i < nc; i += blockDim.x) {
        const uint32_t ab = P.cross_ab[coff + i];
        const double y = P.cross_y[coff + i];
        const float4 A = P.bgeo[off + (ab & 0xffffu)];
        EvRec r;
        r.y = y;
        r.x = (float)xat(A, y);
        r.k = (EV_CROSS << 30) | ab;
        ev[i] = r;
    }
    for (uint32_t i = threadIdx.x; i < m; i += blockDim.x) {
        const float4 g = P.bgeo[off + i];
        if (g.y >= ytop) {
            const uint32_t s = nc + atomicAdd(&c_ev, 1u);
            EvRec r;
            r.y = g.y;
            r.x = g.x;
            r.k = (EV_START << 30) | i;
            ev[s] = r;
        }
        if (g.w < ybot) {
            const uint32_t s = nc + atomicAdd(&c_ev, 1u);
            EvRec r;
            r.y = g.w;
            r.x = g.z;
            r.k = (EV_END << 30) | i;
            ev[s] = r;
        }
    }
    __syncthreads();
    for (uint32_t i = ne + threadIdx.x; i < n2; i += blockDim.x) {
        EvRec r;
        r.y = INFINITY;
        r.x = INFINITY;
        r.k = 0xffffffffu;
        ev[i] = r;
    }
    __syncthreads();
    block_bitonic(ev, n2, [](const EvRec& a, const EvRec& c) { return ev_less(a, c); });
    for (uint32_t i = threadIdx.x; i < ne; i += blockDim.x) {
        const EvRec r = ev[i];
        P.ev_y[eoff + i] = r.y;
        P.ev_x[eoff + i] = r.x;
        P.ev_k[eoff + i] = r.k;
        if (r.y > (double)ytop && (i == 0 || ev[i - 1].y <= (double)ytop)) first_below = r.y;
    }
    __syncthreads();
    // the edges entering at the top, sorted by x at the middle of the first slab
    const double ymid = 0.5 * ((double)ytop + first_below);
    const uint32_t ni = P.b_ninit[b];
    const uint32_t m2 = pow2_at_least(max(ni, 1u));
    struct IK {
        double x;
        uint32_t id;
    };
    // (ni <= M_MAX records of 16 bytes: always in LDS, free again once the events are out)
    IK* ik = reinterpret_cast<IK*>(lds);
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < m; i += blockDim.x) {
        const float4 g = P.bgeo[off + i];
        if (g.y < ytop) {
            const uint32_t s = atomicAdd(&c_init, 1u);
            IK v;
            v.x = xat(g, ymid);
            v.id = i;
            ik[s] = v;
        }
    }
    __syncthreads();
    for (uint32_t i = ni + threadIdx.x; i < m2; i += blockDim.x) {
        IK v;
        v.x = INFINITY;
        v.id = 0xffffffffu;
        ik[i] = v;
    }
    __syncthreads();
    block_bitonic(ik, m2, [](const IK& a, const IK& c) {
        return a.x < c.x || (a.x == c.x && a.id < c.id);
    });
    for (uint32_t i = threadIdx.x; i < ni; i += blockDim.x) P.init_order[off + i] = (uint16_t)ik[i].id;
}

// ---- 5. the sweep ------------------------------------------------------------------------
struct Gap {
    uint32_t l, r;
    int32_t w;
    uint32_t top;  // the event the gap opened at (TOP_EV: the band's top)
    uint32_t matched;
};
// A trapezoid's top and bottom are the heights of events (f64, as sorted): records carry
// event indices so every corner at one event lies on the same exact height, rounded to f32
// only when the triangles are written.  Two values stand for the band's top and bottom.
static constexpr uint32_t TOP_EV = 0xffffffffu, BOT_EV = 0xfffffffeu;

// Events are read through a window of EV_WIN consecutive events staged in LDS (one
// coalesced load per lane), moved forward when the sweep nears its end.
static constexpr uint32_t EV_WIN = 256;
// Bands are swept by the variant with the smallest state that holds their edges, so more
// waves fit a CU's LDS: up to M_SMALL, M_MID, M_MAX edges.
static constexpr uint32_t M_SMALL = 512;
static constexpr uint32_t M_MID = 2048;

template <uint32_t M>
struct SweepState {
    uint16_t act[M];
    uint16_t pos[M];
    int16_t wsum[M];  // winding of the gap right of act[i]
    uint32_t top[M];  // the event that gap opened at
    // one event's rebuild
    uint32_t nb[BLK_MAX];
    double nk[BLK_MAX];
    Gap og[BLK_MAX + 1];
    // the event window
    double ey[EV_WIN];
    float ex[EV_WIN];
    uint32_t ek[EV_WIN];
};

// One wave per band; the variant M takes the bands of (M_lo, M] edges.
template <uint32_t M, uint32_t M_LO>
__global__ void __launch_bounds__(64) sweep_kernel(Params P, const int8_t* bw) {
    const uint32_t b = blockIdx.x;
    if (P.b_status[b] != 0) return;
    const uint32_t m = P.b_nedge[b];
    if (m == 0 || m > M || m <= M_LO) return;
    __shared__ SweepState<M> S;
    const int lane = lane_id();
    const uint32_t off = P.b_edge_off[b];
    const float4* G = P.bgeo + off;
    const uint32_t* gid = P.band_edges + off;
    const int8_t* W8 = bw + off;
    const float ytop = P.b_ytop[b], ybot = P.b_ybot[b];
    const uint32_t rule = opt_of(P, P.b_path[b]).fill_rule;
    const uint32_t ne = P.b_nev[b], eoff = P.b_ev_off[b];
    const uint32_t rbase = P.b_rec_off[b], rcap = P.b_rec_cap[b];
    uint32_t nrec = 0;
    bool overflow = false, block_err = false;

    auto wnd = [&](uint32_t local) -> int32_t { return (int32_t)W8[local]; };
    // the event window: events [wb, wb + EV_WIN) in LDS (wave-uniform wb)
    uint32_t wb = 0;
    auto load_window = [&](uint32_t from) {
        wave_sync();
        wb = from;
        for (uint32_t i = lane; i < EV_WIN; i += 64) {
            const uint32_t e = from + i;
            if (e < ne) {
                S.ey[i] = P.ev_y[eoff + e];
                S.ex[i] = P.ev_x[eoff + e];
                S.ek[i] = P.ev_k[eoff + e];
            }
        }
        wave_sync();
    };
    auto evy = [&](uint32_t e) -> double {
        return e - wb < EV_WIN ? S.ey[e - wb] : P.ev_y[eoff + e];
    };
    auto evx = [&](uint32_t e) -> float {
        return e - wb < EV_WIN ? S.ex[e - wb] : P.ev_x[eoff + e];
    };
    auto evk = [&](uint32_t e) -> uint32_t {
        return e - wb < EV_WIN ? S.ek[e - wb] : P.ev_k[eoff + e];
    };
    auto height = [&](uint32_t ev) -> double {
        return ev == TOP_EV ? (double)ytop : (ev == BOT_EV ? (double)ybot : evy(ev));
    };
    auto emit = [&](uint32_t l, uint32_t r, uint32_t t, uint32_t bt) {
        // uniform call: lane 0 writes
        if (nrec < rcap) {
            if (lane == 0) P.rec[rbase + nrec] = make_uint4(gid[l], gid[r], t, bt);
        } else {
            overflow = true;
        }
        ++nrec;
    };

    // initial active edges, their windings (prefix over the wave) and gap tops
    uint32_t a = P.b_ninit[b];
    for (uint32_t i = lane; i < m; i += 64) S.pos[i] = NO_POS;
    wave_sync();
    int32_t carry = 0;
    for (uint32_t c = 0; c < a; c += 64) {
        const uint32_t i = c + lane;
        uint32_t id = 0;
        int32_t w = 0;
        if (i < a) {
            id = P.init_order[off + i];
            w = wnd(id);
        }
        // inclusive prefix of w (shifted to unsigned with a bias)
        const int32_t sc = (int32_t)incl_scan((uint32_t)(w + 1)) - (lane + 1);
        if (i < a) {
            S.act[i] = (uint16_t)id;
            S.pos[id] = (uint16_t)i;
            S.wsum[i] = (int16_t)(carry + sc);
            S.top[i] = TOP_EV;
        }
        carry += (int32_t)readlane_u((uint32_t)sc, 63);
    }
    wave_sync();

    uint32_t ei = 0;
    if (ne) load_window(0);
    while (ei < ne && !block_err) {
        // keep the window ahead of the sweep (uniform)
        if (ei - wb > EV_WIN - 64 && wb + EV_WIN < ne) load_window(ei);
        // the event (a vertex: the ends and starts at one point; or one crossing)
        const double y = evy(ei);
        const float x = evx(ei);
        const uint32_t k0 = evk(ei);
        uint32_t ej = ei + 1;
        const bool is_cross = (k0 >> 30) == EV_CROSS;
        if (!is_cross) {
            while (ej < ne && evy(ej) == y && evx(ej) == x && (evk(ej) >> 30) != EV_CROSS)
                ++ej;
        }
        // the next height strictly below
        double ynext = (double)ybot;
        for (uint32_t q = ej; q < ne; ++q) {
            const double yq = evy(q);
            if (yq > y) {
                ynext = yq;
                break;
            }
        }
        const double ymid = 0.5 * (y + fmin(ynext, (double)ybot));
        const uint32_t hev = ei;  // this event's height
        // positions the event rebuilds: [lo, lo + nold)
        int32_t lo, hi;
        uint32_t n_end = 0;
        bool skip = false;
        if (is_cross) {
            const uint32_t pa = S.pos[k0 & 0xffffu], pb = S.pos[(k0 >> 16) & 0x3fffu];
            if (pa == NO_POS || pb == NO_POS) {
                skip = true;
                lo = hi = 0;
            } else {
                lo = (int32_t)min(pa, pb) - 1;
                hi = (int32_t)max(pa, pb) + 1;
            }
        } else {
            int32_t pmin = 0x7fffffff, pmax = -1;
            for (uint32_t q = ei; q < ej; ++q) {
                const uint32_t k = evk(q);
                if ((k >> 30) == EV_END) {
                    const uint32_t p = S.pos[k & 0xffffu];
                    if (p != NO_POS) {
                        pmin = min(pmin, (int32_t)p);
                        pmax = max(pmax, (int32_t)p);
                        ++n_end;
                    }
                }
            }
            if (n_end) {
                // with the edges that pass through the point (a collinear bundle, a
                // T-junction) and one neighbour on each side
                while (pmin > 0 && xat(G[S.act[pmin - 1]], y) == (double)x) --pmin;
                while (pmax + 1 < (int32_t)a && xat(G[S.act[pmax + 1]], y) == (double)x) ++pmax;
                lo = pmin - 1;
                hi = pmax + 1;
            } else {
                // the active edges left of the point, and those through it
                uint32_t clt = 0, cle = 0;
                for (uint32_t c = 0; c < a; c += 64) {
                    const uint32_t i = c + lane;
                    double xi = 0.0;
                    if (i < a) xi = xat(G[S.act[i]], y);
                    clt += __popcll(ballot(i < a && xi < (double)x));
                    cle += __popcll(ballot(i < a && xi <= (double)x));
                }
                lo = (int32_t)clt - 1;
                hi = (int32_t)cle;
            }
        }
        if (!skip) {
            lo = max(lo, 0);
            hi = min(hi, (int32_t)a - 1);
            const uint32_t nold = hi >= lo ? (uint32_t)(hi - lo + 1) : 0;
            if (nold > BLK_MAX) {
                block_err = true;
                break;
            }
            // old gaps (lo - 1 .. hi), within 0 .. a - 2
            uint32_t nog = 0;
            for (int32_t g = lo - 1; g <= hi; ++g) {
                if (g < 0 || g > (int32_t)a - 2) continue;
                Gap o;
                o.l = S.act[g];
                o.r = S.act[g + 1];
                o.w = S.wsum[g];
                o.top = S.top[g];
                o.matched = 0;
                if (lane == 0) S.og[nog] = o;
                ++nog;
            }
            // the new block: old positions minus ends, plus starts, sorted at ymid
            uint32_t nn = 0;
            int32_t wold = 0, wnew = 0;
            for (uint32_t i = 0; i < nold; ++i) {
                const uint32_t id = S.act[lo + i];
                wold += wnd(id);
                bool ends = false;
                if (!is_cross)
                    for (uint32_t q = ei; q < ej; ++q) {
                        const uint32_t k = evk(q);
                        if ((k >> 30) == EV_END && (k & 0xffffu) == id) ends = true;
                    }
                if (ends) continue;
                if (nn < BLK_MAX && lane == 0) S.nb[nn] = id;
                ++nn;
            }
            if (!is_cross)
                for (uint32_t q = ei; q < ej; ++q) {
                    const uint32_t k = evk(q);
                    if ((k >> 30) == EV_START) {
                        if (nn < BLK_MAX && lane == 0) S.nb[nn] = k & 0xffffu;
                        ++nn;
                    }
                }
            if (nn > BLK_MAX || nog > BLK_MAX + 1) {
                block_err = true;
                break;
            }
            wave_sync();
            // insertion sort of the block by x at ymid (ties by band-local id)
            if (lane == 0) {
                for (uint32_t i = 0; i < nn; ++i) S.nk[i] = xat(G[S.nb[i]], ymid);
                for (uint32_t i = 1; i < nn; ++i) {
                    const uint32_t id = S.nb[i];
                    const double k = S.nk[i];
                    uint32_t j = i;
                    while (j > 0 && (S.nk[j - 1] > k || (S.nk[j - 1] == k && S.nb[j - 1] > id))) {
                        S.nb[j] = S.nb[j - 1];
                        S.nk[j] = S.nk[j - 1];
                        --j;
                    }
                    S.nb[j] = id;
                    S.nk[j] = k;
                }
            }
            wave_sync();
            for (uint32_t i = 0; i < nn; ++i) wnew += wnd(S.nb[i]);
            const int32_t wdelta = wnew - wold;
            // move the tail [lo + nold, a) to lo + nn
            const int32_t d = (int32_t)nn - (int32_t)nold;
            const uint32_t tail0 = (uint32_t)lo + nold;
            if (wdelta != 0) {
                // an event that does not balance its windings (a vertex whose other edge is
                // horizontal: horizontal edges are not swept) changes the winding of every
                // gap right of it; a gap that crosses the fill rule's boundary there closes
                // and opens again (the other end of the horizontal edge, at the same
                // height, restores the rest)
                const uint32_t tn = a - tail0;
                for (uint32_t c = 0; c < tn; c += 64) {
                    const uint32_t i = c + lane;
                    bool flip = false, em = false;
                    uint32_t l = 0, r = 0, tp = 0;
                    if (i + 1 < tn) {
                        const int32_t w = S.wsum[tail0 + i];
                        flip = inside(rule, w) != inside(rule, w + wdelta);
                        if (flip) {
                            tp = S.top[tail0 + i];
                            l = S.act[tail0 + i];
                            r = S.act[tail0 + i + 1];
                            em = inside(rule, w) && height(tp) < y;
                        }
                    }
                    const uint64_t bal = ballot(em);
                    if (em) {
                        const uint32_t slot = nrec + __popcll(bal & mask_lt(lane));
                        if (slot < rcap) P.rec[rbase + slot] = make_uint4(gid[l], gid[r], tp, hev);
                    }
                    nrec += __popcll(bal);
                    wave_sync();
                    if (flip) S.top[tail0 + i] = hev;
                }
                wave_sync();
            }
            if (d != 0 || wdelta != 0) {
                const uint32_t tn = a - tail0;
                if (d > 0) {
                    for (int32_t c = (int32_t)tn - 64; c > -64; c -= 64) {
                        const int32_t i = c + lane;
                        uint16_t id = 0;
                        int16_t ws = 0;
                        uint32_t tp = 0;
                        if (i >= 0) {
                            id = S.act[tail0 + i];
                            ws = S.wsum[tail0 + i];
                            tp = S.top[tail0 + i];
                        }
                        wave_sync();
                        if (i >= 0) {
                            const uint32_t dst = tail0 + i + d;
                            S.act[dst] = id;
                            S.wsum[dst] = (int16_t)(ws + wdelta);
                            S.top[dst] = tp;
                            S.pos[id] = (uint16_t)dst;
                        }
                        wave_sync();
                    }
                } else {
                    for (uint32_t c = 0; c < tn; c += 64) {
                        const uint32_t i = c + lane;
                        uint16_t id = 0;
                        int16_t ws = 0;
                        uint32_t tp = 0;
                        if (i < tn) {
                            id = S.act[tail0 + i];
                            ws = S.wsum[tail0 + i];
                            tp = S.top[tail0 + i];
                        }
                        wave_sync();
                        if (i < tn) {
                            const uint32_t dst = (uint32_t)((int32_t)(tail0 + i) + d);
                            S.act[dst] = id;
                            S.wsum[dst] = (int16_t)(ws + wdelta);
                            S.top[dst] = tp;
                            S.pos[id] = (uint16_t)dst;
                        }
                        wave_sync();
                    }
                }
            }
            // ends leave the active set
            if (!is_cross && lane == 0)
                for (uint32_t q = ei; q < ej; ++q) {
                    const uint32_t k = evk(q);
                    if ((k >> 30) == EV_END) S.pos[k & 0xffffu] = NO_POS;
                }
            wave_sync();
            // the block
            int32_t wl = lo > 0 ? (int32_t)S.wsum[lo - 1] : 0;
            if (lane == 0) {
                for (uint32_t i = 0; i < nn; ++i) {
                    const uint32_t id = S.nb[i];
                    wl += wnd(id);
                    S.act[lo + i] = (uint16_t)id;
                    S.pos[id] = (uint16_t)(lo + i);
                    S.wsum[lo + i] = (int16_t)wl;
                }
            }
            a = (uint32_t)((int32_t)a + d);
            wave_sync();
            // new gaps (lo - 1 .. lo + nn - 1): keep the top of an unchanged one
            for (int32_t g = lo - 1; g <= lo + (int32_t)nn - 1; ++g) {
                if (g < 0 || g > (int32_t)a - 2) continue;
                const uint32_t l = S.act[g], r = S.act[g + 1];
                const int32_t w = S.wsum[g];
                uint32_t tp = hev;
                for (uint32_t q = 0; q < nog; ++q) {
                    const Gap o = S.og[q];
                    if (!o.matched && o.l == l && o.r == r && inside(rule, o.w) == inside(rule, w)) {
                        tp = o.top;
                        if (lane == 0) S.og[q].matched = 1;
                        wave_sync();
                        break;
                    }
                }
                if (lane == 0) S.top[g] = tp;
            }
            wave_sync();
            for (uint32_t q = 0; q < nog; ++q) {
                const Gap o = S.og[q];
                if (!o.matched && inside(rule, o.w) && height(o.top) < y) emit(o.l, o.r, o.top, hev);
            }
            wave_sync();
        }
        ei = ej;
    }
    // close every gap at the band's bottom
    for (uint32_t c = 0; c + 1 < a; c += 64) {
        const uint32_t g = c + lane;
        bool in = false;
        uint32_t l = 0, r = 0, tp = 0;
        if (g + 1 < a) {
            l = S.act[g];
            r = S.act[g + 1];
            tp = S.top[g];
            in = inside(rule, S.wsum[g]) && height(tp) < (double)ybot;
        }
        const uint64_t bal = ballot(in);
        const uint32_t before = __popcll(bal & mask_lt(lane));
        if (in) {
            const uint32_t slot = nrec + before;
            if (slot < rcap) P.rec[rbase + slot] = make_uint4(gid[l], gid[r], tp, BOT_EV);
        }
        nrec += __popcll(bal);
    }
    if (nrec > rcap) overflow = true;
    if (lane == 0) {
        P.b_nrec[b] = min(nrec, rcap);
        uint32_t st = (overflow ? BS_RECORDS : 0u) | (block_err ? BS_BLOCK : 0u);
        if (st) {
            P.b_status[b] = st;
            atomicOr(&P.p_status[P.b_path[b]], st);
            // the room the band needed
            if (overflow) P.b_rec_cap[b] = nrec;
        }
    }
}

// ---- 6. trapezoids to triangles ------------------------------------------------------------
struct Quad {
    vec2 lt, rt, rb, lb;
    uint32_t nv;  // 0, 3 or 4
    uint32_t shape;  // 0 both widths, 1 top only (triangle lt rt lb), 2 bottom only (lt rb lb)
};
// Where the lines of L and R meet, clamped into [yt, yb] (f64; symmetric in L and R).
__device__ __forceinline__ vec2 apex(float4 L, float4 R, double yt, double yb, vec2 fallback) {
    const double sl = ((double)L.z - L.x) / ((double)L.w - L.y);
    const double sr = ((double)R.z - R.x) / ((double)R.w - R.y);
    if (!(sl != sr)) return fallback;
    double y = (((double)R.x - R.y * sr) - ((double)L.x - L.y * sl)) / (sl - sr);
    y = fmin(fmax(y, yt), yb);
    const double x = 0.5 * (xat(L, y) + xat(R, y));
    return mk((float)x, (float)y);
}

__device__ __forceinline__ Quad corners(const Params& P, uint32_t b, uint4 r) {
    Quad q;
    const float4 L = P.geo[r.x], R = P.geo[r.y];
    const uint32_t eoff = P.b_ev_off[b];
    auto height = [&](uint32_t ev) -> double {
        return ev == TOP_EV ? (double)P.b_ytop[b] : (ev == BOT_EV ? (double)P.b_ybot[b] : P.ev_y[eoff + ev]);
    };
    const double yt = height(r.z), yb = height(r.w);
    const float ft = (float)yt, fb = (float)yb;
    q.lt = mk((float)xat(L, yt), ft);
    q.rt = mk((float)xat(R, yt), ft);
    q.lb = mk((float)xat(L, yb), fb);
    q.rb = mk((float)xat(R, yb), fb);
    const bool wt = q.rt.x > q.lt.x, wb = q.rb.x > q.lb.x;
    q.nv = 0;
    q.shape = 0;
    if (!(fb > ft)) return q;
    if (wt && wb) {
        q.nv = 4;
    } else if (wt) {
        // L and R meet at the bottom (their crossing is an event): the apex is where they
        // meet, so no sliver is added or lost there
        q.nv = 3;
        q.shape = 1;
        q.lb = apex(L, R, yt, yb, q.lb);
    } else if (wb) {
        q.nv = 3;
        q.shape = 2;
        q.lt = apex(L, R, yt, yb, q.lt);
    }
    return q;
}

// Records of band b live in [b_rec_off[b], b_rec_off[b] + b_nrec[b]); the rest of the room
// counts 0.
__global__ void expand_count_kernel(Params P) {
    const uint32_t b = blockIdx.x;
    const uint32_t r0 = P.b_rec_off[b], r1 = P.b_rec_off[b + 1];
    const uint32_t n = P.b_status[b] == 0 ? P.b_nrec[b] : 0;
    for (uint32_t i = r0 + threadIdx.x; i < r1; i += blockDim.x) {
        uint32_t nv = 0;
        if (i < r0 + n) nv = corners(P, b, P.rec[i]).nv;
        P.r_nv[i] = nv;
        P.r_nt[i] = nv == 4 ? 2 : (nv == 3 ? 1 : 0);
    }
}

__global__ void expand_write_kernel(Params P) {
    const uint32_t b = blockIdx.x;
    const uint32_t r0 = P.b_rec_off[b];
    const uint32_t n = P.b_status[b] == 0 ? P.b_nrec[b] : 0;
    const uint32_t p = P.b_path[b];
    const uint32_t vbase = P.p_voff[p];
    const uint32_t idbase = P.vertex_offset + (P.local_ids ? 0u : vbase);
    for (uint32_t i = r0 + threadIdx.x; i < r0 + n; i += blockDim.x) {
        const Quad q = corners(P, b, P.rec[i]);
        if (q.nv == 0) continue;
        const uint32_t v = P.r_voff[i];
        const uint32_t t = P.r_toff[i];
        float* o = P.out_v + 2ull * v;
        uint32_t* ix = P.out_i + 3ull * t;
        const uint32_t id = idbase + (v - vbase);
        if (q.nv == 4) {
            o[0] = q.lt.x; o[1] = q.lt.y; o[2] = q.rt.x; o[3] = q.rt.y;
            o[4] = q.rb.x; o[5] = q.rb.y; o[6] = q.lb.x; o[7] = q.lb.y;
            ix[0] = id; ix[1] = id + 1; ix[2] = id + 2;
            ix[3] = id; ix[4] = id + 2; ix[5] = id + 3;
        } else if (q.shape == 1) {
            o[0] = q.lt.x; o[1] = q.lt.y; o[2] = q.rt.x; o[3] = q.rt.y;
            o[4] = q.lb.x; o[5] = q.lb.y;
            ix[0] = id; ix[1] = id + 1; ix[2] = id + 2;
        } else {
            o[0] = q.lt.x; o[1] = q.lt.y; o[2] = q.rb.x; o[3] = q.rb.y;
            o[4] = q.lb.x; o[5] = q.lb.y;
            ix[0] = id; ix[1] = id + 1; ix[2] = id + 2;
        }
    }
}


// Small per-path steps.  which = 0: y ranges start empty; 1: each path's first edge;
// 2: histogram bins of each path; 3: each path's first vertex and index.
__global__ void path_offsets_kernel(Params P, uint32_t* path_edge_off, uint32_t* p_hbins,
                                    uint32_t* vtx_off, uint32_t* idx_off, int which) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p > P.n_paths) return;
    if (which == 0) {
        if (p < P.n_paths) {
            P.p_yr[2 * p] = 0xffffffffu;
            P.p_yr[2 * p + 1] = 0u;
        }
    } else if (which == 1) {
        path_edge_off[p] = P.v_ebase[P.verb_off[p]];
    } else if (which == 2) {
        if (p < P.n_paths) p_hbins[p] = P.p_nband[p] > 1 ? 64u * P.p_nband[p] : 0u;
    } else {
        const uint32_t r = P.b_rec_off[P.p_band_off[p]];
        vtx_off[p] = P.r_voff[r];
        idx_off[p] = 3u * P.r_toff[r];
    }
}

// After a retry round: a path's status is the union of its bands' statuses again.
__global__ void path_status_kernel(Params P) {
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < P.n_paths) P.p_status[b] &= ~BS_RECORDS;
}


// ---- flattening error ----------------------------------------------------------------
// How far each flattened curve strays from the true curve: for every chord (two
// consecutive flattened points, f32, exactly as the tessellator computes them) the largest
// distance from the curve itself, evaluated in f64 from the f32 control points over the
// chord's parameter range, to the chord.
struct Poly64 {
    double x[4], y[4];
    int deg;
};
__device__ __forceinline__ Poly64 power_basis(vec2 p0, vec2 p1, vec2 p2, vec2 p3, bool cubic) {
    Poly64 b;
    const double x0 = p0.x, y0 = p0.y, x1 = p1.x, y1 = p1.y, x2 = p2.x, y2 = p2.y;
    b.x[0] = x0;
    b.y[0] = y0;
    if (cubic) {
        const double x3 = p3.x, y3 = p3.y;
        b.x[1] = 3.0 * (x1 - x0);
        b.y[1] = 3.0 * (y1 - y0);
        b.x[2] = 3.0 * (x0 - 2.0 * x1 + x2);
        b.y[2] = 3.0 * (y0 - 2.0 * y1 + y2);
        b.x[3] = x3 - x0 + 3.0 * (x1 - x2);
        b.y[3] = y3 - y0 + 3.0 * (y1 - y2);
        b.deg = 3;
    } else {
        b.x[1] = 2.0 * (x1 - x0);
        b.y[1] = 2.0 * (y1 - y0);
        b.x[2] = x0 - 2.0 * x1 + x2;
        b.y[2] = y0 - 2.0 * y1 + y2;
        b.x[3] = b.y[3] = 0.0;
        b.deg = 2;
    }
    return b;
}
__device__ __forceinline__ double chord_error(const Poly64& B, double ta, double tb, vec2 p, vec2 q) {
    const double cx = (double)q.x - p.x, cy = (double)q.y - p.y;
    const double L2 = cx * cx + cy * cy;
    auto seg_dist = [&](double t) {
        const double x = ((B.x[3] * t + B.x[2]) * t + B.x[1]) * t + B.x[0];
        const double y = ((B.y[3] * t + B.y[2]) * t + B.y[1]) * t + B.y[0];
        const double wx = x - p.x, wy = y - p.y;
        double u = L2 > 0.0 ? (wx * cx + wy * cy) / L2 : 0.0;
        u = fmin(fmax(u, 0.0), 1.0);
        const double dx = wx - u * cx, dy = wy - u * cy;
        return sqrt(dx * dx + dy * dy);
    };
    // the distance to the chord's line peaks at a root of its derivative (a quadratic in
    // t); the distance to the chord itself is checked there, at the ends and at 15 evenly
    // spaced parameters between them
    double m = fmax(seg_dist(ta), seg_dist(tb));
    for (int k = 1; k < 16; ++k) m = fmax(m, seg_dist(ta + (tb - ta) * (k / 16.0)));
    const double e1 = cx * B.y[1] - cy * B.x[1];
    const double e2 = cx * B.y[2] - cy * B.x[2];
    const double e3 = cx * B.y[3] - cy * B.x[3];
    const double a = 3.0 * e3, b = 2.0 * e2, c = e1;
    if (a == 0.0) {
        if (b != 0.0) {
            const double t = -c / b;
            if (t > ta && t < tb) m = fmax(m, seg_dist(t));
        }
    } else {
        const double disc = b * b - 4.0 * a * c;
        if (disc >= 0.0) {
            const double sq = sqrt(disc);
            const double qq = -0.5 * (b + (b >= 0.0 ? sq : -sq));
            const double r1 = qq / a;
            const double r2 = qq != 0.0 ? c / qq : r1;
            if (r1 > ta && r1 < tb) m = fmax(m, seg_dist(r1));
            if (r2 > ta && r2 < tb) m = fmax(m, seg_dist(r2));
        }
    }
    return m;
}
// The largest chord error of one flattened curve (lyon_geom's for_each_flattened_with_t:
// Wang count, serial t, Horner in f32) whose t in [0, 1] maps to [t0, t0 + dt] of B.
__device__ double flattened_error(const Poly64& B, double t0, double dt, vec2 f, vec2 a, vec2 b,
                                  vec2 t, bool cubic, float tol) {
    float n = cubic ? cubic_wang(f, a, b, t, tol) : quad_wang(f, a, t, tol);
    float step = 1.0f / n;
    uint32_t cnt = seg_count(n);
    vec2 a0 = f, a1, a2, a3 = mk(0.0f, 0.0f);
    if (cubic) {
        a1 = mul(sub(a, f), 3.0f);
        a2 = add(sub(mul(f, 3.0f), mul(a, 6.0f)), mul(b, 3.0f));
        a3 = add(sub(t, f), mul(sub(a, b), 3.0f));
    } else {
        a1 = mul(sub(a, f), 2.0f);
        a2 = sub(add(f, t), mul(a, 2.0f));
    }
    float u0 = 0.0f;
    vec2 lf = f;
    double worst = 0.0;
    for (uint32_t i = 0; i < cnt; ++i) {
        vec2 lt;
        float u1;
        if (i + 1 == cnt) {
            lt = t;
            u1 = 1.0f;
        } else {
            u1 = u0 + step;
            vec2 q = a0;
            float tt = u1;
            q = add(q, mul(a1, tt));
            tt = tt * u1;
            q = add(q, mul(a2, tt));
            if (cubic) {
                tt = tt * u1;
                q = add(q, mul(a3, tt));
            }
            lt = q;
        }
        worst = fmax(worst, chord_error(B, t0 + dt * u0, t0 + dt * fmin(u1, 1.0f), lf, lt));
        lf = lt;
        u0 = u1;
    }
    return worst;
}

// One thread per verb.  rule 0: the fill's flattening (curves turned to run down the sweep,
// event_queue.rs:758-900); rule 1: the stroker's (quadratics with a sharp turn split first,
// stroke.rs:2755-2779).  path_err gets each path's largest error (f32 bits, atomic max).
__global__ void flat_error_kernel(Params P, const float* tol, uint32_t tol_stride, int rule,
                                  uint32_t* path_err) {
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= P.n_verbs) return;
    Seg s = seg_of(P, v);
    if (s.kind < 2 || s.nan) return;
    const uint32_t p = P.v_path[v];
    const float tl = tol[tol_stride ? p : 0];
    const bool cubic = s.kind == 3;
    double e = 0.0;
    if (rule == 0) {
        vec2 f, a, b, t;
        bool sw;
        turned(s, f, a, b, t, sw);
        const Poly64 B = cubic ? power_basis(f, a, b, t, true) : power_basis(f, a, t, t, false);
        e = flattened_error(B, 0.0, 1.0, f, a, b, t, cubic, tl);
    } else {
        const Poly64 B = cubic ? power_basis(s.from, s.c1, s.c2, s.to, true)
                               : power_basis(s.from, s.c1, s.to, s.to, false);
        float ts = cubic ? -1.0f : find_sharp_turn(s.from, s.c1, s.to);
        if (ts > 0.0f) {
            // QuadraticBezierSegment::split (quadratic_bezier.rs:208-223), as the stroker
            const vec2 sp = quad_sample(s.from, s.c1, s.to, ts);
            e = flattened_error(B, 0.0, ts, s.from, lerp(s.from, s.c1, ts), mk(0, 0), sp, false, tl);
            e = fmax(e, flattened_error(B, ts, 1.0 - ts, sp, lerp(s.c1, s.to, ts), mk(0, 0), s.to,
                                        false, tl));
        } else {
            e = flattened_error(B, 0.0, 1.0, s.from, s.c1, s.c2, s.to, cubic, tl);
        }
    }
    const float ef = (float)e;
    if (ef > 0.0f) atomicMax(&path_err[p], f2u(ef));
}

}  // namespace band
}  // namespace lyg

extern "C" {

using lyg::band::Params;

#define LYG_BAND_LAUNCH_1D(kernel, n, P)                                                     \
    do {                                                                                     \
        if ((n) > 0)                                                                         \
            hipLaunchKernelGGL(kernel, dim3(((n) + 255) / 256), dim3(256), 0, s, P);         \
    } while (0)

hipError_t lyg_band_verbs(const Params* P, hipStream_t s) {
    LYG_BAND_LAUNCH_1D(lyg::band::verb_info_kernel, P->n_verbs, *P);
    return hipGetLastError();
}
hipError_t lyg_band_begins(const Params* P, hipStream_t s) {
    LYG_BAND_LAUNCH_1D(lyg::band::begin_list_kernel, P->n_verbs, *P);
    return hipGetLastError();
}
hipError_t lyg_band_edge_count(const Params* P, hipStream_t s) {
    LYG_BAND_LAUNCH_1D(lyg::band::edge_count_kernel, P->n_verbs, *P);
    return hipGetLastError();
}
hipError_t lyg_band_edge_emit(const Params* P, hipStream_t s) {
    LYG_BAND_LAUNCH_1D(lyg::band::edge_emit_kernel, P->n_verbs, *P);
    return hipGetLastError();
}
hipError_t lyg_band_path_bands(const Params* P, const uint32_t* path_edge_off, hipStream_t s) {
    if (P->n_paths)
        hipLaunchKernelGGL(lyg::band::path_bands_kernel, dim3((P->n_paths + 255) / 256), dim3(256),
                           0, s, *P, path_edge_off);
    return hipGetLastError();
}
hipError_t lyg_band_hist(const Params* P, hipStream_t s) {
    LYG_BAND_LAUNCH_1D(lyg::band::hist_kernel, P->n_edges, *P);
    return hipGetLastError();
}
hipError_t lyg_band_bounds(const Params* P, hipStream_t s) {
    if (P->n_paths)
        hipLaunchKernelGGL(lyg::band::bounds_kernel, dim3(P->n_paths), dim3(256), 0, s, *P);
    return hipGetLastError();
}
hipError_t lyg_band_count(const Params* P, hipStream_t s) {
    LYG_BAND_LAUNCH_1D(lyg::band::band_count_kernel, P->n_edges, *P);
    return hipGetLastError();
}
hipError_t lyg_band_fill(const Params* P, hipStream_t s) {
    LYG_BAND_LAUNCH_1D(lyg::band::band_fill_kernel, P->n_edges, *P);
    return hipGetLastError();
}
hipError_t lyg_band_cross(const Params* P, int emit, hipStream_t s) {
    if (P->n_bands == 0) return hipSuccess;
    if (emit)
        hipLaunchKernelGGL(lyg::band::cross_kernel<1>, dim3(P->n_bands), dim3(256), 0, s, *P);
    else
        hipLaunchKernelGGL(lyg::band::cross_kernel<0>, dim3(P->n_bands), dim3(256), 0, s, *P);
    return hipGetLastError();
}
hipError_t lyg_band_event_count(const Params* P, hipStream_t s) {
    if (P->n_bands)
        hipLaunchKernelGGL(lyg::band::event_count_kernel, dim3((P->n_bands + 255) / 256),
                           dim3(256), 0, s, *P);
    return hipGetLastError();
}
// scratch: 32 * (total events) bytes (events sorted in global memory when a band's do not
// fit in LDS)
hipError_t lyg_band_event_sort(const Params* P, void* scratch, hipStream_t s) {
    if (P->n_bands)
        hipLaunchKernelGGL(lyg::band::event_sort_kernel, dim3(P->n_bands), dim3(256), 0, s, *P,
                           static_cast<lyg::band::EvRec*>(scratch));
    return hipGetLastError();
}
hipError_t lyg_band_sort(const Params* P, int8_t* bw, hipStream_t s) {
    if (P->n_bands)
        hipLaunchKernelGGL(lyg::band::band_sort_kernel, dim3(P->n_bands), dim3(256), 0, s, *P, bw);
    return hipGetLastError();
}
// Each band by the variant of its size class (the others return at once); the three
// launches go to three streams so the classes run at once.
hipError_t lyg_band_sweep(const Params* P, const int8_t* bw, hipStream_t s_small,
                          hipStream_t s_mid, hipStream_t s_big) {
    using namespace lyg::band;
    if (P->n_bands) {
        hipLaunchKernelGGL((sweep_kernel<M_MAX, M_MID>), dim3(P->n_bands), dim3(64), 0, s_big,
                           *P, bw);
        hipLaunchKernelGGL((sweep_kernel<M_MID, M_SMALL>), dim3(P->n_bands), dim3(64), 0, s_mid,
                           *P, bw);
        hipLaunchKernelGGL((sweep_kernel<M_SMALL, 0>), dim3(P->n_bands), dim3(64), 0, s_small,
                           *P, bw);
    }
    return hipGetLastError();
}
hipError_t lyg_band_path_offsets(const Params* P, uint32_t* path_edge_off, uint32_t* p_hbins,
                                 uint32_t* vtx_off, uint32_t* idx_off, int which, hipStream_t s) {
    hipLaunchKernelGGL(lyg::band::path_offsets_kernel, dim3((P->n_paths + 1 + 255) / 256), dim3(256),
                       0, s, *P, path_edge_off, p_hbins, vtx_off, idx_off, which);
    return hipGetLastError();
}
hipError_t lyg_band_path_status(const Params* P, const uint32_t* b_status_in, hipStream_t s) {
    (void)b_status_in;
    hipLaunchKernelGGL(lyg::band::path_status_kernel, dim3((P->n_paths + 255) / 256), dim3(256), 0,
                       s, *P);
    return hipGetLastError();
}
hipError_t lyg_band_flat_error(const Params* P, const float* tol, uint32_t tol_stride, int rule,
                               uint32_t* path_err, hipStream_t s) {
    if (P->n_verbs)
        hipLaunchKernelGGL(lyg::band::flat_error_kernel, dim3((P->n_verbs + 255) / 256), dim3(256),
                           0, s, *P, tol, tol_stride, rule, path_err);
    return hipGetLastError();
}
hipError_t lyg_band_expand_count(const Params* P, hipStream_t s) {
    if (P->n_bands)
        hipLaunchKernelGGL(lyg::band::expand_count_kernel, dim3(P->n_bands), dim3(256), 0, s, *P);
    return hipGetLastError();
}
hipError_t lyg_band_expand_write(const Params* P, hipStream_t s) {
    if (P->n_bands)
        hipLaunchKernelGGL(lyg::band::expand_write_kernel, dim3(P->n_bands), dim3(256), 0, s, *P);
    return hipGetLastError();
}

}  // extern "C"
